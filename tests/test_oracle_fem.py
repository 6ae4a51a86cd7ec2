"""The one-dimensional finite-element oracle (oracle/fem_oracle.c) against
independent forms: the reference's own quadrature tables
(tests/golden/quadrature_rules.json, checked against numpy's Gauss-Legendre
rule), numpy's linspace and the closed forms of the graded meshes, the
partition of unity and polynomial reproduction of the shape functions,
numpy/scipy Lagrange interpolation on each element, exact Gauss integration of
polynomials, and the example's basis functions (tests/golden/example/), which
are orthonormal to 3.4e-5."""
import json
import os

import numpy as np
import pytest
from scipy.interpolate import lagrange

import oracle

HERE = os.path.dirname(__file__)
RULES = json.load(open(os.path.join(HERE, "golden", "quadrature_rules.json")))["rules"]
EXAMPLE = os.path.join(HERE, "golden", "example")


def rule(n):
    return np.array(RULES[str(n)]["points"]), np.array(RULES[str(n)]["weights"])


@pytest.mark.parametrize("n", range(2, 21))
def test_reference_rules_are_gauss_legendre(n):
    """The fixture: a permutation of numpy's rule, integrating x^k exactly up
    to k = 2n - 1 and not beyond (the 16-digit tables agree to 2e-15)."""
    xi, w = rule(n)
    x0, w0 = np.polynomial.legendre.leggauss(n)
    order = np.argsort(xi)
    assert np.abs(xi[order] - x0).max() < 1e-15
    assert np.abs(w[order] - w0).max() < 2e-15
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, xi ** k) - exact) < 1e-14
    # and not x^(2n): the error there is about 2^(3 - 2n), far above rounding
    assert abs(np.dot(w, xi ** (2 * n)) - 2.0 / (2 * n + 1)) > 1e-13


def test_uniform_mesh_is_linspace():
    nodes = oracle.fem_mesh("uniform", 40, 5, -16.0, 16.0)
    assert len(nodes) == 161
    assert np.abs(nodes - np.linspace(-16.0, 16.0, 161)).max() < 1e-14
    assert nodes[0] == -16.0 and nodes[-1] == 16.0


def element_ends(nodes, npe):
    return nodes[::npe - 1]


def test_adaptive_mesh_closed_form():
    """Inner elements of size h about the centre, outer ones growing by the
    coarsening factor; nodes evenly spaced inside each element."""
    nel, npe, inner, h, c = 30, 4, 8.0, 0.5, 1.25
    nodes = oracle.fem_mesh("adaptive", nel, npe, -40.0, 40.0, inner, h, c)
    ends = element_ends(nodes, npe)
    sizes = np.diff(ends)
    n_in = 16
    n_out = (nel - n_in) // 2
    assert np.abs(sizes[n_out:n_out + n_in] - h).max() < 1e-13
    for i in range(1, n_out + 1):
        assert abs(sizes[n_out - i] - h * c ** i) < 1e-12
        if i < n_out:  # the last node is set to the domain end
            assert abs(sizes[n_out + n_in - 1 + i] - h * c ** i) < 1e-12
    assert ends[n_out] == -inner / 2
    assert nodes[-1] == 40.0
    for e in range(nel - 1):
        seg = nodes[e * (npe - 1):(e + 1) * (npe - 1) + 1]
        assert np.abs(np.diff(seg) - (seg[-1] - seg[0]) / (npe - 1)).max() < 1e-12


def test_hartree_outer_mesh_spans_the_domain():
    """Geometric outer elements sized so that the mesh ends at the domain ends."""
    nel, npe, inner, h, c = 40, 5, 10.0, 0.5, 1.3
    nodes = oracle.fem_mesh("hartree_outer", nel, npe, -50.0, 50.0, inner, h, c)
    sizes = np.diff(element_ends(nodes, npe))
    assert abs(nodes[0] + 50.0) < 1e-12 and abs(nodes[-1] - 50.0) < 1e-12
    assert np.abs(sizes[10:30] - h).max() < 1e-13
    assert np.abs(sizes[:9] / sizes[1:10] - c).max() < 1e-12


def test_manual1_mesh_sizes():
    """manual1: outer element sizes 1.5^(i+1), whatever the coarsening factor."""
    nodes = oracle.fem_mesh("manual1", 12, 3, -10.0, 10.0, 4.0, 1.0, 7.0)
    sizes = np.diff(element_ends(nodes, 3))
    assert np.abs(sizes[4:8] - 1.0).max() < 1e-14
    for i in range(1, 5):
        assert abs(sizes[4 - i] - 1.5 ** (i + 1)) < 1e-12
        assert abs(sizes[7 + i] - 1.5 ** (i + 1)) < 1e-12


@pytest.mark.parametrize("npe", [2, 3, 5, 7])
def test_shape_functions(npe):
    xi = np.linspace(-1.0, 1.0, 13)
    S, D = oracle.fem_shape(npe, xi)
    assert np.abs(S.sum(axis=0) - 1.0).max() < 1e-13
    assert np.abs(D.sum(axis=0)).max() < 1e-12
    xs = np.linspace(-1.0, 1.0, npe)
    Sn, _ = oracle.fem_shape(npe, xs)
    assert np.abs(Sn - np.eye(npe)).max() < 1e-14
    # reproduction of a polynomial of degree npe - 1 and of its derivative
    c = np.random.default_rng(npe).standard_normal(npe)
    p = np.polynomial.Polynomial(c)
    assert np.abs(p(xs) @ S - p(xi)).max() < 1e-12
    assert np.abs(p(xs) @ D - p.deriv()(xi)).max() < 1e-11


@pytest.mark.parametrize("npe,nq", [(5, 10), (3, 4), (2, 2), (6, 14)])
def test_interpolation_against_lagrange(npe, nq):
    """Per-element Lagrange interpolation (scipy) of random nodal values; the
    derivative with respect to xi is jac times the physical one."""
    nel = 7
    xi, w = rule(nq)
    nodes = oracle.fem_mesh("uniform", nel, npe, -3.0, 4.0)
    u = np.random.default_rng(nq).standard_normal((2, len(nodes)))
    val, dval, jw, ijw, xq = oracle.fem_interpolate(nel, npe, xi, w, nodes, u)
    for e in range(nel):
        xe = nodes[e * (npe - 1):e * (npe - 1) + npe]
        jac = (xe[-1] - xe[0]) / 2
        pts = 0.5 * (xe[-1] - xe[0]) * xi + 0.5 * (xe[-1] + xe[0])
        sl = slice(e * nq, (e + 1) * nq)
        assert np.abs(xq[sl] - pts).max() < 1e-13
        assert np.abs(jw[sl] - jac * w).max() < 1e-15
        assert np.abs(ijw[sl] - w / jac).max() < 1e-14
        for f in range(2):
            t = np.linspace(-1.0, 1.0, npe)
            poly = np.polynomial.Polynomial(lagrange(t, u[f, e * (npe - 1):e * (npe - 1) + npe]).coef[::-1])
            assert np.abs(val[f, sl] - poly(xi)).max() < 1e-11
            assert np.abs(dval[f, sl] - poly.deriv()(xi)).max() < 1e-10


def test_integration_is_exact_for_polynomials():
    """sum jw f(xq) over an 8-element mesh equals the exact integral for
    degree <= 2 nq - 1."""
    nq = 6
    xi, w = rule(nq)
    nodes = oracle.fem_mesh("uniform", 8, 4, -2.0, 3.0)
    _, _, jw, _, xq = oracle.fem_interpolate(8, 4, xi, w, nodes)
    p = np.polynomial.Polynomial(np.random.default_rng(0).standard_normal(2 * nq))
    P = p.integ()
    assert abs(np.dot(jw, p(xq)) - (P(3.0) - P(-2.0))) < 1e-10 * abs(P(3.0) - P(-2.0)) + 1e-12


def example_basis_at_quad(nq):
    B = np.loadtxt(os.path.join(EXAMPLE, "basis_x.txt"))
    nodal = np.hstack([np.zeros((B.shape[0], 1)), B, np.zeros((B.shape[0], 1))])
    nodes = oracle.fem_mesh("uniform", 40, 5, -16.0, 16.0)
    xi, w = rule(nq)
    return nodal, nodes, oracle.fem_interpolate(40, 5, xi, w, nodes, nodal)


def numpy_matrices(nodal, nodes, npe, weight=lambda x: 1.0 + 0 * x):
    """Overlap, gradient and weighted overlap of piecewise Lagrange functions,
    by numpy's own 8-point rule (exact for the degrees involved) and scipy's
    Lagrange polynomials."""
    x8, w8 = np.polynomial.legendre.leggauss(8)
    nel = (len(nodes) - 1) // (npe - 1)
    r = nodal.shape[0]
    M, K, P = np.zeros((r, r)), np.zeros((r, r)), np.zeros((r, r))
    for e in range(nel):
        xe = nodes[e * (npe - 1):e * (npe - 1) + npe]
        a, b = xe[0], xe[-1]
        x = 0.5 * (b - a) * x8 + 0.5 * (b + a)
        wx = 0.5 * (b - a) * w8
        V, dV = np.zeros((r, 8)), np.zeros((r, 8))
        for f in range(r):
            poly = np.polynomial.Polynomial(lagrange(xe - a, nodal[f, e * (npe - 1):e * (npe - 1) + npe]).coef[::-1])
            V[f], dV[f] = poly(x - a), poly.deriv()(x - a)
        M += (V * wx) @ V.T
        K += (dV * wx) @ dV.T
        P += (V * (wx * weight(x))) @ V.T
    return M, K, P


def test_example_basis_matrices():
    """The example's 40 basis functions: MatX is the identity to the file's
    orthonormality (3.4e-5); MatX, MatDX and MatPot (Veff factor x^2 / 10 and
    1) agree with numpy's Lagrange route."""
    nodal, nodes, (val, dval, jw, ijw, xq) = example_basis_at_quad(10)
    U = np.stack([np.ones_like(xq), xq ** 2 / 10], axis=1)
    MX, MDX, MP = oracle.tucker_basis_matrices(val, dval, jw, ijw, U)
    assert np.abs(MX - np.eye(40)).max() < 4e-5
    M, K, P = numpy_matrices(nodal[:, :], nodes, 5, lambda x: x ** 2 / 10)
    assert np.abs(MX - M).max() < 1e-12
    assert np.abs(MDX - K).max() < 1e-10 * np.abs(K).max()
    assert np.abs(MP[0] - MX).max() < 1e-13
    assert np.abs(MP[1] - P).max() < 1e-11 * np.abs(P).max()
    for A in (MX, MDX, MP[0], MP[1]):
        assert np.array_equal(A, A.T)
