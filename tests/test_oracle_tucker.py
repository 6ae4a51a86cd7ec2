"""The Tucker recompression and kernel-expansion Hartree oracle
(oracle/tucker_oracle.c) against independent forms: numpy SVDs of the
unfoldings, exact recovery of tensors of low multilinear rank, per-term
Gaussian integrals in closed form, and the Coulomb potential of a Gaussian
charge with the reference example's exponential-sum coefficients
(tests/golden/kernel_expansion_k35.json)."""
import json
import os

import numpy as np
import pytest
from scipy.special import erf

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "kernel_expansion_k35.json")


def unfold(T, n):
    return np.moveaxis(T, n, 0).reshape(T.shape[n], -1)


def ttm(T, U, n):
    """T x_n U (U of shape (new, old))."""
    return np.moveaxis(np.tensordot(U, np.moveaxis(T, n, 0), axes=(1, 0)), 0, n)


def numpy_sthosvd(T, ranks):
    """Sequentially truncated HOSVD through SVDs of the unfoldings, with the
    largest |entry| of each factor column positive."""
    Y = T
    Us = []
    for n, r in enumerate(ranks):
        u = np.linalg.svd(unfold(Y, n), full_matrices=False)[0][:, :r]
        idx = np.argmax(np.abs(u), axis=0)
        u = u * np.sign(u[idx, np.arange(r)])
        Us.append(u)
        Y = ttm(Y, u.T, n)
    return Y, Us


def reconstruct(G, Us):
    for n, u in enumerate(Us):
        G = ttm(G, u, n)
    return G


def test_sthosvd_matches_numpy_svd_route():
    rng = np.random.default_rng(1)
    T = rng.standard_normal((14, 11, 9))
    ranks = (5, 4, 3)
    G, Us, ev = oracle.sthosvd(T, ranks)
    G2, Us2 = numpy_sthosvd(T, ranks)
    for u, u2 in zip(Us, Us2):
        assert np.abs(u - u2).max() < 1e-9
    assert np.abs(G - G2).max() < 1e-9 * np.abs(G2).max()
    # kept eigenvalues of mode 0 = squared leading singular values of the unfolding
    s = np.linalg.svd(unfold(T, 0), compute_uv=False)
    assert np.abs(ev[0] - s[:5] ** 2).max() < 1e-10 * s[0] ** 2


def test_sthosvd_recovers_low_multilinear_rank():
    rng = np.random.default_rng(2)
    ranks = (4, 3, 5)
    G0 = rng.standard_normal(ranks)
    Us0 = [np.linalg.qr(rng.standard_normal((n, r)))[0] for n, r in zip((12, 10, 13), ranks)]
    T = reconstruct(G0, Us0)
    G, Us, _ = oracle.sthosvd(T, ranks)
    assert np.abs(reconstruct(G, Us) - T).max() < 1e-12 * np.abs(T).max()
    for u in Us:
        assert np.abs(u.T @ u - np.eye(u.shape[1])).max() < 1e-12


def test_sthosvd_error_is_the_discarded_energy():
    """For orthonormal factors, |T - T_hat|^2 = |T|^2 - |G|^2."""
    rng = np.random.default_rng(3)
    T = rng.standard_normal((9, 8, 7))
    G, Us, _ = oracle.sthosvd(T, (3, 3, 3))
    err = np.sum((T - reconstruct(G, Us)) ** 2)
    assert abs(err - (np.sum(T ** 2) - np.sum(G ** 2))) < 1e-10 * np.sum(T ** 2)


def gauss_legendre(a, b, nel, npt):
    x, w = np.polynomial.legendre.leggauss(npt)
    e = np.linspace(a, b, nel + 1)
    pts = np.concatenate([0.5 * (e1 - e0) * x + 0.5 * (e1 + e0) for e0, e1 in zip(e[:-1], e[1:])])
    wts = np.concatenate([0.5 * (e1 - e0) * w for e0, e1 in zip(e[:-1], e[1:])])
    return pts, wts


def test_hartree_terms_closed_form():
    """Rank-1 Gaussian density exp(-beta |x|^2): every term is a product of
    1-D Gaussian convolutions sqrt(pi/(a+b)) exp(-a b x^2/(a+b))."""
    beta = 0.7
    xe, jw = gauss_legendre(-9.0, 9.0, 60, 10)
    u = np.exp(-beta * xe ** 2)[:, None]
    xq = [np.linspace(-2.0, 2.5, 5), np.linspace(-1.5, 1.0, 4), np.linspace(0.0, 3.0, 3)]
    alpha = np.array([0.05, 0.8, 3.0])
    omega = np.array([1.3, -0.4, 0.25])
    V = oracle.hartree_expansion(np.ones((1, 1, 1)), [u, u, u], [xe] * 3, [jw] * 3, xq,
                                 alpha, omega)
    X, Y, Z = np.meshgrid(*xq, indexing="ij")
    ref = np.zeros_like(X)
    for a, w in zip(alpha, omega):
        c = np.sqrt(np.pi / (a + beta))
        g = a * beta / (a + beta)
        ref += w * c ** 3 * np.exp(-g * (X ** 2 + Y ** 2 + Z ** 2))
    assert np.abs(V - ref).max() < 1e-12 * np.abs(ref).max()
    # add mode accumulates
    V2 = oracle.hartree_expansion(np.ones((1, 1, 1)), [u, u, u], [xe] * 3, [jw] * 3, xq,
                                  alpha, omega, V=V)
    assert np.abs(V2 - 2 * ref).max() < 1e-12 * np.abs(ref).max()


def test_hartree_coulomb_potential_of_gaussian_charge():
    """With the example's 35-term expansion (exponents alphak35 / A^2, weights
    omegak35 / A) the potential of rho = exp(-|x|^2) is the Coulomb one,
    pi^(3/2) erf(r) / r, to the expansion's accuracy."""
    d = json.load(open(GOLDEN))
    A2 = d["Asquare"]
    alpha = np.array(d["alphak35_1e8"]) / A2
    omega = np.array(d["omegak35_1e8"]) / np.sqrt(A2)
    xe, jw = gauss_legendre(-7.0, 7.0, 1400, 10)  # resolves exp(-alpha x^2) up to alpha ~ 2e4
    u = np.exp(-xe ** 2)[:, None]
    xq = [np.array([0.3, 1.1, 2.0]), np.array([-0.7, 0.4]), np.array([0.5, 1.6])]
    V = oracle.hartree_expansion(np.ones((1, 1, 1)), [u, u, u], [xe] * 3, [jw] * 3, xq,
                                 alpha, omega)
    X, Y, Z = np.meshgrid(*xq, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    ref = np.pi ** 1.5 * erf(r) / r
    assert np.abs(V / ref - 1).max() < 2e-5


def test_hartree_matches_dense_reconstruction():
    """General rank: the sum over terms of the reconstructed tensors, formed
    with numpy from the same 1-D integrals."""
    rng = np.random.default_rng(4)
    G = rng.standard_normal((3, 2, 4))
    xe = [np.sort(rng.uniform(-3, 3, n)) for n in (11, 9, 13)]
    jw = [rng.uniform(0.1, 0.5, len(x)) for x in xe]
    U = [rng.standard_normal((len(x), r)) for x, r in zip(xe, G.shape)]
    xq = [rng.uniform(-2, 2, p) for p in (5, 6, 4)]
    alpha = np.array([0.3, 1.7])
    omega = np.array([0.9, -0.2])
    V = oracle.hartree_expansion(G, U, xe, jw, xq, alpha, omega)
    ref = 0.0
    for a, w in zip(alpha, omega):
        F = [np.exp(-a * (q[:, None] - e[None, :]) ** 2) @ (j[:, None] * u)
             for q, e, j, u in zip(xq, xe, jw, U)]
        ref = ref + w * reconstruct(G, F)
    assert np.abs(V - ref).max() < 1e-12 * np.abs(ref).max()
