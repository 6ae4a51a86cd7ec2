"""Tabulated pseudopotentials: the oracle pinned, and the engine's host spline.

* The parabolically terminated spline (alglib spline1dbuildcubic as
  atoms/NonLocalPSPData.cc:382-502 calls it) against the dense global system
  of its defining conditions, and its exact reproduction of quadratics.
* The real harmonics of NonLocalPSPData::computePhi_lm (:527-604) against
  scipy.special.sph_harm_y.
* C_lm of the example's tables on the example's nonlocal mesh (6 quartic
  elements on [-2.7937501, 2.7937501], the 10pt rule: example/input.inp,
  main.cc:358-435) against the 1 / C_lm values the reference printed
  (example/output:48-51).
* The local part's -Z/d tail and knot values (computeLocalPart, :214-275).
* cheby_bounds.txt written and read back through the engine's host calls.
"""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import psp_oracle as P

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "example")


def _table(name):
    return np.loadtxt(os.path.join(EX, name))


def example_tables():
    """(r, vloc, vl (2, nr), rl (2, nr)) of the example's atom type: lmax = 2
    (shell1), V_l from nlpV_AT0 columns 1-2, R_l from nlpWaveFun_AT0 columns
    1-2 (the third, l = 2, is not used with lmax = 2)."""
    V = _table("nlpV_AT0")
    R = _table("nlpWaveFun_AT0")
    L = _table("locPotential_AT0")
    assert np.array_equal(V[:, 0], R[:, 0]) and np.array_equal(V[:, 0], L[:, 0])
    return V[:, 0], L[:, 1], V[:, 1:3].T.copy(), R[:, 1:3].T.copy()


def example_nonlocal_grid():
    """Quadrature points and jac w of one axis of the example's nonlocal mesh."""
    with open(os.path.join(HERE, "golden", "quadrature_rules.json")) as fh:
        rule = json.load(fh)["rules"]["10"]
    a = 2.7937501
    nodes = oracle.fem_mesh("uniform", 6, 5, -a, a)
    _, _, jw, _, xq = oracle.fem_interpolate(6, 5, np.array(rule["points"]),
                                             np.array(rule["weights"]), nodes)
    return xq, jw


def _dense_parabolic_spline(x, y):
    """Coefficients from the 4(n-1) conditions written out as one dense system."""
    n = len(x)
    ns = n - 1
    A = np.zeros((4 * ns, 4 * ns))
    b = np.zeros(4 * ns)
    row = 0
    for s in range(ns):
        h = x[s + 1] - x[s]
        A[row, 4 * s] = 1.0
        b[row] = y[s]
        row += 1
        A[row, 4 * s:4 * s + 4] = [1.0, h, h * h, h ** 3]
        b[row] = y[s + 1]
        row += 1
    for s in range(ns - 1):
        h = x[s + 1] - x[s]
        A[row, 4 * s:4 * s + 4] = [0.0, 1.0, 2 * h, 3 * h * h]
        A[row, 4 * (s + 1) + 1] = -1.0
        row += 1
        A[row, 4 * s:4 * s + 4] = [0.0, 0.0, 2.0, 6 * h]
        A[row, 4 * (s + 1) + 2] = -2.0
        row += 1
    A[row, 3] = 1.0  # third derivative zero on the first segment
    row += 1
    A[row, 4 * (ns - 1) + 3] = 1.0  # and on the last
    row += 1
    assert row == 4 * ns
    return np.linalg.solve(A, b).reshape(ns, 4)


@pytest.mark.parametrize("n", [3, 4, 9, 30])
def test_spline_matches_its_defining_conditions(n):
    rng = np.random.default_rng(n)
    x = np.cumsum(rng.uniform(0.1, 1.0, n))
    y = rng.normal(size=n)
    ref = _dense_parabolic_spline(x, y)
    got = P.spline_parabolic(x, y)
    assert np.max(np.abs(got - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_spline_reproduces_quadratics_and_planted_mistake_fails():
    x = np.array([0.0, 0.3, 0.35, 1.0, 2.5, 2.6, 4.0])
    y = 0.7 - 1.3 * x + 0.4 * x * x
    r = np.linspace(-0.5, 4.5, 101)  # also beyond the table
    c = P.spline_parabolic(x, y)
    exact = 0.7 - 1.3 * r + 0.4 * r * r
    assert np.max(np.abs(P.spline_eval(x, c, r) - exact)) < 1e-12
    # natural end conditions do not reproduce the quadratic: the check sees it
    nat = _natural_spline(x, y)
    assert np.max(np.abs(P.spline_eval(x, nat, r) - exact)) > 1e-3


def _natural_spline(x, y):
    from scipy.interpolate import CubicSpline

    cs = CubicSpline(x, y, bc_type="natural")
    return np.stack([cs.c[3], cs.c[2], cs.c[1], cs.c[0]], axis=1)


def test_engine_host_spline_matches_oracle():
    from ttdft_public_amd import engine as E

    r, vloc, vl, rl = example_tables()
    for y in [vloc, vl[0] - vloc, vl[1] - vloc, rl[0], rl[1]]:
        ref = P.spline_parabolic(r, y)
        got = E.spline_parabolic(r, y)
        scale = np.max(np.abs(ref), axis=0)
        assert np.all(np.abs(got - ref) <= 1e-9 * scale + 1e-300)
        # the values the two forms give agree to rounding along the table
        rr = np.concatenate([r, 0.5 * (r[1:] + r[:-1])])
        assert np.max(np.abs(P.spline_eval(r, got, rr) - P.spline_eval(r, ref, rr))) <= \
            1e-12 * np.max(np.abs(y))
    with pytest.raises(RuntimeError):
        E.spline_parabolic(np.array([0.0, 1.0, 1.0]), np.zeros(3))


def test_real_harmonics_match_scipy():
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(3)
    theta = rng.uniform(0.0, np.pi, 200)
    az = rng.uniform(-np.pi, np.pi, 200)
    Y = P.real_harmonics(4, np.cos(theta), az)
    for l in range(4):
        for m in range(-l, l + 1):
            c = sph_harm_y(l, abs(m), theta, az)
            ref = c.real if m == 0 else np.sqrt(2.0) * (c.imag if m < 0 else c.real)
            assert np.max(np.abs(Y[l * l + l + m] - ref)) < 1e-13, (l, m)


def test_c_lm_matches_reference_output():
    """The reference integrates a rank-25 Tucker form of the same field; its
    printed 1 / C_lm agree with the exact grid sum to 3.1e-8 (s) and 7.0e-9
    (p).  That residual is not the Tucker truncation (a rank-25 STHOSVD of the
    field moves the sum by 6e-12) and is the same for the three p projectors,
    so the bound is set at 1e-7.  A dropped sqrt2, a V_l not differenced with
    V_loc, or swapped l columns move these by 1e-1 or more."""
    with open(os.path.join(HERE, "golden", "example_output_scf.json")) as fh:
        printed = np.array(json.load(fh)["inv_c_lm"])
    r, vloc, vl, rl = example_tables()
    g, w = example_nonlocal_grid()
    got = 1.0 / P.c_lm(r, vloc, vl, rl, g, g, g, w, w, w)
    assert got.shape == printed.shape == (4,)
    assert np.max(np.abs(got / printed - 1.0)) < 1e-7
    # the three p projectors agree by the symmetry of the grid (p_z differs by
    # 3e-12 through the 1e-12 added to r in its polar angle)
    assert np.ptp(got[1:]) < 1e-11 * got[1]
    # a planted mistake (V_l not differenced with the local part) fails
    bad = 1.0 / P.c_lm(r, np.zeros_like(vloc), vl, rl, g, g, g, w, w, w)
    assert np.max(np.abs(bad / printed - 1.0)) > 1e-2


def test_local_part_tail_and_knots():
    r, vloc, _, _ = example_tables()
    # one atom far away: every point is beyond 75.001 -> -Z / d
    atoms = np.array([[3.0, 200.0, 0.0, 0.0, 0.0, 13.0, 0.0],
                      [3.0, 0.0, 0.0, 0.0, 0.0, 13.0, 0.0]])
    gx = np.array([r[10], r[100], r[300], r[450]])
    gy = np.array([0.0])
    gz = np.array([0.0])
    V = P.local_part(atoms, r, vloc, gx, gy, gz)[:, 0, 0]
    far = -3.0 / (np.abs(gx - 200.0) + 1e-12)
    near = vloc[[10, 100, 300, 450]]  # knot distances reproduce the table
    assert np.max(np.abs(V - far - near)) < 1e-12 * np.max(np.abs(V))


def test_restart_bounds_round_trip(tmp_path):
    from ttdft_public_amd import engine as E

    st = E.ChfsiState()
    st.upper, st.lower_unwanted, st.lower_wanted = 1.5, -0.25, -3.0
    st.err_upper, st.occ_energy = 1e-3, -0.1829358112013791671
    p = tmp_path / "cheby_bounds.txt"
    E.restart_write_bounds(str(p), st, -26.75)
    text = p.read_text()
    fields = text.split("\t")
    assert len(fields) == 6 and fields[0] == "1.5000000000000000e+00"
    back = E.ChfsiState()
    e = E.restart_read_bounds(str(p), back)
    assert e == -26.75
    for f in ["upper", "lower_unwanted", "lower_wanted", "err_upper", "occ_energy"]:
        assert getattr(back, f) == getattr(st, f)
    p.write_text("1.0\t2.0\t3.0")
    with pytest.raises(RuntimeError):
        E.restart_read_bounds(str(p), back)
