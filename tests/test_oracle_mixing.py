"""The Anderson mixing oracle (oracle/mixing_oracle.c) against what the
mathematics fixes: the closed form at two history entries, the weighted
least-squares problem the constants solve, and the fixed point of an affine
map, which Anderson mixing reproduces exactly from d + 1 entries in general
position in d dimensions."""
import numpy as np
import pytest

import oracle


def _history(rng, h, n):
    rin = rng.uniform(0.5, 1.5, (h, n))
    rout = rin + 0.1 * rng.standard_normal((h, n))
    jac = rng.uniform(0.5, 2.0, n)
    wts = rng.uniform(0.1, 1.0, n)
    return rin, rout, jac, wts


def test_two_entries_closed_form():
    rin, rout, jac, wts = _history(np.random.default_rng(1), 2, 500)
    f = rout - rin
    d = f[1] - f[0]
    w = jac * wts
    c = oracle.anderson_constants(rin, rout, jac, wts)
    assert c.shape == (1,)
    assert abs(c[0] - np.sum(d * f[1] * w) / np.sum(d * d * w)) < 1e-13 * abs(c[0])


@pytest.mark.parametrize("h", [3, 5, 11])
def test_constants_solve_weighted_least_squares(h):
    """c minimises sum_i w_i (F_n - sum_m c_m (F_n - F_{h-2-m}))_i^2."""
    rin, rout, jac, wts = _history(np.random.default_rng(h), h, 2000)
    f = rout - rin
    D = np.stack([f[h - 1] - f[h - 2 - m] for m in range(h - 1)], axis=1)  # n x (h-1)
    sw = np.sqrt(jac * wts)
    ref = np.linalg.lstsq(sw[:, None] * D, sw * f[h - 1], rcond=None)[0]
    c = oracle.anderson_constants(rin, rout, jac, wts)
    assert np.abs(c - ref).max() < 1e-9 * max(1.0, np.abs(ref).max())
    # normal equations: the weighted residual is orthogonal to every difference
    r = f[h - 1] - D @ c
    g = D.T @ (jac * wts * r)
    assert np.abs(g).max() < 1e-10 * np.abs(D.T @ (jac * wts * f[h - 1])).max()


@pytest.mark.parametrize("alpha", [0.3, 0.5, 1.0])
def test_affine_map_fixed_point(alpha):
    """out = A in + b in d dimensions: d + 1 history entries in general
    position make the mixed residual vanish, so the mix is the fixed point."""
    rng = np.random.default_rng(7)
    d = 4
    A = 0.4 * rng.standard_normal((d, d))
    xstar = rng.uniform(1.0, 2.0, d)
    b = xstar - A @ xstar
    rin = rng.uniform(0.0, 3.0, (d + 1, d))
    rout = rin @ A.T + b
    jac, wts = rng.uniform(0.5, 2.0, d), rng.uniform(0.5, 2.0, d)
    c = oracle.anderson_constants(rin, rout, jac, wts)
    rho = oracle.anderson_mix(rin, rout, c, alpha)
    assert np.abs(rho - xstar).max() < 1e-11


def test_single_entry_is_simple_mixing():
    rin, rout, _, _ = _history(np.random.default_rng(3), 1, 50)
    rho = oracle.anderson_mix(rin, rout, np.zeros(0), 0.25)
    assert np.array_equal(rho, np.abs(0.25 * rout[0] + 0.75 * rin[0]))


def test_mix_takes_absolute_value():
    """A combination that overshoots below zero comes back as its magnitude."""
    rin = np.array([[1.0, 1.0], [2.0, 2.0]])
    rout = np.array([[1.0, 1.0], [2.0, 2.0]])
    rho = oracle.anderson_mix(rin, rout, np.array([3.0]), 0.5)
    # in_mix = out_mix = 3 (1 - 2) + 2 = -1
    assert np.array_equal(rho, np.array([1.0, 1.0]))
