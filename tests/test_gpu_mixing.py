"""Device-resident Anderson mixing (ttdft_mixer_*) against the mixing oracle,
through the C ABI, over whole SCF sequences: first iteration (absolute value),
second (simple mixing), then Anderson steps with a history that fills up,
drops its oldest entries and is reset.

Tolerances: the constants solve a small linear system whose matrix the
device sums in a different order (matrix-core Gram over the grid): relative
1e-9 of the largest constant; mixed densities relative 1e-11."""
import numpy as np
import pytest

import oracle
from ttdft_public_amd import engine as E

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def run_sequence(ctx, n_grid, n_nodal, alpha, max_hist, nscf, clear_at=None, seed=0):
    rng = np.random.default_rng(seed)
    jac = rng.uniform(0.5, 2.0, n_grid)
    wts = rng.uniform(0.1, 1.0, n_grid)
    dj, dw = dev(ctx, jac), dev(ctx, wts)
    mix = E.Mixer(ctx, n_grid, n_nodal, dj.ptr, dw.ptr, alpha, max_hist)
    # a fixed contraction towards a positive density, plus noise, as "the SCF map"
    target_g = rng.uniform(0.5, 1.5, n_grid)
    target_n = rng.uniform(0.5, 1.5, n_nodal)
    rho_g = rng.uniform(-0.2, 1.8, n_grid)  # some negative entries: iteration 0 takes |.|
    rho_n = rng.uniform(-0.2, 1.8, n_nodal)
    hist = []  # (g_in, g_out, n_in, n_out)
    offset = 0
    checked = 0
    for scf in range(nscf):
        dg, dn = dev(ctx, rho_g), dev(ctx, rho_n) if n_nodal else None
        print(f"mixer scf={scf} history={len(hist)} n_grid={n_grid} n_nodal={n_nodal}", flush=True)
        mix.compute_rho_in(scf, dn.ptr if dn else None, dg.ptr)
        off = scf - offset
        if off == 0:
            exp_g, exp_n = np.abs(rho_g), np.abs(rho_n)
        elif off == 1:
            gi, go, ni, no = hist[-1]
            exp_g = alpha * go + (1 - alpha) * gi
            exp_n = alpha * no + (1 - alpha) * ni
        else:
            gin = np.stack([e[0] for e in hist])
            gout = np.stack([e[1] for e in hist])
            c = oracle.anderson_constants(gin, gout, jac, wts)
            got_c = mix.constants()
            assert got_c.shape == c.shape
            if c.size:
                assert np.abs(got_c - c).max() <= 1e-9 * max(1.0, np.abs(c).max()), scf
            exp_g = oracle.anderson_mix(gin, gout, c, alpha)
            exp_n = oracle.anderson_mix(np.stack([e[2] for e in hist]),
                                        np.stack([e[3] for e in hist]), c, alpha) \
                if n_nodal else np.zeros(0)
            checked += 1
        got_g = dg.to_host()
        assert np.abs(got_g - exp_g).max() <= 1e-11 * np.abs(exp_g).max(), scf
        if n_nodal:
            got_n = dn.to_host()
            assert np.abs(got_n - exp_n).max() <= 1e-11 * np.abs(exp_n).max(), scf
        rho_g, rho_n = got_g, dn.to_host() if n_nodal else rho_n
        out_g = target_g + 0.6 * (rho_g - target_g) + 0.01 * rng.standard_normal(n_grid)
        out_n = target_n + 0.6 * (rho_n - target_n) + 0.01 * rng.standard_normal(n_nodal)
        dog, don = dev(ctx, out_g), dev(ctx, out_n) if n_nodal else None
        mix.update(dn.ptr if dn else None, dg.ptr, don.ptr if don else None, dog.ptr, scf)
        hist.append((rho_g.copy(), out_g, rho_n.copy(), out_n))
        hist = hist[-max_hist:]
        assert mix.history_size() == len(hist)
        if clear_at is not None and scf == clear_at:
            mix.clear_history(scf)
            hist = []
            offset = scf + 1
        rho_g, rho_n = out_g, out_n
    mix.free()
    return checked


@pytest.mark.parametrize("n_grid,n_nodal,alpha,max_hist,nscf", [
    (30000, 4000, 0.5, 10, 9),      # history never full
    (30000, 4000, 0.5, 3, 9),       # oldest entries dropped
    (12345, 0, 0.3, 5, 7),          # grid only, ragged chunk
    (700, 9, 0.8, 20, 24),          # h - 1 = 19 > 16: two row and column tiles of the Gram
    (5000, 100, 0.5, 1, 5),         # history of one: Anderson with no constants
])
def test_mixer_sequence_matches_oracle(ctx, n_grid, n_nodal, alpha, max_hist, nscf):
    assert run_sequence(ctx, n_grid, n_nodal, alpha, max_hist, nscf) > 0


def test_mixer_clear_history_restarts_the_sequence(ctx):
    assert run_sequence(ctx, 20000, 3000, 0.5, 6, 11, clear_at=5, seed=3) > 0


def test_mixer_rejects_compute_without_history(ctx):
    jac = np.ones(100)
    dj = dev(ctx, jac)
    mix = E.Mixer(ctx, 100, 0, dj.ptr, dj.ptr, 0.5, 4)
    dg = dev(ctx, np.ones(100))
    with pytest.raises(E.TtdftError):
        mix.compute_rho_in(1, None, dg.ptr)
    with pytest.raises(E.TtdftError):
        E.Mixer(ctx, 100, 0, dj.ptr, dj.ptr, 0.5, 65)


def test_mixer_affine_fixed_point_large_grid(ctx):
    """300^3 grid, 5 history entries of an affine map acting on a 4-dimensional
    subspace: the mixed density is the map's fixed point (size-independent)."""
    n = 300 ** 3
    rng = np.random.default_rng(11)
    d = 4
    xstar = 1.0 + 0.5 * np.sin(np.arange(n) * 1e-3)
    V = [0.1 * rng.standard_normal(n) for _ in range(d)]
    M = 0.5 * rng.standard_normal((d, d))
    a = rng.standard_normal((d + 1, d))
    jac = rng.uniform(0.5, 2.0, n)
    wts = rng.uniform(0.1, 1.0, n)
    dj, dw = dev(ctx, jac), dev(ctx, wts)
    del jac, wts
    mix = E.Mixer(ctx, n, 0, dj.ptr, dw.ptr, 0.5, 8)
    for k in range(d + 1):
        x_in = xstar + sum(a[k, j] * V[j] for j in range(d))
        ma = M @ a[k]
        x_out = xstar + sum(ma[j] * V[j] for j in range(d))
        di, do = dev(ctx, x_in), dev(ctx, x_out)
        mix.update(None, di.ptr, None, do.ptr, k)
        del di, do, x_in, x_out
    drho = dev(ctx, np.zeros(n))
    print("mixer 300^3 affine fixed point", flush=True)
    mix.compute_rho_in(d + 1, None, drho.ptr)
    assert mix.constants().shape == (d,)
    assert np.abs(drho.to_host() - xstar).max() < 1e-9
