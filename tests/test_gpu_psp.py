"""Tabulated pseudopotentials and restart files on the device, through the C ABI.

* ttdft_psp_projectors on the example's nonlocal grid and tables: the fields
  dV_l phi_lm and phi_lm against the oracle, C_lm against the oracle and
  against the 1 / C_lm the reference printed (example/output:48-51).
* the same for l up to 3 on a ragged off-centre grid with synthetic tables.
* ttdft_psp_local for the example's 13 atoms plus one beyond 75.001 bohr.
* ttdft_restart_write / _read: big-endian bytes, strided blocks, size checks.

Tolerances: the device evaluates the same splines and harmonics with the
device math library (sqrt, atan2, sincos within an ulp or two) and sums C_lm
over 256 blocks in another order: fields 1e-12 relative to their largest
value, C_lm 1e-12 relative.  Against the reference's printed values: 1e-7
(its Tucker-form integral; see test_oracle_psp.py)."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import psp_oracle as P
from ttdft_public_amd import engine as E

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "example")


def dev(ctx, a):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def tables():
    V = np.loadtxt(os.path.join(EX, "nlpV_AT0"))
    R = np.loadtxt(os.path.join(EX, "nlpWaveFun_AT0"))
    L = np.loadtxt(os.path.join(EX, "locPotential_AT0"))
    return V[:, 0], L[:, 1], V[:, 1:3].T.copy(), R[:, 1:3].T.copy()


def nonlocal_grid():
    xi, w = E.quadrature_rule(10)
    a = 2.7937501
    nodes = E.fem_mesh("uniform", 6, 5, -a, a)
    _, _, jw, _, xq = oracle.fem_interpolate(6, 5, xi, w, nodes)
    return xq, jw


def run_projectors(ctx, r, vloc, vl, rl, gx, gy, gz, w=None):
    F = vl.shape[0] ** 2
    shape = (len(gx), len(gy), len(gz))
    npts = int(np.prod(shape))
    dg = [dev(ctx, g) for g in (gx, gy, gz)]
    dw = [dev(ctx, x) for x in w] if w is not None else [None, None, None]
    dfield = E.DeviceArray(ctx, (F * npts,))
    dphi = E.DeviceArray(ctx, (F * npts,))
    clm = ctx.psp_projectors(r, vloc, vl, rl, dg[0].ptr, dg[1].ptr, dg[2].ptr, shape,
                             *(x.ptr if x is not None else None for x in dw),
                             d_field=dfield.ptr, d_phi=dphi.ptr)
    # device fields are x fastest: (F, qz, qy, qx) -> (F, qx, qy, qz)
    field = dfield.to_host().reshape(F, shape[2], shape[1], shape[0]).transpose(0, 3, 2, 1)
    phi = dphi.to_host().reshape(F, shape[2], shape[1], shape[0]).transpose(0, 3, 2, 1)
    return field, phi, clm


def close(a, b, tol=1e-12):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


def test_psp_projectors_example_matches_oracle_and_reference(ctx):
    print("psp_projectors: example tables, 60^3 nonlocal grid", flush=True)
    r, vloc, vl, rl = tables()
    g, w = nonlocal_grid()
    field, phi, clm = run_projectors(ctx, r, vloc, vl, rl, g, g, g, (w, w, w))
    f0, p0, _ = P.projector_fields(r, vloc, vl, rl, g, g, g)
    for f in range(4):
        assert close(field[f], f0[f]) and close(phi[f], p0[f]), f
    c0 = P.c_lm(r, vloc, vl, rl, g, g, g, w, w, w)
    assert np.max(np.abs(clm / c0 - 1.0)) < 1e-12
    with open(os.path.join(HERE, "golden", "example_output_scf.json")) as fh:
        printed = np.array(json.load(fh)["inv_c_lm"])
    assert np.max(np.abs((1.0 / clm) / printed - 1.0)) < 1e-7


def test_psp_projectors_higher_l_ragged_grid(ctx):
    print("psp_projectors: lmax 4, 7 x 5 x 9 grid", flush=True)
    r = np.concatenate([[0.0], np.geomspace(1e-3, 12.0, 80)])
    vloc = -2.0 * np.exp(-r * r / 3.0)
    vl = np.stack([np.exp(-r * r / (1.0 + l)) * (1.0 + 0.1 * l) for l in range(4)])
    rl = np.stack([r ** l * np.exp(-r * r / 2.0) for l in range(4)])
    gx = np.array([-3.0, -1.1, -0.2, 0.0, 0.4, 1.7, 2.9])
    gy = np.array([-2.5, 0.0, 0.3, 1.2, 4.0])
    gz = np.array([-3.3, -2.0, -0.7, -0.1, 0.0, 0.5, 1.5, 2.2, 13.0])  # 13: past the table
    wx, wy, wz = (np.linspace(0.1, 1.0, len(a)) for a in (gx, gy, gz))
    field, phi, clm = run_projectors(ctx, r, vloc, vl, rl, gx, gy, gz, (wx, wy, wz))
    f0, p0, _ = P.projector_fields(r, vloc, vl, rl, gx, gy, gz)
    for f in range(16):
        assert close(field[f], f0[f]) and close(phi[f], p0[f]), f
    c0 = P.c_lm(r, vloc, vl, rl, gx, gy, gz, wx, wy, wz)
    assert np.max(np.abs(clm - c0)) <= 1e-12 * np.max(np.abs(c0))
    # no weights: fields only
    field2, _, none = run_projectors(ctx, r, vloc, vl, rl, gx, gy, gz)
    assert none is None and np.array_equal(field2, field)


def test_psp_projectors_rejects_bad_arguments(ctx):
    r, vloc, vl, rl = tables()
    g = dev(ctx, np.zeros(3))
    with pytest.raises(RuntimeError):
        ctx.psp_projectors(r, vloc, np.zeros((5, len(r))), np.zeros((5, len(r))), g.ptr, g.ptr,
                           g.ptr, (3, 1, 1))
    with pytest.raises(RuntimeError):
        ctx.psp_projectors(r[::-1].copy(), vloc, vl, rl, g.ptr, g.ptr, g.ptr, (3, 1, 1))


def test_psp_local_matches_oracle(ctx):
    print("psp_local: 13 atoms + one far atom, 23 x 17 x 11 grid", flush=True)
    r, vloc, _, _ = tables()
    _, _, atoms = E.read_atom_file(os.path.join(EX, "shell1"))
    far = np.array([[3.0, 90.0, 10.0, -5.0, 0.0, 13.0, 0.0]])
    atoms = np.concatenate([atoms, far])
    gx = np.linspace(-16.0, 16.0, 23)
    gy = np.linspace(-15.0, 14.0, 17)
    gz = np.linspace(-16.0, 16.0, 11)
    dg = [dev(ctx, g) for g in (gx, gy, gz)]
    dv = E.DeviceArray(ctx, (len(gx) * len(gy) * len(gz),))
    ctx.psp_local(atoms, r, vloc, dg[0].ptr, dg[1].ptr, dg[2].ptr, (23, 17, 11), dv.ptr)
    got = dv.to_host().reshape(11, 17, 23).transpose(2, 1, 0)
    ref = P.local_part(atoms, r, vloc, gx, gy, gz)
    assert close(got, ref)
    ctx.psp_local(atoms, r, vloc, dg[0].ptr, dg[1].ptr, dg[2].ptr, (23, 17, 11), dv.ptr,
                  add=True)
    assert close(dv.to_host().reshape(11, 17, 23).transpose(2, 1, 0), 2.0 * ref)


def test_restart_block_round_trip(ctx, tmp_path):
    print("restart_write/read: 1000 x 7 basis-major block, ld 9; FP32; a grid field", flush=True)
    n, k, ld = 1000, 7, 9
    rng = np.random.default_rng(11)
    X = rng.standard_normal((n, ld))  # basis-major: X[i, j] = state j of function i
    dX = dev(ctx, X)
    p = str(tmp_path / "wfn_proc0")
    ctx.restart_write(p, dX.ptr, ld, n, k)
    raw = open(p, "rb").read()
    # the reference's MATDENSE order: n values of state 0, then state 1, ... big-endian
    assert raw == np.ascontiguousarray(X[:, :k].T).astype(">f8").tobytes()
    dY = E.DeviceArray(ctx, (n * k,))
    ctx.restart_read(p, dY.ptr, k, n, k)
    assert np.array_equal(dY.to_host().reshape(n, k), X[:, :k])
    # FP32 blocks are widened on the way out and rounded back exactly
    X32 = X[:, :k].astype(np.float32)
    d32 = E.DeviceArray.from_host(ctx, np.ascontiguousarray(X32))
    ctx.restart_write(p, d32.ptr, k, n, k, dtype=E.F32)
    assert open(p, "rb").read() == np.ascontiguousarray(X32.T).astype(">f8").tobytes()
    back = E.DeviceArray(ctx, (n * k,), np.float32)
    ctx.restart_read(p, back.ptr, k, n, k, dtype=E.F32)
    assert np.array_equal(back.to_host().reshape(n, k), X32)
    # a grid field (the local part's potloc file): k = 1, ld = 1
    g = rng.standard_normal(4096)
    dg = dev(ctx, g)
    q = str(tmp_path / "potloc_proc0")
    ctx.restart_write(q, dg.ptr, 1, 4096, 1)
    assert open(q, "rb").read() == g.astype(">f8").tobytes()
    with pytest.raises(RuntimeError):
        ctx.restart_read(p, dY.ptr, k - 1, n, k - 1)  # file holds more
    with pytest.raises(RuntimeError):
        ctx.restart_read(p, dY.ptr, k, n + 1, k)  # file holds fewer
    with pytest.raises(RuntimeError):
        ctx.restart_read(str(tmp_path / "missing"), dY.ptr, k, n, k)
