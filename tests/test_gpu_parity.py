"""HIP engine vs the CPU oracle on identical seeded inputs, through the C ABI.

Tolerances: FP64 paths agree to rounding (relative 1e-12 or tighter; the
summation order differs between the GPU kernels and the oracle); the FP32
SpMM to 1e-5 relative; split-BF16 Gram/rotation to 5e-5 relative.
"""
import numpy as np
import pytest

import oracle
from ttdft_public_amd import engine as E

pytestmark = pytest.mark.gpu


def dev(ctx, a, dtype=np.float64):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=dtype))


def synth(p=5, dims=(8, 8, 8), bands=(2, 2, 3), seed=7):
    s = E.SynthHam(*dims, *bands, p, seed=seed)
    return s, oracle.Ham.from_synth(s)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---------------------------------------------------------------- H apply
@pytest.mark.parametrize("k", [1, 7, 64, 65, 130, 300])
def test_ham_apply_f64(ctx, k):
    s, H = synth()
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32)
    X = np.random.default_rng(k).standard_normal((s.n, k))
    dX, dY = dev(ctx, X), E.DeviceArray(ctx, (s.n, k))
    ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, E.F64)
    assert rel(dY.to_host(), oracle.ham_apply(H, X)) < 1e-13


@pytest.mark.parametrize("k", [5, 64, 300])
def test_ham_apply_f32(ctx, k):
    s, H = synth()
    ham = ctx.ham_upload(s, E.HAM_F32)
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(np.float32)
    dX, dY = dev(ctx, X, np.float32), E.DeviceArray(ctx, (s.n, k), np.float32)
    ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, E.F32)
    assert rel(dY.to_host().astype(np.float64), oracle.ham_apply(H, X.astype(np.float64))) < 1e-5


@pytest.mark.parametrize("k", [1, 70, 300])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_ham_apply_wide_projectors(ctx, k, dtype):
    """Few projectors with ~1300 nonzeros each: the projection is split over blocks."""
    s, H = synth(p=4, dims=(20, 18, 16), bands=(2, 2, 2), seed=1)
    assert s.nnz_c // 4 > 1024
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=True)
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(npdt)
    dX, dY = dev(ctx, X, npdt), E.DeviceArray(ctx, (s.n, k), npdt)
    ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, dtype)
    ref = oracle.ham_apply(H, X.astype(np.float64))
    assert rel(dY.to_host().astype(np.float64), ref) < (1e-13 if dtype == E.F64 else 1e-5)


def test_ham_apply_ragged_and_no_projectors(ctx):
    """Empty rows, a single dense row, p = 0, leading dimension > k."""
    n = 200
    rng = np.random.default_rng(3)
    rows = []
    for i in range(n):
        if i % 7 == 0:
            rows.append([])
        elif i == 5:
            rows.append(list(range(n)))
        else:
            rows.append(sorted(set(rng.integers(0, n, size=rng.integers(1, 90)).tolist())))
    # symmetrise the pattern with random values
    A = np.zeros((n, n))
    for i, cs in enumerate(rows):
        for c in cs:
            A[i, c] = A[c, i] = rng.standard_normal()
    rowptr = np.zeros(n + 1, np.int64)
    cols, vals = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        cols.extend(nz.tolist())
        vals.extend(A[i, nz].tolist())
        rowptr[i + 1] = len(cols)
    ham = E.Hamiltonian(ctx, n, rowptr, np.array(cols, np.int32), np.array(vals), 0, None, None,
                        None, None)
    k, ld = 70, 80
    X = rng.standard_normal((n, ld))
    dX, dY = dev(ctx, X), dev(ctx, np.full((n, ld), 7.0))
    ctx.ham_apply(ham, dX.ptr, ld, dY.ptr, ld, k, E.F64)
    Y = dY.to_host()
    assert rel(Y[:, :k], A @ X[:, :k]) < 1e-13
    assert np.all(Y[:, k:] == 7.0)  # padding columns untouched


def test_ham_upload_rejects_bad_csr(ctx):
    with pytest.raises(E.TtdftError) as ei:
        E.Hamiltonian(ctx, 3, np.array([0, 1, 2, 3]), np.array([0, 5, 2], np.int32),
                      np.ones(3), 0, None, None, None, None)
    assert ei.value.status == E.EBADARG


# ---------------------------------------------------------------- band format
@pytest.mark.parametrize("dims,bands,k", [((11, 9, 7), (3, 2, 2), 16), ((20, 10, 6), (2, 3, 1), 64),
                                          ((12, 12, 12), (6, 6, 7), 100),
                                          ((50, 6, 5), (6, 2, 2), 300),
                                          ((70, 5, 4), (3, 1, 1), 40),    # two x-chunks
                                          ((6, 7, 5), (1, 2, 2), 33),     # padded narrow chunk
                                          ((17, 9, 8), (7, 3, 2), 81),    # widest x band
                                          ((22, 6, 5), (2, 2, 1), 50),    # two split x-bricks
                                          ((12, 12, 12), (6, 6, 7), 900),  # 900 states: 3 / 12 groups
                                          ((17, 9, 8), (7, 3, 2), 900)])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_band_spmm_matches_oracle(ctx, dims, bands, k, dtype):
    s = E.SynthHam(*dims, *bands, 4, seed=sum(dims))
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=True)
    fmt, b = ham.format()
    assert fmt == 1 and b == tuple(bands)
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(npdt)
    dX, dY = dev(ctx, X, npdt), E.DeviceArray(ctx, (s.n, k), npdt)
    ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, dtype)
    ref = oracle.ham_apply(H, X.astype(np.float64))
    assert rel(dY.to_host().astype(np.float64), ref) < (1e-13 if dtype == E.F64 else 1e-5)


@pytest.mark.parametrize("dims,bands", [((20, 18, 16), (6, 6, 7)), ((13, 11, 9), (7, 3, 2)),
                                         ((9, 10, 7), (1, 2, 2))])
@pytest.mark.parametrize("k", [257, 300, 400])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_band_spmm_state_groups_in_one_workgroup(ctx, monkeypatch, dims, bands, k, dtype):
    """More than 4 state groups' worth of states (FP64 > 256, FP32 > 240) runs
    all state groups of a row block in one workgroup (x-chunks of 3 or 4
    bricks, the last one ragged); it must give the same bits as one workgroup
    per state group, and match the oracle."""
    s = E.SynthHam(*dims, *bands, 4, seed=sum(dims) + k)
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=True)
    assert ham.format()[0] == 1
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(npdt)
    dX, dY = dev(ctx, X, npdt), E.DeviceArray(ctx, (s.n, k), npdt)
    out = {}
    for co in ("1", "0"):
        monkeypatch.setenv("TTDFT_BRICK_COGROUPS", co)
        ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, dtype)
        out[co] = dY.to_host()
    assert np.array_equal(out["1"], out["0"])
    ref = oracle.ham_apply(H, X.astype(np.float64))
    assert rel(out["1"].astype(np.float64), ref) < (1e-13 if dtype == E.F64 else 1e-5)


@pytest.mark.parametrize("dims,bands,k", [((20, 18, 16), (6, 6, 7), 300),
                                          ((13, 11, 9), (7, 3, 2), 320),
                                          ((9, 10, 7), (1, 2, 2), 257),
                                          ((17, 9, 5), (7, 3, 2), 80),
                                          ((22, 6, 3), (2, 2, 1), 50)])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_band_spmm_two_planes_per_workgroup(ctx, monkeypatch, dims, bands, k, dtype):
    """Workgroups over two z planes (a staged line serves both planes in its
    z band) give the same bits as one plane per workgroup, odd plane counts
    included (the last workgroup has one plane), and match the oracle."""
    s = E.SynthHam(*dims, *bands, 4, seed=sum(dims) + k)
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=True)
    assert ham.format()[0] == 1
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(npdt)
    dX, dY = dev(ctx, X, npdt), E.DeviceArray(ctx, (s.n, k), npdt)
    out = {}
    for zb in ("1", "2"):
        monkeypatch.setenv("TTDFT_BRICK_ZB", zb)
        print(f"band two planes dims={dims} k={k} dtype={dtype} zb={zb}", flush=True)
        ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, dtype)
        out[zb] = dY.to_host()
    assert np.array_equal(out["1"], out["2"])
    ref = oracle.ham_apply(H, X.astype(np.float64))
    assert rel(out["2"].astype(np.float64), ref) < (1e-13 if dtype == E.F64 else 1e-5)


def test_band_filter_with_holes_matches_oracle(ctx):
    """Truncation leaves holes inside the band: stored as zeros, same result."""
    s = E.SynthHam(10, 10, 10, 4, 4, 4, 6, truncation=2e-5, seed=3)
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, tucker=True)
    assert ham.format()[0] == 1
    assert s.nnz < 1000 * 9 ** 3  # some entries were truncated
    k, m = 48, 9
    X = np.random.default_rng(2).standard_normal((s.n, k))
    dX = dev(ctx, X)
    ctx.cheb_filter(ham, dX.ptr, k, k, m, -0.5, 30.0, -3.0, 2, E.F64)
    ref = oracle.cheb_filter(H, X, m, -0.5, 30.0, -3.0, 2)
    assert rel(dX.to_host(), ref) < 1e-12


def test_band_rejects_sparse_band_keeps_csr(ctx):
    """A band with too few entries stays in CSR (and stays correct)."""
    n = 6 * 6 * 6
    rows = [[i, (i + 7 * 36) % n] for i in range(n)]  # far coupling -> wide, empty band
    A = np.zeros((n, n))
    for i, cs in enumerate(rows):
        for c in cs:
            A[i, c] = A[c, i] = 1.0 + 0.01 * i
    rowptr = np.zeros(n + 1, np.int64)
    cols, vals = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        cols += nz.tolist()
        vals += A[i, nz].tolist()
        rowptr[i + 1] = len(cols)
    ham = E.Hamiltonian(ctx, n, rowptr, np.array(cols, np.int32), np.array(vals), 0, None, None,
                        None, None, dims=(6, 6, 6))
    assert ham.format()[0] == 0
    X = np.random.default_rng(1).standard_normal((n, 32))
    dX, dY = dev(ctx, X), E.DeviceArray(ctx, (n, 32))
    ctx.ham_apply(ham, dX.ptr, 32, dY.ptr, 32, 32)
    assert rel(dY.to_host(), A @ X) < 1e-13


# ---------------------------------------------------------------- filter
@pytest.mark.parametrize("convention", [E.DEGREE_REFERENCE, E.DEGREE_EXACT])
@pytest.mark.parametrize("m,nb", [(1, 1), (2, 1), (5, 1), (8, 3)])
def test_cheb_filter_f64(ctx, convention, m, nb):
    s, H = synth()
    ham = ctx.ham_upload(s)
    k = 24
    X = np.random.default_rng(m).standard_normal((s.n, k))
    a, b, a0 = -5.0, 20.0, -30.0
    dX = dev(ctx, X)
    ctx.cheb_filter(ham, dX.ptr, k, k, m, a, b, a0, nb, E.F64, convention)
    ref = oracle.cheb_filter(H, X, m, a, b, a0, nb, convention)
    assert rel(dX.to_host(), ref) < 1e-12


def test_cheb_filter_f32(ctx):
    s, H = synth()
    ham = ctx.ham_upload(s, E.HAM_F32)
    k, m = 40, 10
    X = np.random.default_rng(9).standard_normal((s.n, k)).astype(np.float32)
    dX = dev(ctx, X, np.float32)
    ctx.cheb_filter(ham, dX.ptr, k, k, m, -5.0, 20.0, -30.0, 1, E.F32)
    ref = oracle.cheb_filter(H, X.astype(np.float64), m, -5.0, 20.0, -30.0)
    assert rel(dX.to_host().astype(np.float64), ref) < 5e-5


# ---------------------------------------------------------------- dense
@pytest.mark.parametrize("n,k", [(512, 24), (3000, 65), (5000, 300)])
def test_gram_f64(ctx, n, k):
    rng = np.random.default_rng(n + k)
    X, Y = rng.standard_normal((n, k)), rng.standard_normal((n, k))
    dS = E.DeviceArray(ctx, (k, k))
    dX, dY = dev(ctx, X), dev(ctx, Y)  # keep the buffers alive through the call
    ctx.gram(dX.ptr, k, dY.ptr, k, n, k, dS.ptr)
    S = dS.to_host().T  # column-major on device
    assert rel(S, oracle.gram(X, Y)) < 1e-13


@pytest.mark.parametrize("n,k,ld,same", [(5000, 300, 300, True), (5000, 300, 300, False),
                                          (3001, 161, 163, True), (3001, 161, 163, False),
                                          (777, 333, 333, True), (100, 17, 17, False),
                                          (20000, 900, 900, True), (40, 160, 161, False)])
def test_gram_f64_shapes(ctx, n, k, ld, same):
    """FP64 Gram over its tilings (160-wide blocks, 16-row steps staged in LDS):
    ragged k and n, odd leading dimension (scalar loads), fewer rows than
    one step.  X^T X takes the upper-triangle form and is exactly
    symmetric."""
    rng = np.random.default_rng(n + k + same)
    X = rng.standard_normal((n, ld))
    Y = X if same else rng.standard_normal((n, ld))
    dX = dev(ctx, X)
    dY = dX if same else dev(ctx, Y)
    dS = E.DeviceArray(ctx, (k, k))
    print(f"gram f64 n={n} k={k} ld={ld} same={same}", flush=True)
    ctx.gram(dX.ptr, ld, dY.ptr, ld, n, k, dS.ptr)
    S = dS.to_host().reshape(k, k).T
    assert rel(S, oracle.gram(np.ascontiguousarray(X[:, :k]), np.ascontiguousarray(Y[:, :k]))) \
        < 1e-13
    if same:
        assert np.array_equal(S, S.T)


@pytest.mark.parametrize("n,k,ld", [(1000, 24, 24), (5001, 131, 133), (3000, 300, 301),
                                     (2000, 300, 300), (129, 17, 20), (700, 900, 900)])
def test_rotate_f64_shapes(ctx, n, k, ld):
    """FP64 rotation from pre-arranged Q fragments: several groups of eight
    column tiles, k not a multiple of 16 or 4, odd leading dimension (scalar
    loads), n not a multiple of the 128-row workgroup.  Columns past k are
    left alone."""
    rng = np.random.default_rng(n + k + ld)
    X = rng.standard_normal((n, ld))
    Q = rng.standard_normal((k, k)) / np.sqrt(k)
    dX = dev(ctx, X)
    dQ = dev(ctx, Q.T.copy())  # column-major
    print(f"rotate f64 n={n} k={k} ld={ld}", flush=True)
    ctx.rotate(dX.ptr, ld, n, k, dQ.ptr, E.F64, E.GEMM_F64)
    got = dX.to_host().reshape(n, ld)
    assert rel(got[:, :k], oracle.rotate(np.ascontiguousarray(X[:, :k]), Q)) < 1e-13
    assert np.array_equal(got[:, k:], X[:, k:])


@pytest.mark.parametrize("n,k", [(512, 24), (5000, 300)])
def test_gram_bf16x3(ctx, n, k):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, k)).astype(np.float32)
    Y = rng.standard_normal((n, k)).astype(np.float32)
    dS = E.DeviceArray(ctx, (k, k))
    dX, dY = dev(ctx, X, np.float32), dev(ctx, Y, np.float32)
    ctx.gram(dX.ptr, k, dY.ptr, k, n, k, dS.ptr, E.F32, E.GEMM_BF16X3)
    ref = X.astype(np.float64).T @ Y.astype(np.float64)
    # error relative to the natural scale sqrt(n) of a Gram entry
    assert np.abs(dS.to_host().T - ref).max() / np.sqrt(n) < 5e-5


@pytest.mark.parametrize("n,k,ld,same", [(5000, 300, 300, True), (5000, 300, 300, False),
                                          (3001, 161, 163, True), (3001, 161, 163, False),
                                          (777, 333, 336, True), (100, 17, 17, False),
                                          (20000, 900, 900, True), (40, 160, 160, False)])
def test_gram_bf16x3_shapes(ctx, n, k, ld, same):
    """Split-BF16 Gram over its tilings: one to six 160-wide blocks per
    dimension, k not a multiple of 16 or of 4, rows not 16-byte aligned (odd
    leading dimension: scalar loads), n below one 16-row step and not a
    multiple of it.  X^T X (same pointer) runs the upper-triangle form and
    must come out exactly symmetric.

    The bar is elementwise against |X|^T |Y|: the three-term split drops
    lo*lo' (up to 2^-18 of each product) and rounds x = hi + lo to ~2^-17,
    which on the diagonal of X^T X is a one-signed bias growing with n (a
    numpy emulation of the split on the 5000 x 300 case gives 3.1e-6 of
    |X|^T |X|, the same 0.0155 maximum error the kernel shows), so the
    sqrt(n) scale of a random-sign sum does not bound it."""
    rng = np.random.default_rng(n + k + same)
    X = rng.standard_normal((n, ld)).astype(np.float32)
    Y = X if same else rng.standard_normal((n, ld)).astype(np.float32)
    dX = dev(ctx, X, np.float32)
    dY = dX if same else dev(ctx, Y, np.float32)
    dS = E.DeviceArray(ctx, (k, k))
    print(f"gram bf16x3 n={n} k={k} ld={ld} same={same}", flush=True)
    ctx.gram(dX.ptr, ld, dY.ptr, ld, n, k, dS.ptr, E.F32, E.GEMM_BF16X3)
    S = dS.to_host().reshape(k, k).T
    X64, Y64 = X[:, :k].astype(np.float64), Y[:, :k].astype(np.float64)
    ref = X64.T @ Y64
    scale = np.abs(X64).T @ np.abs(Y64)
    assert np.all(np.abs(S - ref) <= 1e-5 * scale + 1e-12)
    if same:
        assert np.array_equal(S, S.T)


@pytest.mark.parametrize("n,k,ld", [(1000, 24, 24), (5001, 131, 133), (3000, 300, 301),
                                     (2000, 300, 300), (129, 17, 20)])
def test_rotate_bf16x3_shapes(ctx, n, k, ld):
    """Split-BF16 rotation over its tilings: several groups of eight column
    tiles (k = 300), k not a multiple of 16, rows not 16-byte aligned
    (odd leading dimension: scalar loads), n not a multiple of the 128-row
    workgroup.  Columns past k are left alone."""
    rng = np.random.default_rng(n + k)
    X = rng.standard_normal((n, ld)).astype(np.float32)
    Q = rng.standard_normal((k, k)) / np.sqrt(k)
    dX = dev(ctx, X, np.float32)
    dQ = dev(ctx, Q.T.copy())  # column-major
    print(f"rotate bf16x3 n={n} k={k} ld={ld}", flush=True)
    ctx.rotate(dX.ptr, ld, n, k, dQ.ptr, E.F32, E.GEMM_BF16X3)
    got = dX.to_host().reshape(n, ld).astype(np.float64)
    ref = oracle.rotate(X[:, :k].astype(np.float64), Q)
    assert rel(got[:, :k], ref) < 5e-5
    assert np.array_equal(got[:, k:], X[:, k:].astype(np.float64))


@pytest.mark.parametrize("mode", [E.GEMM_F64, E.GEMM_BF16X3])
def test_rotate_and_orth(ctx, mode):
    n, k = 4000, 100
    rng = np.random.default_rng(mode)
    f32 = mode == E.GEMM_BF16X3
    dt, npdt = (E.F32, np.float32) if f32 else (E.F64, np.float64)
    X = rng.standard_normal((n, k)).astype(npdt)
    Q = rng.standard_normal((k, k))
    dX = dev(ctx, X, npdt)
    dQ = dev(ctx, Q.T.copy())  # column-major
    ctx.rotate(dX.ptr, k, n, k, dQ.ptr, dt, mode)
    ref = oracle.rotate(X.astype(np.float64), Q)
    tol = 1e-13 if not f32 else 5e-5
    assert rel(dX.to_host().astype(np.float64), ref) < tol
    ctx.orth(dX.ptr, k, n, k, dt, mode)
    Xo = dX.to_host().astype(np.float64)
    # Cholesky-QR leaves an orthogonality error of order eps * cond(X)^2
    assert np.abs(Xo.T @ Xo - np.eye(k)).max() < (1e-10 if not f32 else 1e-5)
    # same span as the input
    assert np.abs(Xo @ (Xo.T @ ref) - ref).max() / np.abs(ref).max() < (1e-11 if not f32 else 1e-4)


@pytest.mark.parametrize("k", [40, 64, 80])  # k <= 64: one-workgroup Cholesky inverse
def test_orth_matches_oracle_f64(ctx, k):
    n = 2000
    X = np.random.default_rng(1).standard_normal((n, k))
    dX = dev(ctx, X)
    ctx.orth(dX.ptr, k, n, k)
    assert rel(dX.to_host(), oracle.orth(X)) < 1e-11


def test_orth_not_spd_reports_error(ctx):
    n, k = 300, 8
    X = np.random.default_rng(2).standard_normal((n, k))
    X[:, 5] = 0.0  # singular overlap: pivot 5 is exactly zero
    dX = dev(ctx, X)
    with pytest.raises(E.TtdftError) as ei:
        ctx.orth(dX.ptr, k, n, k)
    assert ei.value.status in (E.ENOTSPD,)


@pytest.mark.parametrize("k", [1, 7, 24, 64, 65, 300])  # k <= 64: one-workgroup Jacobi
def test_rr_eig(ctx, k):
    A = np.random.default_rng(k).standard_normal((k, k))
    A = A + A.T
    dA = dev(ctx, A.T.copy())
    dE, dQ = E.DeviceArray(ctx, (k,)), E.DeviceArray(ctx, (k, k))
    ctx.rr_eig(dA.ptr, k, dE.ptr, dQ.ptr)
    w = dE.to_host()
    Q = dQ.to_host().T
    wo, Qo = oracle.syev(A)
    assert np.abs(w - wo).max() < 1e-12 * np.abs(wo).max()
    assert np.abs(A @ Q - Q * w).max() < 1e-11
    # same eigenvectors up to sign
    assert np.abs(np.abs(np.sum(Q * Qo, axis=0)) - 1).max() < 1e-9


# ---------------------------------------------------------------- Lanczos
@pytest.mark.parametrize("dims,p", [((8, 8, 8), 5), ((20, 18, 16), 4)])
def test_lanczos_matches_oracle(ctx, dims, p):
    s, H = synth(p=p, dims=dims, bands=(2, 2, 2))
    ham = ctx.ham_upload(s)
    up, lo = ctx.lanczos_bounds(ham, 20, seed=11)
    upo, loo = oracle.lanczos_bounds(H, 20, seed=11)
    assert abs(up - upo) < 1e-9 * abs(upo) and abs(lo - loo) < 1e-9 * abs(loo)


# ---------------------------------------------------------------- density
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_density(ctx, dtype):
    rng = np.random.default_rng(5)
    rx, ry, rz = 6, 5, 7
    px, py, pz = 70, 33, 20
    Bx, By, Bz = (rng.standard_normal((p, r)) for p, r in ((px, rx), (py, ry), (pz, rz)))
    k = 21
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = rng.standard_normal((rx * ry * rz, k)).astype(npdt)
    occ = np.linspace(1.0, 0.0, k)  # last state has zero weight
    dRho = E.DeviceArray(ctx, (pz, py, px))
    dX = dev(ctx, X, npdt)
    dBx, dBy, dBz = dev(ctx, Bx.T.copy()), dev(ctx, By.T.copy()), dev(ctx, Bz.T.copy())
    ctx.density(dX.ptr, k, k, occ, (rx, ry, rz), dBx.ptr, dBy.ptr, dBz.ptr, (px, py, pz),
                dRho.ptr, dtype)
    ref = oracle.density(X.astype(np.float64), occ, (rx, ry, rz), Bx, By, Bz)
    got = dRho.to_host()
    # an FP32 block is contracted in FP32 up to the y-factor density matrices
    tol = 1e-12 if dtype == E.F64 else 1e-5
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_density_chunked_with_sparse_weights(ctx, monkeypatch, dtype):
    """Several qz chunks, zero weights scattered among the states, ry > rx, rz."""
    monkeypatch.setenv("TTDFT_DENSITY_PLANES", "4")
    rng = np.random.default_rng(11)
    rx, ry, rz = 9, 17, 5
    px, py, pz = 40, 31, 23
    Bx, By, Bz = (rng.standard_normal((p, r)) for p, r in ((px, rx), (py, ry), (pz, rz)))
    k = 37
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = rng.standard_normal((rx * ry * rz, k + 3)).astype(npdt)  # ld > k
    occ = rng.uniform(0.0, 1.0, k)
    occ[[0, 5, 6, 20, 36]] = 0.0
    dRho = E.DeviceArray(ctx, (pz, py, px))
    dX = dev(ctx, X, npdt)
    dBx, dBy, dBz = dev(ctx, Bx.T.copy()), dev(ctx, By.T.copy()), dev(ctx, Bz.T.copy())
    ctx.density(dX.ptr, k + 3, k, occ, (rx, ry, rz), dBx.ptr, dBy.ptr, dBz.ptr, (px, py, pz),
                dRho.ptr, dtype)
    ref = oracle.density(X[:, :k].astype(np.float64), occ, (rx, ry, rz), Bx, By, Bz)
    got = dRho.to_host()
    tol = 1e-12 if dtype == E.F64 else 1e-5
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("dtype", [E.F64, E.F32])
@pytest.mark.parametrize("ry", [33, 50, 70, 97, 128])
def test_density_wide_y(ctx, ry, dtype):
    """y ranks spanning 3..8 MFMA row tiles (the y contraction's tile groups)."""
    rng = np.random.default_rng(ry)
    rx, rz = 3, 2
    px, py, pz = 7, 19, 5
    Bx, By, Bz = (rng.standard_normal((p, r)) for p, r in ((px, rx), (py, ry), (pz, rz)))
    k = 22
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = rng.standard_normal((rx * ry * rz, k)).astype(npdt)
    occ = rng.uniform(0.0, 1.0, k)
    dRho = E.DeviceArray(ctx, (pz, py, px))
    dX = dev(ctx, X, npdt)
    dBx, dBy, dBz = dev(ctx, Bx.T.copy()), dev(ctx, By.T.copy()), dev(ctx, Bz.T.copy())
    ctx.density(dX.ptr, k, k, occ, (rx, ry, rz), dBx.ptr, dBy.ptr, dBz.ptr, (px, py, pz),
                dRho.ptr, dtype)
    ref = oracle.density(X.astype(np.float64), occ, (rx, ry, rz), Bx, By, Bz)
    tol = 1e-12 if dtype == E.F64 else 1e-5
    assert np.linalg.norm(dRho.to_host() - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("dtype", [E.F64, E.F32])
@pytest.mark.parametrize("rx,ry,k", [(13, 16, 9), (30, 20, 17), (50, 40, 8), (61, 64, 33),
                                     (4, 50, 3), (70, 70, 11), (80, 72, 5)])
def test_density_fused_shapes(ctx, monkeypatch, rx, ry, k, dtype):
    """The fused x + y-factor step over its tile classes (a steps of 4, 8, 13, 16, 20; 1-5
    blocks of 16 b; ranks past 64 are fused in FP32 only), a ragged last qx tile and ragged
    state chunks, against the oracle and against the two-step form (x GEMM to memory, then
    the packed y-factor products)."""
    rng = np.random.default_rng(rx * 100 + ry)
    rz = 2
    px, py, pz = 37, 11, 3
    Bx, By, Bz = (rng.standard_normal((p, r)) for p, r in ((px, rx), (py, ry), (pz, rz)))
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = rng.standard_normal((rx * ry * rz, k)).astype(npdt)
    occ = rng.uniform(0.1, 1.0, k)
    dX = dev(ctx, X, npdt)
    dBx, dBy, dBz = dev(ctx, Bx.T.copy()), dev(ctx, By.T.copy()), dev(ctx, Bz.T.copy())
    got = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("TTDFT_DENSITY_FUSED", fused)
        dRho = E.DeviceArray(ctx, (pz, py, px))
        ctx.density(dX.ptr, k, k, occ, (rx, ry, rz), dBx.ptr, dBy.ptr, dBz.ptr, (px, py, pz),
                    dRho.ptr, dtype)
        got[fused] = dRho.to_host()
    ref = oracle.density(X.astype(np.float64), occ, (rx, ry, rz), Bx, By, Bz)
    tol = 1e-12 if dtype == E.F64 else 1e-5
    for f in ("1", "0"):
        assert np.linalg.norm(got[f] - ref) / np.linalg.norm(ref) < tol, f
    assert np.linalg.norm(got["1"] - got["0"]) / np.linalg.norm(ref) < tol


def test_initial_density_and_integrals(ctx):
    r = np.linspace(0.0, 9.0, 901)
    rho_r = np.exp(-r**2) * (1 + 0.3 * r)
    xyz = np.array([[0.3, -0.2, 0.1], [1.7, 0.4, -0.9], [-1.1, 1.2, 0.6]])
    x, w = np.polynomial.legendre.leggauss(40)
    gx, wx = 7.0 * x, 7.0 * w
    gy = np.linspace(-8, 8, 35)
    wy = np.full(35, 16 / 34)
    gz = gx[::2].copy()
    wz = wx[::2].copy()
    dims = (len(gx), len(gy), len(gz))
    dR = E.DeviceArray(ctx, (dims[2], dims[1], dims[0]))
    dgx, dgy, dgz = dev(ctx, gx), dev(ctx, gy), dev(ctx, gz)
    ctx.initial_density(xyz, r, rho_r, dgx.ptr, dgy.ptr, dgz.ptr, dims, dR.ptr)
    ref = oracle.initial_density(xyz, r, rho_r, gx, gy, gz)
    got = dR.to_host()
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
    dwx, dwy, dwz = dev(ctx, wx), dev(ctx, wy), dev(ctx, wz)
    q = ctx.grid_integral(dR.ptr, dims, dwx.ptr, dwy.ptr, dwz.ptr)
    assert abs(q - oracle.grid_integral(ref, wx, wy, wz)) < 1e-11 * abs(q)
    dR2 = dev(ctx, ref * 1.01)
    res = ctx.grid_residual(dR.ptr, dR2.ptr, dims, dwx.ptr, dwy.ptr, dwz.ptr)
    assert abs(res - oracle.grid_integral((0.01 * ref) ** 2, wx, wy, wz)) < 1e-10 * res
    ctx.grid_scale(dR.ptr, int(np.prod(dims)), 2.0)
    assert np.abs(dR.to_host() - 2 * ref).max() < 1e-13 * np.abs(ref).max()


# ---------------------------------------------------------------- driver
def test_chfsi_step_matches_oracle(ctx):
    s, H = synth(p=5, seed=21)
    ham = ctx.ham_upload(s)
    k, nocc = 16, 10
    X0 = oracle.orth(oracle.rng_normal(5, s.n * k).reshape(s.n, k))
    Xo = X0.copy()
    evo = np.zeros(k)
    sto = oracle.new_state()
    dX = dev(ctx, X0)
    ev = np.zeros(k)
    st = E.new_state()
    for scf in range(6):
        rt = 10 if scf == 0 else 3
        oracle.chfsi_step(H, Xo, evo, sto, m=12, scf_iter=scf, occ_index=nocc, restart_times=rt)
        ctx.chfsi_step(ham, dX.ptr, k, k, 12, ev, st, scf, nocc, rt)
        assert st.passes == sto.passes and st.lanczos_runs == sto.lanczos_runs
        # occupied + stop-test states agree to rounding; the unconverged top of
        # the block amplifies the different summation orders through the filter
        assert np.abs(ev[: nocc + 1] - evo[: nocc + 1]).max() < 1e-9, scf
        assert np.abs(ev - evo).max() < 1e-6, scf
        assert abs(st.upper - sto.upper) < 1e-8 * abs(sto.upper)
    assert np.abs(ev - evo).max() < 1e-8
    lam = np.linalg.eigvalsh(H.dense())
    assert np.abs(ev[:nocc] - lam[:nocc]).max() < 1e-8
    # band energy and density from the converged subspace
    ef, occ = E.fermi_occupancy(ev, 2 * nocc - 1)
    efo = oracle.fermi_energy(evo, 2 * nocc - 1)
    assert abs(ef - efo) < 1e-8
    X = dX.to_host()
    # subspaces agree (eigenvector signs are free)
    assert np.abs(np.abs(np.sum(X[:, :nocc] * Xo[:, :nocc], axis=0)) - 1).max() < 1e-6


@pytest.mark.parametrize("tucker", [False, True])
def test_dense_storage_matches_oracle(ctx, tucker):
    """A small, nearly full H_loc (every 1-D factor couples all basis
    functions) is stored dense and applied by one matrix product plus the
    fused epilogue: H X, the filter and both precisions against the oracle."""
    s = E.SynthHam(9, 8, 7, 8, 7, 6, 5, seed=17)
    assert s.nnz > 0.2 * s.n ** 2
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=tucker)
    assert ham.format()[0] == 2
    for dtype, tol in ((E.F64, 1e-13), (E.F32, 1e-5)):
        npdt = np.float64 if dtype == E.F64 else np.float32
        X = np.random.default_rng(4).standard_normal((s.n, 40)).astype(npdt)
        dX, dY = dev(ctx, X, npdt), E.DeviceArray(ctx, (s.n, 40), npdt)
        ctx.ham_apply(ham, dX.ptr, 40, dY.ptr, 40, 40, dtype)
        assert rel(dY.to_host().astype(np.float64), oracle.ham_apply(H, X.astype(np.float64))) < tol
    X = np.random.default_rng(5).standard_normal((s.n, 24))
    dX = dev(ctx, X)
    ctx.cheb_filter(ham, dX.ptr, 24, 24, 8, -0.5, 30.0, -3.0, 1, E.F64)
    assert rel(dX.to_host(), oracle.cheb_filter(H, X, 8, -0.5, 30.0, -3.0, 1)) < 1e-12
    X32 = X.astype(np.float32)
    dX = dev(ctx, X32, np.float32)
    ctx.cheb_filter(ham, dX.ptr, 24, 24, 8, -0.5, 30.0, -3.0, 1, E.F32)
    ref = oracle.cheb_filter(H, X32.astype(np.float64), 8, -0.5, 30.0, -3.0, 1)
    assert rel(dX.to_host().astype(np.float64), ref) < 5e-5


@pytest.mark.parametrize("dims", [(4, 3, 5), (9, 8, 7), (11, 10, 13)])
@pytest.mark.parametrize("k", [1, 8, 16, 24, 33, 64, 80, 96, 100, 150, 192])
def test_dense_product_tilings(ctx, dims, k):
    """The dense product over its state tilings (16 to 96 states per
    workgroup, one or two workgroup columns, partial tiles), leading dimension
    of the block above k, and n
    not a multiple of 32 (n = 60: six of the eight waves get no l range;
    n = 1430: the last wave's range is shorter)."""
    rx, ry, rz = dims
    s = E.SynthHam(rx, ry, rz, rx - 1, ry - 1, rz - 1, 3, seed=k)
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=False)
    assert ham.format()[0] == 2
    print(f"dense product n={s.n} k={k}", flush=True)
    # an odd and an even leading dimension (96-state multiples with an even
    # one take the six-consecutive-states layout)
    for ld, dtype, npdt, tol in ((k + 3, E.F64, np.float64, 1e-13), (k + 3, E.F32, np.float32, 1e-5),
                                 (k + 4, E.F64, np.float64, 1e-13), (k + 4, E.F32, np.float32, 1e-5)):
        X = np.random.default_rng(k).standard_normal((s.n, ld)).astype(npdt)
        dX, dY = dev(ctx, X, npdt), dev(ctx, np.full((s.n, ld), 7.0), npdt)
        ctx.ham_apply(ham, dX.ptr, ld, dY.ptr, ld, k, dtype)
        Y = dY.to_host().reshape(s.n, ld).astype(np.float64)
        ref = oracle.ham_apply(H, X[:, :k].astype(np.float64))
        assert rel(Y[:, :k], ref) < tol
        assert np.all(Y[:, k:] == 7.0)  # columns past k untouched


def test_dense_storage_reuse_across_hamiltonians(ctx):
    """Hamiltonians of different sizes built and freed in turn on one context:
    the dense storage and the projector fold come from the context's block
    cache (a freed block of the same size class is reused), and every product
    still matches the oracle."""
    for dims in [(9, 8, 7), (4, 3, 5), (11, 10, 13), (9, 8, 7), (8, 8, 8), (4, 3, 5)]:
        s = E.SynthHam(*dims, dims[0] - 1, dims[1] - 1, dims[2] - 1, 5, seed=sum(dims))
        H = oracle.Ham.from_synth(s)
        ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=False)
        assert ham.format()[0] == 2
        X = np.random.default_rng(sum(dims)).standard_normal((s.n, 40))
        dX, dY = dev(ctx, X), E.DeviceArray(ctx, (s.n, 40))
        print(f"reuse n={s.n}", flush=True)
        ctx.ham_apply(ham, dX.ptr, 40, dY.ptr, 40, 40, E.F64)
        assert rel(dY.to_host(), oracle.ham_apply(H, X)) < 1e-13
        ham.free()


@pytest.mark.parametrize("dims,bands,k", [((50, 6, 5), (6, 2, 2), 300), ((22, 9, 7), (3, 2, 3), 81),
                                          ((50, 8, 9), (6, 3, 2), 96)])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_band_spmm_light_chunks_last_same_bits(ctx, monkeypatch, dims, bands, k, dtype):
    """TTDFT_BRICK_ORDER=1 dispatches the ragged last x-chunk's workgroups
    after all full ones on every XCD; every output is still computed once, by
    the same code: bitwise the default order's result."""
    s = E.SynthHam(*dims, *bands, 4, seed=sum(dims) + k)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=True)
    assert ham.format()[0] == 1
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(k).standard_normal((s.n, k)).astype(npdt)
    dX = dev(ctx, X, npdt)
    out = {}
    for order in ("0", "1"):
        monkeypatch.setenv("TTDFT_BRICK_ORDER", order)
        dY = dev(ctx, np.full((s.n, k), np.nan), npdt)  # every output must be written
        ctx.ham_apply(ham, dX.ptr, k, dY.ptr, k, k, dtype)
        out[order] = dY.to_host()
    assert np.array_equal(out["0"], out["1"])
    if dtype == E.F64:
        H = oracle.Ham.from_synth(s)
        assert rel(out["1"], oracle.ham_apply(H, X)) < 1e-13


@pytest.mark.parametrize("n,k,ld", [(3000, 40, 40), (5000, 300, 301)])
@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_random_subspace_matches_oracle(ctx, n, k, ld, dtype):
    """The initial block (main.cc:807-829 with a seeded generator): the
    library's N(0,1) sequence is the oracle's bit for bit, and the block is
    two Cholesky-QR passes of it -- in FP64 the oracle's twice-orthonormalized
    block to 1e-11, in FP32 (split-BF16 Gram, whose XᵀX diagonal carries a
    2^-18 one-signed bias) orthonormal to 3e-5."""
    seed = 1234 + k
    assert np.array_equal(E.rng_normal(seed, 1000), oracle.rng_normal(seed, 1000))
    npdt = np.float64 if dtype == E.F64 else np.float32
    dX = E.DeviceArray(ctx, (n, ld), npdt)
    ctx.random_subspace(dX.ptr, ld, n, k, seed=seed, dtype=dtype)
    X = dX.to_host().reshape(n, ld)[:, :k].astype(np.float64)
    ref = oracle.orth(oracle.orth(oracle.rng_normal(seed, n * k).reshape(n, k)))
    if dtype == E.F64:
        assert rel(X, ref) < 1e-11
    else:
        assert rel(X, ref) < 1e-4
    assert np.abs(X.T @ X - np.eye(k)).max() < (1e-12 if dtype == E.F64 else 3e-5)
