"""Context lifetime, stream switches and block alignment, through the C ABI.

* A context destroyed while Hamiltonians / mixers created on it are alive
  stays alive until the last of them is freed (their storage goes back to
  its block cache).
* Switching the context's stream waits for the outgoing stream, so storage
  released on it and reused on the new stream is never overwritten early.
* Trimming the context's block cache between Hamiltonians.
* A filter split into column blocks whose second block starts at an odd
  column takes the dense product's unpaired path and still matches the
  oracle (ChebyShevFilter.cc:57-60 chunking).
"""
import numpy as np
import pytest

import oracle
from ttdft_public_amd import engine as E

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_ctx_destroy_before_children_are_freed():
    c = E.Context(0)
    s = E.SynthHam(8, 8, 8, 2, 2, 3, 5, seed=3)
    H = oracle.Ham.from_synth(s)
    ham = c.ham_upload(s)
    X = np.random.default_rng(1).standard_normal((s.n, 24))
    dX, dY = E.DeviceArray.from_host(c, X), E.DeviceArray(c, (s.n, 24))
    c.ham_apply(ham, dX.ptr, 24, dY.ptr, 24, 24, E.F64)
    assert rel(dY.to_host(), oracle.ham_apply(H, X)) < 1e-13
    dX.free()
    dY.free()
    jw = E.DeviceArray.from_host(c, np.ones(1000))
    mix = E.Mixer(c, 1000, 0, jw.ptr, jw.ptr, 0.5, 4)
    jw.free()
    h = c.h
    c.close()  # two children alive: deferred
    ham.free()  # one left
    mix.free()  # the last child completes the destruction
    assert c.h is None and h


def _hip():
    """The HIP runtime the engine links (already loaded by it): a second
    stream is made there, at the default priority."""
    import ctypes

    E.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_stream_switch_then_rebuild():
    import ctypes

    hip = _hip()
    c = E.Context(0)
    s1 = E.SynthHam(10, 9, 8, 3, 3, 3, 4, seed=11)
    ham = c.ham_upload(s1, E.HAM_F64 | E.HAM_F32)
    X1 = np.random.default_rng(2).standard_normal((s1.n, 40))
    dX, dY = E.DeviceArray.from_host(c, X1), E.DeviceArray(c, (s1.n, 40))
    c.ham_apply(ham, dX.ptr, 40, dY.ptr, 40, 40, E.F64)
    ham.free()  # its storage goes back to the cache, work may still be queued
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    c.use_stream(st.value)
    s2 = E.SynthHam(10, 9, 8, 3, 3, 3, 4, seed=12)  # same sizes: reuses the blocks
    H2 = oracle.Ham.from_synth(s2)
    ham2 = c.ham_upload(s2, E.HAM_F64 | E.HAM_F32)
    X2 = np.random.default_rng(3).standard_normal((s2.n, 40))
    dX.copy_from(X2)
    c.ham_apply(ham2, dX.ptr, 40, dY.ptr, 40, 40, E.F64)
    assert rel(dY.to_host(), oracle.ham_apply(H2, X2)) < 1e-13
    c.use_stream(None)
    c.ham_apply(ham2, dX.ptr, 40, dY.ptr, 40, 40, E.F64)
    assert rel(dY.to_host(), oracle.ham_apply(H2, X2)) < 1e-13
    ham2.free()
    dX.free()
    dY.free()
    c.close()
    assert hip.hipStreamDestroy(st) == 0


def test_trim_between_hamiltonians():
    """ttdft_ctx_trim gives the cached blocks back; the next Hamiltonian allocates anew and
    its product still matches the oracle."""
    c = E.Context(0)
    s1 = E.SynthHam(10, 9, 8, 3, 3, 3, 4, seed=21)
    ham = c.ham_upload(s1, E.HAM_F64)
    X = np.random.default_rng(4).standard_normal((s1.n, 24))
    dX, dY = E.DeviceArray.from_host(c, X), E.DeviceArray(c, (s1.n, 24))
    c.ham_apply(ham, dX.ptr, 24, dY.ptr, 24, 24, E.F64)  # queued work on the storage
    ham.free()
    c.trim()
    s2 = E.SynthHam(10, 9, 8, 3, 3, 3, 4, seed=22)
    ham2 = c.ham_upload(s2, E.HAM_F64)
    c.ham_apply(ham2, dX.ptr, 24, dY.ptr, 24, 24, E.F64)
    assert rel(dY.to_host(), oracle.ham_apply(oracle.Ham.from_synth(s2), X)) < 1e-13
    ham2.free()
    dX.free()
    dY.free()
    c.close()


@pytest.mark.parametrize("dtype", [E.F64, E.F32])
def test_dense_filter_second_block_at_odd_column(ctx, dtype):
    """k = 193 in two column blocks: 97 then 96 states, the second starting
    at column 97, so its rows are not pair-aligned although the leading
    dimension is even."""
    s = E.SynthHam(9, 8, 7, 8, 7, 6, 5, seed=17)
    H = oracle.Ham.from_synth(s)
    ham = ctx.ham_upload(s, E.HAM_F64 | E.HAM_F32, tucker=False)
    assert ham.format()[0] == 2
    k, ld = 193, 194
    npdt = np.float64 if dtype == E.F64 else np.float32
    X = np.random.default_rng(9).standard_normal((s.n, ld)).astype(npdt)
    dX = E.DeviceArray.from_host(ctx, X, npdt)
    ctx.cheb_filter(ham, dX.ptr, ld, k, 6, -0.5, 30.0, -3.0, 2, dtype)
    got = dX.to_host().reshape(s.n, ld).astype(np.float64)
    ref = oracle.cheb_filter(H, X[:, :k].astype(np.float64), 6, -0.5, 30.0, -3.0, 2)
    assert rel(got[:, :k], ref) < (1e-12 if dtype == E.F64 else 5e-5)
    assert np.array_equal(got[:, k:], X[:, k:].astype(np.float64))
    ham.free()
