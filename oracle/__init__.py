"""CPU FP64 oracle for the ChFSI hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The engine never routes through it.

The C restatement lives in ``chfsi_oracle.c``; each function there cites the
reference file:line it follows.  This module is a thin ctypes/numpy wrapper.

Parity pinning (what checks each function independently of its own formula):
  * orc_fermi_energy / orc_occupancy / orc_band_energy: bit-exact against the
    24 SCF iterations printed in the reference's example/output.
  * orc_syev: numpy.linalg.eigh, 1-D Laplacian closed form.
  * orc_cheb_filter: closed-form Chebyshev polynomial on eigenvectors.
  * orc_ham_apply / orc_csr_spmm / orc_gram / orc_rotate: dense brute force.
  * orc_orth: X^T X = I and span preservation.
  * orc_lanczos_bounds: containment of the dense spectrum.
  * orc_density: separable closed form, charge sum, brute force.
  * orc_spline_*: exact reproduction of clamped cubics; orc_initial_density:
    closed-form Gaussian superposition.
  * orc_chfsi_step: converged Ritz values against dense eigenvalues.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc_chfsi.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class OrcHam(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("rowptr", _i64p),
        ("col", _i32p),
        ("val", _dp),
        ("p", ctypes.c_int32),
        ("c_rowptr", _i64p),
        ("c_col", _i32p),
        ("c_val", _dp),
        ("inv_c_lm", _dp),
    ]


class OrcChfsiState(ctypes.Structure):
    _fields_ = [
        ("upper", ctypes.c_double),
        ("lower_unwanted", ctypes.c_double),
        ("lower_wanted", ctypes.c_double),
        ("err_upper", ctypes.c_double),
        ("occ_energy", ctypes.c_double),
        ("passes", ctypes.c_int32),
        ("lanczos_runs", ctypes.c_int32),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_NATIVE_PATH = os.path.join(_HERE, "liborc_chfsi_native.so")


def use_native_build(timeout=300) -> bool:
    """Rebuild the oracle for this host (-O3 -march=native, OpenMP) and route
    every later call through that build.  For timing the CPU baseline only:
    parity checks use the default build.  False if the build failed (the
    default build stays in use)."""
    global _lib
    try:
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True, timeout=timeout,
                       capture_output=True)
    except (subprocess.SubprocessError, OSError):
        return False
    _lib = _bind(ctypes.CDLL(_NATIVE_PATH))
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    """Declare the argument types of the oracle's C functions on library L."""
    L.orc_csr_spmm.argtypes = [ctypes.c_int64, _i64p, _i32p, _dp, _dp, ctypes.c_int64,
                               ctypes.c_int, _dp, ctypes.c_int64]
    L.orc_ham_apply.argtypes = [ctypes.POINTER(OrcHam), _dp, ctypes.c_int64, ctypes.c_int,
                                _dp, ctypes.c_int64]
    L.orc_cheb_filter.argtypes = [ctypes.POINTER(OrcHam), _dp, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.orc_gram.argtypes = [ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int64, _dp,
                           ctypes.c_int64, _dp]
    L.orc_cholesky_lower.argtypes = [ctypes.c_int, _dp]
    L.orc_orth.argtypes = [ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int64]
    L.orc_syev.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    L.orc_rotate.argtypes = [ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int64, _dp]
    L.orc_lanczos_bounds.argtypes = [ctypes.POINTER(OrcHam), ctypes.c_int, ctypes.c_uint64,
                                     _dp, _dp]
    L.orc_fermi_energy.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int]
    L.orc_fermi_energy.restype = ctypes.c_double
    L.orc_occupancy.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _dp]
    L.orc_band_energy.argtypes = [_dp, _dp, ctypes.c_int]
    L.orc_band_energy.restype = ctypes.c_double
    L.orc_density.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, _dp]
    L.orc_spline_build.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    L.orc_spline_eval.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double]
    L.orc_spline_eval.restype = ctypes.c_double
    L.orc_initial_density.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                      _dp, _dp]
    L.orc_grid_integral.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                    _dp, _dp]
    L.orc_grid_integral.restype = ctypes.c_double
    L.orc_chfsi_step.argtypes = [ctypes.POINTER(OrcHam), _dp, ctypes.c_int64, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_int, _dp, ctypes.POINTER(OrcChfsiState)]
    L.orc_assemble_hloc.argtypes = [ctypes.c_int] * 6 + [_dp] * 10 + [ctypes.c_double,
                                                                       _i64p, _i32p, _dp]
    L.orc_assemble_hloc.restype = ctypes.c_int64
    L.orc_assemble_hloc_rows.argtypes = ([ctypes.c_int] * 6 + [_dp] * 10
                                         + [ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                            _i64p, _i32p, _dp])
    L.orc_assemble_hloc_rows.restype = ctypes.c_int64
    L.orc_assemble_projectors.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double,
                                          _i64p, _i32p, _dp]
    L.orc_assemble_projectors.restype = ctypes.c_int64
    L.orc_anderson_constants.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp,
                                         _dp]
    L.orc_anderson_constants.restype = ctypes.c_int
    L.orc_anderson_mix.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp,
                                   ctypes.c_double, _dp]
    L.orc_anderson_mix.restype = None
    L.orc_sthosvd.argtypes = [ctypes.c_int] * 6 + [_dp] * 6
    L.orc_sthosvd.restype = ctypes.c_int
    _dpa = ctypes.POINTER(_dp)
    _ip = ctypes.POINTER(ctypes.c_int)
    L.orc_hartree_expansion.argtypes = [ctypes.c_int] * 3 + [_dp, _dpa, _ip, _dpa, _dpa, _ip,
                                                             _dpa, ctypes.c_int, _dp, _dp,
                                                             _dp, ctypes.c_int]
    L.orc_hartree_expansion.restype = None
    L.orc_fem_mesh.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [_dp]
    L.orc_fem_mesh.restype = ctypes.c_int64
    L.orc_fem_shape.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    L.orc_fem_shape.restype = None
    L.orc_fem_interpolate.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.c_int] + [_dp] * 6
    L.orc_fem_interpolate.restype = None
    L.orc_tucker_basis_matrices.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 4 + [
        ctypes.c_int] + [_dp] * 4
    L.orc_tucker_basis_matrices.restype = None
    L.orc_rng_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, _dp]
    L.orc_rng_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, _dp]
    return L


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i64(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)


def _i32(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


class Ham:
    """Holds numpy arrays of H_loc (CSR) and C (CSR, unscaled) alive for C."""

    def __init__(self, n, rowptr, col, val, c_rowptr=None, c_col=None, c_val=None,
                 inv_c_lm=None):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        if c_rowptr is None:
            c_rowptr = np.zeros(1, np.int64)
            c_col = np.zeros(1, np.int32)
            c_val = np.zeros(1, np.float64)
            inv_c_lm = np.zeros(1, np.float64)
            self.p = 0
        else:
            self.p = len(c_rowptr) - 1
        self.c_rowptr = np.ascontiguousarray(c_rowptr, dtype=np.int64)
        self.c_col = np.ascontiguousarray(c_col, dtype=np.int32)
        self.c_val = np.ascontiguousarray(c_val, dtype=np.float64)
        self.inv_c_lm = np.ascontiguousarray(inv_c_lm, dtype=np.float64)
        self.s = OrcHam(self.n, _i64(self.rowptr), _i32(self.col), _d(self.val), self.p,
                        _i64(self.c_rowptr), _i32(self.c_col), _d(self.c_val),
                        _d(self.inv_c_lm))

    @classmethod
    def from_synth(cls, s):
        """From a ``ttdft_public_amd.synth.SynthHam`` (numpy arrays only)."""
        return cls(s.n, s.rowptr, s.col, s.val, s.c_rowptr, s.c_col, s.c_val, s.inv_c_lm)

    def dense(self):
        H = np.zeros((self.n, self.n))
        for i in range(self.n):
            a, b = self.rowptr[i], self.rowptr[i + 1]
            H[i, self.col[a:b]] += self.val[a:b]
        if self.p:
            C = np.zeros((self.p, self.n))
            for r in range(self.p):
                a, b = self.c_rowptr[r], self.c_rowptr[r + 1]
                C[r, self.c_col[a:b]] += self.c_val[a:b]
            H += C.T @ (self.inv_c_lm[:, None] * C)
        return H


def ham_apply(H: Ham, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, k = X.shape
    Y = np.empty_like(X)
    lib().orc_ham_apply(ctypes.byref(H.s), _d(X), k, k, _d(Y), k)
    return Y


def csr_spmm(H: Ham, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, k = X.shape
    Y = np.empty_like(X)
    lib().orc_csr_spmm(H.n, _i64(H.rowptr), _i32(H.col), _d(H.val), _d(X), k, k, _d(Y), k)
    return Y


def cheb_filter(H: Ham, X, m, a, b, a0, num_blocks=1, convention=0):
    X = np.array(X, dtype=np.float64, order="C", copy=True)
    n, k = X.shape
    rc = lib().orc_cheb_filter(ctypes.byref(H.s), _d(X), k, k, m, a, b, a0, num_blocks,
                               convention)
    if rc != 0:
        raise ValueError("orc_cheb_filter failed")
    return X


def gram(X, Y):
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    n, k = X.shape
    S = np.empty(k * k)
    lib().orc_gram(n, k, _d(X), k, _d(Y), k, _d(S))
    return S.reshape(k, k, order="F")


def cholesky_lower(S):
    k = S.shape[0]
    A = np.asfortranarray(S, dtype=np.float64).ravel(order="F").copy()
    rc = lib().orc_cholesky_lower(k, _d(A))
    if rc != 0:
        raise np.linalg.LinAlgError(f"not SPD at minor {-rc - 1}")
    return A.reshape(k, k, order="F")


def orth(X):
    X = np.array(X, dtype=np.float64, order="C", copy=True)
    n, k = X.shape
    rc = lib().orc_orth(n, k, _d(X), k)
    if rc != 0:
        raise np.linalg.LinAlgError("orc_orth: Gram matrix not SPD")
    return X


def syev(A):
    k = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    w = np.empty(k)
    Q = np.empty(k * k)
    rc = lib().orc_syev(k, _d(a), _d(w), _d(Q))
    if rc < 0:
        raise np.linalg.LinAlgError("orc_syev did not converge")
    return w, Q.reshape(k, k, order="F")


def rotate(X, Q):
    X = np.array(X, dtype=np.float64, order="C", copy=True)
    n, k = X.shape
    q = np.asfortranarray(Q, dtype=np.float64).ravel(order="F").copy()
    lib().orc_rotate(n, k, _d(X), k, _d(q))
    return X


def lanczos_bounds(H: Ham, maxit=20, seed=20261202):
    up = ctypes.c_double()
    lo = ctypes.c_double()
    rc = lib().orc_lanczos_bounds(ctypes.byref(H.s), maxit, seed, ctypes.byref(up),
                                  ctypes.byref(lo))
    if rc != 0:
        raise RuntimeError("orc_lanczos_bounds failed")
    return up.value, lo.value


def fermi_energy(evals, nelec, kb=3.166811429e-06, T=500.0, maxit=20):
    e = np.ascontiguousarray(evals, dtype=np.float64)
    return lib().orc_fermi_energy(_d(e), len(e), int(nelec), kb, T, maxit)


def occupancy(evals, ef, kb=3.166811429e-06, T=500.0):
    e = np.ascontiguousarray(evals, dtype=np.float64)
    occ = np.empty_like(e)
    lib().orc_occupancy(_d(e), len(e), ef, kb, T, _d(occ))
    return occ


def band_energy(evals, occ):
    e = np.ascontiguousarray(evals, dtype=np.float64)
    o = np.ascontiguousarray(occ, dtype=np.float64)
    return lib().orc_band_energy(_d(e), _d(o), len(e))


def density(X, occ, rdims, Bx, By, Bz):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, k = X.shape
    rx, ry, rz = rdims
    Bx = np.asfortranarray(Bx, dtype=np.float64)
    By = np.asfortranarray(By, dtype=np.float64)
    Bz = np.asfortranarray(Bz, dtype=np.float64)
    px, py, pz = Bx.shape[0], By.shape[0], Bz.shape[0]
    o = np.ascontiguousarray(occ, dtype=np.float64)
    rho = np.empty(px * py * pz)
    bx, by, bz = (b.ravel(order="F").copy() for b in (Bx, By, Bz))
    lib().orc_density(_d(X), k, k, _d(o), rx, ry, rz, _d(bx), _d(by), _d(bz), px, py, pz,
                      _d(rho))
    return rho.reshape(pz, py, px)


def spline_build(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.empty(4 * (len(x) - 1))
    lib().orc_spline_build(len(x), _d(x), _d(y), _d(c))
    return c


def spline_eval(x, coef, r):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_spline_eval(len(x), _d(x), _d(coef), float(r))


def initial_density(xyz, r, rho_r, gx, gy, gz):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
    r = np.ascontiguousarray(r, dtype=np.float64)
    rho_r = np.ascontiguousarray(rho_r, dtype=np.float64)
    gx, gy, gz = (np.ascontiguousarray(g, dtype=np.float64) for g in (gx, gy, gz))
    rho = np.empty(len(gx) * len(gy) * len(gz))
    lib().orc_initial_density(len(xyz) // 3, _d(xyz), len(r), _d(r), _d(rho_r), len(gx),
                              len(gy), len(gz), _d(gx), _d(gy), _d(gz), _d(rho))
    return rho.reshape(len(gz), len(gy), len(gx))


def grid_integral(f, wx, wy, wz):
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    wx, wy, wz = (np.ascontiguousarray(w, dtype=np.float64) for w in (wx, wy, wz))
    return lib().orc_grid_integral(_d(f), len(wx), len(wy), len(wz), _d(wx), _d(wy), _d(wz))


def chfsi_step(H: Ham, X, evals, state: OrcChfsiState, m, scf_iter, occ_index,
               restart_times, num_blocks=1, start_iter=0, maxit_lanczos=20,
               seed=20261202, convention=0):
    """One SCF iteration's ChFSI eigensolve (main.cc:977-1237).  X and evals are
    updated in place (X must be C-contiguous float64)."""
    assert X.flags.c_contiguous and X.dtype == np.float64
    n, k = X.shape
    rc = lib().orc_chfsi_step(ctypes.byref(H.s), _d(X), k, k, m, num_blocks, restart_times,
                              occ_index, scf_iter, start_iter, maxit_lanczos, seed, convention,
                              _d(evals), ctypes.byref(state))
    if rc != 0:
        raise RuntimeError(f"orc_chfsi_step failed rc={rc}")
    return state


def new_state():
    """Initial ChFSI state of main.cc:733-734 and :685-688."""
    return OrcChfsiState(1e16, -1e16, 0.0, 1e16, 1.0e9, 0, 0)


def rng_uniform(seed, count):
    out = np.empty(count)
    lib().orc_rng_uniform(seed, count, _d(out))
    return out


def rng_normal(seed, count):
    out = np.empty(count)
    lib().orc_rng_normal(seed, count, _d(out))
    return out


def assemble_hloc(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, truncation):
    """H_loc rows (dft/ProjectHamiltonianSparse.cc:330-575) as (rowptr, col, val).

    X..DZ are r x r (row-major [I][J]); PX is (rvx, rx, rx) = MatPotX[p][I][J];
    G is the Veff Tucker core with shape (rvx, rvy, rvz), stored mode 0 fastest."""
    rx, ry, rz = X.shape[0], Y.shape[0], Z.shape[0]
    rvx, rvy, rvz = PX.shape[0], PY.shape[0], PZ.shape[0]
    mats = [np.ascontiguousarray(m, dtype=np.float64) for m in (X, DX, Y, DY, Z, DZ, PX, PY, PZ)]
    g = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(rvx, rvy, rvz)
                             .transpose(2, 1, 0)).reshape(-1)  # p fastest
    args = [rx, ry, rz, rvx, rvy, rvz] + [_d(m) for m in mats] + [_d(g), float(truncation)]
    L = lib()
    nnz = L.orc_assemble_hloc(*args, None, None, None)
    n = rx * ry * rz
    rowptr = np.zeros(n + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    L.orc_assemble_hloc(*args, _i64(rowptr), _i32(col), _d(val))
    return rowptr, col[:nnz], val[:nnz]


def assemble_hloc_rows(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, truncation, row0, row1):
    """Rows row0 <= i < row1 of assemble_hloc (rowptr relative to row0)."""
    rx, ry, rz = X.shape[0], Y.shape[0], Z.shape[0]
    rvx, rvy, rvz = PX.shape[0], PY.shape[0], PZ.shape[0]
    mats = [np.ascontiguousarray(m, dtype=np.float64) for m in (X, DX, Y, DY, Z, DZ, PX, PY, PZ)]
    g = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(rvx, rvy, rvz)
                             .transpose(2, 1, 0)).reshape(-1)
    args = [rx, ry, rz, rvx, rvy, rvz] + [_d(m) for m in mats] + [_d(g), float(truncation),
                                                                  int(row0), int(row1)]
    L = lib()
    nnz = L.orc_assemble_hloc_rows(*args, None, None, None)
    rowptr = np.zeros(row1 - row0 + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    L.orc_assemble_hloc_rows(*args, _i64(rowptr), _i32(col), _d(val))
    return rowptr, col[:nnz], val[:nnz]


def assemble_projectors(cores, mx, my, mz, truncation):
    """Projector rows (dft/ProjectHamiltonianSparse.cc:196-278) as (rowptr, col, val).

    cores[i] has shape (r_i, r_i, r_i) indexed [a, b, c]; mx[i] is rx x r_i."""
    p = len(cores)
    ranks = np.array([c.shape[0] for c in cores], np.int32)
    rx, ry, rz = mx[0].shape[0], my[0].shape[0], mz[0].shape[0]
    cc = np.concatenate([np.asarray(c, np.float64).transpose(2, 1, 0).reshape(-1) for c in cores])
    fx = np.concatenate([np.asarray(m, np.float64).T.reshape(-1) for m in mx])  # column-major
    fy = np.concatenate([np.asarray(m, np.float64).T.reshape(-1) for m in my])
    fz = np.concatenate([np.asarray(m, np.float64).T.reshape(-1) for m in mz])
    L = lib()
    args = [p, _i32(ranks), rx, ry, rz, _d(cc), _d(fx), _d(fy), _d(fz), float(truncation)]
    nnz = L.orc_assemble_projectors(*args, None, None, None)
    rowptr = np.zeros(p + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    L.orc_assemble_projectors(*args, _i64(rowptr), _i32(col), _d(val))
    return rowptr, col[:nnz], val[:nnz]


def anderson_constants(rho_in, rho_out, jacob, weight):
    """Mixing constants of AndersonMixing.cc:150-262.  rho_in / rho_out are
    (h, n) arrays, oldest history entry first."""
    rin = np.ascontiguousarray(rho_in, dtype=np.float64)
    rout = np.ascontiguousarray(rho_out, dtype=np.float64)
    h, n = rin.shape
    j = np.ascontiguousarray(jacob, dtype=np.float64)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    c = np.zeros(max(h - 1, 1))
    info = lib().orc_anderson_constants(h, n, _d(rin), _d(rout), _d(j), _d(w), _d(c))
    if info != 0:
        raise np.linalg.LinAlgError(f"singular mixing system (pivot {info})")
    return c[: h - 1]


def anderson_mix(rho_in, rho_out, constants, alpha):
    """Mixed density of AndersonMixing.cc:264-323 for (h, n) histories."""
    rin = np.ascontiguousarray(rho_in, dtype=np.float64)
    rout = np.ascontiguousarray(rho_out, dtype=np.float64)
    h, n = rin.shape
    c = np.ascontiguousarray(constants, dtype=np.float64)
    if c.size == 0:
        c = np.zeros(1)
    rho = np.empty(n)
    lib().orc_anderson_mix(h, n, _d(rin), _d(rout), _d(c), float(alpha), _d(rho))
    return rho


def sthosvd(T, ranks):
    """Fixed-rank STHOSVD of a 3-D array T (indexed [i0, i1, i2]): returns
    (core [a, b, c], [U0, U1, U2] with U_n of shape (n_n, r_n), eigenvalues
    per mode)."""
    T = np.asarray(T, dtype=np.float64)
    n0, n1, n2 = T.shape
    r0, r1, r2 = ranks
    tf = np.ascontiguousarray(T.transpose(2, 1, 0)).reshape(-1)  # first index fastest
    core = np.zeros(r0 * r1 * r2)
    U = [np.zeros(n * r) for n, r in zip(T.shape, ranks)]
    ev = np.zeros(r0 + r1 + r2)
    rc = lib().orc_sthosvd(n0, n1, n2, r0, r1, r2, _d(tf), _d(core), _d(U[0]), _d(U[1]),
                           _d(U[2]), _d(ev))
    if rc != 0:
        raise RuntimeError("orc_sthosvd: eigen-solver did not converge")
    G = core.reshape(r2, r1, r0).transpose(2, 1, 0)
    Us = [u.reshape(r, n).T for u, n, r in zip(U, T.shape, ranks)]
    return G, Us, [ev[:r0], ev[r0:r0 + r1], ev[r0 + r1:]]


def hartree_expansion(G, U, xe, jw, xq, alpha, omega, V=None):
    """Kernel-expansion Hartree potential (KSDFTPotential.cc:269-340,440-668).
    G: core [a, b, c]; U[d]: (ne_d, r_d) factor values at the electrostatic
    quadrature points xe[d] with jacobian x weight jw[d]; xq[d]: target points.
    Returns V indexed [x, y, z] (added to V when given)."""
    G = np.asarray(G, dtype=np.float64)
    r0, r1, r2 = G.shape
    gf = np.ascontiguousarray(G.transpose(2, 1, 0)).reshape(-1)
    keep = []
    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _d(a)
    Up = (_dp * 3)(*[arr(np.asarray(u).T.reshape(-1)) for u in U])  # column-major
    xep = (_dp * 3)(*[arr(x) for x in xe])
    jwp = (_dp * 3)(*[arr(x) for x in jw])
    xqp = (_dp * 3)(*[arr(x) for x in xq])
    ne = (ctypes.c_int * 3)(*[len(x) for x in xe])
    p = (ctypes.c_int * 3)(*[len(x) for x in xq])
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    w = np.ascontiguousarray(omega, dtype=np.float64)
    out = np.zeros(p[0] * p[1] * p[2]) if V is None else \
        np.ascontiguousarray(np.asarray(V, np.float64).transpose(2, 1, 0)).reshape(-1)
    lib().orc_hartree_expansion(r0, r1, r2, _d(gf), Up, ne, xep, jwp, p, xqp, len(a), _d(a),
                                _d(w), _d(out), 0 if V is None else 1)
    return out.reshape(p[2], p[1], p[0]).transpose(2, 1, 0)


MESH_TYPES = {"uniform": 0, "adaptive": 1, "manual1": 3, "hartree_outer": 4}


def fem_mesh(mesh_type, nel, npe, start, end, inner_size=0.0, inner_h=0.0, coarsening=0.0):
    """Global node coordinates (fem/FEM.cc:159-375)."""
    nodes = np.zeros(nel * (npe - 1) + 1)
    n = lib().orc_fem_mesh(MESH_TYPES[mesh_type], nel, npe, start, end, inner_size, inner_h,
                           coarsening, _d(nodes))
    if n < 0:
        raise ValueError(mesh_type)
    return nodes


def fem_shape(npe, xi):
    """Shape functions and their xi-derivatives at the points xi: arrays
    (npe, len(xi)) (fem/FEM.cc:418-528)."""
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    S = np.zeros((npe, len(xi)))
    D = np.zeros((npe, len(xi)))
    lib().orc_fem_shape(npe, len(xi), _d(xi), _d(S), _d(D))
    return S, D


def fem_interpolate(nel, npe, xi, w, nodes, nodal=None):
    """Values and xi-derivatives of the nodal fields (rows of nodal) at the
    quadrature points, with jw, inv_jw and the point positions
    (fem/FEM.cc:552-571, 897-946).  Returns (val, dval, jw, inv_jw, xq)."""
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    nq = len(xi)
    npts = nel * nq
    jw, ijw, xq = np.zeros(npts), np.zeros(npts), np.zeros(npts)
    if nodal is None:
        nodal = np.zeros((0, len(nodes)))
    nodal = np.ascontiguousarray(np.atleast_2d(nodal), dtype=np.float64)
    nf = nodal.shape[0]
    val = np.zeros((max(nf, 1), npts))
    dval = np.zeros((max(nf, 1), npts))
    lib().orc_fem_interpolate(nel, npe, nq, _d(xi), _d(w), _d(nodes), nf, _d(nodal), _d(val),
                              _d(dval), _d(jw), _d(ijw), _d(xq))
    return val[:nf], dval[:nf], jw, ijw, xq


def tucker_basis_matrices(B, DB, jw, inv_jw, U=None):
    """MatX, MatDX and MatPot (rv, r, r) of one axis
    (dft/ProjectHamiltonianSparse.cc:577-722).  B, DB: (r, nq) basis values
    and xi-derivatives at the quadrature points; U: (nq, rv) or None."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    DB = np.ascontiguousarray(DB, dtype=np.float64)
    r, nq = B.shape
    jw = np.ascontiguousarray(jw, dtype=np.float64)
    inv_jw = np.ascontiguousarray(inv_jw, dtype=np.float64)
    MX, MDX = np.zeros((r, r)), np.zeros((r, r))
    rv = 0 if U is None else U.shape[1]
    Uc = np.ascontiguousarray(np.asarray(U if U is not None else np.zeros((nq, 1)),
                                         dtype=np.float64).T)  # column-major nq x rv
    MP = np.zeros((max(rv, 1), r, r))
    lib().orc_tucker_basis_matrices(nq, r, _d(B), _d(DB), _d(jw), _d(inv_jw), rv, _d(Uc), _d(MX),
                                    _d(MDX), _d(MP) if rv else None)
    return MX, MDX, (MP if rv else None)
