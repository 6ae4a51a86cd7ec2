"""ctypes binding of the engine's C ABI (include/ttdft_chfsi.h).

Device buffers are plain integers (device addresses); callers that use torch
pass ``tensor.data_ptr()`` and make the context issue work on torch's current
stream with :meth:`Context.use_stream`.  Every call raises
:class:`TtdftError` on a nonzero status, with the library's message.

The library must have been built (``__graft_entry__.build()``); loading fails
loudly otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libttdft_chfsi.so")

OK, EBADARG, EHIP, ENOTSPD, ENOCONV, ENOMEM, ELIB = 0, -1, -2, -3, -4, -5, -6
F64, F32 = 0, 1
GEMM_F64, GEMM_BF16X3 = 0, 1
DEGREE_REFERENCE, DEGREE_EXACT = 0, 1
HAM_F64, HAM_F32 = 1, 2
(KSTAT_SPMM, KSTAT_NLPROJ, KSTAT_GRAM, KSTAT_ROTATE, KSTAT_DENSITY, KSTAT_ASSEMBLY, KSTAT_MIXING,
 KSTAT_STHOSVD, KSTAT_HARTREE) = range(9)

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_d = ctypes.c_double


class TtdftError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ttdft status {status}: {msg}")
        self.status = status


class ChfsiState(ctypes.Structure):
    _fields_ = [
        ("upper", ctypes.c_double),
        ("lower_unwanted", ctypes.c_double),
        ("lower_wanted", ctypes.c_double),
        ("err_upper", ctypes.c_double),
        ("occ_energy", ctypes.c_double),
        ("passes", ctypes.c_int32),
        ("lanczos_runs", ctypes.c_int32),
        ("occ_tol", ctypes.c_double),
    ]


class SynthParams(ctypes.Structure):
    _fields_ = [
        ("rx", ctypes.c_int32), ("ry", ctypes.c_int32), ("rz", ctypes.c_int32),
        ("bx", ctypes.c_int32), ("by", ctypes.c_int32), ("bz", ctypes.c_int32),
        ("p", ctypes.c_int32),
        ("h", ctypes.c_double),
        ("truncation", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "ttdft_last_error": (ctypes.c_char_p, []),
    "ttdft_get_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "ttdft_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "ttdft_ctx_destroy": (ctypes.c_int, [_vp]),
    "ttdft_ctx_set_stream": (ctypes.c_int, [_vp, _vp]),
    "ttdft_ctx_synchronize": (ctypes.c_int, [_vp]),
    "ttdft_ctx_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ttdft_ctx_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_int, _c_i64p, _c_dp, _c_dp, _c_dp]),
    "ttdft_ctx_reset_stats": (ctypes.c_int, [_vp]),
    "ttdft_malloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "ttdft_free": (ctypes.c_int, [_vp, _vp]),
    "ttdft_memcpy_h2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "ttdft_memcpy_d2h": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "ttdft_memcpy_d2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "ttdft_block_upload_colmajor": (ctypes.c_int, [_vp, _c_dp, _i64, _i64, _i32, _vp, _i64,
                                                   ctypes.c_int]),
    "ttdft_block_download_colmajor": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _i64, _i32,
                                                     _c_dp, _i64]),
    "ttdft_block_convert": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _vp, _i64, ctypes.c_int,
                                           _i64, _i32]),
    "ttdft_ham_upload": (ctypes.c_int, [_vp, _i64, _c_i64p, _c_i32p, _c_dp, _i32, _c_i64p,
                                        _c_i32p, _c_dp, _c_dp, ctypes.c_int,
                                        ctypes.POINTER(_vp)]),
    "ttdft_ham_upload_tucker": (ctypes.c_int, [_vp, _i32, _i32, _i32, _c_i64p, _c_i32p, _c_dp,
                                               _i32, _c_i64p, _c_i32p, _c_dp, _c_dp, ctypes.c_int,
                                               ctypes.POINTER(_vp)]),
    "ttdft_ham_format": (ctypes.c_int, [_vp, _c_i32p, _c_i32p]),
    "ttdft_ham_free": (ctypes.c_int, [_vp]),
    "ttdft_ham_info": (ctypes.c_int, [_vp, _c_i64p, _c_i64p, _c_i32p]),
    "ttdft_ham_apply": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _i64, _i32, ctypes.c_int]),
    "ttdft_ham_apply_noalloc": (ctypes.c_int,
                                [_vp, _vp, _vp, _i64, _vp, _i64, _i32, ctypes.c_int]),
    "ttdft_cheb_filter": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _d, _d, _d, _i32,
                                         ctypes.c_int, ctypes.c_int]),
    "ttdft_gram": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, _i64, _i32, ctypes.c_int,
                                  ctypes.c_int, _vp]),
    "ttdft_orth": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i32, ctypes.c_int, ctypes.c_int]),
    "ttdft_rr_eig": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "ttdft_rotate": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i32, _vp, ctypes.c_int,
                                    ctypes.c_int]),
    "ttdft_lanczos_bounds": (ctypes.c_int, [_vp, _vp, _i32, ctypes.c_uint64, _c_dp, _c_dp]),
    "ttdft_fermi_occupancy": (ctypes.c_int, [_c_dp, _i32, _i32, _d, _d, _i32, _c_dp, _c_dp]),
    "ttdft_band_energy": (ctypes.c_double, [_c_dp, _c_dp, _i32]),
    "ttdft_density": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _i32, _c_dp, _i32, _i32, _i32,
                                     _vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    "ttdft_initial_density": (ctypes.c_int, [_vp, _i32, _c_dp, _i32, _c_dp, _c_dp, _i32, _i32,
                                             _i32, _vp, _vp, _vp, _vp]),
    "ttdft_grid_integral": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _c_dp]),
    "ttdft_grid_residual": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp,
                                           _c_dp]),
    "ttdft_grid_scale": (ctypes.c_int, [_vp, _vp, _i64, _d]),
    "ttdft_fill_uniform": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_uint64, _d, _d]),
    "ttdft_chfsi_state_init": (ctypes.c_int, [ctypes.POINTER(ChfsiState)]),
    "ttdft_chfsi_step": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i32,
                                        _i32, _i32, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                        _c_dp, ctypes.POINTER(ChfsiState)]),
    "ttdft_random_subspace": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i32, ctypes.c_uint64,
                                             ctypes.c_int]),
    "ttdft_synth_hamiltonian": (ctypes.c_int, [ctypes.POINTER(SynthParams), _c_i64p, _c_i64p,
                                               _c_i64p, _c_i32p, _c_dp, _c_i64p, _c_i32p, _c_dp,
                                               _c_dp]),
    "ttdft_assemble_hloc": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _i32, _i32] + [_c_dp] * 10
                            + [_d, _c_i64p, _c_i64p, _c_i32p, _c_dp]),
    "ttdft_assemble_projectors": (ctypes.c_int, [_vp, _i32, _c_i32p, _i32, _i32, _i32, _c_dp,
                                                 _c_dp, _c_dp, _c_dp, _d, _c_i64p, _c_i64p,
                                                 _c_i32p, _c_dp]),
    "ttdft_ham_assemble_tucker": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, _i32, _i32]
                                  + [_c_dp] * 10 + [_d, _i32, _c_i32p, _c_dp, _c_dp, _c_dp,
                                                    _c_dp, _d, _c_dp, ctypes.c_int,
                                                    ctypes.POINTER(_vp)]),
    "ttdft_mixer_create": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _d, _i32,
                                          ctypes.POINTER(_vp)]),
    "ttdft_mixer_destroy": (ctypes.c_int, [_vp]),
    "ttdft_mixer_update": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32]),
    "ttdft_mixer_compute_rho_in": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "ttdft_mixer_constants": (ctypes.c_int, [_vp, _c_dp, _c_i32p]),
    "ttdft_mixer_clear_history": (ctypes.c_int, [_vp, _i32]),
    "ttdft_mixer_history_size": (ctypes.c_int, [_vp, _c_i32p]),
    "ttdft_sthosvd": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp,
                                     _vp, _vp, _c_dp]),
    "ttdft_sthosvd_refine": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp,
                                            _vp, _vp, _vp, _i32, _c_dp]),
    "ttdft_hartree_expansion": (ctypes.c_int, [_vp, _c_i32p, _vp, ctypes.POINTER(_vp), _c_i32p,
                                               ctypes.POINTER(_vp), ctypes.POINTER(_vp), _c_i32p,
                                               ctypes.POINTER(_vp), _i32, _c_dp, _c_dp, _vp,
                                               ctypes.c_int]),
    "ttdft_fem_gauss_legendre": (ctypes.c_int, [_i32, _c_dp, _c_dp]),
    "ttdft_fem_quadrature_rule": (ctypes.c_int, [_i32, _c_dp, _c_dp]),
    "ttdft_fem_mesh": (ctypes.c_int, [_i32, _i32, _i32, _d, _d, _d, _d, _d, _c_dp]),
    "ttdft_fem_interpolate": (ctypes.c_int, [_vp, _i32, _i32, _i32, _c_dp, _c_dp, _c_dp, _i32,
                                             _vp, _vp, _vp, _c_dp, _c_dp, _c_dp]),
    "ttdft_tucker_basis_matrices": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _i32,
                                                   _vp, _vp, _vp, _vp]),
    "ttdft_read_basis_file": (ctypes.c_int, [ctypes.c_char_p, _c_i32p, _c_i32p, _c_dp]),
    "ttdft_read_atom_file": (ctypes.c_int, [ctypes.c_char_p, _c_i32p, _c_i32p, _c_i32p, _c_i32p,
                                            _c_dp]),
    "ttdft_read_table": (ctypes.c_int, [ctypes.c_char_p, _c_i32p, _c_i32p, _c_dp]),
    "ttdft_spline_parabolic": (ctypes.c_int, [_i32, _c_dp, _c_dp, _c_dp]),
    "ttdft_psp_local": (ctypes.c_int, [_vp, _i32, _c_dp, _i32, _c_dp, _c_dp, _i32, _i32, _i32,
                                       _vp, _vp, _vp, _vp, ctypes.c_int]),
    "ttdft_psp_projectors": (ctypes.c_int, [_vp, _i32, _c_dp, _c_dp, _c_dp, _c_dp, _i32, _i32,
                                            _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _c_dp]),
    "ttdft_restart_write": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp, _i64, ctypes.c_int, _i64,
                                           _i32]),
    "ttdft_restart_read": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp, _i64, ctypes.c_int, _i64,
                                          _i32]),
    "ttdft_ctx_trim": (ctypes.c_int, [_vp]),
    "ttdft_restart_write_bounds": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ChfsiState),
                                                  _d]),
    "ttdft_restart_read_bounds": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ChfsiState),
                                                 _c_dp]),
    "ttdft_lda_xc": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "ttdft_local_pseudopotential": (ctypes.c_int, [_vp, _i32, _c_dp, _d, _i32, _i32, _i32, _vp,
                                                   _vp, _vp, _vp, ctypes.c_int]),
    "ttdft_grid_weights": (ctypes.c_int, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "ttdft_grid_dot": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _c_dp]),
    "ttdft_grid_resample": (ctypes.c_int, [_vp, _vp, _c_i32p, _c_dp, _c_dp, _i32, _i32, _i32,
                                           _vp, _vp, _vp, _vp, ctypes.c_int]),
    "ttdft_gaussian_splat": (ctypes.c_int, [_vp, _i32, _c_dp, _d, _c_i32p, _c_dp, _c_dp, _vp]),
    "ttdft_grid_clamp": (ctypes.c_int, [_vp, _vp, _i64, _d]),
    "ttdft_grid_axpby": (ctypes.c_int, [_vp, _vp, _d, _vp, _d, _i64]),
    "ttdft_nn_set_math": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ttdft_nn_conv3d_forward": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp, _vp, _i32, _i32,
                                               _vp, ctypes.c_int]),
    "ttdft_nn_conv3d_backward_data": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp, _i32, _i32,
                                                     _vp, ctypes.c_int]),
    "ttdft_nn_conv3d_backward_weight": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp, _i32, _i32,
                                                       _vp, _vp, ctypes.c_int]),
    "ttdft_nn_leaky_relu": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_float]),
    "ttdft_nn_leaky_relu_backward": (ctypes.c_int, [_vp, _vp, _vp, _i64, ctypes.c_float]),
    "ttdft_nn_avgpool2": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp]),
    "ttdft_nn_avgpool2_backward": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp, ctypes.c_int]),
    "ttdft_nn_upsample2": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp]),
    "ttdft_nn_upsample2_backward": (ctypes.c_int, [_vp, _vp, _i32, _c_i32p, _vp, ctypes.c_int]),
    "ttdft_nn_mse": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _c_dp]),
    "ttdft_nn_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float, _i32]),
    "ttdft_nn_to_f32": (ctypes.c_int, [_vp, _vp, _vp, _i64, _d]),
    "ttdft_nn_to_f64": (ctypes.c_int, [_vp, _vp, _vp, _i64, _d]),
    "ttdft_rng_uniform": (ctypes.c_int, [ctypes.c_uint64, _i64, _c_dp]),
    "ttdft_rng_normal": (ctypes.c_int, [ctypes.c_uint64, _i64, _c_dp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """Load libttdft_chfsi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        msg = lib().ttdft_last_error()
        raise TtdftError(rc, msg.decode() if msg else "")


def _dptr(a):
    return a.ctypes.data_as(_c_dp)


def _i64ptr(a):
    return a.ctypes.data_as(_c_i64p)


def _i32ptr(a):
    return a.ctypes.data_as(_c_i32p)


def device_count():
    n = ctypes.c_int(0)
    _check(lib().ttdft_get_device_count(ctypes.byref(n)))
    return n.value


def rng_uniform(seed, count):
    out = np.empty(count)
    _check(lib().ttdft_rng_uniform(seed, count, _dptr(out)))
    return out


def rng_normal(seed, count):
    out = np.empty(count)
    _check(lib().ttdft_rng_normal(seed, count, _dptr(out)))
    return out


MESH_UNIFORM, MESH_ADAPTIVE, MESH_MANUAL1, MESH_HARTREE_OUTER = 0, 1, 3, 4
LDA_PZ81, LDA_PZ81_REFERENCE = 0, 1
NN_FP32, NN_BF16 = 0, 1
MESH_TYPES = {"uniform": MESH_UNIFORM, "adaptive": MESH_ADAPTIVE, "manual1": MESH_MANUAL1,
              "hartree_outer": MESH_HARTREE_OUTER}


def gauss_legendre(n):
    """n-point Gauss-Legendre rule on [-1, 1], points ascending."""
    xi, w = np.empty(n), np.empty(n)
    _check(lib().ttdft_fem_gauss_legendre(n, _dptr(xi), _dptr(w)))
    return xi, w


def quadrature_rule(name):
    """The reference's "<n>pt" rule (fem/QuadratureRule.cc) in its point
    order: name is "10pt" or the number of points."""
    n = int(str(name).rstrip("pt"))
    xi, w = np.empty(n), np.empty(n)
    _check(lib().ttdft_fem_quadrature_rule(n, _dptr(xi), _dptr(w)))
    return xi, w


def fem_mesh(mesh_type, nel, npe, start, end, inner_size=0.0, inner_h=0.0, coarsening=0.0):
    """Global node coordinates of FEM::generateNodes (fem/FEM.cc:159-375)."""
    t = MESH_TYPES[mesh_type] if isinstance(mesh_type, str) else int(mesh_type)
    nodes = np.empty(nel * (npe - 1) + 1)
    _check(lib().ttdft_fem_mesh(t, nel, npe, start, end, inner_size, inner_h, coarsening,
                                _dptr(nodes)))
    return nodes


def read_basis_file(path):
    """Basis functions of a file, zero padded at both ends: (nfun, nnodes)."""
    nf, nn = ctypes.c_int32(), ctypes.c_int32()
    p = os.fsencode(path)
    _check(lib().ttdft_read_basis_file(p, ctypes.byref(nf), ctypes.byref(nn), None))
    out = np.empty((nf.value, nn.value))
    _check(lib().ttdft_read_basis_file(p, ctypes.byref(nf), ctypes.byref(nn), _dptr(out)))
    return out


def read_atom_file(path):
    """(number of electrons, lmax per type, atoms (natoms, 7) grouped by type)."""
    nt, na, ne = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    p = os.fsencode(path)
    _check(lib().ttdft_read_atom_file(p, ctypes.byref(nt), ctypes.byref(na), ctypes.byref(ne),
                                      None, None))
    lmax = np.empty(nt.value, dtype=np.int32)
    atoms = np.empty((na.value, 7))
    _check(lib().ttdft_read_atom_file(p, ctypes.byref(nt), ctypes.byref(na), ctypes.byref(ne),
                                      _i32ptr(lmax), _dptr(atoms)))
    return ne.value, lmax, atoms


def read_table(path):
    """Whitespace-separated numeric table: (nrows, ncols)."""
    nr, nc = ctypes.c_int32(), ctypes.c_int32()
    p = os.fsencode(path)
    _check(lib().ttdft_read_table(p, ctypes.byref(nr), ctypes.byref(nc), None))
    out = np.empty((nr.value, nc.value))
    _check(lib().ttdft_read_table(p, ctypes.byref(nr), ctypes.byref(nc), _dptr(out)))
    return out


def spline_parabolic(x, y):
    """Cubic spline with parabolic end segments through (x, y): (n - 1, 4)
    power-basis coefficients per segment in t = x - x_s."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    coef = np.empty((max(len(x) - 1, 1), 4))
    _check(lib().ttdft_spline_parabolic(len(x), _dptr(x), _dptr(y), _dptr(coef)))
    return coef


def restart_write_bounds(path, state, energy):
    """cheby_bounds.txt of main.cc:1150-1157 from a ChfsiState and the energy."""
    _check(lib().ttdft_restart_write_bounds(os.fsencode(path), ctypes.byref(state), energy))


def restart_read_bounds(path, state):
    """Sets the five bound fields of state from cheby_bounds.txt; returns the
    energy stored with them."""
    e = ctypes.c_double()
    _check(lib().ttdft_restart_read_bounds(os.fsencode(path), ctypes.byref(state),
                                           ctypes.byref(e)))
    return e.value


def fermi_occupancy(evals, nelec, kb=3.166811429e-06, temperature=500.0, maxit=20):
    e = np.ascontiguousarray(evals, dtype=np.float64)
    occ = np.empty_like(e)
    ef = ctypes.c_double()
    _check(lib().ttdft_fermi_occupancy(_dptr(e), len(e), int(nelec), kb, temperature, maxit,
                                       ctypes.byref(ef), _dptr(occ)))
    return ef.value, occ


def band_energy(evals, occ):
    e = np.ascontiguousarray(evals, dtype=np.float64)
    o = np.ascontiguousarray(occ, dtype=np.float64)
    return lib().ttdft_band_energy(_dptr(e), _dptr(o), len(e))


class SynthHam:
    """Synthetic Kronecker-banded Hamiltonian (CSR arrays in numpy)."""

    def __init__(self, rx, ry, rz, bx, by, bz, p, h=2.0, truncation=1e-6, seed=20261202):
        prm = SynthParams(rx, ry, rz, bx, by, bz, p, h, truncation, seed)
        nnz = ctypes.c_int64()
        nnz_c = ctypes.c_int64()
        _check(lib().ttdft_synth_hamiltonian(ctypes.byref(prm), ctypes.byref(nnz),
                                             ctypes.byref(nnz_c), None, None, None, None, None,
                                             None, None))
        self.dims = (rx, ry, rz)
        self.n = rx * ry * rz
        self.p = p
        self.rowptr = np.empty(self.n + 1, np.int64)
        self.col = np.empty(max(nnz.value, 1), np.int32)
        self.val = np.empty(max(nnz.value, 1), np.float64)
        self.c_rowptr = np.zeros(p + 1, np.int64)
        self.c_col = np.empty(max(nnz_c.value, 1), np.int32)
        self.c_val = np.empty(max(nnz_c.value, 1), np.float64)
        self.inv_c_lm = np.ones(max(p, 1), np.float64)
        _check(lib().ttdft_synth_hamiltonian(ctypes.byref(prm), ctypes.byref(nnz),
                                             ctypes.byref(nnz_c), _i64ptr(self.rowptr),
                                             _i32ptr(self.col), _dptr(self.val),
                                             _i64ptr(self.c_rowptr), _i32ptr(self.c_col),
                                             _dptr(self.c_val), _dptr(self.inv_c_lm)))
        self.nnz = nnz.value
        self.nnz_c = nnz_c.value
        self.col = self.col[: self.nnz]
        self.val = self.val[: self.nnz]
        self.c_col = self.c_col[: self.nnz_c]
        self.c_val = self.c_val[: self.nnz_c]
        self.inv_c_lm = self.inv_c_lm[:p]


def _hloc_args(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G):
    """(arrays to keep alive, C arguments rx..rvz, X..G) of the H_loc assembly."""
    rx, ry, rz = X.shape[0], Y.shape[0], Z.shape[0]
    rvx, rvy, rvz = PX.shape[0], PY.shape[0], PZ.shape[0]
    mats = [np.ascontiguousarray(m, dtype=np.float64) for m in (X, DX, Y, DY, Z, DZ, PX, PY, PZ)]
    g = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(rvx, rvy, rvz)
                             .transpose(2, 1, 0)).reshape(-1)  # first index fastest
    keep = mats + [g]
    return keep, [rx, ry, rz, rvx, rvy, rvz] + [_dptr(m) for m in keep]


def _proj_args(cores, mx, my, mz):
    """(arrays to keep alive, C arguments ranks, cores, mx, my, mz) of the
    projector assembly: cores[i] is (r_i, r_i, r_i) indexed [a, b, c] (stored
    first index fastest), factors column-major."""
    ranks = np.array([c.shape[0] for c in cores], np.int32)
    cc = np.concatenate([np.asarray(c, np.float64).transpose(2, 1, 0).reshape(-1)
                         for c in cores])
    fx, fy, fz = (np.concatenate([np.asarray(m, np.float64).T.reshape(-1) for m in ms])
                  for ms in (mx, my, mz))
    keep = [ranks, cc, fx, fy, fz]
    return keep, [_i32ptr(ranks), _dptr(cc), _dptr(fx), _dptr(fy), _dptr(fz)]


class Context:
    """One device + one stream (ttdft_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        _check(lib().ttdft_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _check(lib().ttdft_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_stream(self, stream_handle):
        _check(lib().ttdft_ctx_set_stream(self.h, stream_handle))

    def synchronize(self):
        _check(lib().ttdft_ctx_synchronize(self.h))

    def trim(self):
        """Give the context's cached storage blocks back to the HIP runtime."""
        _check(lib().ttdft_ctx_trim(self.h))

    def set_profiling(self, on=True):
        _check(lib().ttdft_ctx_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self, cls):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        b = ctypes.c_double()
        f = ctypes.c_double()
        _check(lib().ttdft_ctx_kernel_stats(self.h, cls, ctypes.byref(n), ctypes.byref(ms),
                                            ctypes.byref(b), ctypes.byref(f)))
        return {"launches": n.value, "ms": ms.value, "bytes": b.value, "flops": f.value}

    def reset_stats(self):
        _check(lib().ttdft_ctx_reset_stats(self.h))

    # ---- Hamiltonian ----
    def ham_upload(self, s: SynthHam, copies=HAM_F64, tucker=True):
        """Upload a synthetic H; with tucker=True its ranks are passed so the
        engine can use the Kronecker band format."""
        return Hamiltonian(self, s.n, s.rowptr, s.col, s.val, s.p, s.c_rowptr, s.c_col, s.c_val,
                           s.inv_c_lm, copies, dims=s.dims if tucker else None)

    def ham_apply(self, ham, X, ldx, Y, ldy, k, dtype=F64):
        _check(lib().ttdft_ham_apply(self.h, ham.h, X, ldx, Y, ldy, k, dtype))

    def ham_apply_noalloc(self, ham, X, ldx, Y, ldy, k, dtype=F64):
        """ham_apply with the projections in the library's static work space
        (allocates nothing)."""
        _check(lib().ttdft_ham_apply_noalloc(self.h, ham.h, X, ldx, Y, ldy, k, dtype))

    def cheb_filter(self, ham, X, ldx, k, m, a, b, a0, num_blocks=1, dtype=F64,
                    convention=DEGREE_REFERENCE):
        _check(lib().ttdft_cheb_filter(self.h, ham.h, X, ldx, k, m, a, b, a0, num_blocks, dtype,
                                       convention))

    def gram(self, X, ldx, Y, ldy, n, k, dS, dtype=F64, mode=GEMM_F64):
        _check(lib().ttdft_gram(self.h, X, ldx, Y, ldy, n, k, dtype, mode, dS))

    def orth(self, X, ldx, n, k, dtype=F64, mode=GEMM_F64):
        _check(lib().ttdft_orth(self.h, X, ldx, n, k, dtype, mode))

    def rr_eig(self, dA, k, d_evals, dQ):
        _check(lib().ttdft_rr_eig(self.h, dA, k, d_evals, dQ))

    def rotate(self, X, ldx, n, k, dQ, dtype=F64, mode=GEMM_F64):
        _check(lib().ttdft_rotate(self.h, X, ldx, n, k, dQ, dtype, mode))

    def lanczos_bounds(self, ham, maxit=20, seed=20261202):
        up = ctypes.c_double()
        lo = ctypes.c_double()
        _check(lib().ttdft_lanczos_bounds(self.h, ham.h, maxit, seed, ctypes.byref(up),
                                          ctypes.byref(lo)))
        return up.value, lo.value

    def density(self, X, ldx, k, occ, rdims, dBx, dBy, dBz, pdims, d_rho, dtype=F64):
        o = np.ascontiguousarray(occ, dtype=np.float64)
        _check(lib().ttdft_density(self.h, X, ldx, dtype, k, _dptr(o), rdims[0], rdims[1],
                                   rdims[2], dBx, dBy, dBz, pdims[0], pdims[1], pdims[2], d_rho))

    def initial_density(self, xyz, r, rho_r, dgx, dgy, dgz, pdims, d_rho):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(r, dtype=np.float64)
        rho_r = np.ascontiguousarray(rho_r, dtype=np.float64)
        _check(lib().ttdft_initial_density(self.h, len(xyz) // 3, _dptr(xyz), len(r), _dptr(r),
                                           _dptr(rho_r), pdims[0], pdims[1], pdims[2], dgx, dgy,
                                           dgz, d_rho))

    def psp_local(self, atoms, r, vloc, dgx, dgy, dgz, pdims, d_v, add=False):
        """Local pseudopotential of the atoms (rows of 7: Z, x, y, z, ...) from
        the radial table (r, vloc) on the device grid; written or added."""
        atoms = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 7)
        r = np.ascontiguousarray(r, dtype=np.float64)
        vloc = np.ascontiguousarray(vloc, dtype=np.float64)
        _check(lib().ttdft_psp_local(self.h, len(atoms), _dptr(atoms), len(r), _dptr(r),
                                     _dptr(vloc), pdims[0], pdims[1], pdims[2], dgx, dgy, dgz,
                                     d_v, 1 if add else 0))

    def psp_projectors(self, r, vloc, vl, rl, dgx, dgy, dgz, pdims, dwx=None, dwy=None,
                       dwz=None, d_field=None, d_phi=None):
        """dV_l phi_lm and phi_lm fields of one atom type on an atom-centred
        device grid (written where given); returns the C_lm integrals when the
        weights are given, else None."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        vloc = np.ascontiguousarray(vloc, dtype=np.float64)
        vl = np.ascontiguousarray(np.atleast_2d(vl), dtype=np.float64)
        rl = np.ascontiguousarray(np.atleast_2d(rl), dtype=np.float64)
        lmax = vl.shape[0]
        clm = np.empty(lmax * lmax) if dwx is not None else None
        _check(lib().ttdft_psp_projectors(self.h, len(r), _dptr(r), _dptr(vloc), _dptr(vl),
                                          _dptr(rl), lmax, pdims[0], pdims[1], pdims[2], dgx,
                                          dgy, dgz, dwx, dwy, dwz, d_field, d_phi,
                                          _dptr(clm) if clm is not None else None))
        return clm

    def restart_write(self, path, X, ldx, n, k, dtype=F64):
        """Basis-major device block to the reference's column-major big-endian
        FP64 restart file."""
        _check(lib().ttdft_restart_write(self.h, os.fsencode(path), X, ldx, dtype, n, k))

    def restart_read(self, path, X, ldx, n, k, dtype=F64):
        _check(lib().ttdft_restart_read(self.h, os.fsencode(path), X, ldx, dtype, n, k))

    def grid_integral(self, d_f, pdims, dwx, dwy, dwz):
        out = ctypes.c_double()
        _check(lib().ttdft_grid_integral(self.h, d_f, pdims[0], pdims[1], pdims[2], dwx, dwy, dwz,
                                         ctypes.byref(out)))
        return out.value

    def grid_residual(self, d_f, d_g, pdims, dwx, dwy, dwz):
        out = ctypes.c_double()
        _check(lib().ttdft_grid_residual(self.h, d_f, d_g, pdims[0], pdims[1], pdims[2], dwx, dwy,
                                         dwz, ctypes.byref(out)))
        return out.value

    def sthosvd(self, d_T, dims, ranks, d_core, d_U):
        """Fixed-rank STHOSVD of a device tensor (first index fastest) into the
        device core and factors d_U[0..2]; returns the kept Gram eigenvalues."""
        ev = np.zeros(sum(ranks))
        _check(lib().ttdft_sthosvd(self.h, d_T, *dims, *ranks, d_core, d_U[0], d_U[1], d_U[2],
                                   _dptr(ev)))
        return ev

    def sthosvd_refine(self, d_T, dims, ranks, d_core, d_U, iters=2):
        """STHOSVD from the factors already in d_U (refined in place)."""
        ev = np.zeros(sum(ranks))
        _check(lib().ttdft_sthosvd_refine(self.h, d_T, *dims, *ranks, d_core, d_U[0], d_U[1],
                                          d_U[2], iters, _dptr(ev)))
        return ev

    def hartree_expansion(self, ranks, d_core, d_U, d_xe, d_jw, ne, d_xq, p, alpha, omega, d_V,
                          add=False):
        """Kernel-expansion Hartree potential into the device grid d_V."""
        r = (ctypes.c_int32 * 3)(*ranks)
        nes = (ctypes.c_int32 * 3)(*ne)
        ps = (ctypes.c_int32 * 3)(*p)
        U = (_vp * 3)(*d_U)
        xe = (_vp * 3)(*d_xe)
        jw = (_vp * 3)(*d_jw)
        xq = (_vp * 3)(*d_xq)
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        w = np.ascontiguousarray(omega, dtype=np.float64)
        _check(lib().ttdft_hartree_expansion(self.h, r, d_core, U, nes, xe, jw, ps, xq, len(a),
                                             _dptr(a), _dptr(w), d_V, 1 if add else 0))

    def fem_interpolate(self, nel, npe, xi, w, nodes, nfun=0, d_nodal=None, d_val=None,
                        d_dval=None):
        """Values and xi-derivatives at the quadrature points of nfun device
        nodal fields (rows of d_nodal) into d_val / d_dval; returns the host
        arrays (jw, inv_jw, positions) of the nel len(xi) points."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        npts = nel * len(xi)
        jw, ijw, xq = np.empty(npts), np.empty(npts), np.empty(npts)
        _check(lib().ttdft_fem_interpolate(self.h, nel, npe, len(xi), _dptr(xi), _dptr(w),
                                           _dptr(nodes), nfun, d_nodal, d_val, d_dval,
                                           _dptr(jw), _dptr(ijw), _dptr(xq)))
        return jw, ijw, xq

    def tucker_basis_matrices(self, nq, r, d_B, d_DB, d_jw, d_inv_jw, rv=0, d_U=None,
                              d_MatX=None, d_MatDX=None, d_MatPot=None):
        """MatX, MatDX, MatPot of one axis into device buffers (r x r each)."""
        _check(lib().ttdft_tucker_basis_matrices(self.h, nq, r, d_B, d_DB, d_jw, d_inv_jw, rv,
                                                 d_U, d_MatX, d_MatDX, d_MatPot))

    def lda_xc(self, d_rho, count, d_vxc=None, d_exc=None, add=False, variant=LDA_PZ81):
        _check(lib().ttdft_lda_xc(self.h, d_rho, count, d_vxc, d_exc, 1 if add else 0, variant))

    def local_pseudopotential(self, atoms, rc, pdims, d_gx, d_gy, d_gz, d_V, add=False):
        """atoms: (natoms, 4) host array of x, y, z, Z."""
        a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
        _check(lib().ttdft_local_pseudopotential(self.h, a.shape[0], _dptr(a), rc, *pdims, d_gx,
                                                 d_gy, d_gz, d_V, 1 if add else 0))

    def grid_weights(self, pdims, d_wx, d_wy, d_wz, d_w):
        _check(lib().ttdft_grid_weights(self.h, *pdims, d_wx, d_wy, d_wz, d_w))

    def grid_dot(self, d_f, d_g, pdims, dwx, dwy, dwz):
        out = ctypes.c_double()
        _check(lib().ttdft_grid_dot(self.h, d_f, d_g, *pdims, dwx, dwy, dwz, ctypes.byref(out)))
        return out.value

    def grid_resample(self, d_src, n, origin, h, pdims, d_gx, d_gy, d_gz, d_dst, add=False):
        nn = (ctypes.c_int32 * 3)(*n)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        hh = np.ascontiguousarray(h, dtype=np.float64)
        _check(lib().ttdft_grid_resample(self.h, d_src, nn, _dptr(o), _dptr(hh), *pdims, d_gx,
                                         d_gy, d_gz, d_dst, 1 if add else 0))

    def gaussian_splat(self, atoms, sigma, n, origin, h, d_out):
        """atoms: (natoms, 4) host array of x, y, z, amplitude."""
        a = np.ascontiguousarray(atoms, dtype=np.float64).reshape(-1, 4)
        nn = (ctypes.c_int32 * 3)(*n)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        hh = np.ascontiguousarray(h, dtype=np.float64)
        _check(lib().ttdft_gaussian_splat(self.h, a.shape[0], _dptr(a), sigma, nn, _dptr(o),
                                          _dptr(hh), d_out))

    def grid_axpby(self, d_y, a, d_x, b, count):
        """y = a y + b x on the device."""
        _check(lib().ttdft_grid_axpby(self.h, d_y, a, d_x, b, count))

    def copy(self, d_dst, d_src, nbytes):
        _check(lib().ttdft_memcpy_d2d(self.h, d_dst, d_src, nbytes))

    def grid_clamp(self, d_f, count, lo=0.0):
        _check(lib().ttdft_grid_clamp(self.h, d_f, count, lo))

    def fill_uniform(self, d_f, count, seed, lo=0.0, hi=1.0):
        _check(lib().ttdft_fill_uniform(self.h, d_f, count, seed, float(lo), float(hi)))

    def grid_scale(self, d_f, count, s):
        _check(lib().ttdft_grid_scale(self.h, d_f, count, s))

    def chfsi_step(self, ham, X, ldx, k, m, evals, state, scf_iter, occ_index, restart_times,
                   num_blocks=1, start_iter=0, maxit_lanczos=20, seed=20261202,
                   convention=DEGREE_REFERENCE, precision=F64):
        assert evals.dtype == np.float64 and evals.flags.c_contiguous and len(evals) == k
        _check(lib().ttdft_chfsi_step(self.h, ham.h, X, ldx, k, m, num_blocks, restart_times,
                                      occ_index, scf_iter, start_iter, maxit_lanczos, seed,
                                      convention, precision, _dptr(evals), ctypes.byref(state)))
        return state

    def assemble_hloc(self, X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, truncation):
        """H_loc rows as (rowptr, col, val) numpy arrays, assembled on the device
        (dft/ProjectHamiltonianSparse.cc:330-575).  X..DZ are r x r with [I][J]
        indexing, PX is (rvx, rx, rx) = MatPotX[p][I][J], G has shape (rvx, rvy, rvz)."""
        keep, hargs = _hloc_args(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G)
        rx, ry, rz = hargs[:3]
        args = [self.h] + hargs + [float(truncation)]
        nnz = ctypes.c_int64()
        _check(lib().ttdft_assemble_hloc(*args, ctypes.byref(nnz), None, None, None))
        n = rx * ry * rz
        rowptr = np.empty(n + 1, np.int64)
        col = np.empty(max(nnz.value, 1), np.int32)
        val = np.empty(max(nnz.value, 1), np.float64)
        _check(lib().ttdft_assemble_hloc(*args, ctypes.byref(nnz), _i64ptr(rowptr),
                                         _i32ptr(col), _dptr(val)))
        return rowptr, col[: nnz.value], val[: nnz.value]

    def assemble_projectors(self, cores, mx, my, mz, truncation):
        """Projector rows (dft/ProjectHamiltonianSparse.cc:196-278) as (rowptr, col, val),
        assembled on the device.  cores[i] has shape (r_i, r_i, r_i) indexed [a, b, c];
        mx[i] is rx x r_i, my[i] ry x r_i, mz[i] rz x r_i."""
        p = len(cores)
        keep, pargs = _proj_args(cores, mx, my, mz)
        args = [self.h, p, pargs[0], mx[0].shape[0], my[0].shape[0], mz[0].shape[0]] + \
            pargs[1:] + [float(truncation)]
        nnz = ctypes.c_int64()
        _check(lib().ttdft_assemble_projectors(*args, ctypes.byref(nnz), None, None, None))
        rowptr = np.empty(p + 1, np.int64)
        col = np.empty(max(nnz.value, 1), np.int32)
        val = np.empty(max(nnz.value, 1), np.float64)
        _check(lib().ttdft_assemble_projectors(*args, ctypes.byref(nnz), _i64ptr(rowptr),
                                               _i32ptr(col), _dptr(val)))
        return rowptr, col[: nnz.value], val[: nnz.value]

    def ham_assemble_tucker(self, hloc, truncation, proj=None, c_truncation=0.0,
                            inv_c_lm=None, copies=HAM_F64):
        """Device-resident H = H_loc + C^T D C from the assembly inputs:
        hloc = (X, DX, Y, DY, Z, DZ, PX, PY, PZ, G) as for assemble_hloc,
        proj = (cores, mx, my, mz) as for assemble_projectors (or None)."""
        keep, hargs = _hloc_args(*hloc)
        rx, ry, rz = hargs[:3]
        if proj:
            p = len(proj[0])
            keep2, pargs = _proj_args(*proj)
            inv = np.ascontiguousarray(inv_c_lm, np.float64)
            assert inv.shape == (p,)
            tail = [p] + pargs + [float(c_truncation), _dptr(inv)]
        else:
            keep2, inv = None, None
            tail = [0, None, None, None, None, None, 0.0, None]
        h = _vp()
        _check(lib().ttdft_ham_assemble_tucker(self.h, *hargs, float(truncation), *tail, copies,
                                               ctypes.byref(h)))
        return Hamiltonian.adopt(self, h, rx * ry * rz)

    def random_subspace(self, X, ldx, n, k, seed=20261202, dtype=F64):
        _check(lib().ttdft_random_subspace(self.h, X, ldx, n, k, seed, dtype))

    def block_upload_colmajor(self, host_cm, dX, ldx, dtype=F64):
        a = np.asfortranarray(host_cm, dtype=np.float64)
        n, k = a.shape
        _check(lib().ttdft_block_upload_colmajor(self.h, a.ctypes.data_as(_c_dp), n, n, k, dX,
                                                 ldx, dtype))

    def block_download_colmajor(self, dX, ldx, n, k, dtype=F64):
        out = np.empty((n, k), order="F")
        _check(lib().ttdft_block_download_colmajor(self.h, dX, ldx, dtype, n, k,
                                                   out.ctypes.data_as(_c_dp), n))
        return out


def new_state():
    st = ChfsiState()
    _check(lib().ttdft_chfsi_state_init(ctypes.byref(st)))
    return st


class Hamiltonian:
    """Device copy of H = H_loc + C^T D C (ttdft_ham)."""

    def __init__(self, ctx, n, rowptr, col, val, p, c_rowptr, c_col, c_val, inv_c_lm,
                 copies=HAM_F64, dims=None):
        self.ctx = ctx
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        h = _vp()
        if p > 0:
            c_rowptr = np.ascontiguousarray(c_rowptr, np.int64)
            c_col = np.ascontiguousarray(c_col, np.int32)
            c_val = np.ascontiguousarray(c_val, np.float64)
            inv_c_lm = np.ascontiguousarray(inv_c_lm, np.float64)
            args = (_i64ptr(c_rowptr), _i32ptr(c_col), _dptr(c_val), _dptr(inv_c_lm))
        else:
            args = (None, None, None, None)
        if dims is None:
            _check(lib().ttdft_ham_upload(ctx.h, n, _i64ptr(rowptr), _i32ptr(col), _dptr(val), p,
                                          *args, copies, ctypes.byref(h)))
        else:
            assert dims[0] * dims[1] * dims[2] == n
            _check(lib().ttdft_ham_upload_tucker(ctx.h, dims[0], dims[1], dims[2],
                                                 _i64ptr(rowptr), _i32ptr(col), _dptr(val), p,
                                                 *args, copies, ctypes.byref(h)))
        self.h = h
        self.n = n

    @classmethod
    def adopt(cls, ctx, handle, n):
        """Wrap a ttdft_ham created on the device (ttdft_ham_assemble_tucker)."""
        obj = cls.__new__(cls)
        obj.ctx = ctx
        obj.h = handle
        obj.n = n
        return obj

    def format(self):
        """(0 = CSR | 1 = Kronecker band, (bx, by, bz))"""
        f = ctypes.c_int32()
        b = (ctypes.c_int32 * 3)()
        _check(lib().ttdft_ham_format(self.h, ctypes.byref(f), b))
        return f.value, tuple(b)

    def info(self):
        n = ctypes.c_int64()
        nnz = ctypes.c_int64()
        p = ctypes.c_int32()
        _check(lib().ttdft_ham_info(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(p)))
        return n.value, nnz.value, p.value

    def free(self):
        if self.h:
            _check(lib().ttdft_ham_free(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """Device buffer owned through the library's allocator (ttdft_malloc)."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _check(lib().ttdft_malloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(ctx, a.shape, a.dtype)
        _check(lib().ttdft_memcpy_h2d(ctx.h, d.ptr, a.ctypes.data_as(_vp), a.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().ttdft_memcpy_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().ttdft_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(_vp), a.nbytes))

    def free(self):
        if self.ptr:
            _check(lib().ttdft_free(self.ctx.h, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mixer:
    """Anderson density mixing with the history in device memory (ttdft_mixer,
    the AndersonMixing object of main.cc:698-707).  Densities are device
    pointers: grid (n_grid values) and nodal (n_nodal, may be 0)."""

    def __init__(self, ctx, n_grid, n_nodal, d_jacob, d_weight, alpha, max_history):
        h = _vp()
        _check(lib().ttdft_mixer_create(ctx.h, n_grid, n_nodal, d_jacob, d_weight, float(alpha),
                                        max_history, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx

    def update(self, nodal_in, grid_in, nodal_out, grid_out, scf_iter):
        _check(lib().ttdft_mixer_update(self.h, nodal_in, grid_in, nodal_out, grid_out, scf_iter))

    def compute_rho_in(self, scf_iter, nodal_in, grid_in):
        _check(lib().ttdft_mixer_compute_rho_in(self.h, scf_iter, nodal_in, grid_in))

    def constants(self):
        n = ctypes.c_int32()
        _check(lib().ttdft_mixer_constants(self.h, None, ctypes.byref(n)))
        c = np.zeros(max(n.value, 1))
        _check(lib().ttdft_mixer_constants(self.h, _dptr(c), ctypes.byref(n)))
        return c[: n.value]

    def clear_history(self, scf_iter):
        _check(lib().ttdft_mixer_clear_history(self.h, scf_iter))

    def history_size(self):
        n = ctypes.c_int32()
        _check(lib().ttdft_mixer_history_size(self.h, ctypes.byref(n)))
        return n.value

    def free(self):
        if self.h:
            _check(lib().ttdft_mixer_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
