"""The C-ABI library loads without a GPU and exports every function the
public header declares; host-only entry points agree with the oracle."""
import ctypes
import os
import re

import numpy as np

import oracle
from ttdft_public_amd import engine as E

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "ttdft_chfsi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ttdft_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ttdft_ctx_create", "ttdft_ham_upload", "ttdft_ham_apply", "ttdft_cheb_filter",
                 "ttdft_gram", "ttdft_orth", "ttdft_rr_eig", "ttdft_rotate",
                 "ttdft_lanczos_bounds", "ttdft_density", "ttdft_initial_density",
                 "ttdft_chfsi_step", "ttdft_fermi_occupancy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(E.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) == set(E.EXPORTED)


def test_host_rng_matches_oracle():
    assert np.array_equal(E.rng_uniform(99, 1000), oracle.rng_uniform(99, 1000))
    assert np.array_equal(E.rng_normal(99, 1001), oracle.rng_normal(99, 1001))


def test_synthetic_hamiltonian_structure():
    s = E.SynthHam(6, 7, 5, 2, 1, 2, 4, truncation=0.0)
    H = oracle.Ham.from_synth(s)
    D = H.dense()
    assert np.abs(D - D.T).max() < 1e-14
    # without truncation every in-range Kronecker-band entry is present
    def cnt(R, b):
        return np.array([min(R - 1, i + b) - max(0, i - b) + 1 for i in range(R)])
    cx, cy, cz = cnt(6, 2), cnt(7, 1), cnt(5, 2)
    per_row = (cz[:, None, None] * cy[None, :, None] * cx[None, None, :]).reshape(-1)
    assert np.array_equal(np.diff(s.rowptr), per_row)
    # columns ascending within each row
    for i in range(s.n):
        c = s.col[s.rowptr[i]:s.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)
    # truncation only removes entries
    t = E.SynthHam(6, 7, 5, 2, 1, 2, 4, truncation=1e-3)
    assert t.nnz < s.nnz


def test_host_errors_are_reported_not_fatal():
    try:
        E.fermi_occupancy(np.zeros(3), nelec=50)
    except E.TtdftError as e:
        assert e.status == E.EBADARG and "electrons" in str(e)
    else:
        raise AssertionError("expected TtdftError")
