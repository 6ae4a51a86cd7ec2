"""The assembly oracle (oracle/assembly_oracle.c) against independent forms.

Inputs are deliberately non-symmetric, so a transposed index (MatPotX[p][I][J]
read as [p][J][I], a swapped mode) fails the comparison.  Row and column order
is the basis order x + rx y + rx ry z, so the dense operator is a numpy kron
with the z factor outermost.
"""
import numpy as np

import oracle


def _inputs(rng, rx=4, ry=3, rz=5, rvx=2, rvy=3, rvz=2):
    m = lambda r: rng.standard_normal((r, r))
    X, DX, Y, DY, Z, DZ = m(rx), m(rx), m(ry), m(ry), m(rz), m(rz)
    PX = rng.standard_normal((rvx, rx, rx))
    PY = rng.standard_normal((rvy, ry, ry))
    PZ = rng.standard_normal((rvz, rz, rz))
    G = rng.standard_normal((rvx, rvy, rvz))
    return X, DX, Y, DY, Z, DZ, PX, PY, PZ, G


def _dense(rowptr, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    return A


def _formula(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G):
    """H[(I,J,K),(i,j,k)] from the defining sums, as an n x n matrix."""
    kin = 0.5 * (np.einsum("Ii,Jj,Kk->KJIkji", DX, Y, Z) + np.einsum("Ii,Jj,Kk->KJIkji", X, DY, Z)
                 + np.einsum("Ii,Jj,Kk->KJIkji", X, Y, DZ))
    pot = np.einsum("pqr,pIi,qJj,rKk->KJIkji", G, PX, PY, PZ)
    n = X.shape[0] * Y.shape[0] * Z.shape[0]
    return (kin + pot).reshape(n, n)


def test_hloc_matches_defining_formula():
    args = _inputs(np.random.default_rng(1))
    n = args[0].shape[0] * args[2].shape[0] * args[4].shape[0]
    rowptr, col, val = oracle.assemble_hloc(*args, 0.0)
    ref = _formula(*args)
    assert rowptr[-1] == np.count_nonzero(ref)
    assert np.abs(_dense(rowptr, col, val, n) - ref).max() < 1e-12 * np.abs(ref).max()


def test_hloc_kinetic_kronecker_closed_form():
    X, DX, Y, DY, Z, DZ, PX, PY, PZ, G = _inputs(np.random.default_rng(2))
    G[:] = 0.0
    n = X.shape[0] * Y.shape[0] * Z.shape[0]
    ref = 0.5 * (np.kron(Z, np.kron(Y, DX)) + np.kron(Z, np.kron(DY, X)) + np.kron(DZ, np.kron(Y, X)))
    rowptr, col, val = oracle.assemble_hloc(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, 0.0)
    assert np.abs(_dense(rowptr, col, val, n) - ref).max() < 1e-13 * np.abs(ref).max()


def test_hloc_single_core_entry_is_one_kronecker_product():
    X, DX, Y, DY, Z, DZ, PX, PY, PZ, G = _inputs(np.random.default_rng(3))
    for M in (X, DX, Y, DY, Z, DZ):
        M[:] = 0.0
    G[:] = 0.0
    G[1, 2, 0] = 1.7
    n = X.shape[0] * Y.shape[0] * Z.shape[0]
    ref = 1.7 * np.kron(PZ[0], np.kron(PY[2], PX[1]))
    rowptr, col, val = oracle.assemble_hloc(X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, 0.0)
    assert np.abs(_dense(rowptr, col, val, n) - ref).max() < 1e-13 * np.abs(ref).max()


def test_hloc_truncation_keeps_exactly_the_large_entries_in_column_order():
    args = _inputs(np.random.default_rng(4))
    ref = _formula(*args)
    trunc = 0.7 * np.median(np.abs(ref))
    rowptr, col, val = oracle.assemble_hloc(*args, trunc)
    n = ref.shape[0]
    for i in range(n):
        cols = col[rowptr[i]:rowptr[i + 1]]
        want = np.nonzero(np.abs(ref[i]) > trunc)[0]
        assert np.array_equal(cols, want)
        assert np.all(np.diff(cols) > 0)
    assert np.abs(val - ref[np.abs(ref) > trunc]).max() < 1e-12 * np.abs(ref).max()


def test_projector_rows_match_tucker_reconstruction():
    rng = np.random.default_rng(5)
    rx, ry, rz = 5, 4, 3
    ranks = [2, 3, 1]
    cores = [rng.standard_normal((r, r, r)) for r in ranks]
    mx = [rng.standard_normal((rx, r)) for r in ranks]
    my = [rng.standard_normal((ry, r)) for r in ranks]
    mz = [rng.standard_normal((rz, r)) for r in ranks]
    dense = np.array([np.einsum("abc,xa,yb,zc->zyx", S, A, B, C).reshape(-1)
                      for S, A, B, C in zip(cores, mx, my, mz)])
    trunc = 0.3 * np.median(np.abs(dense))
    rowptr, col, val = oracle.assemble_projectors(cores, mx, my, mz, trunc)
    for i in range(len(ranks)):
        want = np.nonzero(np.abs(dense[i]) > trunc)[0]
        assert np.array_equal(col[rowptr[i]:rowptr[i + 1]], want)
        assert np.allclose(val[rowptr[i]:rowptr[i + 1]], dense[i, want], rtol=1e-13, atol=0)


def test_hloc_row_range_is_a_slice_of_the_full_assembly():
    args = _inputs(np.random.default_rng(9), rx=5, ry=4, rz=3)
    rp, c, v = oracle.assemble_hloc(*args, 0.3)
    r0, r1 = 7, 41
    rp2, c2, v2 = oracle.assemble_hloc_rows(*args, 0.3, r0, r1)
    assert np.array_equal(rp2, rp[r0:r1 + 1] - rp[r0])
    assert np.array_equal(c2, c[rp[r0]:rp[r1]])
    assert np.array_equal(v2, v[rp[r0]:rp[r1]])
