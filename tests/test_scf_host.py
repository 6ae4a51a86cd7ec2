"""Host-side pieces of the SCF driver and the learned start (no device work):
cluster generators, piecewise-Lagrange point evaluation against the oracle's
finite-element interpolation, the Coulomb expansion data, the network's
parameter layout and the fallback probe's bar."""
import math

import numpy as np

import oracle
from oracle import unet_oracle as U
from ttdft_public_amd import learned as Ln
from ttdft_public_amd import scf as S


def test_cuboctahedron_geometry():
    xyz = S.cuboctahedron(5.4, 0.0)
    assert xyz.shape == (13, 3)
    d = np.linalg.norm(xyz[1:], axis=1)
    assert np.allclose(d, 5.4)  # twelve neighbours at the nearest-neighbour distance
    dd = np.linalg.norm(xyz[1:, None] - xyz[None, 1:], axis=-1)
    assert np.isclose(np.min(dd[dd > 0]), 5.4)  # and 5.4 apart among themselves
    j = S.cuboctahedron(5.4, 0.1, seed=3) - xyz
    assert np.abs(j).max() <= 0.1 and np.abs(j).max() > 0.05


def test_fcc_cluster():
    xyz = S.fcc_cluster(172, lattice=7.6, jitter=0.0)
    assert xyz.shape == (172, 3)
    assert len({tuple(p) for p in np.round(xyz, 9)}) == 172
    dd = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    assert np.isclose(np.min(dd[dd > 0]), 7.6 / math.sqrt(2))
    assert np.allclose(xyz[0], 0.0)


def test_random_cluster_is_seeded():
    a = Ln.random_cluster(np.random.default_rng(4))
    b = Ln.random_cluster(np.random.default_rng(4))
    assert np.array_equal(a, b) and 13 <= len(a) <= 19
    assert np.abs(a.mean(axis=0)).max() < 0.31  # centred before the jitter


def test_eval_nodal_matches_fem_interpolation():
    """Point evaluation of nodal functions at the quadrature points equals the
    oracle's per-element interpolation."""
    nodes = oracle.fem_mesh("hartree_outer", 12, 5, -20.0, 20.0, 6.0, 1.0, 1.3)
    vals = np.random.default_rng(0).standard_normal((3, len(nodes)))
    xi, w = np.polynomial.legendre.leggauss(6)
    v0, _, _, _, xq = oracle.fem_interpolate(12, 5, xi, w, nodes, vals)
    assert np.abs(S.eval_nodal(nodes, 5, vals, xq) - v0).max() < 1e-12


def test_coulomb_expansion_fits_one_over_r():
    a, w = S.coulomb_expansion()
    assert len(a) == len(w) == 35
    r = np.geomspace(0.05, 50.0, 200)
    approx = (w[None, :] * np.exp(-a[None, :] * r[:, None] ** 2)).sum(1)
    assert np.abs(approx * r - 1).max() < 1e-4


def test_network_layout_matches_oracle():
    assert Ln.layer_specs() == U.layer_specs()
    n = sum(co * ci * ks ** 3 + co for _, ci, co, ks in Ln.layer_specs())
    assert n == 1460017


def test_probe_bar_from_report(tmp_path):
    """The lazy-probe bar is a quarter of the smallest superposition first
    residual in the evaluation report, below every learned first residual
    that report shows; a missing or malformed report gives None (always
    probe)."""
    bar = Ln.default_probe_below()
    import json
    import os
    rep = json.load(open(os.path.join(os.path.dirname(Ln.DEFAULT_WEIGHTS),
                                      "unet_initial_density_report.json")))
    sup = [c["superposition"]["first_residual"] for c in rep["test"]]
    lrn = [c["learned"]["first_residual"] for c in rep["test"]]
    assert bar == 0.25 * min(sup)
    assert max(lrn) < bar < min(sup)
    p = tmp_path / "r.json"
    p.write_text(json.dumps({"test": [{"superposition": {"first_residual": 8.0}},
                                      {"superposition": {"first_residual": 4.0}}]}))
    assert Ln.default_probe_below(str(p)) == 1.0
    p.write_text("{}")
    assert Ln.default_probe_below(str(p)) is None
    assert Ln.default_probe_below(str(tmp_path / "absent.json")) is None


def test_windowed_basis_is_compact_orthonormal_and_banded():
    """The windowed localized basis (172-atom shape: 80-bohr box, 60 quartic
    elements, 50 functions over [-23, 23] bohr, windows of 20 nodes): exactly
    orthonormal in the FEM mass matrix, every function zero outside its
    window, and every 1-D matrix (overlap, kinetic, potential) zero beyond
    the neighbours whose windows meet -- the band that makes H_loc a
    Kronecker band."""
    import scipy.linalg

    xyz = S.fcc_cluster(172, lattice=7.6, jitter=0.1, seed=20261202)
    nodes = oracle.fem_mesh("hartree_outer", 60, 5, -40.0, 40.0, 34.0, 1.0, 1.2)
    xi, w = np.polynomial.legendre.leggauss(10)
    nn = len(nodes)
    Bn, DBn, jw, ijw, xq = oracle.fem_interpolate(60, 5, xi, w, nodes, np.eye(nn))
    v1 = -0.1 * np.exp(-0.5 * (xq[:, None] - xyz[:, 0][None, :]) ** 2).sum(1)
    M, K, P = oracle.tucker_basis_matrices(Bn, DBn, jw, ijw, v1[:, None])
    b = S.windowed_basis(nodes, M, 0.5 * K + P[0], 50, 20, 23.0)
    assert b.shape == (50, nn)
    assert np.abs(b @ M @ b.T - np.eye(50)).max() < 1e-12
    support = [np.nonzero(r)[0] for r in b]
    assert all(len(s) <= 20 and s[-1] - s[0] < 20 for s in support)
    assert all(s[0] >= 1 and s[-1] <= nn - 2 for s in support)  # Dirichlet ends
    assert all(support[i][0] <= support[i + 1][0] for i in range(49))  # ordered
    T, V = b @ K @ b.T, b @ P[0] @ b.T
    d = np.abs(np.subtract.outer(np.arange(50), np.arange(50)))
    for A in (T, V):
        assert np.abs(A[d > 6]).max() == 0.0
    # the lowest state of the 1-D model is within the basis' reach
    lam = scipy.linalg.eigh(0.5 * K[1:-1, 1:-1] + P[0][1:-1, 1:-1], M[1:-1, 1:-1],
                            eigvals_only=True)[0]
    assert np.linalg.eigvalsh(0.5 * T + V)[0] - lam < 0.25 * abs(lam)
