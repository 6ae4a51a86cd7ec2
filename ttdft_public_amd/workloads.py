"""Synthetic workloads of the reference configurations, shared by the
benchmark and the tests.

The reference ships inputs only for its 13-atom example; the 172- and
500-atom shapes exist as sizes (basis ranks, states, degree, density grid)
and these seeded stand-ins reproduce them:

* ``CLUSTER172`` -- Tucker basis 50 x 50 x 50 (N = 125,000), Kronecker
  half-bandwidths (6, 6, 7): up to 2,535 nonzeros per row of H_loc after
  truncation at 1e-6, 172 s-type projectors, 300 states, Chebyshev degree 30,
  516 electrons, density on 60 elements x 10 Gauss points = 600 points per
  axis of an 80-bohr cube.
* ``CLUSTER500`` -- Tucker basis 70 x 70 x 70 (N = 343,000), half-bandwidths
  (7, 7, 7) (about 2,900 nonzeros per row), 500 projectors, 900 states,
  degree 40, 1,500 electrons.

``synth_hamiltonian`` builds the seeded H (``ttdft_synth_hamiltonian``: a
sinc-DVR kinetic stencil, a shallow well with a few hundred bound states and
a banded coupling filling the Kronecker band).  ``basis_tables`` gives 1-D
basis functions at the Gauss points, orthonormal under the quadrature as the
Tucker basis is, so the density of an orthonormal block integrates to
sum_i 2 f_i = N_e.
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Workload:
    dims: tuple
    bands: tuple
    nproj: int
    states: int
    degree: int
    nelec: int
    quad: int          # density points per axis (elements x 10 Gauss points)
    box: float         # cube edge (bohr)
    h: float           # grid spacing of the synthetic kinetic stencil (bohr)
    seed: int

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    @property
    def natoms(self):
        return self.nelec // 3  # Z = 3 valence electrons per atom

    def occ_index(self):
        """eig[min(Ne/2 + 10, k - 1)]: the state whose change stops the
        restart loop (main.cc:192, 1130-1145)."""
        return min(self.nelec // 2 + 10, self.states - 1)


# h = 0.559 bohr puts the top of the spectrum near 48 Ha, the Lanczos upper
# bound the reference prints for its example run
CLUSTER172 = Workload(dims=(50, 50, 50), bands=(6, 6, 7), nproj=172, states=300, degree=30,
                      nelec=516, quad=600, box=80.0, h=0.559, seed=20261202)
CLUSTER500 = Workload(dims=(70, 70, 70), bands=(7, 7, 7), nproj=500, states=900, degree=40,
                      nelec=1500, quad=850, box=110.0, h=2.0, seed=4)


def synth_hamiltonian(w: Workload):
    from . import engine as E

    return E.SynthHam(*w.dims, *w.bands, w.nproj, h=w.h, truncation=1e-6, seed=w.seed)


def basis_tables(R, P, box):
    """R localized smooth 1-D functions sampled at the P Gauss points (10 per
    element) of [-box/2, box/2], orthonormal under the quadrature.  Returns
    (points, weights, B) with B column-major P x R (B[:, j] = function j)."""
    x, w = np.polynomial.legendre.leggauss(10)
    half = 0.5 * box
    edges = np.linspace(-half, half, P // 10 + 1)
    pts = np.concatenate([0.5 * (b - a) * x + 0.5 * (b + a) for a, b in zip(edges[:-1], edges[1:])])
    wts = np.concatenate([0.5 * (b - a) * w for a, b in zip(edges[:-1], edges[1:])])
    centers = np.linspace(-(half - 2.0), half - 2.0, R)
    width = 1.2 * (centers[1] - centers[0])
    B = np.exp(-((pts[:, None] - centers[None, :]) ** 2) / (2 * width**2))
    Q, _ = np.linalg.qr(np.sqrt(wts)[:, None] * B)
    return pts, wts, np.asfortranarray(Q / np.sqrt(wts)[:, None])
