"""Self-consistent field loop for the synthetic pseudopotential clusters, on
the device through the C ABI (main.cc:600-1512 restated for one GPU).

One SCF iteration, as main.cc orders it:
  1. rho_in from the Anderson mixer (computeRhoIn, main.cc:862-864);
  2. V_eff = V_loc + V_H[rho_in] + V_xc[rho_in] on the quadrature grid
     (KSDFTPotential: Tucker recompression of rho, kernel-expansion Hartree
     potential, LDA);
  3. Tucker recompression of V_eff, its 1-D matrices, and H = H_loc + C^T D C
     assembled in device memory (Create_Hloc / Create_Cnloc);
  4. one ChFSI step (filter, Cholesky-QR, Rayleigh-Ritz, rotation);
  5. Fermi level and occupancies, rho_out at the quadrature points;
  6. residual int (rho_out - rho_in)^2 and the mixer update (updateRho).

The system: atoms with valence charge Z, local pseudopotential
-Z erf(r / rc) / r and one normalised s-type Kleinman-Bylander projector
pi^(-3/4) exp(-r^2 / 2) of energy E_KB per atom.  The 1-D Tucker basis of
each axis is the lowest R eigenfunctions of a 1-D model Hamiltonian on the
finite-element mesh (Gaussian wells at the atoms' coordinates), orthonormal
in the FEM mass matrix.  Host work is set-up only: the 1-D generalised
eigenproblems (size = nodes per axis), the projector factors, and point
evaluation of the basis on a uniform grid.
"""
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg

from . import PKG_DIR
from . import engine as E

KB_HARTREE = 3.166811429e-06


@dataclass
class ScfConfig:
    box: float = 40.0          # cube edge (bohr), centred at the origin
    nel: int = 30              # finite elements per axis
    npe: int = 5               # nodes per element (quartic)
    inner_size: float = 14.0   # evenly meshed inner region (bohr)
    inner_h: float = 1.0       # its element size
    coarsening: float = 1.2    # growth of the outer elements
    quad: int = 10             # Gauss points per element
    rank: int = 20             # Tucker basis rank per axis
    nstates: int = 24
    degree: int = 20
    rho_rank: int = 30         # Tucker ranks of rho for the Hartree potential
    veff_rank: int = 30        # Tucker ranks of V_eff for H_loc
    temperature: float = 500.0
    alpha: float = 0.5
    history: int = 10
    tol: float = 1e-6          # int (rho_out - rho_in)^2
    max_iter: int = 50
    rc: float = 0.8            # local pseudopotential core radius
    kb_energy: float = 1.0
    well_depth: float = 1.0    # 1-D model wells for the basis (Ha)
    basis: str = "eigen"       # "eigen": lowest eigenvectors of the 1-D model; "windowed":
                               # compact, orthonormal, localized (H_loc a Kronecker band)
    basis_window: int = 20     # mesh nodes per windowed basis function
    basis_span: float = 23.0   # windows cover [-span, span] (bohr)
    truncation: float = 1e-10
    tucker_refine: int = 2     # STHOSVD from the previous iteration's factors (0: exact each time)
    restart_first: int = 10
    restart_other: int = 12
    occ_tol: float = 1e-3      # restart-loop stop on eig[occ] (main.cc:1130-1145 uses 1e-3)
    lda_variant: int = E.LDA_PZ81
    mixed: bool = False        # FP32 filter + split-BF16 GEMMs, then FP64 refinement
    refine_below: float = 1e-4  # residual at which a mixed run switches to FP64
    seed: int = 20261202
    # total energy: "harris" (Harris-Foulkes functional of rho_in, one Hartree
    # solve per iteration), "reference" (the reference's evaluation at
    # rho_out, main.cc:1368-1401, which needs V_H[rho_out]: a second solve), or
    # "both" (Harris drives `energies`, the rho_out form fills `energies_out`)
    energy_form: str = "harris"


@dataclass
class ScfResult:
    converged: bool
    iterations: int
    residuals: list
    energies: list
    energy: float
    fermi: float
    evals: np.ndarray
    seconds: float
    phase_seconds: dict = field(default_factory=dict)
    fell_back: bool = False
    probed: bool = False
    energies_out: list = field(default_factory=list)  # rho_out form, per iteration


def coulomb_expansion():
    """Exponents and weights of the 35-term exponential-sum expansion of 1/r
    shipped with the reference's example (exponents from alphak35_1e8 / A^2,
    weights from omegak35_1e8 / A)."""
    with open(os.path.join(PKG_DIR, "data", "coulomb_expansion_k35.json")) as fh:
        d = json.load(fh)
    a2 = d["Asquare"]
    return np.array(d["alphak35_1e8"]) / a2, np.array(d["omegak35_1e8"]) / math.sqrt(a2)


def cuboctahedron(spacing=5.4, jitter=0.1, seed=0):
    """13-atom cuboctahedral cluster (centre + 12 neighbours at the given
    nearest-neighbour distance), with seeded uniform jitter."""
    a = spacing / math.sqrt(2.0)
    pts = [(0.0, 0.0, 0.0)]
    for s1 in (-1, 1):
        for s2 in (-1, 1):
            pts += [(s1 * a, s2 * a, 0.0), (s1 * a, 0.0, s2 * a), (0.0, s1 * a, s2 * a)]
    xyz = np.array(pts)
    return xyz + np.random.default_rng(seed).uniform(-jitter, jitter, xyz.shape)


def fcc_cluster(natoms, lattice=7.6, jitter=0.1, seed=0):
    """The natoms FCC sites nearest the origin (ties broken by a fixed order),
    with seeded uniform jitter."""
    h = lattice / 2.0
    n = int(math.ceil((natoms * 2.0) ** (1.0 / 3.0))) + 2
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    g = g[(g.sum(axis=1) % 2) == 0] * h
    d = np.round((g ** 2).sum(axis=1), 9)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], d))
    xyz = g[order[:natoms]].astype(np.float64)
    return xyz + np.random.default_rng(seed).uniform(-jitter, jitter, xyz.shape)


def _dev(ctx, a):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def eval_nodal(nodes, npe, values, points):
    """Piecewise-Lagrange functions (rows of values, at the mesh nodes) at
    arbitrary points (host; set-up only)."""
    nel = (len(nodes) - 1) // (npe - 1)
    ends = nodes[::npe - 1]
    e = np.clip(np.searchsorted(ends, points, side="right") - 1, 0, nel - 1)
    a, b = ends[e], ends[e + 1]
    xi = 2.0 * (points - a) / (b - a) - 1.0
    xs = np.linspace(-1.0, 1.0, npe)
    out = np.zeros((values.shape[0], len(points)))
    for i in range(npe):
        li = np.ones_like(xi)
        for k in range(npe):
            if k != i:
                li *= (xi - xs[k]) / (xs[i] - xs[k])
        out += values[:, e * (npe - 1) + i] * li
    return out


def windowed_basis(nodes, M, H1, rank, window, span):
    """Localized 1-D basis with compact support, orthonormal in the mass
    matrix M (nodal coefficients, rank x nodes).

    Function k lives on a window of `window` consecutive mesh nodes; the
    windows start at evenly spaced nodes over [-span, span] and are taken left
    to right.  In its window, function k is the lowest-energy function (1-D
    model Hamiltonian H1, generalised eigenproblem with M) that is
    M-orthogonal to the earlier functions whose windows reach into it.
    Functions whose windows do not meet share no element, so they are
    orthogonal without constraint: the basis is exactly orthonormal, and two
    functions couple in any 1-D overlap, kinetic or potential matrix only
    when their windows meet.  The Tucker-product H_loc is then a Kronecker
    band whose half-width is the number of neighbouring windows reached.

    This is the localization of basis/BasisLocalization.cc:110-168, 247-308
    (per-centre lowest eigenvectors of a localizing operator, functions
    ordered by centre) restated for exact compact support: the reference
    localizes by a |x - R|^p penalty over the whole axis, and its
    localized functions are neither compact nor mutually orthogonal."""
    nn = len(nodes)
    interior = np.arange(1, nn - 1)  # zero Dirichlet ends
    cand = interior[(nodes[interior] > -span) & (nodes[interior] < span)]
    if len(cand) < window + rank // 2:
        raise ValueError("windowed basis: too few mesh nodes inside the span")
    starts = np.round(np.linspace(cand[0], cand[-1] - window + 1, rank)).astype(int)
    # how far (in nodes) the mass matrix couples: nodes of one element
    reach = max(int(np.nonzero(M[i])[0].max()) - i for i in range(nn))
    phi = np.zeros((nn, rank))
    for k, s in enumerate(starts):
        idx = np.arange(s, s + window)
        prev = [i for i in range(k) if starts[i] + window - 1 + reach >= s]
        A, Hw = M[np.ix_(idx, idx)], H1[np.ix_(idx, idx)]
        if prev:
            Z = scipy.linalg.null_space((M[idx, :] @ phi[:, prev]).T)
        else:
            Z = np.eye(window)
        _, v = scipy.linalg.eigh(Z.T @ Hw @ Z, Z.T @ A @ Z, subset_by_index=[0, 0])
        c = Z @ v[:, 0]
        phi[idx, k] = c / math.sqrt(c @ A @ c)
    return phi.T


class Axis:
    """Mesh, quadrature, 1-D basis and its matrices for one axis."""

    def __init__(self, ctx, cfg, centers):
        self.ctx = ctx
        half = cfg.box / 2.0
        self.nodes = E.fem_mesh("hartree_outer", cfg.nel, cfg.npe, -half, half, cfg.inner_size,
                                cfg.inner_h, cfg.coarsening)
        self.xi, self.w = E.quadrature_rule(cfg.quad)
        nn = len(self.nodes)
        self.np = cfg.nel * cfg.quad
        # shape functions of every node at the quadrature points (device)
        d_eye = _dev(ctx, np.eye(nn))
        d_bn = E.DeviceArray(ctx, (nn * self.np,))
        d_dbn = E.DeviceArray(ctx, (nn * self.np,))
        self.jw, self.inv_jw, self.xq = ctx.fem_interpolate(cfg.nel, cfg.npe, self.xi, self.w,
                                                            self.nodes, nn, d_eye.ptr, d_bn.ptr,
                                                            d_dbn.ptr)
        self.d_jw, self.d_inv_jw, self.d_xq = (_dev(ctx, self.jw), _dev(ctx, self.inv_jw),
                                               _dev(ctx, self.xq))
        # 1-D model Hamiltonian: 1/2 K + V1, V1 = Gaussian wells at the atoms
        v1 = -cfg.well_depth * np.exp(-0.5 * (self.xq[:, None] - centers[None, :]) ** 2).sum(1)
        d_v1 = _dev(ctx, v1)
        d_m, d_k, d_p = (E.DeviceArray(ctx, (nn * nn,)) for _ in range(3))
        ctx.tucker_basis_matrices(self.np, nn, d_bn.ptr, d_dbn.ptr, self.d_jw.ptr,
                                  self.d_inv_jw.ptr, 1, d_v1.ptr, d_m.ptr, d_k.ptr, d_p.ptr)
        M, K, P = (d.to_host().reshape(nn, nn) for d in (d_m, d_k, d_p))
        if cfg.basis == "windowed":
            self.basis = windowed_basis(self.nodes, M, 0.5 * K + P, cfg.rank, cfg.basis_window,
                                        cfg.basis_span)
        else:
            inner = slice(1, nn - 1)  # zero Dirichlet ends
            lam, c = scipy.linalg.eigh(0.5 * K[inner, inner] + P[inner, inner], M[inner, inner],
                                       subset_by_index=[0, cfg.rank - 1])
            self.basis = np.zeros((cfg.rank, nn))
            self.basis[:, 1:-1] = c.T
        # basis values and derivatives at the quadrature points (device)
        d_nodal = _dev(ctx, self.basis)
        self.d_B = E.DeviceArray(ctx, (cfg.rank * self.np,))
        self.d_DB = E.DeviceArray(ctx, (cfg.rank * self.np,))
        ctx.fem_interpolate(cfg.nel, cfg.npe, self.xi, self.w, self.nodes, cfg.rank, d_nodal.ptr,
                            self.d_B.ptr, self.d_DB.ptr)
        d_mx, d_mdx = E.DeviceArray(ctx, (cfg.rank ** 2,)), E.DeviceArray(ctx, (cfg.rank ** 2,))
        ctx.tucker_basis_matrices(self.np, cfg.rank, self.d_B.ptr, self.d_DB.ptr, self.d_jw.ptr,
                                  self.d_inv_jw.ptr, 0, None, d_mx.ptr, d_mdx.ptr, None)
        self.MatX = d_mx.to_host().reshape(cfg.rank, cfg.rank)
        self.MatDX = d_mdx.to_host().reshape(cfg.rank, cfg.rank)
        self.B = self.d_B.to_host().reshape(cfg.rank, self.np)
        self.cfg = cfg

    def mat_pot(self, d_U, rv):
        """MatPot (rv, R, R) of the V_eff factor d_U (np x rv column-major)."""
        r = self.cfg.rank
        if getattr(self, "_mp_rank", None) != rv:
            self._d_mp = E.DeviceArray(self.ctx, (rv * r * r,))
            self._mp_rank = rv
        d_mp = self._d_mp
        self.ctx.tucker_basis_matrices(self.np, r, self.d_B.ptr, self.d_DB.ptr, self.d_jw.ptr,
                                       self.d_inv_jw.ptr, rv, d_U.ptr, None, None, d_mp.ptr)
        return d_mp.to_host().reshape(rv, r, r)

    def projector_factor(self, x0):
        """sum_q B_i(x_q) pi^(-1/4) exp(-(x_q - x0)^2 / 2) jw_q for every basis function."""
        g = math.pi ** -0.25 * np.exp(-0.5 * (self.xq - x0) ** 2)
        return self.B @ (g * self.jw)

    def at_points(self, points):
        return eval_nodal(self.nodes, self.cfg.npe, self.basis, points)


class Scf:
    """SCF of one cluster (xyz: natoms x 3, bohr; zval per atom) on one device."""

    def __init__(self, ctx, xyz, zval=3.0, cfg=None):
        self.ctx = ctx
        self.cfg = cfg = cfg or ScfConfig()
        self.xyz = np.asarray(xyz, dtype=np.float64)
        self.natoms = len(self.xyz)
        self.zval = np.full(self.natoms, float(zval))
        self.nelec = int(round(self.zval.sum()))
        t0 = time.perf_counter()
        self.axes = [Axis(ctx, cfg, self.xyz[:, d]) for d in range(3)]
        self.p = tuple(a.np for a in self.axes)
        self.ngrid = int(np.prod(self.p))
        g = self.p
        self.d_vloc = E.DeviceArray(ctx, (self.ngrid,))
        atoms = np.column_stack([self.xyz, self.zval])
        ctx.local_pseudopotential(atoms, cfg.rc, g, *(a.d_xq.ptr for a in self.axes),
                                  self.d_vloc.ptr)
        self.d_jw3 = E.DeviceArray(ctx, (self.ngrid,))
        ctx.grid_weights(g, *(a.d_jw.ptr for a in self.axes), self.d_jw3.ptr)
        self.d_one = E.DeviceArray(ctx, (self.ngrid,))
        ctx.fill_uniform(self.d_one.ptr, self.ngrid, 0, 1.0, 1.0)
        # projectors: rank-1 Tucker rows, D = E_KB
        self.proj = ([np.ones((1, 1, 1))] * self.natoms,
                     [self.axes[0].projector_factor(x)[:, None] for x in self.xyz[:, 0]],
                     [self.axes[1].projector_factor(y)[:, None] for y in self.xyz[:, 1]],
                     [self.axes[2].projector_factor(z)[:, None] for z in self.xyz[:, 2]])
        self.inv_c_lm = np.full(self.natoms, cfg.kb_energy)
        self.alpha, self.omega = coulomb_expansion()
        d = self.xyz[:, None, :] - self.xyz[None, :, :]
        rr = np.sqrt((d ** 2).sum(-1)) + np.eye(self.natoms)
        self.e_nn = 0.5 * float(((self.zval[:, None] * self.zval[None, :]) / rr)
                                [~np.eye(self.natoms, dtype=bool)].sum())
        self.setup_seconds = time.perf_counter() - t0

    # ---- densities ---------------------------------------------------------------
    def superposition(self, d_rho, width=1.0, points=None, n=None):
        """Atomic superposition start: a Gaussian of the valence charge per atom
        (the radial table is its spline, as InitialGuess reads Density_AT*),
        renormalised to the electron count on the quadrature grid
        (main.cc:674-680).  With points (device 1-D coordinates, n per axis)
        the same renormalised superposition is evaluated there instead.
        Returns the renormalisation factor."""
        r = np.linspace(0.0, 12.0, 481)
        rho_r = self.zval[0] * (math.pi * width ** 2) ** -1.5 * np.exp(-(r / width) ** 2)
        ax = [a.d_xq.ptr for a in self.axes]
        if points is None:
            self.ctx.initial_density(self.xyz, r, rho_r, *ax, self.p, d_rho.ptr)
            return self.renormalise(d_rho)
        tmp = E.DeviceArray(self.ctx, (self.ngrid,))
        self.ctx.initial_density(self.xyz, r, rho_r, *ax, self.p, tmp.ptr)
        f = self.nelec / self.ctx.grid_integral(tmp.ptr, self.p, *(a.d_jw.ptr for a in self.axes))
        self.ctx.initial_density(self.xyz, r, rho_r, points.ptr, points.ptr, points.ptr,
                                 (n, n, n), d_rho.ptr)
        self.ctx.grid_scale(d_rho.ptr, n ** 3, f)
        return f

    def renormalise(self, d_rho):
        q = self.ctx.grid_integral(d_rho.ptr, self.p, *(a.d_jw.ptr for a in self.axes))
        self.ctx.grid_scale(d_rho.ptr, self.ngrid, self.nelec / q)
        return self.nelec / q

    def integral(self, d_f, d_g=None):
        w = [a.d_jw.ptr for a in self.axes]
        if d_g is None:
            return self.ctx.grid_integral(d_f.ptr, self.p, *w)
        return self.ctx.grid_dot(d_f.ptr, d_g.ptr, self.p, *w)

    # ---- potential -----------------------------------------------------------------
    def tucker(self, d_field, rank, key, refine=False):
        """STHOSVD of a grid field into buffers kept per field (reused every
        iteration); with refine, from the field's previous factors
        (ttdft_sthosvd_refine, cfg.tucker_refine iterations)."""
        if not hasattr(self, "_tucker_bufs"):
            self._tucker_bufs = {}
        if key not in self._tucker_bufs:
            self._tucker_bufs[key] = (E.DeviceArray(self.ctx, (rank ** 3,)),
                                      [E.DeviceArray(self.ctx, (n * rank,)) for n in self.p])
            refine = False
        d_core, d_U = self._tucker_bufs[key]
        if refine and self.cfg.tucker_refine > 0:
            self.ctx.sthosvd_refine(d_field.ptr, self.p, (rank,) * 3, d_core.ptr,
                                    [u.ptr for u in d_U], self.cfg.tucker_refine)
        else:
            self.ctx.sthosvd(d_field.ptr, self.p, (rank,) * 3, d_core.ptr, [u.ptr for u in d_U])
        return d_core, d_U

    def effective_potential(self, d_rho, d_veff, d_vh, d_vxc, d_exc, refine=False):
        ctx, cfg = self.ctx, self.cfg
        core, U = self.tucker(d_rho, cfg.rho_rank, "rho", refine)
        ax = self.axes
        ctx.hartree_expansion((cfg.rho_rank,) * 3, core.ptr, [u.ptr for u in U],
                              [a.d_xq.ptr for a in ax], [a.d_jw.ptr for a in ax], self.p,
                              [a.d_xq.ptr for a in ax], self.p, self.alpha, self.omega, d_vh.ptr)
        ctx.lda_xc(d_rho.ptr, self.ngrid, d_vxc.ptr, d_exc.ptr, variant=cfg.lda_variant)
        ctx.copy(d_veff.ptr, self.d_vloc.ptr, 8 * self.ngrid)
        for d_src in (d_vh, d_vxc):
            ctx.grid_axpby(d_veff.ptr, 1.0, d_src.ptr, 1.0, self.ngrid)

    def energy_at_output(self, d_out, d_vh_in, d_vxc_in, e_band, refine=False):
        """The reference's total energy (main.cc:1368-1401): with rho_out and
        its Hartree potential,
          E = int[eps_xc(rho_out) rho_out + 1/2 rho_out V_H[rho_out]
                  - rho_out (V_eff,in - V_loc)] + E_nn + E_band,
        V_eff,in - V_loc = V_H[rho_in] + V_xc[rho_in].  V_H[rho_out] comes
        from the same Tucker compression and kernel expansion as V_H[rho_in]
        (main.cc:1288-1302), with factors kept for rho_out."""
        ctx, cfg = self.ctx, self.cfg
        if not hasattr(self, "_out_bufs"):
            self._out_bufs = [E.DeviceArray(ctx, (self.ngrid,)) for _ in range(3)]
        d_vh, d_vxc, d_exc = self._out_bufs
        core, U = self.tucker(d_out, cfg.rho_rank, "rho_out", refine)
        ax = self.axes
        ctx.hartree_expansion((cfg.rho_rank,) * 3, core.ptr, [u.ptr for u in U],
                              [a.d_xq.ptr for a in ax], [a.d_jw.ptr for a in ax], self.p,
                              [a.d_xq.ptr for a in ax], self.p, self.alpha, self.omega, d_vh.ptr)
        ctx.lda_xc(d_out.ptr, self.ngrid, d_vxc.ptr, d_exc.ptr, variant=cfg.lda_variant)
        return (self.integral(d_out, d_exc) + 0.5 * self.integral(d_out, d_vh)
                - self.integral(d_out, d_vh_in) - self.integral(d_out, d_vxc_in)
                + self.e_nn + e_band)

    def hamiltonian(self, d_veff, copies=E.HAM_F64, refine=False):
        cfg = self.cfg
        rv = cfg.veff_rank
        core, U = self.tucker(d_veff, rv, "veff", refine)
        PX, PY, PZ = (a.mat_pot(u, rv) for a, u in zip(self.axes, U))
        G = core.to_host().reshape(rv, rv, rv).transpose(2, 1, 0)  # [a, b, c]
        a = self.axes
        hloc = (a[0].MatX, a[0].MatDX, a[1].MatX, a[1].MatDX, a[2].MatX, a[2].MatDX, PX, PY, PZ, G)
        return self.ctx.ham_assemble_tucker(hloc, cfg.truncation, self.proj, 0.0, self.inv_c_lm,
                                            copies)

    # ---- the loop ------------------------------------------------------------------
    def run(self, d_rho_start=None, callback=None, fallback_residual=None):
        """SCF from d_rho_start (default: the atomic superposition).  With
        fallback_residual, a first residual above it restarts the loop from
        the superposition (the wasted iteration is counted).  fallback_residual
        may be a callable: given the first residual, it returns the bar, or
        None for no fallback."""
        ctx, cfg = self.ctx, self.cfg
        n, k, R = self.ngrid, cfg.nstates, cfg.rank
        N = R ** 3
        t_start = time.perf_counter()
        phases = {k_: 0.0 for k_ in ("potential", "hamiltonian", "chfsi", "density", "mixing")}
        d_in = E.DeviceArray(ctx, (n,))
        if d_rho_start is None:
            self.superposition(d_in)
        else:
            ctx.copy(d_in.ptr, d_rho_start.ptr, 8 * n)
        d_out, d_veff, d_vh, d_vxc, d_exc = (E.DeviceArray(ctx, (n,)) for _ in range(5))
        mixer = E.Mixer(ctx, n, 0, self.d_jw3.ptr, self.d_one.ptr, cfg.alpha, cfg.history)
        dX = E.DeviceArray(ctx, (N * k,), np.float64)
        ctx.random_subspace(dX.ptr, k, N, k, seed=cfg.seed, dtype=E.F64)
        dt = E.F64
        if cfg.mixed:  # the same start block, rounded to FP32
            dX32 = E.DeviceArray(ctx, (N * k,), np.float32)
            E._check(E.lib().ttdft_block_convert(ctx.h, dX.ptr, k, E.F64, dX32.ptr, k, E.F32, N,
                                                 k))
            dt = E.F32
        ev = np.zeros(k)
        st = E.new_state()
        st.occ_tol = cfg.occ_tol
        occ_index = min(self.nelec // 2 + 10, k - 1)
        self.fp32_iterations = 0
        self.chfsi_passes = []  # filter passes of every iteration's ChFSI step
        residuals, energies, energies_out = [], [], []
        converged, ef = False, 0.0
        it = 0
        for it in range(cfg.max_iter):
            t = time.perf_counter()
            mixer.compute_rho_in(it, None, d_in.ptr)
            self.effective_potential(d_in, d_veff, d_vh, d_vxc, d_exc, refine=it > 0)
            ctx.synchronize()
            phases["potential"] += time.perf_counter() - t
            t = time.perf_counter()
            if dt == E.F32 and residuals and residuals[-1] < cfg.refine_below:
                # FP64 refinement from here on (the FP32 block widened to FP64)
                E._check(E.lib().ttdft_block_convert(ctx.h, dX32.ptr, k, E.F32, dX.ptr, k, E.F64,
                                                     N, k))
                dt = E.F64
            # the Lanczos bounds always run on the FP64 copy
            ham = self.hamiltonian(d_veff, E.HAM_F64 if dt == E.F64 else E.HAM_F64 | E.HAM_F32,
                                   refine=it > 0)
            phases["hamiltonian"] += time.perf_counter() - t
            t = time.perf_counter()
            dXc = dX if dt == E.F64 else dX32
            ctx.chfsi_step(ham, dXc.ptr, k, k, cfg.degree, ev, st, it, occ_index,
                           cfg.restart_first if it == 0 else cfg.restart_other,
                           precision=dt)
            self.fp32_iterations += dt == E.F32
            self.chfsi_passes.append(st.passes)
            ham.free()
            phases["chfsi"] += time.perf_counter() - t
            t = time.perf_counter()
            ef, occ = E.fermi_occupancy(ev, self.nelec, KB_HARTREE, cfg.temperature)
            ctx.density(dXc.ptr, k, k, occ, (R, R, R), *(a.d_B.ptr for a in self.axes), self.p,
                        d_out.ptr, dtype=dt)
            e_band = float(E.band_energy(ev, occ))
            self.energy_out = None
            if cfg.energy_form in ("reference", "both"):
                self.energy_out = self.energy_at_output(d_out, d_vh, d_vxc, e_band,
                                                        refine=it > 0)
                energies_out.append(self.energy_out)
            if cfg.energy_form == "reference":
                energy = self.energy_out
            else:
                # Harris-Foulkes energy of the input density
                e_h = 0.5 * self.integral(d_in, d_vh)
                e_xc = self.integral(d_in, d_exc)
                e_dc = self.integral(d_in, d_vh) + self.integral(d_in, d_vxc)
                energy = e_band - e_dc + e_h + e_xc + self.e_nn
            energies.append(energy)
            res = ctx.grid_residual(d_out.ptr, d_in.ptr, self.p,
                                    *(a.d_jw.ptr for a in self.axes))
            residuals.append(res)
            phases["density"] += time.perf_counter() - t
            t = time.perf_counter()
            mixer.update(None, d_in.ptr, None, d_out.ptr, it)
            phases["mixing"] += time.perf_counter() - t
            if callback:
                callback(it, res, energy)
            bar = None
            if it == 0 and fallback_residual is not None:
                bar = fallback_residual(res) if callable(fallback_residual) else fallback_residual
            if bar is not None and res > bar:
                r = self.run(None, callback)
                r.iterations += 1
                r.residuals = residuals + r.residuals
                r.energies = energies + r.energies
                r.energies_out = energies_out + r.energies_out
                r.seconds = time.perf_counter() - t_start
                r.fell_back = True
                return r
            if res < cfg.tol:
                converged = True
                break
        ctx.synchronize()
        self.d_rho = d_out
        self.X, self.X_dtype, self.occ = dXc, dt, occ
        return ScfResult(converged, it + 1, residuals, energies, energies[-1], ef, ev.copy(),
                         time.perf_counter() - t_start, phases, energies_out=energies_out)

    def density_on_uniform(self, n, d_out=None, box=None):
        """rho of the last iteration's states on an n^3 uniform grid spanning a
        cube centred on the origin, by default the SCF box (point i at
        -box/2 + i box/(n-1))."""
        box = self.cfg.box if box is None else box
        pts = np.linspace(-box / 2, box / 2, n)
        R, k = self.cfg.rank, self.cfg.nstates
        d_b = [_dev(self.ctx, a.at_points(pts)) for a in self.axes]
        d_out = d_out or E.DeviceArray(self.ctx, (n ** 3,))
        self.ctx.density(self.X.ptr, k, k, self.occ, (R, R, R), *(b.ptr for b in d_b),
                         (n, n, n), d_out.ptr, dtype=self.X_dtype)
        return d_out
