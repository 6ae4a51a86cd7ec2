"""The reference's example deck run end to end on the device, through the C ABI.

The deck (example/ of the reference: input.inp, shell1, basis_{x,y,z}.txt,
locPotential_AT0, nlpV_AT0, nlpWaveFun_AT0, Density_AT0) is a 13-atom cluster
of valence-3 atoms with tabulated pseudopotentials (lmax = 2: s and p
projectors), a fixed Tucker basis of rank 40 per axis read from files, and
main.cc's SCF settings.  One iteration follows main.cc:846-1449:

  1. rho_in on the quadrature grid and at the mesh nodes from the Anderson
     mixer (both fields carry the same coefficients, AndersonMixing.cc);
  2. V_H from the nodal rho_in: STHOSVD at the nodes (rank 40), factors
     interpolated onto the 14-point electrostatic rule, kernel expansion onto
     the quadrature grid (TensorUtils::computeSTHOSVDonQuadMPI,
     KSDFTPotential::computeHartreePotNew); V_xc by LDA with the reference's
     1/33 (KSDFTPotential.cc:262); V_eff = V_H + V_xc + V_loc;
  3. STHOSVD of V_eff (rank 50), its 1-D matrices, H = H_loc + C^T D C
     assembled in device memory;
  4. one ChFSI step; Fermi level and occupancies;
  5. rho_out at the quadrature points and the nodes, the total energy at
     rho_out (main.cc:1368-1401) with V_H[rho_out] from the nodal rho_out;
  6. residual int (rho_out - rho_in)^2 and the mixer update.

Set-up (main.cc:440-592, 674-680), for every atom type with its own tables:
V_loc by ttdft_psp_local; the projectors of each atom from the Tucker form
(rank rankNloc) of its type's dV_l phi_lm on the nonlocal mesh, contracted with the basis functions evaluated at the atom's
shifted nonlocal quadrature points (Create_Cnloc,
ProjectHamiltonianSparse.cc:83-328), D = 1 / C_lm; the initial density from
Density_AT0 renormalised to the electron count.

Differences from the reference that change numbers slightly: rho_out is exact
at the quadrature points (the reference recompresses it by STHOSVD with a
1e-4 tolerance), C_lm is the exact grid sum, and the H_loc / C_nloc
truncations of the recorded run (command-line arguments) are not known.
"""
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import engine as E
from .scf import KB_HARTREE, coulomb_expansion, eval_nodal


@dataclass
class DeckConfig:
    """example/input.inp (values of the recorded run)."""
    nel: int = 40                  # numEleX/Y/Z
    npe: int = 5                   # eleType: quartic elements
    start: float = -16.0           # domainStart / domainElectroStart
    end: float = 16.0
    quad: int = 10                 # quadrtureRule=10pt
    quad_electro: int = 14         # quadRuleElectro=14pt
    nl_nel: int = 6                # number elements of nonlocal X/Y/Z
    nl_radius: float = 2.7937501   # nonlocal radius delta x/y/z
    rho_rank: int = 40             # decomposition rank rho
    veff_rank: int = 50            # decomposition rank veff
    rank_nloc: int = 25            # rankNloc
    temperature: float = 500.0
    degree: int = 20               # polynomial degree for Chebyshevfilter
    nstates: int = 24              # number eigenvalues
    alpha: float = 0.5
    history: int = 10
    tol: float = 5.0e-7            # SCF tolerance
    max_iter: int = 50
    restart_first: int = 10
    restart_other: int = 12
    lanczos_iter: int = 20
    hloc_truncation: float = 1e-8  # command-line arguments of the recorded run: not recorded
    cnloc_truncation: float = 1e-8
    seed: int = 20261202
    mixed: bool = False            # FP32 filter + split-BF16 GEMMs, then FP64 refinement
    refine_below: float = 1e-4     # residual at which a mixed run switches to FP64


def _dev(ctx, a):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


class DeckAxis:
    """Mesh, rules, file basis and 1-D matrices of one axis."""

    def __init__(self, ctx, cfg, basis_path):
        self.ctx, self.cfg = ctx, cfg
        self.nodes = E.fem_mesh("uniform", cfg.nel, cfg.npe, cfg.start, cfg.end)
        self.basis = E.read_basis_file(basis_path)  # (R, nodes), zero at both ends
        self.R, nn = self.basis.shape
        if nn != len(self.nodes):
            raise ValueError(f"{basis_path}: {nn} nodal values for a mesh of {len(self.nodes)}")
        self.xi, self.w = E.quadrature_rule(cfg.quad)
        self.np = cfg.nel * cfg.quad
        self.d_nodal = _dev(ctx, self.basis)
        self.d_B = E.DeviceArray(ctx, (self.R * self.np,))
        self.d_DB = E.DeviceArray(ctx, (self.R * self.np,))
        self.jw, self.inv_jw, self.xq = ctx.fem_interpolate(cfg.nel, cfg.npe, self.xi, self.w,
                                                            self.nodes, self.R, self.d_nodal.ptr,
                                                            self.d_B.ptr, self.d_DB.ptr)
        self.d_jw, self.d_inv_jw, self.d_xq = (_dev(ctx, self.jw), _dev(ctx, self.inv_jw),
                                               _dev(ctx, self.xq))
        d_mx, d_mdx = E.DeviceArray(ctx, (self.R ** 2,)), E.DeviceArray(ctx, (self.R ** 2,))
        ctx.tucker_basis_matrices(self.np, self.R, self.d_B.ptr, self.d_DB.ptr, self.d_jw.ptr,
                                  self.d_inv_jw.ptr, 0, None, d_mx.ptr, d_mdx.ptr, None)
        self.MatX = d_mx.to_host().reshape(self.R, self.R)
        self.MatDX = d_mdx.to_host().reshape(self.R, self.R)
        # the electrostatic rule on the same mesh (sources of the Hartree potential)
        self.xe_rule, self.we_rule = E.quadrature_rule(cfg.quad_electro)
        self.ne = cfg.nel * cfg.quad_electro
        self.jw_e, _, self.xq_e = ctx.fem_interpolate(cfg.nel, cfg.npe, self.xe_rule,
                                                      self.we_rule, self.nodes)
        self.d_jw_e, self.d_xq_e = _dev(ctx, self.jw_e), _dev(ctx, self.xq_e)
        self.d_nodes = _dev(ctx, self.nodes)

    def mat_pot(self, d_U, rv):
        if getattr(self, "_mp_rank", None) != rv:
            self._d_mp = E.DeviceArray(self.ctx, (rv * self.R * self.R,))
            self._mp_rank = rv
        self.ctx.tucker_basis_matrices(self.np, self.R, self.d_B.ptr, self.d_DB.ptr,
                                       self.d_jw.ptr, self.d_inv_jw.ptr, rv, d_U.ptr, None, None,
                                       self._d_mp.ptr)
        return self._d_mp.to_host().reshape(rv, self.R, self.R)

    def to_electro(self, d_U_nodal, rank, d_U_e):
        """Nodal factor columns (nodes x rank, column-major) at the electrostatic points."""
        self.ctx.fem_interpolate(self.cfg.nel, self.cfg.npe, self.xe_rule, self.we_rule,
                                 self.nodes, rank, d_U_nodal.ptr, d_U_e.ptr)


class ExampleDeck:
    """Set-up and SCF of the deck found in `deck` (a directory): the atom file `system`,
    basis_{x,y,z}.txt and the four tables of each atom type."""

    def __init__(self, ctx, deck, cfg=None, system="shell1"):
        self.ctx = ctx
        self.cfg = cfg = cfg or DeckConfig()
        t0 = time.perf_counter()
        f = lambda name: os.path.join(deck, name)  # noqa: E731
        self.nelec, lmax, atoms = E.read_atom_file(f(system))
        self.lmax = [int(v) for v in lmax]
        self.atoms = atoms  # grouped by type, as all_nuclei
        self.types = atoms[:, 6].astype(int)
        self.xyz = atoms[:, 1:4].copy()
        self.zval = atoms[:, 0].copy()
        self.natoms = len(atoms)
        self.axes = [DeckAxis(ctx, cfg, f(n)) for n in ("basis_x.txt", "basis_y.txt",
                                                        "basis_z.txt")]
        self.R = tuple(a.R for a in self.axes)
        self.p = tuple(a.np for a in self.axes)
        self.pn = tuple(len(a.nodes) for a in self.axes)
        self.ngrid = int(np.prod(self.p))
        self.nnodal = int(np.prod(self.pn))
        ax = self.axes
        # tables of each atom type (main.cc:501-519: locPotential_AT<i>, nlpV_AT<i>,
        # nlpWaveFun_AT<i>; Density_AT<i> for the initial density)
        self.tables = []
        for t in range(len(self.lmax)):
            V = E.read_table(f(f"nlpV_AT{t}"))
            Rt = E.read_table(f(f"nlpWaveFun_AT{t}"))
            L = E.read_table(f(f"locPotential_AT{t}"))
            if not (np.array_equal(V[:, 0], Rt[:, 0]) and np.array_equal(V[:, 0], L[:, 0])):
                raise ValueError(f"type {t}: pseudopotential tables on different radial grids")
            lm = self.lmax[t]
            self.tables.append((L[:, 0], L[:, 1], V[:, 1:1 + lm].T.copy(),
                                Rt[:, 1:1 + lm].T.copy(), E.read_table(f(f"Density_AT{t}"))))
        # local part on the quadrature grid (computeLocalPart, summed over the types)
        self.d_vloc = E.DeviceArray(ctx, (self.ngrid,))
        first = True
        for t, (r, vloc, _, _, _) in enumerate(self.tables):
            sel = self.types == t
            if sel.any():
                ctx.psp_local(atoms[sel], r, vloc, *(a.d_xq.ptr for a in ax), self.p,
                              self.d_vloc.ptr, add=not first)
                first = False
        self.proj, self.inv_c_lm, self.c_lm_by_type = self._projectors()
        self.c_lm = self.c_lm_by_type[0]
        # weights
        self.d_jw3 = E.DeviceArray(ctx, (self.ngrid,))
        ctx.grid_weights(self.p, *(a.d_jw.ptr for a in ax), self.d_jw3.ptr)
        self.d_one = E.DeviceArray(ctx, (self.ngrid,))
        ctx.fill_uniform(self.d_one.ptr, self.ngrid, 0, 1.0, 1.0)
        self.alpha, self.omega = coulomb_expansion()
        d = self.xyz[:, None, :] - self.xyz[None, :, :]
        rr = np.sqrt((d ** 2).sum(-1)) + np.eye(self.natoms)
        self.e_nn = 0.5 * float(((self.zval[:, None] * self.zval[None, :]) / rr)
                                [~np.eye(self.natoms, dtype=bool)].sum())
        self.setup_seconds = time.perf_counter() - t0

    # ---- set-up pieces -----------------------------------------------------------------
    def _projectors(self):
        """Tucker form (rank rankNloc) of every dV_l phi_lm of each type on the nonlocal
        mesh, and each atom's projector factors mx[I, a] = sum_q B_I(R_x + x_q) U_a(x_q) jw_q,
        atoms in all_nuclei order, (l, m) within an atom."""
        ctx, cfg = self.ctx, self.cfg
        xi, w = E.quadrature_rule(cfg.quad)
        nl_nodes = E.fem_mesh("uniform", cfg.nl_nel, cfg.npe, -cfg.nl_radius, cfg.nl_radius)
        jw_nl, _, xq_nl = ctx.fem_interpolate(cfg.nl_nel, cfg.npe, xi, w, nl_nodes)
        m = len(xq_nl)
        npts = m ** 3
        d_g, d_w = _dev(ctx, xq_nl), _dev(ctx, jw_nl)
        rk = cfg.rank_nloc
        d_core = E.DeviceArray(ctx, (rk ** 3,))
        d_U = [E.DeviceArray(ctx, (m * rk,)) for _ in range(3)]
        forms, c_lms = [], []
        for t, (r, vloc, vl, rl, _) in enumerate(self.tables):
            F = self.lmax[t] ** 2
            d_field = E.DeviceArray(ctx, (F * npts,))
            c_lms.append(ctx.psp_projectors(r, vloc, vl, rl, d_g.ptr, d_g.ptr, d_g.ptr,
                                            (m, m, m), d_w.ptr, d_w.ptr, d_w.ptr,
                                            d_field=d_field.ptr))
            ft = []
            for fi in range(F):
                ctx.sthosvd(d_field.ptr + 8 * fi * npts, (m, m, m), (rk,) * 3, d_core.ptr,
                            [u.ptr for u in d_U])
                G = d_core.to_host().reshape(rk, rk, rk).transpose(2, 1, 0)  # [a, b, c]
                U = [u.to_host().reshape(rk, m) for u in d_U]  # rows: factor functions
                ft.append((G, U))
            forms.append(ft)
        cores, mx, my, mz, inv = [], [], [], [], []
        for a in range(self.natoms):
            t = self.types[a]
            Bq = [eval_nodal(ax.nodes, self.cfg.npe, ax.basis, self.xyz[a, d] + xq_nl)
                  for d, ax in enumerate(self.axes)]  # (R, m) per axis
            for G, U in forms[t]:
                cores.append(G)
                mx.append(Bq[0] @ (U[0] * jw_nl).T)
                my.append(Bq[1] @ (U[1] * jw_nl).T)
                mz.append(Bq[2] @ (U[2] * jw_nl).T)
            inv.extend(1.0 / c_lms[t])
        return (cores, mx, my, mz), np.array(inv), c_lms

    def initial_density(self, d_grid, d_nodal):
        """Superposition of each type's Density_AT<i> over its atoms at the quadrature
        points and the nodes, both scaled so that the quadrature grid holds the electron
        count (main.cc:674-680)."""
        ctx, ax = self.ctx, self.axes
        targets = ((d_grid, [a.d_xq for a in ax], self.p, self.ngrid),
                   (d_nodal, [a.d_nodes for a in ax], self.pn, self.nnodal))
        first, tmp = True, None
        for t, (_, _, _, _, dens) in enumerate(self.tables):
            sel = self.types == t
            if not sel.any():
                continue
            if not first and tmp is None:
                tmp = [E.DeviceArray(ctx, (count,)) for _, _, _, count in targets]
            for i, (d_out, pts, dims, count) in enumerate(targets):
                dst = d_out if first else tmp[i]  # the first type writes, later ones add
                ctx.initial_density(self.xyz[sel], dens[:, 0], dens[:, 1],
                                    *(g.ptr for g in pts), dims, dst.ptr)
                if not first:
                    ctx.grid_axpby(d_out.ptr, 1.0, dst.ptr, 1.0, count)
            first = False
        q = self.integral(d_grid)
        s = self.nelec / q
        ctx.grid_scale(d_grid.ptr, self.ngrid, s)
        ctx.grid_scale(d_nodal.ptr, self.nnodal, s)
        return s

    # ---- grid helpers --------------------------------------------------------------------
    def integral(self, d_f, d_g=None):
        w = [a.d_jw.ptr for a in self.axes]
        if d_g is None:
            return self.ctx.grid_integral(d_f.ptr, self.p, *w)
        return self.ctx.grid_dot(d_f.ptr, d_g.ptr, self.p, *w)

    def hartree(self, d_rho_nodal, d_vh):
        """V_H on the quadrature grid from the nodal density."""
        ctx, cfg, ax = self.ctx, self.cfg, self.axes
        rk = cfg.rho_rank
        if not hasattr(self, "_h_bufs"):
            self._h_bufs = (E.DeviceArray(ctx, (rk ** 3,)),
                            [E.DeviceArray(ctx, (n * rk,)) for n in self.pn],
                            [E.DeviceArray(ctx, (a.ne * rk,)) for a in ax])
        d_core, d_Un, d_Ue = self._h_bufs
        ctx.sthosvd(d_rho_nodal.ptr, self.pn, (rk,) * 3, d_core.ptr, [u.ptr for u in d_Un])
        for a, un, ue in zip(ax, d_Un, d_Ue):
            a.to_electro(un, rk, ue)
        ctx.hartree_expansion((rk,) * 3, d_core.ptr, [u.ptr for u in d_Ue],
                              [a.d_xq_e.ptr for a in ax], [a.d_jw_e.ptr for a in ax],
                              tuple(a.ne for a in ax), [a.d_xq.ptr for a in ax], self.p,
                              self.alpha, self.omega, d_vh.ptr)

    def hamiltonian(self, d_veff, copies=E.HAM_F64):
        ctx, cfg = self.ctx, self.cfg
        rv = cfg.veff_rank
        if not hasattr(self, "_v_bufs"):
            self._v_bufs = (E.DeviceArray(ctx, (rv ** 3,)),
                            [E.DeviceArray(ctx, (n * rv,)) for n in self.p])
        d_core, d_U = self._v_bufs
        ctx.sthosvd(d_veff.ptr, self.p, (rv,) * 3, d_core.ptr, [u.ptr for u in d_U])
        PX, PY, PZ = (a.mat_pot(u, rv) for a, u in zip(self.axes, d_U))
        G = d_core.to_host().reshape(rv, rv, rv).transpose(2, 1, 0)
        a = self.axes
        hloc = (a[0].MatX, a[0].MatDX, a[1].MatX, a[1].MatDX, a[2].MatX, a[2].MatDX, PX, PY, PZ,
                G)
        return ctx.ham_assemble_tucker(hloc, cfg.hloc_truncation, self.proj,
                                       cfg.cnloc_truncation, self.inv_c_lm, copies)

    # ---- the loop --------------------------------------------------------------------------
    def run(self, callback=None, max_iter=None):
        """SCF from the superposition; returns a dict of per-iteration records."""
        ctx, cfg = self.ctx, self.cfg
        n, nn, k = self.ngrid, self.nnodal, cfg.nstates
        N = int(np.prod(self.R))
        t_start = time.perf_counter()
        d_in, d_out, d_veff, d_vh, d_vxc, d_exc = (E.DeviceArray(ctx, (n,)) for _ in range(6))
        d_vh_out, d_vxc_out, d_exc_out = (E.DeviceArray(ctx, (n,)) for _ in range(3))
        d_in_n, d_out_n = E.DeviceArray(ctx, (nn,)), E.DeviceArray(ctx, (nn,))
        self.initial_density(d_in, d_in_n)
        mixer = E.Mixer(ctx, n, nn, self.d_jw3.ptr, self.d_one.ptr, cfg.alpha, cfg.history)
        dX = E.DeviceArray(ctx, (N * k,))
        ctx.random_subspace(dX.ptr, k, N, k, seed=cfg.seed)
        dt = E.F64
        if cfg.mixed:  # the same start block, rounded to FP32
            dX32 = E.DeviceArray(ctx, (N * k,), np.float32)
            E._check(E.lib().ttdft_block_convert(ctx.h, dX.ptr, k, E.F64, dX32.ptr, k, E.F32, N,
                                                 k))
            dt = E.F32
        ev = np.zeros(k)
        st = E.new_state()
        occ_index = min(self.nelec // 2 + 10, k - 1)
        recs = []
        converged = False
        for it in range(max_iter or cfg.max_iter):
            t = time.perf_counter()
            mixer.compute_rho_in(it, d_in_n.ptr, d_in.ptr)
            self.hartree(d_in_n, d_vh)
            ctx.lda_xc(d_in.ptr, n, d_vxc.ptr, d_exc.ptr, variant=E.LDA_PZ81_REFERENCE)
            ctx.copy(d_veff.ptr, self.d_vloc.ptr, 8 * n)
            ctx.grid_axpby(d_veff.ptr, 1.0, d_vh.ptr, 1.0, n)
            ctx.grid_axpby(d_veff.ptr, 1.0, d_vxc.ptr, 1.0, n)
            if dt == E.F32 and recs and recs[-1]["residual"] < cfg.refine_below:
                # FP64 refinement from here on (the FP32 block widened to FP64)
                E._check(E.lib().ttdft_block_convert(ctx.h, dX32.ptr, k, E.F32, dX.ptr, k, E.F64,
                                                     N, k))
                dt = E.F64
            # the Lanczos bounds always run on the FP64 copy
            ham = self.hamiltonian(d_veff, E.HAM_F64 if dt == E.F64 else E.HAM_F64 | E.HAM_F32)
            h_n, h_nnz, h_p = ham.info()
            dXc = dX if dt == E.F64 else dX32
            ctx.chfsi_step(ham, dXc.ptr, k, k, cfg.degree, ev, st, it, occ_index,
                           cfg.restart_first if it == 0 else cfg.restart_other,
                           maxit_lanczos=cfg.lanczos_iter, precision=dt)
            passes = st.passes
            ham.free()
            ef, occ = E.fermi_occupancy(ev, self.nelec, KB_HARTREE, cfg.temperature)
            e_band = float(E.band_energy(ev, occ))
            ctx.density(dXc.ptr, k, k, occ, self.R, *(a.d_B.ptr for a in self.axes), self.p,
                        d_out.ptr, dtype=dt)
            ctx.density(dXc.ptr, k, k, occ, self.R, *(a.d_nodal.ptr for a in self.axes),
                        self.pn, d_out_n.ptr, dtype=dt)
            charge = self.integral(d_out)
            # total energy at rho_out (main.cc:1368-1401)
            self.hartree(d_out_n, d_vh_out)
            ctx.lda_xc(d_out.ptr, n, d_vxc_out.ptr, d_exc_out.ptr, variant=E.LDA_PZ81_REFERENCE)
            energy = (self.integral(d_out, d_exc_out) + 0.5 * self.integral(d_out, d_vh_out)
                      - self.integral(d_out, d_vh) - self.integral(d_out, d_vxc)
                      + self.e_nn + e_band)
            res = ctx.grid_residual(d_out.ptr, d_in.ptr, self.p,
                                    *(a.d_jw.ptr for a in self.axes))
            mixer.update(d_in_n.ptr, d_in.ptr, d_out_n.ptr, d_out.ptr, it)
            rec = {"iter": it, "evals": ev.copy().tolist(), "fermi": ef, "band_energy": e_band,
                   "total_energy": energy, "charge_out": charge, "residual": res,
                   "passes": passes, "nnz": h_nnz, "fp32": dt == E.F32,
                   "seconds": time.perf_counter() - t, "elapsed": time.perf_counter() - t_start}
            recs.append(rec)
            if callback:
                callback(rec)
            if res < cfg.tol:
                converged = True
                break
        ctx.synchronize()
        return {"converged": converged, "iterations": len(recs), "records": recs,
                "seconds": time.perf_counter() - t_start, "setup_seconds": self.setup_seconds,
                "c_lm": self.c_lm.tolist(), "e_nn": self.e_nn}


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="run the reference's example deck on the device")
    ap.add_argument("deck", help="directory with shell1, basis_*.txt and the tables")
    ap.add_argument("--max-iter", type=int, default=None)
    ap.add_argument("--hloc-truncation", type=float, default=DeckConfig.hloc_truncation)
    ap.add_argument("--cnloc-truncation", type=float, default=DeckConfig.cnloc_truncation)
    ap.add_argument("--mixed", action="store_true",
                    help="FP32 filter and split-BF16 GEMMs, FP64 below residual 1e-4")
    ap.add_argument("--out", default=None, help="write the records as JSON here")
    a = ap.parse_args(argv)
    cfg = DeckConfig(hloc_truncation=a.hloc_truncation, cnloc_truncation=a.cnloc_truncation,
                     mixed=a.mixed)
    ctx = E.Context(0)
    deck = ExampleDeck(ctx, a.deck, cfg)
    print(f"set-up {deck.setup_seconds:.1f} s; 1/C_lm = {np.round(1.0 / deck.c_lm, 9).tolist()}",
          flush=True)

    def show(rec):
        print(f"scf {rec['iter']:2d}: E = {rec['total_energy']:.10f}  band = "
              f"{rec['band_energy']:.10f}  ef = {rec['fermi']:.8f}  res = {rec['residual']:.3e}  "
              f"eps0 = {rec['evals'][0]:.10f}  q = {rec['charge_out']:.8f}  passes {rec['passes']}"
              f"  nnz/row {rec['nnz'] / math.prod(deck.R):.0f}  {rec['seconds']:.2f} s", flush=True)

    out = deck.run(show, a.max_iter)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(out, fh, indent=1)
    return out


if __name__ == "__main__":
    main()
