/*
 * CPU FP64 oracle for the ChFSI hot path (test infrastructure only).
 *
 * This is a plain-C restatement of the reference algorithms that the
 * MI355X engine replaces.  It exists to check the HIP path: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and never as the thing measured or shipped.
 *
 * Layout conventions (shared with the engine's C ABI):
 *   - A wavefunction block X holds k states over n Tucker basis functions.
 *     It is stored basis-major: X[i*ld + j] is the coefficient of basis
 *     function i in state j (ld >= k).  The reference keeps the same block
 *     column-major in a PETSc dense Mat; the two are transposes of each other.
 *   - Basis index I = x + Rx*y + Rx*Ry*z
 *     (reference dft/ProjectHamiltonianSparse.cc:420-433).
 *   - Small k x k matrices are column-major: S[a + k*b].
 *   - 1-D basis tables B are column-major (points x functions): B[q + P*a].
 *   - 3-D grids are x-fastest: rho[qx + Px*qy + Px*Py*qz].
 *
 * Pinning status of each function is recorded next to its declaration.
 */
#ifndef TTDFT_CHFSI_ORACLE_H
#define TTDFT_CHFSI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sparse projected Hamiltonian H = H_loc + C^T D C.
 *   H_loc : n x n CSR (reference Mat H_loc, dft/ProjectHamiltonianSparse.h:97)
 *   C     : p x n CSR, NOT row-scaled (reference C_nloc before
 *           MatDiagonalScale, dft/ProjectHamiltonianSparse.cc:316-325)
 *   inv_c_lm : p entries, D = diag(inv_c_lm)                                  */
typedef struct {
    int64_t n;
    const int64_t *rowptr;
    const int32_t *col;
    const double *val;
    int32_t p;
    const int64_t *c_rowptr;
    const int32_t *c_col;
    const double *c_val;
    const double *inv_c_lm;
} orc_ham;

/* Y = H_loc X (CSR times basis-major block).  Pinned by brute-force dense
 * products in tests. */
void orc_csr_spmm(int64_t n, const int64_t *rowptr, const int32_t *col,
                  const double *val, const double *X, int64_t ldx, int k,
                  double *Y, int64_t ldy);

/* Y = C^T (D (C X)) + H_loc X, the operator applied in
 * eigensolver/ChebyShevFilter.cc:102-126 and EigenSolver.cc:224-232.
 * Pinned: brute-force dense H, symmetry (HX)^T Y = X^T (HY). */
void orc_ham_apply(const orc_ham *H, const double *X, int64_t ldx, int k,
                   double *Y, int64_t ldy);

/* Zhou-Saad scaled Chebyshev filter, eigensolver/ChebyShevFilter.cc:34-188.
 * convention 0 (reference): returns Y_{m-1} (the reference's Xp aliases the
 * caller's X, so the iterate written back is the previous one,
 * ChebyShevFilter.cc:79-85,163-176).  convention 1: returns Y_m.
 * Pinned: closed form p(lambda) v on eigenvectors (tests). Returns 0 or -1. */
int orc_cheb_filter(const orc_ham *H, double *X, int64_t ldx, int k, int m,
                    double a, double b, double a0, int num_blocks,
                    int convention);

/* S = X^T Y (k x k col-major), eigensolver/MultTm.cu:86-224 (intended math).
 * Pinned: brute force + symmetry. */
void orc_gram(int64_t n, int k, const double *X, int64_t ldx, const double *Y,
              int64_t ldy, double *S);

/* In-place lower Cholesky of a k x k col-major SPD matrix (upper part zeroed).
 * Returns 0, or -(j+1) if the leading minor j is not positive. */
int orc_cholesky_lower(int k, double *S);

/* Cholesky-QR X <- X L^{-T} with S = X^T X = L L^T,
 * eigensolver/MultTm.cu:226-467.  Pinned: X^T X = I, span preserved. */
int orc_orth(int64_t n, int k, double *X, int64_t ldx);

/* Symmetric eigen-decomposition (cyclic Jacobi) of the lower triangle of a
 * k x k col-major matrix: ascending eigenvalues, Q col-major with column i the
 * eigenvector of evals[i].  Replaces SLEPc EPSLAPACK in
 * eigensolver/EigenSolver.cc:30-184.  Pinned: numpy.linalg.eigh, 1-D Laplacian
 * closed form, residual/orthogonality. Returns sweeps used or -1. */
int orc_syev(int k, const double *A, double *evals, double *Q);

/* X <- X Q (k x k col-major Q), main.cc:1185-1207.  Pinned: brute force. */
void orc_rotate(int64_t n, int k, double *X, int64_t ldx, const double *Q);

/* Lanczos bound estimator, eigensolver/EigenSolver.cc:186-388, with the start
 * vector drawn from orc_rng (uniform [0,1), like PETSc VecSetRandom) instead
 * of PETSc's generator.  upper = |lambda_max(T)| + ||f||, lower = lambda_min(T).
 * Pinned: bounds contain the dense spectrum; exact on small diagonal H. */
int orc_lanczos_bounds(const orc_ham *H, int maxit, uint64_t seed,
                       double *upper, double *lower);

/* Fermi level by Newton iteration, dft/KSDFTEnergyFunctional.cc:76-124.
 * Pinned bit-exactly by example/output (24 SCF iterations). */
double orc_fermi_energy(const double *evals, int k, int nelec, double kb,
                        double T, int maxit);
/* Fermi-Dirac occupancies, main.cc:1253-1261.  Pinned by example/output. */
void orc_occupancy(const double *evals, int k, double ef, double kb, double T,
                   double *occ);
/* Band energy sum 2 f_i e_i, main.cc:1315-1318.  Pinned by example/output. */
double orc_band_energy(const double *evals, const double *occ, int k);

/* Density rho = sum_i 2 f_i psi_i^2 on a tensor grid of 1-D points,
 * psi_i = C_i x_z Bz x_x Bx x_y By (contraction order of
 * dft/ProjectHamiltonianSparse.cc:1052-1074).  Bx is px x rx col-major etc.
 * rho is written (not accumulated).  Pinned: separable closed forms, charge
 * sum, brute force. */
void orc_density(const double *X, int64_t ldx, int k, const double *occ,
                 int rx, int ry, int rz, const double *Bx, const double *By,
                 const double *Bz, int px, int py, int pz, double *rho);

/* Clamped cubic spline with zero end slopes (alglib spline1dbuildcubic with
 * boundary type 1, value 0, as called in dft/InitialGuess.cc:625-633).
 * coef holds 4*(n-1) values: segment s is
 * y = c0 + c1 t + c2 t^2 + c3 t^3 with t = x - x_s.
 * Pinned: reproduces cubic polynomials with zero end slopes exactly. */
int orc_spline_build(int n, const double *x, const double *y, double *coef);
/* Evaluation; outside [x_0, x_{n-1}] the end segment's cubic is used. */
double orc_spline_eval(int n, const double *x, const double *coef, double r);

/* Superposition of radial atomic densities on a tensor grid,
 * dft/InitialGuess.cc:573-681.  xyz holds 3*natoms coordinates.
 * Pinned: Gaussian radial table gives a closed-form superposition. */
void orc_initial_density(int natoms, const double *xyz, int nr,
                         const double *r, const double *rho_r, int px, int py,
                         int pz, const double *gx, const double *gy,
                         const double *gz, double *rho);

/* Sum f[i,j,k] wx[i] wy[j] wz[k] (tensor quadrature with separable weights,
 * the 3-D integral of dft/KSDFTEnergyFunctional.cc:126-192).  Pinned by
 * polynomial integrands under Gauss rules. */
double orc_grid_integral(const double *f, int px, int py, int pz,
                         const double *wx, const double *wy, const double *wz);

/* State carried between ChFSI calls by main.cc (lines 730-734, 977-1237). */
typedef struct {
    double upper;           /* upperBoundUnWantedSpectrum   */
    double lower_unwanted;  /* lowerBoundUnWantedSpectrum   */
    double lower_wanted;    /* lowerBoundWantedSpectrum     */
    double err_upper;       /* error_in_upper_bound         */
    double occ_energy;      /* occupied_orbital_energy      */
    int32_t passes;         /* filter passes run by the last call (output) */
    int32_t lanczos_runs;   /* Lanczos calls made by the last call (output) */
} orc_chfsi_state;

/* One SCF iteration's eigensolve, the control flow of main.cc:977-1237:
 * bound update, restart loop (filter, orth, Rayleigh-Ritz, stop test on the
 * relative change of eig[occ_index]), then rotation X <- X Q.
 * evals (k) is read (previous iteration) and overwritten.  Pinned: converged
 * Ritz values equal the lowest eigenvalues of dense brute force. */
int orc_chfsi_step(const orc_ham *H, double *X, int64_t ldx, int k, int m,
                   int num_blocks, int restart_times, int occ_index,
                   int scf_iter, int start_iter, int maxit_lanczos,
                   uint64_t seed, int convention, double *evals,
                   orc_chfsi_state *st);

/* Random numbers shared with the engine's host code: splitmix64 seeding and
 * xoshiro256** (Blackman and Vigna).  orc_rng_uniform fills [0,1) doubles. */
void orc_rng_uniform(uint64_t seed, int64_t count, double *out);
void orc_rng_normal(uint64_t seed, int64_t count, double *out);

/* Hamiltonian assembly (assembly_oracle.c).
 * orc_assemble_hloc: H_loc of dft/ProjectHamiltonianSparse.cc:330-575 from the
 * 1-D overlap (X, Y, Z) and kinetic (DX, DY, DZ) matrices, row-major r x r
 * (the reference's MatX[I][J]), the 1-D potential matrices PX[(p rx + I) rx + J]
 * (MatPotX[p][I][J]) and the Veff Tucker core G[p + rvx (q + rvy r)].
 * orc_assemble_projectors: the rows of C_nloc (:196-278).
 * Both return the number of kept entries and fill rowptr/col/val when rowptr
 * is not NULL.  Pinned: Kronecker-product closed forms and a dense einsum of
 * the defining formula (test_oracle_assembly.py). */
int64_t orc_assemble_hloc(int rx, int ry, int rz, int rvx, int rvy, int rvz,
                          const double *X, const double *DX, const double *Y,
                          const double *DY, const double *Z, const double *DZ,
                          const double *PX, const double *PY, const double *PZ,
                          const double *G, double truncation, int64_t *rowptr,
                          int32_t *col, double *val);
/* rows row0 <= row < row1 only; rowptr has row1 - row0 + 1 entries */
int64_t orc_assemble_hloc_rows(int rx, int ry, int rz, int rvx, int rvy, int rvz,
                               const double *X, const double *DX, const double *Y,
                               const double *DY, const double *Z, const double *DZ,
                               const double *PX, const double *PY, const double *PZ,
                               const double *G, double truncation, int64_t row0,
                               int64_t row1, int64_t *rowptr, int32_t *col, double *val);
int64_t orc_assemble_projectors(int p, const int32_t *ranks, int rx, int ry, int rz,
                                const double *cores, const double *mx, const double *my,
                                const double *mz, double truncation, int64_t *rowptr,
                                int32_t *col, double *val);

/* ---- Anderson mixing (mixing_oracle.c) ---------------------------------------
 * History entries oldest first: in[j*n + i], out[j*n + i], j < h.
 * orc_anderson_constants: the h-1 constants of AndersonMixing.cc:150-262
 * (returns the LU info, 0 on success).  orc_anderson_mix: the mixed density
 * of :264-323.  Pinned: closed form at h = 2, the normal equations of the
 * weighted least-squares problem, and the exact fixed point of an affine map
 * (test_oracle_mixing.py). */
int orc_anderson_constants(int h, int64_t n, const double *in, const double *out,
                           const double *jacob, const double *weight, double *constants);
void orc_anderson_mix(int h, int64_t n, const double *in, const double *out,
                      const double *constants, double alpha, double *rho);

/* ---- Tucker recompression and Hartree potential (tucker_oracle.c) -----------
 * orc_sthosvd: fixed-rank STHOSVD (TuckerMPI's algorithm), modes 0, 1, 2;
 * tensors first index fastest, factors column-major, largest |entry| of each
 * factor column positive.  orc_hartree_expansion: the kernel-expansion
 * Hartree potential of KSDFTPotential.cc:269-340,440-668.  Pinned: exact
 * recovery of multilinear-rank tensors, numpy SVD of the unfoldings, core
 * all-orthogonality; per-term Gaussian closed forms and the Coulomb potential
 * of a Gaussian charge (test_oracle_tucker.py). */
int orc_sthosvd(int n0, int n1, int n2, int r0, int r1, int r2, const double *T, double *core,
                double *U0, double *U1, double *U2, double *evals);
void orc_hartree_expansion(int r0, int r1, int r2, const double *G, const double *const U[3],
                           const int ne[3], const double *const xe[3], const double *const jw[3],
                           const int p[3], const double *const xq[3], int nterms,
                           const double *alpha, const double *omega, double *V, int add);

/* ---- one-dimensional finite elements (fem_oracle.c) -------------------------
 * orc_fem_mesh: FEM::generateNodes for the uniform, adaptive, manual1 and
 * hartree_outer meshes.  orc_fem_shape / orc_fem_interpolate: equispaced
 * Lagrange shape functions (values in product form, derivatives from monomial
 * coefficients as the reference forms them) and the per-element interpolation
 * with the per-point weights.  orc_tucker_basis_matrices: MatX, MatDX and
 * MatPot of ProjectHamiltonianSparse.cc:577-722.  Pinned: numpy's uniform
 * linspace and the closed forms of the graded meshes (end points, growth
 * ratios); partition of unity, Kronecker property and exact reproduction of
 * polynomials with their derivatives; exact integration of polynomials of
 * degree 2 nq - 1; the example's basis functions orthonormal to 3.4e-5 and
 * their kinetic and potential matrices against numpy's Lagrange route
 * (test_oracle_fem.py). */
int64_t orc_fem_mesh(int type, int nel, int npe, double start, double end, double inner_size,
                     double inner_h, double coarsening, double *nodes);
void orc_fem_shape(int npe, int nq, const double *xi, double *S, double *D);
void orc_fem_interpolate(int nel, int npe, int nq, const double *xi, const double *w,
                         const double *nodes, int nfun, const double *nodal, double *val,
                         double *dval, double *jw, double *inv_jw, double *xq);
void orc_tucker_basis_matrices(int nq, int r, const double *B, const double *DB,
                               const double *jw, const double *inv_jw, int rv, const double *U,
                               double *MatX, double *MatDX, double *MatPot);

#ifdef __cplusplus
}
#endif

#endif
