/*
 * ttdft_chfsi.h -- C ABI of the MI355X ChFSI engine (libttdft_chfsi.so).
 *
 * Drop-in boundary for the reference's Chebyshev-filtered subspace iteration
 * path.  Every entry point takes plain pointers and sizes; nothing here names
 * a torch, PETSc or HIP type.  Each declaration cites the reference interface
 * it replaces (paths relative to the reference repository).
 *
 * Conventions
 *   - Status: every function returns int; 0 = success, negative = error class
 *     (TTDFT_E*).  The message of the last failure on the calling thread is
 *     returned by ttdft_last_error().  No entry point aborts the process
 *     (the reference prints and calls std::terminate, MultCuda.cu:57-60).
 *   - Context: one ttdft_ctx = one device + one HIP stream.  Calls on a
 *     context are stream-ordered and not thread-safe; distinct contexts are
 *     independent.  Functions that return host-visible results synchronise
 *     the stream before returning.
 *   - Wavefunction blocks ("X blocks") live in device memory and are stored
 *     basis-major: X[i*ld + j] is the coefficient of basis function i in
 *     state j, ld >= k.  (The reference stores the same block column-major
 *     in a PETSc MATDENSE; ttdft_block_upload_colmajor converts.)  Element
 *     type is TTDFT_F64 (double) or TTDFT_F32 (float).
 *   - Basis index I = x + Rx*y + Rx*Ry*z (ProjectHamiltonianSparse.cc:420-433).
 *   - Small k x k matrices are column-major, FP64.
 *   - 1-D basis tables are column-major (points x functions), FP64, device.
 *   - 3-D grids are x-fastest, FP64, device: rho[qx + Px*qy + Px*Py*qz].
 */
#ifndef TTDFT_CHFSI_H
#define TTDFT_CHFSI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes, dtypes, modes ---------------------------------------- */
enum {
    TTDFT_OK = 0,
    TTDFT_EBADARG = -1,   /* invalid argument (null pointer, size, mode)   */
    TTDFT_EHIP = -2,      /* HIP runtime / kernel launch failure           */
    TTDFT_ENOTSPD = -3,   /* overlap matrix not positive definite          */
    TTDFT_ENOCONV = -4,   /* dense eigensolver did not converge            */
    TTDFT_ENOMEM = -5,    /* device allocation failed                      */
    TTDFT_ELIB = -6       /* rocBLAS / rocSOLVER failure                   */
};

enum { TTDFT_F64 = 0, TTDFT_F32 = 1 };

/* Dense contraction modes for the Gram / projection / rotation GEMMs. */
enum {
    TTDFT_GEMM_F64 = 0,    /* FP64 data, FP64 MFMA                              */
    TTDFT_GEMM_BF16X3 = 1  /* FP32 data, split-BF16 3-term MFMA, FP32 accumulate */
};

/* Which Chebyshev iterate ttdft_cheb_filter returns. */
enum {
    TTDFT_DEGREE_REFERENCE = 0, /* Y_{m-1}: what ChebyShevFilter.cc:163-176 writes back */
    TTDFT_DEGREE_EXACT = 1      /* Y_m: textbook degree-m filter                       */
};

/* Which device copies ttdft_ham_upload keeps (bit mask). */
enum { TTDFT_HAM_F64 = 1, TTDFT_HAM_F32 = 2 };

/* Kernel classes whose launch times the context records when profiling. */
enum {
    TTDFT_KSTAT_SPMM = 0,    /* H_loc SpMM with fused recurrence epilogue */
    TTDFT_KSTAT_NLPROJ = 1,  /* W = D C X (nonlocal projection)           */
    TTDFT_KSTAT_GRAM = 2,    /* X^T Y                                     */
    TTDFT_KSTAT_ROTATE = 3,  /* X Q                                       */
    TTDFT_KSTAT_DENSITY = 4, /* fused density contraction                 */
    TTDFT_KSTAT_ASSEMBLY = 5, /* H_loc assembly (device part of each pass) */
    TTDFT_KSTAT_MIXING = 6,  /* Anderson mixing: Gram pass and combination */
    TTDFT_KSTAT_STHOSVD = 7, /* Tucker recompression of a 3-D field        */
    TTDFT_KSTAT_HARTREE = 8, /* Hartree potential by kernel expansion      */
    TTDFT_KSTAT_COUNT = 9
};

typedef struct ttdft_ctx_s *ttdft_ctx;
typedef struct ttdft_ham_s *ttdft_ham;
typedef struct ttdft_mixer_s *ttdft_mixer;

/* ---- errors ---------------------------------------------------------------- */
/* Message of the last failure on this thread ("" if none). */
const char *ttdft_last_error(void);

/* ---- device / context ------------------------------------------------------ */
/* replaces device_get_device_count, eigensolver/DeviceUtils.cu:26-28 */
int ttdft_get_device_count(int *count);

/* Binds `device`, creates a stream and the dense-solver handle.
 * replaces device_set_device (DeviceUtils.cu:30-31), the device binding of
 * DeviceCommUtils (DeviceCommUtils.cc:69-71) and MultTM::MultTM (MultTm.cu:31-84) */
int ttdft_ctx_create(int device, ttdft_ctx *out);
/* Ends the context.  Hamiltonians and mixers created on it keep it alive:
 * with any of them not yet freed, the call synchronises the stream and
 * returns, and the last ttdft_ham_free / ttdft_mixer_destroy completes the
 * destruction.  No other call may use the context after this one.
 * replaces MultTM::~MultTM (MultTm.cu:469-473) */
int ttdft_ctx_destroy(ttdft_ctx ctx);
/* Make the context issue its work on an existing HIP stream (e.g. the
 * caller's); NULL restores the context's own stream.  The outgoing stream is
 * synchronised first, since the context's scratch and released storage
 * blocks are reused in the order of one stream. */
int ttdft_ctx_set_stream(ttdft_ctx ctx, void *hip_stream);
int ttdft_ctx_synchronize(ttdft_ctx ctx);
/* Returns the context's released storage blocks to the HIP runtime (up to
 * 64 GiB are kept for reuse by later Hamiltonians and mixers), after the
 * context's stream has finished with them: for a caller that shares the
 * device with another allocator (torch's, say).  No counterpart in the
 * reference, whose storage is PETSc's. */
int ttdft_ctx_trim(ttdft_ctx ctx);
/* Enable per-launch HIP-event timing of the kernel classes TTDFT_KSTAT_*. */
int ttdft_ctx_set_profiling(ttdft_ctx ctx, int enable);
/* Launch count, summed milliseconds and algorithmic bytes / flops recorded
 * for one kernel class since the last reset (synchronises the stream). */
int ttdft_ctx_kernel_stats(ttdft_ctx ctx, int kernel_class, int64_t *launches,
                           double *total_ms, double *bytes, double *flops);
int ttdft_ctx_reset_stats(ttdft_ctx ctx);

/* ---- device memory helpers (for callers without their own allocator) ------ */
int ttdft_malloc(ttdft_ctx ctx, size_t bytes, void **dptr);
int ttdft_free(ttdft_ctx ctx, void *dptr);
int ttdft_memcpy_h2d(ttdft_ctx ctx, void *dst, const void *src, size_t bytes);
int ttdft_memcpy_d2h(ttdft_ctx ctx, void *dst, const void *src, size_t bytes);
int ttdft_memcpy_d2d(ttdft_ctx ctx, void *dst, const void *src, size_t bytes);

/* Upload a host column-major n x k FP64 block (PETSc MATDENSE layout, lda)
 * into a device basis-major block of dtype `dtype` (ld >= k). */
int ttdft_block_upload_colmajor(ttdft_ctx ctx, const double *host, int64_t lda,
                                int64_t n, int32_t k, void *dX, int64_t ldx,
                                int dtype);
/* Inverse of the above: device basis-major block -> host column-major FP64. */
int ttdft_block_download_colmajor(ttdft_ctx ctx, const void *dX, int64_t ldx,
                                  int dtype, int64_t n, int32_t k, double *host,
                                  int64_t lda);
/* Element-type conversion of a device block (FP64 <-> FP32; same dtype copies). */
int ttdft_block_convert(ttdft_ctx ctx, const void *src, int64_t lds, int src_dtype,
                        void *dst, int64_t ldd, int dst_dtype, int64_t n, int32_t k);

/* ---- Hamiltonian ------------------------------------------------------------ */
/* Upload H = H_loc + C^T D C.
 *   H_loc: n x n CSR (host arrays; rowptr int64[n+1], colind int32[nnz],
 *          val double[nnz]); must be symmetric (as the reference requires,
 *          Mult.h:37-42).
 *   C    : p x n CSR of the projectors, NOT row-scaled; D = diag(inv_c_lm).
 *          p = 0 means no nonlocal part (pointers may be NULL).
 *   copies: TTDFT_HAM_F64 | TTDFT_HAM_F32 (device copies to keep).
 * replaces AX::setup_system + setup_mata_on_device (MultCuda.cu:105-371) and
 * the C_nloc / C_nloc_trans / 1/C_lm scaling of Create_Cnloc
 * (dft/ProjectHamiltonianSparse.cc:316-325). */
int ttdft_ham_upload(ttdft_ctx ctx, int64_t n, const int64_t *rowptr,
                     const int32_t *colind, const double *val, int32_t p,
                     const int64_t *c_rowptr, const int32_t *c_colind,
                     const double *c_val, const double *inv_c_lm, int copies,
                     ttdft_ham *out);
/* As ttdft_ham_upload, for H_loc in a Tucker product basis with ranks
 * (rx, ry, rz), n = rx*ry*rz, basis index x + rx*y + rx*ry*z.  The ranks let the
 * engine detect the Kronecker band structure of H_loc (ProjectHamiltonianSparse.cc
 * :420-491: |x-x'| <= bx, |y-y'| <= by, |z-z'| <= bz) and store it as
 * index-free band blocks for the register-blocked band SpMM; when the band is
 * too sparse or too wide the CSR kernel is used.  The reference passes the
 * same ranks to Create_Hloc (ProjectHamiltonianSparse.cc:330-340). */
int ttdft_ham_upload_tucker(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz,
                            const int64_t *rowptr, const int32_t *colind,
                            const double *val, int32_t p, const int64_t *c_rowptr,
                            const int32_t *c_colind, const double *c_val,
                            const double *inv_c_lm, int copies, ttdft_ham *out);

/* Storage chosen for H_loc: *format = 0 (CSR) or 1 (Kronecker band), and the
 * half-bandwidths found (bands[3], zero for CSR).  Introspection only; no
 * reference counterpart. */
int ttdft_ham_format(ttdft_ham ham, int32_t *format, int32_t *bands);

/* replaces AX::free_mata_on_device / AX::~AX (MultCuda.cu:673-717) */
int ttdft_ham_free(ttdft_ham ham);
/* Basis size n, H_loc nnz, projector count p. */
int ttdft_ham_info(ttdft_ham ham, int64_t *n, int64_t *nnz, int32_t *p);

/* Y = H X over k states (device blocks of dtype `dtype`).
 * replaces AX::perform_ax (MultCuda.cu:719-820) together with the two
 * MatMatMult nonlocal products at ChebyShevFilter.cc:103-114 and
 * main.cc:1079-1092; unlike the published perform_ax it writes Y. */
int ttdft_ham_apply(ttdft_ctx ctx, ttdft_ham ham, const void *X, int64_t ldx,
                    void *Y, int64_t ldy, int32_t k, int dtype);

/* Y = H X as ttdft_ham_apply (same kernels, same result), but allocating
 * nothing: no hipMalloc / hipFree, no library allocation.  No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The nonlocal projections go to a static
 * device work space of 32 MiB per GPU (held until the process ends): the
 * call fails with TTDFT_EBADARG when p * k * (1 + splits) * sizeof(dtype)
 * exceeds it (k = 300 and p = 172 need at most 2.5 MB in FP64).  A product on
 * dense storage also keeps its partial sums there (4 n k doubles) when they
 * fit (n k up to about 1e6); above that it runs in one kernel, which agrees
 * with ttdft_ham_apply up to the order of the additions.  Every call on one
 * GPU shares that work space, so calls that overlap on one GPU (different
 * contexts or streams) race; calls on one context are stream-ordered and
 * safe.  Not recorded by the context's kernel timing. */
int ttdft_ham_apply_noalloc(ttdft_ctx ctx, ttdft_ham ham, const void *X, int64_t ldx,
                            void *Y, int64_t ldy, int32_t k, int dtype);

/* In-place Chebyshev filter of the k states of X (device, dtype).
 * Filter interval [a, b], a0 = lowest wanted bound, degree m, the columns
 * processed in num_blocks chunks, `degree_convention` TTDFT_DEGREE_*.
 * Needs k*ldx*sizeof(dtype) bytes of context scratch.
 * replaces ChebyShevFilter::computeFilteredSubspace (ChebyShevFilter.h:31-39,
 * ChebyShevFilter.cc:34-188). */
int ttdft_cheb_filter(ttdft_ctx ctx, ttdft_ham ham, void *X, int64_t ldx,
                      int32_t k, int32_t m, double a, double b, double a0,
                      int32_t num_blocks, int dtype, int degree_convention);

/* ---- dense subspace operations -------------------------------------------- */
/* S = X^T Y (k x k, column-major FP64, device).  X, Y: n x k device blocks.
 * mode TTDFT_GEMM_F64 needs dtype F64; TTDFT_GEMM_BF16X3 needs dtype F32.
 * replaces MultTM::mult (MultTm.h:35, MultTm.cu:86-224). */
int ttdft_gram(ttdft_ctx ctx, const void *X, int64_t ldx, const void *Y,
               int64_t ldy, int64_t n, int32_t k, int dtype, int mode,
               double *dS);

/* Cholesky-QR in place: S = X^T X = L L^T (FP64 on the device),
 * X <- X L^{-T}.  replaces MultTM::orth (MultTm.h:37, MultTm.cu:226-467). */
int ttdft_orth(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k,
               int dtype, int mode);

/* All eigenpairs of the symmetric k x k matrix dA (device, column-major, lower
 * triangle referenced; dA is not modified), ascending: d_evals (device, k),
 * dQ (device, k x k column-major, column i <-> evals[i]).
 * replaces EigenSolver::computeEigenPairsSeq (EigenSolver.h:33-39,
 * EigenSolver.cc:30-184). */
int ttdft_rr_eig(ttdft_ctx ctx, const double *dA, int32_t k, double *d_evals,
                 double *dQ);

/* X <- X Q in place (Q: k x k column-major FP64, device).
 * replaces the rotation of main.cc:1185-1207. */
int ttdft_rotate(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k,
                 const double *dQ, int dtype, int mode);

/* Lanczos spectrum bounds with `maxit` steps, start vector drawn from the
 * engine's seeded generator (splitmix64 + xoshiro256**, uniform [0,1)).
 * upper = |lambda_max(T)| + ||f||, lower = lambda_min(T).
 * replaces EigenSolver::computeUpperBoundWithLanczos (EigenSolver.h:41-45,
 * EigenSolver.cc:186-388). */
int ttdft_lanczos_bounds(ttdft_ctx ctx, ttdft_ham ham, int32_t maxit,
                         uint64_t seed, double *upper, double *lower);

/* ---- occupancies ------------------------------------------------------------ */
/* Fermi level (Newton, tol 1e-8, <= maxit steps) and Fermi-Dirac occupancies
 * for k eigenvalues (host arrays).
 * replaces KSDFTEnergyFunctional::computeFermiEnergy
 * (dft/KSDFTEnergyFunctional.cc:76-124) and main.cc:1239-1271. */
int ttdft_fermi_occupancy(const double *evals, int32_t k, int32_t nelec,
                          double kb, double temperature, int32_t maxit,
                          double *fermi, double *occ);
/* Band energy sum_i 2 f_i e_i, main.cc:1315-1318. */
double ttdft_band_energy(const double *evals, const double *occ, int32_t k);

/* ---- density ---------------------------------------------------------------- */
/* rho = sum_i 2 occ_i psi_i^2 on the px x py x pz tensor grid of 1-D points,
 * psi_i = sum_abc X[abc][i] Bx[qx,a] By[qy,b] Bz[qz,c]; occ is a host array,
 * Bx/By/Bz device (px x rx, ... column-major), rho device (written).
 * An FP64 block is contracted in FP64; an FP32 block (the mixed-precision
 * form) in FP32 with FP32 accumulation throughout, rho widened to FP64 at
 * the end (about 1e-6 relative).
 * replaces ProjectHamiltonianSparse::computeRhoOut
 * (dft/ProjectHamiltonianSparse.cc:911-1221, direct form of :956-983,1081-1104)
 * and the RhoComputer stub (dft/RhoComputer.cc:28-53). */
int ttdft_density(ttdft_ctx ctx, const void *X, int64_t ldx, int dtype,
                  int32_t k, const double *occ, int32_t rx, int32_t ry,
                  int32_t rz, const double *dBx, const double *dBy,
                  const double *dBz, int32_t px, int32_t py, int32_t pz,
                  double *d_rho);

/* ---- initial density and grid integrals ------------------------------------ */
/* rho(q) = sum_atoms spline(|q - R_a|) on a tensor grid (gx, gy, gz device
 * 1-D coordinates; rho device, written).  The radial table (r, rho_r; host)
 * is interpolated by a clamped cubic spline with zero end slopes, evaluated
 * beyond the table with the end segments.
 * replaces InitialGuess::initializeRhoFromRadiusFile
 * (dft/InitialGuess.cc:573-681) as called by initialize_rho (:683-762). */
int ttdft_initial_density(ttdft_ctx ctx, int32_t natoms, const double *xyz,
                          int32_t nr, const double *r, const double *rho_r,
                          int32_t px, int32_t py, int32_t pz, const double *dgx,
                          const double *dgy, const double *dgz, double *d_rho);

/* sum f[q] wx[qx] wy[qy] wz[qz] (device f, device weights), result on host.
 * replaces KSDFTEnergyFunctional::compute3DIntegral
 * (dft/KSDFTEnergyFunctional.cc:165-192) for separable weights. */
int ttdft_grid_integral(ttdft_ctx ctx, const double *d_f, int32_t px,
                        int32_t py, int32_t pz, const double *dwx,
                        const double *dwy, const double *dwz, double *result);
/* sum (f - g)^2 wx wy wz: the SCF density residual of main.cc:1410-1428. */
int ttdft_grid_residual(ttdft_ctx ctx, const double *d_f, const double *d_g,
                        int32_t px, int32_t py, int32_t pz, const double *dwx,
                        const double *dwy, const double *dwz, double *result);
/* f *= s in place (renormalisation of main.cc:674-680). */
int ttdft_grid_scale(ttdft_ctx ctx, double *d_f, int64_t count, double s);
/* f[i] = lo + (hi - lo) u_i, u_i uniform in [0, 1) from a hash of (seed, i):
 * synthetic device fields for benchmarks and tests (no reference counterpart). */
int ttdft_fill_uniform(ttdft_ctx ctx, double *d_f, int64_t count, uint64_t seed, double lo,
                       double hi);

/* ---- the ChFSI driver block --------------------------------------------------- */
/* State main.cc carries between SCF iterations (lines 730-734, 685-688). */
typedef struct {
    double upper;          /* upperBoundUnWantedSpectrum          */
    double lower_unwanted; /* lowerBoundUnWantedSpectrum          */
    double lower_wanted;   /* lowerBoundWantedSpectrum            */
    double err_upper;      /* error_in_upper_bound                */
    double occ_energy;     /* occupied_orbital_energy             */
    int32_t passes;        /* out: filter passes in the last call */
    int32_t lanczos_runs;  /* out: Lanczos calls in the last call */
    double occ_tol;        /* in: relative change of eig[occ_index] that ends
                              the restart loop (1e-3 from the init, the value
                              of main.cc:1130-1145); no counterpart field */
} ttdft_chfsi_state;

/* Fills the initial state of main.cc:733-734 and :685.  */
int ttdft_chfsi_state_init(ttdft_chfsi_state *st);

/* One SCF iteration's eigensolve on a device-resident block X:
 * bound update (Lanczos when err_upper > 0.01), restart loop of
 * filter -> orth -> Rayleigh-Ritz with the stop test on eig[occ_index]
 * (relative change < st->occ_tol, 1e-3 as main.cc uses), then the rotation
 * X <- X Q.
 * evals: host array of k (previous iteration in, new out).
 * precision: TTDFT_F64 (X FP64, FP64 GEMMs) or TTDFT_F32 (X FP32, FP32
 * filter, split-BF16 GEMMs; Cholesky and Rayleigh-Ritz stay FP64).
 * replaces the ChFSI block of main.cc:977-1237. */
int ttdft_chfsi_step(ttdft_ctx ctx, ttdft_ham ham, void *X, int64_t ldx,
                     int32_t k, int32_t m, int32_t num_blocks,
                     int32_t restart_times, int32_t occ_index, int32_t scf_iter,
                     int32_t start_iter, int32_t maxit_lanczos, uint64_t seed,
                     int degree_convention, int precision, double *evals,
                     ttdft_chfsi_state *st);

/* Random starting subspace: X ~ N(0,1) from the seeded generator, then
 * Cholesky-QR twice.  replaces MatSetRandom + BVOrthogonalize
 * (main.cc:807-829); PETSc's random stream is not reproduced. */
int ttdft_random_subspace(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n,
                          int32_t k, uint64_t seed, int dtype);

/* ---- synthetic Hamiltonians (host) ------------------------------------------ */
/* Kronecker-banded symmetric test Hamiltonian of the reference's structure
 * (ProjectHamiltonianSparse.cc:464-548):
 *   H = 1/(2h^2) (T_x (x) I (x) I + I (x) T_y (x) I + I (x) I (x) T_z)
 *       + sum_r g_r P_x^r (x) P_y^r (x) P_z^r,
 * T the sinc-DVR kinetic stencil cut to half-bandwidth b, P^r banded
 * Gaussian-coupled potential factors, entries with |h| <= truncation
 * dropped; plus p s-type projectors (products of 1-D Gaussians over the
 * basis index) at seeded sites, truncated the same way, inv_c_lm = 1.
 * Two calls: first with NULL arrays to get nnz / nnz_c, then to fill. */
typedef struct {
    int32_t rx, ry, rz;  /* Tucker ranks per axis (N = rx*ry*rz)      */
    int32_t bx, by, bz;  /* half-bandwidths per axis                  */
    int32_t p;           /* projector count                           */
    double h;            /* grid spacing of the kinetic stencil (bohr) */
    double truncation;   /* drop |h_ij| <= truncation                 */
    uint64_t seed;
} ttdft_synth_params;

int ttdft_synth_hamiltonian(const ttdft_synth_params *prm, int64_t *nnz,
                            int64_t *nnz_c, int64_t *rowptr, int32_t *colind,
                            double *val, int64_t *c_rowptr, int32_t *c_colind,
                            double *c_val, double *inv_c_lm);

/* ---- Hamiltonian assembly (GPU) ------------------------------------------
 * ttdft_assemble_hloc replaces the host-side row loops of
 * dft/ProjectHamiltonianSparse.cc:330-575 (Create_Hloc) and :1223-1592
 * (Compute_nnz_loc):
 *   H[(xI,yI,zI),(x,y,z)] = 1/2 (DX Y Z + X DY Z + X Y DZ)
 *                          + sum_{pqr} G[p,q,r] PX[p][xI][x] PY[q][yI][y] PZ[r][zI][z],
 * row/column index x + rx (y + ry z), entries with |h| > truncation kept,
 * columns ascending.  X..DZ are row-major r x r (MatX[I][J]); PX is
 * [rvx][rx][rx] (MatPotX[p][I][J]); G is the Veff Tucker core with p fastest,
 * G[p + rvx (q + rvy r)].
 * ttdft_assemble_projectors replaces :83-328 (Create_Cnloc, rows :196-278):
 * row i is the Tucker reconstruction of projector i, core cores_i (r_i^3,
 * first index fastest) and column-major factors mx_i (rx x r_i), my_i, mz_i,
 * each concatenated over i, truncated the same way.
 * Both are called twice: with rowptr NULL to get *nnz, then with host arrays
 * rowptr[n+1] (p+1), col[nnz], val[nnz] to fill.  Inputs and outputs are host
 * memory; the work runs on the context's device. */
int ttdft_assemble_hloc(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz, int32_t rvx,
                        int32_t rvy, int32_t rvz, const double *X, const double *DX,
                        const double *Y, const double *DY, const double *Z, const double *DZ,
                        const double *PX, const double *PY, const double *PZ, const double *G,
                        double truncation, int64_t *nnz, int64_t *rowptr, int32_t *col,
                        double *val);
int ttdft_assemble_projectors(ttdft_ctx ctx, int32_t p, const int32_t *ranks, int32_t rx,
                              int32_t ry, int32_t rz, const double *cores, const double *mx,
                              const double *my, const double *mz, double truncation,
                              int64_t *nnz, int64_t *rowptr, int32_t *col, double *val);

/* Assemble H = H_loc + C^T D C directly into device memory: ttdft_assemble_hloc
 * and ttdft_assemble_projectors (same inputs and truncation rules) feeding
 * ttdft_ham_upload_tucker, without the host round trip.  H_loc stays in HBM
 * as CSR (and as Kronecker band fragments when its band is narrow enough);
 * D = diag(inv_c_lm) scales C as in ProjectHamiltonianSparse.cc:316-325.
 * replaces Create_Hloc + Create_Cnloc (ProjectHamiltonianSparse.cc:83-575)
 * followed by AX::setup_system (MultCuda.cu:105-371). */
int ttdft_ham_assemble_tucker(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz, int32_t rvx,
                              int32_t rvy, int32_t rvz, const double *X, const double *DX,
                              const double *Y, const double *DY, const double *Z,
                              const double *DZ, const double *PX, const double *PY,
                              const double *PZ, const double *G, double truncation, int32_t p,
                              const int32_t *ranks, const double *cores, const double *mx,
                              const double *my, const double *mz, double c_truncation,
                              const double *inv_c_lm, int copies, ttdft_ham *out);

/* ---- Anderson density mixing (history in device memory) --------------------
 * One mixer = the AndersonMixing object of main.cc:698-707: densities on the
 * quadrature grid (n_grid values) and at the FEM nodes (n_nodal, may be 0),
 * device arrays.  The history stays in device memory (no files); restart from
 * files is not provided.  jacobian and weights are grid fields (jacob3DMat,
 * weight3DMat) copied at creation; max_history <= 64.
 * replaces AndersonMixing::AndersonMixing (dft/AndersonMixing.cc:51-147) */
int ttdft_mixer_create(ttdft_ctx ctx, int64_t n_grid, int64_t n_nodal, const double *d_jacob,
                       const double *d_weight, double alpha, int32_t max_history,
                       ttdft_mixer *out);
int ttdft_mixer_destroy(ttdft_mixer m);
/* replaces AndersonMixing::updateRho (:376-443): appends copies of the four
 * densities, dropping the oldest entry beyond max_history. */
int ttdft_mixer_update(ttdft_mixer m, const double *d_rho_nodal_in, const double *d_rho_grid_in,
                       const double *d_rho_nodal_out, const double *d_rho_grid_out,
                       int32_t scf_iter);
/* replaces AndersonMixing::computeRhoIn(scfIter, rhoNodalIn, rhoGridIn)
 * (:325-374, with computeMixingConstants :150-262 and the mix :264-323):
 * overwrites both densities in place. */
int ttdft_mixer_compute_rho_in(ttdft_mixer m, int32_t scf_iter, double *d_rho_nodal_in,
                               double *d_rho_grid_in);
/* The constants of the last Anderson step (count = history size - 1). */
int ttdft_mixer_constants(ttdft_mixer m, double *constants, int32_t *count);
/* replaces AndersonMixing::clearHistory (:445-461) */
int ttdft_mixer_clear_history(ttdft_mixer m, int32_t scf_iter);
int ttdft_mixer_history_size(ttdft_mixer m, int32_t *size);

/* ---- Tucker recompression and Hartree potential ------------------------------
 * ttdft_sthosvd: fixed-rank sequentially truncated HOSVD of a device tensor T
 * (n0 x n1 x n2, first index fastest), modes 0, 1, 2 in order, the algorithm
 * of TuckerMPI::STHOSVD that tensor/TensorUtils.cc:245-336 calls on the nodal
 * density.  Outputs (device): core r0 x r1 x r2 (first index fastest),
 * factors U_n (n_n x r_n, column-major; the largest |entry| of each column is
 * positive).  evals (host, r0 + r1 + r2, may be NULL): the kept Gram
 * eigenvalues per mode, descending.  The interpolation of the factors from
 * FEM nodes to quadrature points (TensorUtils.cc:282-318) is the caller's.
 * ttdft_hartree_expansion replaces KSDFTPotential::computeHartreePotNew
 * (dft/KSDFTPotential.cc:269-340, with :440-668):
 *   V (+)= sum_i omega_i  core x_0 F0_i x_1 F1_i x_2 F2_i,
 *   F_d,i[q][a] = sum_e exp(-alpha_i (xq_d[q] - xe_d[e])^2) U_d[e][a] jw_d[e]
 * with U_d (ne_d x r_d, column-major) at the electrostatic quadrature points
 * xe_d (jacobian x weight jw_d), targets xq_d (p_d points), V p0 x p1 x p2
 * first index fastest.  alpha / omega are host arrays used as given (the
 * reference scales them by 1/A^2 and 1/A after reading, :281-284).  add = 0
 * overwrites V (INSERT_VALUES), 1 accumulates.  All other arrays are device
 * memory; the [3] arrays are host arrays of per-axis values or pointers. */
int ttdft_sthosvd(ttdft_ctx ctx, const double *d_T, int32_t n0, int32_t n1, int32_t n2,
                  int32_t r0, int32_t r1, int32_t r2, double *d_core, double *d_U0, double *d_U1,
                  double *d_U2, double *evals);
/* The same decomposition from the factors of a nearby tensor (d_U0..2 in/out,
 * orthonormal columns on input, e.g. the previous SCF iteration's): for each
 * mode, a block of min(r + 16, 64) vectors starting from the old factors is
 * multiplied by the mode Gram matrix iters times (orthonormalised after each
 * product) and Rayleigh-Ritz picks the leading r.  Repeated calls on a fixed
 * tensor converge to ttdft_sthosvd's factors; a mode with r >= 64 or
 * n min(r + 16, 64) > 18432 uses the full eigensolve instead.  No
 * reference counterpart: it replaces the per-iteration TuckerMPI call of
 * main.cc:882-899 inside an SCF loop. */
int ttdft_sthosvd_refine(ttdft_ctx ctx, const double *d_T, int32_t n0, int32_t n1, int32_t n2,
                         int32_t r0, int32_t r1, int32_t r2, double *d_core, double *d_U0,
                         double *d_U1, double *d_U2, int32_t iters, double *evals);
int ttdft_hartree_expansion(ttdft_ctx ctx, const int32_t r[3], const double *d_core,
                            const double *const d_U[3], const int32_t ne[3],
                            const double *const d_xe[3], const double *const d_jw[3],
                            const int32_t p[3], const double *const d_xq[3], int32_t nterms,
                            const double *alpha, const double *omega, double *d_V, int add);

/* ---- one-dimensional finite elements and input files -------------------------- */
/* Mesh types of FEM::generateNodes (fem/FEM.cc:166-375).  "adaptive2"
 * (:219-254) is not offered: its construction writes and reads one entry past
 * the end of its own element array. */
enum {
    TTDFT_MESH_UNIFORM = 0,
    TTDFT_MESH_ADAPTIVE = 1,
    TTDFT_MESH_MANUAL1 = 3,
    TTDFT_MESH_HARTREE_OUTER = 4
};
/* n-point Gauss-Legendre rule on [-1, 1], points ascending (any n in 1..256). */
int ttdft_fem_gauss_legendre(int32_t n, double *xi, double *w);
/* The "<n>pt" rule of QuadratureRule::QuadratureRule (fem/QuadratureRule.cc:
 * 25-560), 2 <= n <= 20, with the points in the order its tables list them. */
int ttdft_fem_quadrature_rule(int32_t n, double *xi, double *w);
/* Global node coordinates of nel elements of npe nodes, nel (npe - 1) + 1 of
 * them (FEM::generateNumberNodes and FEM::generateNodes, fem/FEM.cc:159-375).
 * inner_size, inner_h and coarsening are used by the non-uniform types only;
 * the caller adjusts start and end as main.cc:128-151 does. */
int ttdft_fem_mesh(int32_t type, int32_t nel, int32_t npe, double start, double end,
                   double inner_size, double inner_h, double coarsening, double *nodes);
/* Values and xi-derivatives at the quadrature points of nfun nodal fields
 * (d_nodal: nfun x nnodes, one field per row; d_val, d_dval: nfun x (nel nq),
 * device), plus the per-point weights jac w and w / jac and the positions
 * (host, nel nq each, any may be NULL).  xi, w: the rule (host); nodes: the
 * mesh (host).  replaces FEM::computeFieldAndDiffFieldAtAllQuadPoints
 * (fem/FEM.cc:897-946), FEM::generateShapeFunction / ...Gradient (:418-528)
 * and FEM::generateQuadPointData (:552-571). */
int ttdft_fem_interpolate(ttdft_ctx ctx, int32_t nel, int32_t npe, int32_t nq, const double *xi,
                          const double *w, const double *nodes, int32_t nfun,
                          const double *d_nodal, double *d_val, double *d_dval, double *jw,
                          double *inv_jw, double *xq);
/* One axis of the projected Hamiltonian's 1-D matrices from r basis functions
 * at nq quadrature points (d_B, d_DB: r x nq row-major = one function per row;
 * d_jw, d_inv_jw: nq; d_U: nq x rv column-major Veff factor; all device):
 *   MatX[I][J]      = sum_q B_I B_J jw
 *   MatDX[I][J]     = sum_q B'_I B'_J inv_jw
 *   MatPot[p][I][J] = sum_q U_p B_I B_J jw       (r x r each, row-major)
 * exactly symmetric (upper triangle mirrored); NULL outputs are skipped.
 * replaces ProjectHamiltonianSparse::computeOverlapKineticTuckerBasis and
 * computeOverlapPotentialTuckerBasis (dft/ProjectHamiltonianSparse.cc:577-722). */
int ttdft_tucker_basis_matrices(ttdft_ctx ctx, int32_t nq, int32_t r, const double *d_B,
                                const double *d_DB, const double *d_jw, const double *d_inv_jw,
                                int32_t rv, const double *d_U, double *d_MatX, double *d_MatDX,
                                double *d_MatPot);
/* Basis functions at the interior nodes, one per line, padded with a zero at
 * both ends: nfun x nnodes row-major (values may be NULL to get the sizes).
 * replaces BasisFunctionReader::copy_basis_function_from_file
 * (basis/BasisFunctionReader.h:35-52). */
int ttdft_read_basis_file(const char *path, int32_t *nfun, int32_t *nnodes, double *values);
/* Atom file: "ntype natoms nelectrons", the lmax of each type, then one line
 * "Z, x, y, z, _, Zatomic, type" per atom.  atoms: natoms x 7, grouped by
 * type in type order as all_nuclei (atoms/AtomInformation.cc:28-88); lmax:
 * ntype entries; either may be NULL. */
int ttdft_read_atom_file(const char *path, int32_t *ntype, int32_t *natoms, int32_t *nelec,
                         int32_t *lmax, double *atoms);
/* Whitespace-separated numeric table (pseudopotential tables nlpV_*,
 * nlpWaveFun_*, Density_*, locPotential_*): nrows x ncols row-major. */
int ttdft_read_table(const char *path, int32_t *nrows, int32_t *ncols, double *values);

/* ---- tabulated pseudopotentials --------------------------------------------- */
/* Cubic spline through (x_i, y_i), x strictly ascending, whose first and last
 * segments are parabolas (S''' = 0 there): alglib spline1dbuildcubic with no
 * end condition given, as atoms/NonLocalPSPData.cc:382-502 calls it.  coef:
 * 4 (n - 1) values, segment s is c0 + c1 t + c2 t^2 + c3 t^3, t = x - x_s;
 * beyond the table the end segments extend.  Host only. */
int ttdft_spline_parabolic(int32_t n, const double *x, const double *y, double *coef);
/* Local pseudopotential on a tensor grid (gx, gy, gz: device coordinates,
 * d_v: px py pz device values, x fastest; written, or added to when add != 0):
 *   V(q) = sum_a  d_a > 75.001 ? -Z_a / d_a : vloc(|q - R_a|),  d_a = |q - R_a| + 1e-12,
 * vloc the spline above through the table (r, vloc; host, nr rows).  atoms:
 * natoms rows of 7 as ttdft_read_atom_file writes them (Z, x, y, z, ...).
 * replaces NonLocalPSPData::computeLocalPart (atoms/NonLocalPSPData.cc:214-275)
 * with createSplineObject (:382-408). */
int ttdft_psp_local(ttdft_ctx ctx, int32_t natoms, const double *atoms, int32_t nr,
                    const double *r, const double *vloc, int32_t px, int32_t py, int32_t pz,
                    const double *d_gx, const double *d_gy, const double *d_gz, double *d_v,
                    int add);
/* Projector data of one atom type on an atom-centred tensor grid (gx, gy, gz:
 * device coordinates relative to the atom).  Tables (host, nr rows sharing the
 * ascending radii r): vloc, vl (lmax rows: V_l), rl (lmax rows: R_l).
 *   dV_l = spline(V_l - V_loc), the difference formed on the table (:441-463);
 *   phi_lm = spline(R_l) Y_lm, real harmonics of polar angle acos(z/(r+1e-12))
 *   and azimuth atan2(y, x): m < 0: sqrt2 Im Y_l^|m|, m > 0: sqrt2 Re Y_l^m,
 *   m = 0: Y_l^0, Condon-Shortley phase included (:527-604);
 *   fields f = l^2 + l + m (l = 0..lmax-1, m = -l..l), each px py pz, x fastest.
 * d_field (device, lmax^2 fields; may be NULL): dV_l phi_lm.
 * d_phi (device, lmax^2 fields; may be NULL): phi_lm.
 * clm (host, lmax^2; may be NULL): C_lm = sum phi_lm^2 dV_l wx wy wz with the
 *   device weights d_wx, d_wy, d_wz (jac w of each axis); the reference prints
 *   1 / C_lm and scales the projections by it.  The sum is exact on the
 *   quadrature grid where the reference integrates a rank-rankNloc Tucker
 *   form of the same field.
 * 1 <= lmax <= 4.  replaces the constructor NonLocalPSPData::NonLocalPSPData
 * (:33-212) up to its Tucker decompositions, computePhi_lm (:527-604) and
 * computeC_lm (:609-645). */
int ttdft_psp_projectors(ttdft_ctx ctx, int32_t nr, const double *r, const double *vloc,
                         const double *vl, const double *rl, int32_t lmax, int32_t px,
                         int32_t py, int32_t pz, const double *d_gx, const double *d_gy,
                         const double *d_gz, const double *d_wx, const double *d_wy,
                         const double *d_wz, double *d_field, double *d_phi, double *clm);

/* ---- restart files ----------------------------------------------------------- */
/* An X block (device, basis-major, ld >= k, TTDFT_F64 or TTDFT_F32) to or
 * from a file of its n k values in the reference's column-major order (the n
 * coefficients of state 0, then of state 1, ...) as big-endian FP64 with no
 * header: the files PetscBinaryWrite / PetscBinaryRead of PETSC_DOUBLE data
 * write and read from a MATDENSE array (wavefunctions wfn_proc<rank>,
 * main.cc:748-790, 1146-1170).  A grid field is the case k = 1, ldx = 1 (local
 * part potloc_proc<rank>, :541-592).  A file of another size is an error. */
int ttdft_restart_write(ttdft_ctx ctx, const char *path, const void *dX, int64_t ldx, int dtype,
                        int64_t n, int32_t k);
int ttdft_restart_read(ttdft_ctx ctx, const char *path, void *dX, int64_t ldx, int dtype,
                       int64_t n, int32_t k);
/* cheby_bounds.txt: upper, lower_unwanted, lower_wanted, err_upper, occ_energy
 * and the total energy, tab-separated, %.16e (main.cc:1150-1157, read back at
 * :755-760).  Reading sets those five state fields and *energy (may be NULL). */
int ttdft_restart_write_bounds(const char *path, const ttdft_chfsi_state *st, double energy);
int ttdft_restart_read_bounds(const char *path, ttdft_chfsi_state *st, double *energy);

/* ---- effective potential and grid utilities of the SCF loop ------------------ */
enum {
    TTDFT_LDA_PZ81 = 0,          /* Slater exchange + Perdew-Zunger 1981 correlation   */
    TTDFT_LDA_PZ81_REFERENCE = 1 /* the same, with the reference's 1/33 in place of 1/3
                                    in the high-density correlation potential
                                    (dft/KSDFTPotential.cc:262)                       */
};
/* V_xc(rho) (d_vxc, written or added to) and e_xc(rho), the energy per
 * particle (d_exc, written; may be NULL), pointwise; rho <= 1e-30 gives 0.
 * replaces KSDFTPotential::computeLDAExchangePot and computeLDACorrelationPot
 * (dft/KSDFTPotential.cc:127-267). */
int ttdft_lda_xc(ttdft_ctx ctx, const double *d_rho, int64_t n, double *d_vxc, double *d_exc,
                 int add, int variant);
/* V(q) = sum_a -Z_a erf(|q - R_a| / rc) / |q - R_a| on the px x py x pz tensor
 * grid of 1-D coordinates (device); atoms: host natoms x 4 (x, y, z, Z).  The
 * local pseudopotential of the synthetic clusters; it takes the place of
 * KSDFTPotential::computeEvanescentPSPOnGrid (dft/KSDFTPotential.cc:342-392). */
int ttdft_local_pseudopotential(ttdft_ctx ctx, int32_t natoms, const double *atoms, double rc,
                                int32_t px, int32_t py, int32_t pz, const double *d_gx,
                                const double *d_gy, const double *d_gz, double *d_V, int add);
/* w[q] = wx[qx] wy[qy] wz[qz] (the 3-D jacobian x weight tensor of main.cc:698-707). */
int ttdft_grid_weights(ttdft_ctx ctx, int32_t px, int32_t py, int32_t pz, const double *d_wx,
                       const double *d_wy, const double *d_wz, double *d_w);
/* sum f g wx wy wz (energy integrals of KSDFTEnergyFunctional, dft/
 * KSDFTEnergyFunctional.cc:165-192, for a product of two fields). */
int ttdft_grid_dot(ttdft_ctx ctx, const double *d_f, const double *d_g, int32_t px, int32_t py,
                   int32_t pz, const double *d_wx, const double *d_wy, const double *d_wz,
                   double *result);
/* Trilinear interpolation of a uniform-grid field (n[0] x n[1] x n[2], x
 * fastest, point i at origin + i h) onto a tensor grid; points outside the
 * uniform box get 0.  add = 1 accumulates.  For the learned initial density
 * (no reference counterpart). */
int ttdft_grid_resample(ttdft_ctx ctx, const double *d_src, const int32_t n[3],
                        const double origin[3], const double h[3], int32_t px, int32_t py,
                        int32_t pz, const double *d_gx, const double *d_gy, const double *d_gz,
                        double *d_dst, int add);
/* out = sum_a amp_a exp(-|r - R_a|^2 / (2 sigma^2)) on a uniform grid; atoms:
 * host natoms x 4 (x, y, z, amp).  The network input of the learned initial
 * density (no reference counterpart). */
int ttdft_gaussian_splat(ttdft_ctx ctx, int32_t natoms, const double *atoms, double sigma,
                         const int32_t n[3], const double origin[3], const double h[3],
                         double *d_out);
/* y = a y + b x in place (x may be NULL when b = 0). */
int ttdft_grid_axpby(ttdft_ctx ctx, double *d_y, double a, const double *d_x, double b,
                     int64_t n);
/* f = max(f, lo) in place. */
int ttdft_grid_clamp(ttdft_ctx ctx, double *d_f, int64_t n, double lo);

/* ---- layers of the learned initial-density network ----------------------------
 * No reference counterpart: the network predicts a correction to the atomic
 * superposition of InitialGuess::initialize_rho (dft/InitialGuess.cc:683-762).
 * Tensors: one sample, [C][D][H][W] FP32 on the device (x fastest); dims =
 * {D, H, W}.  Convolutions: kernel ks = 1 or 3, stride 1, zero padding ks/2,
 * weights [co][ci][ks^3] (taps z-major), bias [co] (may be NULL).
 * accumulate = 1 adds into the output instead of writing it. */
/* Arithmetic of the convolutions on this context: FP32 vector math, or BF16
 * MFMA (operands rounded to BF16, FP32 accumulation).  Tensors stay FP32. */
enum { TTDFT_NN_FP32 = 0, TTDFT_NN_BF16 = 1 };
int ttdft_nn_set_math(ttdft_ctx ctx, int math);
int ttdft_nn_conv3d_forward(ttdft_ctx ctx, const float *d_x, int32_t ci, const int32_t dims[3],
                            const float *d_w, const float *d_b, int32_t co, int32_t ks,
                            float *d_y, int accumulate);
int ttdft_nn_conv3d_backward_data(ttdft_ctx ctx, const float *d_dy, int32_t co,
                                  const int32_t dims[3], const float *d_w, int32_t ci,
                                  int32_t ks, float *d_dx, int accumulate);
/* dW (and db when not NULL) from the layer input and the output gradient. */
int ttdft_nn_conv3d_backward_weight(ttdft_ctx ctx, const float *d_x, int32_t ci,
                                    const int32_t dims[3], const float *d_dy, int32_t co,
                                    int32_t ks, float *d_dw, float *d_db, int accumulate);
/* y = max(y, slope y) in place; backward: dy *= (y > 0 ? 1 : slope). */
int ttdft_nn_leaky_relu(ttdft_ctx ctx, float *d_y, int64_t n, float slope);
int ttdft_nn_leaky_relu_backward(ttdft_ctx ctx, const float *d_y, float *d_dy, int64_t n,
                                 float slope);
/* 2x2x2 average pooling and nearest 2x upsampling; dims are the fine grid's. */
int ttdft_nn_avgpool2(ttdft_ctx ctx, const float *d_fine, int32_t c, const int32_t fine_dims[3],
                      float *d_coarse);
int ttdft_nn_avgpool2_backward(ttdft_ctx ctx, const float *d_dcoarse, int32_t c,
                               const int32_t fine_dims[3], float *d_dfine, int accumulate);
int ttdft_nn_upsample2(ttdft_ctx ctx, const float *d_coarse, int32_t c,
                       const int32_t fine_dims[3], float *d_fine);
int ttdft_nn_upsample2_backward(ttdft_ctx ctx, const float *d_dfine, int32_t c,
                                const int32_t fine_dims[3], float *d_dcoarse, int accumulate);
/* loss = mean (y - t)^2 (host); dy = 2 (y - t) / n when not NULL. */
int ttdft_nn_mse(ttdft_ctx ctx, const float *d_y, const float *d_t, int64_t n, float *d_dy,
                 double *loss);
/* One Adam step (step >= 1 counts from 1) over n parameters. */
int ttdft_nn_adam(ttdft_ctx ctx, float *d_p, const float *d_g, float *d_m, float *d_v, int64_t n,
                  float lr, float beta1, float beta2, float eps, int32_t step);
/* dst = scale src, FP64 -> FP32 and FP32 -> FP64. */
int ttdft_nn_to_f32(ttdft_ctx ctx, const double *d_src, float *d_dst, int64_t n, double scale);
int ttdft_nn_to_f64(ttdft_ctx ctx, const float *d_src, double *d_dst, int64_t n, double scale);

/* Host random streams (same generator the engine uses internally). */
int ttdft_rng_uniform(uint64_t seed, int64_t count, double *out);
int ttdft_rng_normal(uint64_t seed, int64_t count, double *out);

#ifdef __cplusplus
}
#endif

#endif /* TTDFT_CHFSI_H */
