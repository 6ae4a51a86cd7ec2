// Internal declarations of the MI355X ChFSI engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <map>
#include <unordered_map>
#include <vector>

#include "../../include/ttdft_chfsi.h"

namespace ttdft {

// ---- error reporting ------------------------------------------------------
void set_error(const char *fmt, ...);
void clear_error();

#define TT_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::ttdft::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,         \
                               hipGetErrorString(_e));                            \
            return TTDFT_EHIP;                                                    \
        }                                                                         \
    } while (0)

#define TT_CHECK(cond, code, ...)                                                 \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::ttdft::set_error(__VA_ARGS__);                                      \
            return (code);                                                        \
        }                                                                         \
    } while (0)

#define TT_TRY(expr)                                                              \
    do {                                                                          \
        int _rc = (expr);                                                         \
        if (_rc != TTDFT_OK) return _rc;                                          \
    } while (0)

#define TT_LAUNCH_CHECK()                                                         \
    do {                                                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) {                                                   \
            ::ttdft::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,     \
                               hipGetErrorString(_e));                            \
            return TTDFT_EHIP;                                                    \
        }                                                                         \
    } while (0)

// ---- grow-only device scratch buffer ---------------------------------------
// owner: the context whose released blocks are given back to the device
// (dev_cache_clear) and the allocation retried when hipMalloc fails.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    ttdft_ctx owner = nullptr;
    int ensure(size_t need);
    void release();
};

// ---- kernel-class timing --------------------------------------------------
struct KStat {
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
    double flops = 0.0;
};

struct PendingTiming {
    int cls;
    hipEvent_t start, stop;
    double bytes, flops;
};

}  // namespace ttdft

struct ttdft_ctx_s {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    int profiling = 0;
    ttdft::KStat stats[TTDFT_KSTAT_COUNT];
    std::vector<ttdft::PendingTiming> *pending = nullptr;
    // scratch: [0] second X block for the filter / rotation, [1] small k x k
    // work, [2] Gram partials, [3] nonlocal projections, [4..7] misc, [8] partial
    // sums of the split dense product
    ttdft::Scratch scratch[9];
    // network-layer work: [0] transformed weights, [1] weight-gradient partials
    ttdft::Scratch nn_scratch[2];
    int nn_math = 0;  // TTDFT_NN_FP32 or TTDFT_NN_BF16 for the network's convolutions
    int *d_info = nullptr;  // rocSOLVER info
    // device blocks handed out by dev_alloc (pointer -> size class) and the
    // released ones kept for reuse on this context's stream (size class ->
    // (release number, block)); see dev_alloc
    std::unordered_map<void *, size_t> blk_size;
    std::multimap<size_t, std::pair<uint64_t, void *>> blk_free;
    size_t blk_free_bytes = 0;
    uint64_t blk_seq = 0;
    // Hamiltonians and mixers created on this context and not yet freed.
    // ttdft_ctx_destroy with any alive only marks the context; the last
    // ttdft_ham_free / ttdft_mixer_destroy then destroys it.
    int live_children = 0;
    bool destroy_pending = false;
};

struct ttdft_ham_s {
    ttdft_ctx ctx = nullptr;
    int64_t n = 0, nnz = 0;
    int32_t p = 0;
    int64_t nnz_c = 0;
    int copies = 0;
    int32_t max_row_nnz = 0;
    // H_loc CSR
    int64_t *rowptr = nullptr;
    int32_t *col = nullptr;
    double *val64 = nullptr;
    float *val32 = nullptr;
    // D C (row-scaled projector rows, p x n) and C^T (n x p, unscaled)
    int64_t *c_rowptr = nullptr;
    int32_t *c_col = nullptr;
    double *c_val64 = nullptr;
    float *c_val32 = nullptr;
    int64_t *ct_rowptr = nullptr;
    int32_t *ct_col = nullptr;
    double *ct_val64 = nullptr;
    float *ct_val32 = nullptr;
    // Kronecker band storage of H_loc (format 1): 16x16x4 MFMA A fragments of
    // 4x4 row bricks against each neighbouring x-line (layout in band.hip)
    int format = 0;
    int32_t rx = 0, ry = 0, rz = 0;
    int32_t bx = 0, by = 0, bz = 0;
    int32_t ks = 0;             // 4-column MFMA steps per x' window
    int32_t gxn = 0, hyn = 0;   // x-bricks, y-bricks
    int32_t gxb = 0, nxc = 0;   // x-bricks per workgroup, x-chunks
    double *mf64 = nullptr;
    float *mf32 = nullptr;
    // dense storage of H_loc (format 2) for small, nearly full matrices:
    // row-major, leading dimension dense_ld (n rounded up to 32), rows padded
    // to a multiple of 16, padding zero
    double *dense64 = nullptr;
    float *dense32 = nullptr;
    int64_t dense_ld = 0;
    bool dense_nl = false;  // C^T D C is folded into the dense storage
};

namespace ttdft {

// ---- timing helpers (no-ops unless ctx->profiling) --------------------------
int timing_begin(ttdft_ctx ctx, int cls, PendingTiming *t);
int timing_end(ttdft_ctx ctx, PendingTiming *t, double bytes, double flops);
int timing_flush(ttdft_ctx ctx);

// ---- kernels (launchers return TTDFT_* status) -------------------------------
// W[p x k] = (D C) X  (nonlocal projection; D folded into c_val)
// W needs nl_workspace(H, k) elements (W itself, then partial sums).
int launch_nl_project(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                      int32_t k, int dtype, void *W);
int64_t nl_workspace(const ttdft_ham_s *H, int32_t k);

// Kronecker band storage of H_loc (format 1) from host CSR, or from the device
// CSR already in H (H->rowptr / col / val64) with known half-bandwidths;
// either keeps CSR only when the band is too wide or too sparse.
int build_band(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz,
               const int64_t *rowptr, const int32_t *colind, const double *val, int copies);
int build_band_device(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz, int bx,
                      int by, int bz, int copies);
// Y = alpha*(H_loc X + C^T W) + beta*X + delta*Z   (Z may alias Y; flags select terms)
// use_beta / use_delta select whether X / Z are read.
// One-workgroup k x k factorisations for k <= 64 (small_dense.hip)
bool small_dense_enabled(int32_t k);
int small_chol_inv(ttdft_ctx ctx, const double *S, int32_t k, double *QT);
int small_syev(ttdft_ctx ctx, const double *A, int32_t k, double *evals, double *Q);
// Orthonormal basis of an n x b column-major block in place (n b <= 18432, b <= 64)
bool small_orth_fits(int32_t n, int32_t b);
// f[i] = lo + (hi - lo) u_i, seeded hash (misc.hip)
int dev_fill_uniform(ttdft_ctx ctx, double *f, int64_t n, uint64_t seed, double lo, double hi);
int small_orth(ttdft_ctx ctx, double *Y, int32_t n, int32_t b);
// Dense storage (format 2) of H_loc when it is nearly full and small enough
// (no-op otherwise); drop_dense frees it.
int maybe_build_dense(ttdft_ctx ctx, ttdft_ham_s *H, int copies);
int finish_dense(ttdft_ctx ctx, ttdft_ham_s *H, int copies);
// Device blocks that live and die with Hamiltonians, rebuilt every SCF
// iteration at multi-GB sizes: dev_alloc takes a released block of the same
// size class (sizes rounded up by at most 1/8) or calls hipMalloc;
// dev_release returns a dev_alloc block to the context (reused in the order
// of its stream, so no synchronisation is needed) and hipFrees any other
// pointer.  dev_cache_clear frees the released blocks.
int dev_alloc(ttdft_ctx ctx, void **p, size_t bytes);
void dev_release(ttdft_ctx ctx, void *p);
void dev_cache_clear(ttdft_ctx ctx);
// hipMalloc that, when it fails, gives the context's released blocks back to
// the device and tries once more (plain, uncached memory: free with hipFree)
int dev_malloc_retry(ttdft_ctx ctx, void **p, size_t bytes);
// A Hamiltonian or mixer was created on / released from ctx (the release
// may complete a deferred ttdft_ctx_destroy, after which ctx is gone).
void ctx_attach(ttdft_ctx ctx);
void ctx_detach(ttdft_ctx ctx);
// Frees the STHOSVD / Hartree work blocks released on (device, stream).
void free_tucker_blocks(int device, hipStream_t stream);
void drop_dense(ttdft_ham_s *H);
int launch_spmm_fused(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                      const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                      int32_t k, int dtype, double alpha, double beta, double delta,
                      int use_beta, int use_delta);
// y = H x for one FP64 vector (Lanczos): shuffle-reduced row SpMV.
int ham_spmv(ttdft_ctx ctx, const ttdft_ham_s *H, const double *x, double *y);
// Band-format SpMM with the same epilogue (format 1 Hamiltonians).
int launch_band_spmm(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                     const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                     int32_t k, int dtype, double alpha, double beta, double delta,
                     int use_beta, int use_delta);
// Full H application (nl projection + fused SpMM) with the given epilogue.
// The projections go to ws (ws_bytes long; an error if they do not fit) when
// ws is given, else to the context's scratch.
int ham_step(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx, void *Y,
             int64_t ldy, const void *Z, int64_t ldz, int32_t k, int dtype, double alpha,
             double beta, double delta, int use_beta, int use_delta, void *ws = nullptr,
             size_t ws_bytes = 0);

// S = X^T Y, k x k column-major FP64 (device)
int launch_gram(ttdft_ctx ctx, const void *X, int64_t ldx, const void *Y, int64_t ldy,
                int64_t n, int32_t k, int dtype, int mode, double *dS);
// Z = X Q (Z separate from X), Q k x k column-major FP64
int launch_rotate(ttdft_ctx ctx, const void *X, int64_t ldx, void *Z, int64_t ldz,
                  int64_t n, int32_t k, const double *dQ, int dtype, int mode);

// small helpers
int launch_copy_block(ttdft_ctx ctx, const void *src, int64_t lds, int src_dtype, void *dst,
                      int64_t ldd, int dst_dtype, int64_t n, int32_t k);
int launch_transpose_small(ttdft_ctx ctx, const double *A, double *AT, int32_t k);

// host random streams (splitmix64 + xoshiro256**)
void rng_uniform(uint64_t seed, int64_t count, double *out);
void rng_normal(uint64_t seed, int64_t count, double *out);

// host small symmetric eigensolver (cyclic Jacobi), ascending, for Lanczos T
int host_syev_small(int k, const double *A, double *evals);

inline size_t dtype_size(int dtype) { return dtype == TTDFT_F32 ? 4 : 8; }

}  // namespace ttdft
