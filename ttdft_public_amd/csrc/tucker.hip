// Tucker recompression of the density and the Hartree potential by kernel
// expansion, on the device (tensor/TensorUtils.cc:245-336,
// dft/KSDFTPotential.cc:269-340,440-668).
//
// STHOSVD (fixed ranks, modes 0, 1, 2; the algorithm of the third-party
// TuckerMPI library the reference calls): for each mode the Gram matrix of
// the mode unfolding (the row-chunked MFMA Gram kernel for modes 0 and 1, a
// rocBLAS product for mode 2; modes 1 and 2 see the tensor already shrunk in
// the earlier modes), rocSOLVER dsyevd, the leading
// eigenvectors with the largest |entry| of each made positive, and the
// product that shrinks the tensor.  Tensors are first index fastest.
//
// Hartree potential: V = sum_i omega_i G x_0 F0_i x_1 F1_i x_2 F2_i with
// F_d,i = K_d,i U_d, K_d,i[q][e] = exp(-alpha_i (xq[q] - xe[e])^2) jw[e].
//  1. K for all terms of one axis as one (terms p) x ne matrix, F = K U: one
//     product per axis.
//  2. T1 = F0 G_(0) for all terms at once: ((terms p0) x r0) (r0 x r1 r2).
//  3. per term, batched over c: T2[:, :, c, i] = T1_i[:, :, c] F1_i^T.
//  4. V = T2 W^T with W[z][(c, i)] = omega_i F2_i[z][c]: one product of depth
//     r2 x terms (35 x 40 = 1400 at the workload), which carries 97% of the
//     flops on the FP64 MFMA.  Terms are chunked when T2 would pass 4 GB.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

#define TT_BLAS(call)                                                                     \
    do {                                                                                  \
        const rocblas_status st_ = (call);                                                \
        TT_CHECK(st_ == rocblas_status_success, TTDFT_ELIB, "%s: rocBLAS status %d", #call, \
                 (int)st_);                                                               \
    } while (0)

// Work buffers of STHOSVD and the Hartree expansion come from a cache of
// blocks by (device, stream, size): the SCF loop calls both with the same
// shapes every iteration, and hipMalloc / hipFree of hundreds of MB (hipFree
// synchronises the device) would otherwise dominate.  A returned block is
// reused only on the stream it was released from, in that stream's order (an
// early error return may release blocks that queued work still uses); the
// cache keeps at most 8 GiB per process.
struct BlockKey {
    int dev;
    hipStream_t stream;
    size_t sz;
    bool operator==(const BlockKey &o) const {
        return dev == o.dev && stream == o.stream && sz == o.sz;
    }
};
struct BlockCache {
    std::mutex mu;  // contexts may run on several host threads
    std::vector<std::pair<BlockKey, void *>> blocks;
    size_t bytes = 0;
    void *take(const BlockKey &k) {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < blocks.size(); ++i)
            if (blocks[i].first == k) {
                void *p = blocks[i].second;
                blocks.erase(blocks.begin() + i);
                bytes -= k.sz;
                return p;
            }
        return nullptr;
    }
    void give(const BlockKey &k, void *p) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (bytes + k.sz <= (size_t(8) << 30)) {
                blocks.push_back({k, p});
                bytes += k.sz;
                return;
            }
        }
        hipError_t e = hipFree(p);
        (void)e;
    }
    // frees every block released on (dev, stream); the caller has
    // synchronised that stream
    void drop(int dev, hipStream_t stream) {
        std::vector<void *> gone;
        {
            std::lock_guard<std::mutex> lk(mu);
            size_t w = 0;
            for (size_t i = 0; i < blocks.size(); ++i) {
                if (blocks[i].first.dev == dev && blocks[i].first.stream == stream) {
                    gone.push_back(blocks[i].second);
                    bytes -= blocks[i].first.sz;
                } else {
                    blocks[w++] = blocks[i];
                }
            }
            blocks.resize(w);
        }
        for (void *p : gone) {
            hipError_t e = hipFree(p);
            (void)e;
        }
    }
};
BlockCache g_blocks;
// the stream of the entry point running on this thread (set on entry)
thread_local hipStream_t g_stream = nullptr;

struct Dev {
    void *p = nullptr;
    BlockKey key{0, nullptr, 0};
    ~Dev() {
        if (p) g_blocks.give(key, p);
    }
    int alloc(size_t bytes) {
        key.sz = (std::max<size_t>(bytes, 8) + 255) / 256 * 256;
        key.stream = g_stream;
        hipError_t e = hipGetDevice(&key.dev);
        if (e != hipSuccess) key.dev = 0;
        const size_t sz = key.sz;
        p = g_blocks.take(key);
        if (p) return TTDFT_OK;
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return TTDFT_ENOMEM;
        }
        return TTDFT_OK;
    }
    double *d() const { return static_cast<double *>(p); }
};

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65535)); }

// U[:, j] = s_j Q[:, ncols - 1 - j] (leading eigenvectors of an ascending
// eigensolver result with ncols columns), s_j making the largest |entry| positive (first on ties).
__global__ __launch_bounds__(256) void leading_vectors_kernel(const double *__restrict__ Q,
                                                              int n, int ncols, int r,
                                                              double *__restrict__ U) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int j = blockIdx.x;
    const double *q = Q + (int64_t)n * (ncols - 1 - j);
    double best = -1.0;
    int bidx = n;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = fabs(q[i]);
        if (a > best) {  // strided scan keeps the first index of its maximum
            best = a;
            bidx = i;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double a = bv[threadIdx.x + o];
            const int ia = bi[threadIdx.x + o];
            if (a > bv[threadIdx.x] || (a == bv[threadIdx.x] && ia < bi[threadIdx.x])) {
                bv[threadIdx.x] = a;
                bi[threadIdx.x] = ia;
            }
        }
        __syncthreads();
    }
    const double s = q[bi[0]] < 0.0 ? -1.0 : 1.0;
    for (int i = threadIdx.x; i < n; i += 256) U[(int64_t)n * j + i] = s * q[i];
}

// out[i1 + n1 (a + r0 i2)] = in[a + r0 (i1 + n1 i2)] (swap of the two fastest modes)
__global__ __launch_bounds__(256) void swap01_kernel(const double *__restrict__ in, int d0,
                                                     int d1, int64_t d2,
                                                     double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)d0 * d1 * d2;
    for (int64_t e = t; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(e % d0);
        const int64_t rest = e / d0;
        const int i1 = (int)(rest % d1);
        const int64_t i2 = rest / d1;
        out[i1 + (int64_t)d1 * (a + (int64_t)d0 * i2)] = in[e];
    }
}

// K[(q + p i) + (terms p) e] = exp(-alpha_i (xq[q] - xe[e])^2) jw[e]
__global__ __launch_bounds__(256) void kernel_matrix_kernel(const double *__restrict__ xq, int p,
                                                            const double *__restrict__ xe,
                                                            const double *__restrict__ jw,
                                                            int ne, const double *__restrict__ alpha,
                                                            int nt, double *__restrict__ K) {
    const int64_t rows = (int64_t)nt * p, total = rows * ne;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = t % rows;
        const int e = (int)(t / rows);
        const int q = (int)(row % p), i = (int)(row / p);
        const double dx = xq[q] - xe[e];
        K[t] = exp(-alpha[i] * dx * dx) * jw[e];
    }
}

// W[z + p2 (c + r2 (i - i0))] = omega_i F2[(z + p2 i) + (terms p2) c], i in [i0, i0 + ni)
__global__ __launch_bounds__(256) void stack_last_kernel(const double *__restrict__ F2, int p2,
                                                         int r2, int nt, int i0, int ni,
                                                         const double *__restrict__ omega,
                                                         double *__restrict__ W) {
    const int64_t total = (int64_t)p2 * r2 * ni;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(t % p2);
        const int64_t ci = t / p2;
        const int c = (int)(ci % r2), il = (int)(ci / r2);
        const int i = i0 + il;
        W[t] = omega[i] * F2[(z + (int64_t)p2 * i) + (int64_t)nt * p2 * c];
    }
}

// Leading r eigenvectors of the n x n Gram matrix in S (destroyed), with
// their eigenvalues (descending) appended to ev.
int gram_eigvecs(ttdft_ctx ctx, double *S, int n, int r, double *U, std::vector<double> &ev) {
    Dev w, e;
    TT_TRY(w.alloc(sizeof(double) * n));
    TT_TRY(e.alloc(sizeof(double) * n));
    const rocblas_status st = rocsolver_dsyevd(ctx->blas, rocblas_evect_original,
                                               rocblas_fill_lower, n, S, n, w.d(), e.d(),
                                               ctx->d_info);
    TT_CHECK(st == rocblas_status_success, TTDFT_ELIB, "rocsolver_dsyevd: status %d", (int)st);
    hipLaunchKernelGGL(leading_vectors_kernel, dim3(r), dim3(256), 0, ctx->stream, S, n, n, r, U);
    TT_LAUNCH_CHECK();
    int info = 0;
    std::vector<double> wh(n);
    TT_HIP(hipMemcpyAsync(&info, ctx->d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipMemcpyAsync(wh.data(), w.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(info == 0, TTDFT_ENOCONV, "rocsolver_dsyevd did not converge (info %d)", info);
    for (int j = 0; j < r; ++j) ev.push_back(wh[n - 1 - j]);
    return TTDFT_OK;
}

// Leading r eigenvectors of the symmetric n x n G from the previous ones (U,
// in place): a block of b = min(r + 16, 64) vectors (U and seeded random
// columns) is orthonormalised, multiplied by G iters times with an
// orthonormalisation after each product, and Rayleigh-Ritz on the block gives
// the Ritz vectors (largest |entry| of each made positive) and values.  At a
// fixed G repeated calls converge to the exact leading eigenvectors; the
// oversampling makes each step contract by lambda_{b+1} / lambda_r.
int refine_mode(ttdft_ctx ctx, const double *G, int n, int r, int iters, double *U,
                std::vector<double> &ev) {
    const int b = std::min(n, std::min(64, r + 16));
    const double one = 1.0, zero = 0.0;
    Dev Y, Z, B, V, W;
    TT_TRY(Y.alloc(sizeof(double) * (size_t)n * b));
    TT_TRY(Z.alloc(sizeof(double) * (size_t)n * b));
    TT_TRY(B.alloc(sizeof(double) * (size_t)b * b));
    TT_TRY(V.alloc(sizeof(double) * (size_t)b * b));
    TT_TRY(W.alloc(sizeof(double) * (size_t)b));
    TT_HIP(hipMemcpyAsync(Y.d(), U, sizeof(double) * (size_t)n * r, hipMemcpyDeviceToDevice,
                          ctx->stream));
    if (b > r) TT_TRY(dev_fill_uniform(ctx, Y.d() + (size_t)n * r, (int64_t)n * (b - r), 977,
                                       -1.0, 1.0));
    TT_TRY(small_orth(ctx, Y.d(), n, b));
    for (int it = 0; it < iters; ++it) {
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, b, n,
                              &one, G, n, Y.d(), n, &zero, Z.d(), n));
        TT_TRY(small_orth(ctx, Z.d(), n, b));
        std::swap(Y.p, Z.p);
    }
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, b, n,
                          &one, G, n, Y.d(), n, &zero, Z.d(), n));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, b, b,
                          n, &one, Y.d(), n, Z.d(), n, &zero, B.d(), b));
    TT_TRY(small_syev(ctx, B.d(), b, W.d(), V.d()));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, n, b, b,
                          &one, Y.d(), n, V.d(), b, &zero, Z.d(), n));
    hipLaunchKernelGGL(leading_vectors_kernel, dim3(r), dim3(256), 0, ctx->stream, Z.d(), n, b,
                       r, U);
    TT_LAUNCH_CHECK();
    std::vector<double> wh(b);
    TT_HIP(hipMemcpyAsync(wh.data(), W.p, sizeof(double) * b, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < r; ++j) ev.push_back(wh[b - 1 - j]);
    return TTDFT_OK;
}

// the exact eigensolve, or refinement from the incoming factors when it fits
int mode_factors(ttdft_ctx ctx, double *G, int n, int r, int iters, double *U,
                 std::vector<double> &ev) {
    const int b = std::min(n, std::min(64, r + 16));
    if (iters > 0 && b > r && small_orth_fits(n, b)) return refine_mode(ctx, G, n, r, iters, U, ev);
    return gram_eigvecs(ctx, G, n, r, U, ev);
}

}  // namespace

void free_tucker_blocks(int device, hipStream_t stream) { g_blocks.drop(device, stream); }

}  // namespace ttdft

using namespace ttdft;

extern "C" {

static int sthosvd_impl(ttdft_ctx ctx, const double *d_T, int32_t n0, int32_t n1, int32_t n2,
                        int32_t r0, int32_t r1, int32_t r2, double *d_core, double *d_U0,
                        double *d_U1, double *d_U2, double *evals, int iters) {
    TT_CHECK(ctx && d_T && d_core && d_U0 && d_U1 && d_U2, TTDFT_EBADARG,
             "sthosvd: NULL argument");
    TT_CHECK(n0 > 0 && n1 > 0 && n2 > 0 && r0 > 0 && r1 > 0 && r2 > 0 && r0 <= n0 && r1 <= n1 &&
                 r2 <= n2,
             TTDFT_EBADARG, "sthosvd: ranks must lie in [1, n]");
    TT_CHECK((int64_t)n1 * n2 < ((int64_t)1 << 31), TTDFT_EBADARG, "sthosvd: n1 n2 too large");
    TT_HIP(hipSetDevice(ctx->device));
    PendingTiming tm;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_STHOSVD, &tm));
    const double one = 1.0, zero = 0.0;
    std::vector<double> ev;
    const int64_t n12 = (int64_t)n1 * n2;
    Dev S0, Y1, Z, S1, W, Y2, S2;
    // mode 0: S0 = T_(0) T_(0)^T, U0, Y1 = U0^T T_(0)  (r0 x n1 n2)
    TT_TRY(S0.alloc(sizeof(double) * n0 * n0));
    // T_(0) is the n12 x n0 row-major block with ld n0: its Gram on the
    // engine's row-chunked MFMA kernel (a library product here would run on
    // a handful of 300 x 300 output tiles with depth n12)
    TT_TRY(launch_gram(ctx, d_T, n0, d_T, n0, n12, n0, TTDFT_F64, TTDFT_GEMM_F64, S0.d()));
    TT_TRY(mode_factors(ctx, S0.d(), n0, r0, iters, d_U0, ev));
    TT_TRY(Y1.alloc(sizeof(double) * r0 * n12));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, r0,
                          (rocblas_int)n12, n0, &one, d_U0, n0, d_T, n0, &zero, Y1.d(), r0));
    // mode 1: Z = Y1 with modes 0, 1 swapped (n1 x r0 n2); S1 = Z Z^T; W = U1^T Z
    TT_TRY(Z.alloc(sizeof(double) * r0 * n12));
    hipLaunchKernelGGL(swap01_kernel, dim3(nblk((int64_t)r0 * n12)), dim3(256), 0, ctx->stream,
                       Y1.d(), r0, n1, (int64_t)n2, Z.d());
    TT_LAUNCH_CHECK();
    TT_TRY(S1.alloc(sizeof(double) * n1 * n1));
    TT_TRY(launch_gram(ctx, Z.d(), n1, Z.d(), n1, (int64_t)r0 * n2, n1, TTDFT_F64,
                       TTDFT_GEMM_F64, S1.d()));
    TT_TRY(mode_factors(ctx, S1.d(), n1, r1, iters, d_U1, ev));
    TT_TRY(W.alloc(sizeof(double) * (int64_t)r1 * r0 * n2));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, r1,
                          r0 * n2, n1, &one, d_U1, n1, Z.d(), n1, &zero, W.d(), r1));
    // back to [a + r0 (b + r1 i2)]: (r0 r1) x n2
    TT_TRY(Y2.alloc(sizeof(double) * (int64_t)r0 * r1 * n2));
    hipLaunchKernelGGL(swap01_kernel, dim3(nblk((int64_t)r0 * r1 * n2)), dim3(256), 0,
                       ctx->stream, W.d(), r1, r0, (int64_t)n2, Y2.d());
    TT_LAUNCH_CHECK();
    // mode 2: S2 = Y2^T Y2, U2, core = Y2 U2
    TT_TRY(S2.alloc(sizeof(double) * n2 * n2));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, n2, n2,
                          r0 * r1, &one, Y2.d(), r0 * r1, Y2.d(), r0 * r1, &zero, S2.d(), n2));
    TT_TRY(mode_factors(ctx, S2.d(), n2, r2, iters, d_U2, ev));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, r0 * r1, r2,
                          n2, &one, Y2.d(), r0 * r1, d_U2, n2, &zero, d_core, r0 * r1));
    const double flops = 2.0 * n0 * n0 * n12 + 2.0 * r0 * n0 * n12 + 2.0 * n1 * n1 * r0 * n2 +
                         2.0 * r1 * n1 * r0 * n2 + 2.0 * n2 * n2 * r0 * r1 +
                         2.0 * r0 * r1 * n2 * r2;
    TT_TRY(timing_end(ctx, &tm, 8.0 * n0 * n12, flops));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    if (evals) std::copy(ev.begin(), ev.end(), evals);
    return TTDFT_OK;
}

int ttdft_sthosvd(ttdft_ctx ctx, const double *d_T, int32_t n0, int32_t n1, int32_t n2,
                  int32_t r0, int32_t r1, int32_t r2, double *d_core, double *d_U0, double *d_U1,
                  double *d_U2, double *evals) {
    clear_error();
    if (ctx) g_stream = ctx->stream;
    return sthosvd_impl(ctx, d_T, n0, n1, n2, r0, r1, r2, d_core, d_U0, d_U1, d_U2, evals, 0);
}

int ttdft_sthosvd_refine(ttdft_ctx ctx, const double *d_T, int32_t n0, int32_t n1, int32_t n2,
                         int32_t r0, int32_t r1, int32_t r2, double *d_core, double *d_U0,
                         double *d_U1, double *d_U2, int32_t iters, double *evals) {
    clear_error();
    if (ctx) g_stream = ctx->stream;
    TT_CHECK(iters >= 1, TTDFT_EBADARG, "sthosvd_refine: iters must be >= 1");
    return sthosvd_impl(ctx, d_T, n0, n1, n2, r0, r1, r2, d_core, d_U0, d_U1, d_U2, evals,
                        iters);
}

int ttdft_hartree_expansion(ttdft_ctx ctx, const int32_t r[3], const double *d_core,
                            const double *const d_U[3], const int32_t ne[3],
                            const double *const d_xe[3], const double *const d_jw[3],
                            const int32_t p[3], const double *const d_xq[3], int32_t nterms,
                            const double *alpha, const double *omega, double *d_V, int add) {
    clear_error();
    if (ctx) g_stream = ctx->stream;
    TT_CHECK(ctx && r && d_core && d_U && ne && d_xe && d_jw && p && d_xq && alpha && omega &&
                 d_V,
             TTDFT_EBADARG, "hartree_expansion: NULL argument");
    TT_CHECK(nterms >= 1, TTDFT_EBADARG, "hartree_expansion: no expansion terms");
    for (int d = 0; d < 3; ++d) {
        TT_CHECK(r[d] > 0 && ne[d] > 0 && p[d] > 0 && d_U[d] && d_xe[d] && d_jw[d] && d_xq[d],
                 TTDFT_EBADARG, "hartree_expansion: bad axis %d", d);
        TT_CHECK((int64_t)nterms * p[d] < ((int64_t)1 << 31), TTDFT_EBADARG,
                 "hartree_expansion: terms x points too large");
    }
    const int64_t p01 = (int64_t)p[0] * p[1];
    TT_CHECK(p01 < ((int64_t)1 << 31), TTDFT_EBADARG, "hartree_expansion: p0 p1 too large");
    TT_HIP(hipSetDevice(ctx->device));
    const double one = 1.0, zero = 0.0;
    Dev dal, dom;
    TT_TRY(dal.alloc(sizeof(double) * nterms));
    TT_TRY(dom.alloc(sizeof(double) * nterms));
    TT_HIP(hipMemcpyAsync(dal.p, alpha, sizeof(double) * nterms, hipMemcpyHostToDevice,
                          ctx->stream));
    TT_HIP(hipMemcpyAsync(dom.p, omega, sizeof(double) * nterms, hipMemcpyHostToDevice,
                          ctx->stream));
    PendingTiming tm;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_HARTREE, &tm));
    // 1. F_d = K_d U_d, (terms p_d) x r_d, column-major
    Dev F[3];
    for (int d = 0; d < 3; ++d) {
        const int64_t rows = (int64_t)nterms * p[d];
        Dev K;
        TT_TRY(K.alloc(sizeof(double) * rows * ne[d]));
        TT_TRY(F[d].alloc(sizeof(double) * rows * r[d]));
        hipLaunchKernelGGL(kernel_matrix_kernel, dim3(nblk(rows * ne[d])), dim3(256), 0,
                           ctx->stream, d_xq[d], p[d], d_xe[d], d_jw[d], ne[d], dal.d(), nterms,
                           K.d());
        TT_LAUNCH_CHECK();
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                              (rocblas_int)rows, r[d], ne[d], &one, K.d(), (rocblas_int)rows,
                              d_U[d], ne[d], &zero, F[d].d(), (rocblas_int)rows));
        TT_HIP(hipStreamSynchronize(ctx->stream));  // K is freed at scope end
    }
    // 2. T1 = F0 G_(0): (terms p0) x (r1 r2)
    const int64_t tp0 = (int64_t)nterms * p[0];
    Dev T1;
    TT_TRY(T1.alloc(sizeof(double) * tp0 * r[1] * r[2]));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                          (rocblas_int)tp0, r[1] * r[2], r[0], &one, F[0].d(), (rocblas_int)tp0,
                          d_core, r[0], &zero, T1.d(), (rocblas_int)tp0));
    // 3-4 in chunks of terms: T2 (p0 p1) x (r2 tc), W p2 x (r2 tc)
    const int64_t budget = (int64_t)4 << 30;
    int tc = (int)std::max<int64_t>(1, std::min<int64_t>(nterms, budget / (8 * p01 * r[2])));
    if (const char *env = std::getenv("TTDFT_HARTREE_TERMS")) {  // smaller chunks of terms
        const long v = std::strtol(env, nullptr, 10);
        if (v > 0 && v < tc) tc = (int)v;
    }
    Dev T2, W;
    TT_TRY(T2.alloc(sizeof(double) * p01 * r[2] * tc));
    TT_TRY(W.alloc(sizeof(double) * (int64_t)p[2] * r[2] * tc));
    const int64_t tp1 = (int64_t)nterms * p[1];
    for (int i0 = 0; i0 < nterms; i0 += tc) {
        const int ni = std::min(tc, nterms - i0);
        for (int il = 0; il < ni; ++il) {
            const int i = i0 + il;
            // T2[:, (c, il)] = T1_i[:, :, c] F1_i^T for c = 0 .. r2-1
            TT_BLAS(rocblas_dgemm_strided_batched(
                ctx->blas, rocblas_operation_none, rocblas_operation_transpose, p[0], p[1],
                r[1], &one, T1.d() + (int64_t)p[0] * i, (rocblas_int)tp0, tp0 * r[1],
                F[1].d() + (int64_t)p[1] * i, (rocblas_int)tp1, 0, &zero,
                T2.d() + p01 * r[2] * il, p[0], p01, r[2]));
        }
        hipLaunchKernelGGL(stack_last_kernel, dim3(nblk((int64_t)p[2] * r[2] * ni)), dim3(256), 0,
                           ctx->stream, F[2].d(), p[2], r[2], nterms, i0, ni, dom.d(), W.d());
        TT_LAUNCH_CHECK();
        const double beta = (i0 == 0 && !add) ? 0.0 : 1.0;
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose,
                              (rocblas_int)p01, p[2], r[2] * ni, &one, T2.d(), (rocblas_int)p01,
                              W.d(), p[2], &beta, d_V, (rocblas_int)p01));
    }
    double flops = 2.0 * p01 * p[2] * r[2] * nterms +
                   2.0 * p01 * r[1] * r[2] * nterms + 2.0 * tp0 * r[0] * r[1] * r[2];
    for (int d = 0; d < 3; ++d) flops += 2.0 * nterms * p[d] * ne[d] * r[d];
    TT_TRY(timing_end(ctx, &tm, 8.0 * p01 * p[2] * (add ? 2.0 : 1.0), flops));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

}  // extern "C"
