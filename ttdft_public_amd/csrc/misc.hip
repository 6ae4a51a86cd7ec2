// Grid-level kernels: atomic-superposition initial density, tensor-quadrature
// integrals and residuals, scaling, and the vector kernels of the Lanczos
// bound estimator.  Reductions are two-stage with a fixed block count and a
// fixed-order host sum, so results are deterministic run to run.
#include "internal.h"

namespace ttdft {
namespace {

constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(256) void initial_density_kernel(
    int32_t natoms, const double *__restrict__ xyz, int32_t nr, const double *__restrict__ r,
    const double *__restrict__ coef, int32_t px, int32_t py, int32_t pz,
    const double *__restrict__ gx, const double *__restrict__ gy,
    const double *__restrict__ gz, double *__restrict__ rho) {
    const int64_t total = (int64_t)px * py * pz;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        const int qy = (int)(rest % py);
        const int qz = (int)(rest / py);
        const double x = gx[qx], y = gy[qy], z = gz[qz];
        double s = 0.0;
        for (int a = 0; a < natoms; ++a) {
            const double dx = x - xyz[3 * a + 0];
            const double dy = y - xyz[3 * a + 1];
            const double dz = z - xyz[3 * a + 2];
            const double rr = sqrt(dx * dx + dy * dy + dz * dz);
            // interval search over knots 0..nr-2; the end segments extend
            int lo = 0, hi = nr - 1;
            while (lo != hi - 1) {
                const int m = (lo + hi) >> 1;
                if (r[m] >= rr)
                    hi = m;
                else
                    lo = m;
            }
            const double t = rr - r[lo];
            const double *c = coef + 4 * lo;
            s += c[0] + t * (c[1] + t * (c[2] + t * c[3]));
        }
        rho[idx] = s;
    }
}

// partial[b] = sum over a grid-stride slice of (f - g?)^p * wx wy wz
__global__ __launch_bounds__(256) void grid_reduce_kernel(const double *__restrict__ f,
                                                          const double *__restrict__ g,
                                                          int32_t px, int32_t py, int32_t pz,
                                                          const double *__restrict__ wx,
                                                          const double *__restrict__ wy,
                                                          const double *__restrict__ wz,
                                                          int squared,
                                                          double *__restrict__ partial) {
    const int64_t total = (int64_t)px * py * pz;
    double s = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        const int qy = (int)(rest % py);
        const int qz = (int)(rest / py);
        double v = f[idx];
        if (squared == 2) {
            v = v * g[idx];
        } else {
            if (g) v = v - g[idx];
            if (squared) v = v * v;
        }
        s += v * (wx[qx] * wy[qy] * wz[qz]);
    }
    const double t = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void dot_kernel(const double *__restrict__ a,
                                                  const double *__restrict__ b, int64_t n,
                                                  double *__restrict__ partial) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    const double t = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y = a*y + b*x  (x may be null when b == 0)
__global__ void axpby_kernel(double *__restrict__ y, double a, const double *__restrict__ x,
                             double b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = y[i] * a;
        if (x) v = v + b * x[i];
        y[i] = v;
    }
}

__global__ void scale_kernel(double *__restrict__ f, int64_t n, double s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        f[i] *= s;
}

// column-major FP64 (n x k, lda) -> basis-major (n x k, ldd) of type D, tiled
// through LDS so both sides are coalesced
template <typename D>
__global__ __launch_bounds__(256) void cm_to_bm_kernel(const double *__restrict__ src,
                                                       int64_t lda, int64_t n, int32_t k,
                                                       D *__restrict__ dst, int64_t ldd) {
    __shared__ double tile[32][33];
    const int64_t i0 = (int64_t)blockIdx.x * 32;
    const int j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + tx;
        const int j = j0 + r;
        tile[r][tx] = (i < n && j < k) ? src[i + lda * j] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + r;
        const int j = j0 + tx;
        if (i < n && j < k) dst[i * ldd + j] = (D)tile[tx][r];
    }
}

template <typename S>
__global__ __launch_bounds__(256) void bm_to_cm_kernel(const S *__restrict__ src, int64_t lds,
                                                       int64_t n, int32_t k,
                                                       double *__restrict__ dst, int64_t lda) {
    __shared__ double tile[32][33];
    const int64_t i0 = (int64_t)blockIdx.x * 32;
    const int j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + r;
        const int j = j0 + tx;
        tile[r][tx] = (i < n && j < k) ? (double)src[i * lds + j] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + tx;
        const int j = j0 + r;
        if (i < n && j < k) dst[i + lda * j] = tile[tx][r];
    }
}

// LT = L^T for a lower-triangular L (entries above the diagonal of L ignored)
__global__ void lower_transpose_kernel(const double *__restrict__ L, double *__restrict__ LT,
                                       int32_t k) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)k * k) return;
    const int a = (int)(idx % k), b = (int)(idx / k);
    LT[a + (int64_t)k * b] = (b >= a) ? L[b + (int64_t)k * a] : 0.0;
}

}  // namespace

int launch_initial_density(ttdft_ctx ctx, int32_t natoms, const double *d_xyz, int32_t nr,
                           const double *d_r, const double *d_coef, int32_t px, int32_t py,
                           int32_t pz, const double *gx, const double *gy, const double *gz,
                           double *rho) {
    const int64_t total = (int64_t)px * py * pz;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(initial_density_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       natoms, d_xyz, nr, d_r, d_coef, px, py, pz, gx, gy, gz, rho);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int grid_reduce(ttdft_ctx ctx, const double *f, const double *g, int32_t px, int32_t py,
                int32_t pz, const double *wx, const double *wy, const double *wz, int squared,
                double *result) {
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * kRedBlocks));
    double *partial = (double *)ctx->scratch[6].ptr;
    hipLaunchKernelGGL(grid_reduce_kernel, dim3(kRedBlocks), dim3(256), 0, ctx->stream, f, g, px,
                       py, pz, wx, wy, wz, squared, partial);
    TT_LAUNCH_CHECK();
    std::vector<double> h(kRedBlocks);
    TT_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < kRedBlocks; ++i) s += h[i];
    *result = s;
    return TTDFT_OK;
}

int dev_dot(ttdft_ctx ctx, const double *a, const double *b, int64_t n, double *result) {
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * kRedBlocks));
    double *partial = (double *)ctx->scratch[6].ptr;
    hipLaunchKernelGGL(dot_kernel, dim3(kRedBlocks), dim3(256), 0, ctx->stream, a, b, n, partial);
    TT_LAUNCH_CHECK();
    std::vector<double> h(kRedBlocks);
    TT_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < kRedBlocks; ++i) s += h[i];
    *result = s;
    return TTDFT_OK;
}

int dev_axpby(ttdft_ctx ctx, double *y, double a, const double *x, double b, int64_t n) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, y, a, x, b,
                       n);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

// f[i] = lo + (hi - lo) u_i, u_i in [0, 1) from a splitmix64 hash of (seed, i):
// the same numbers for any grid shape or launch.
__global__ __launch_bounds__(256) void fill_uniform_kernel(double *__restrict__ f, int64_t n,
                                                           uint64_t seed, double lo,
                                                           double hi) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        f[i] = lo + (hi - lo) * ((double)(z >> 11) * 0x1.0p-53);
    }
}

int dev_fill_uniform(ttdft_ctx ctx, double *f, int64_t n, uint64_t seed, double lo, double hi) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, f,
                       n, seed, lo, hi);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int dev_scale(ttdft_ctx ctx, double *f, int64_t n, double s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, f, n, s);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int launch_cm_to_bm(ttdft_ctx ctx, const double *src, int64_t lda, int64_t n, int32_t k,
                    void *dst, int64_t ldd, int dtype) {
    const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((k + 31) / 32));
    if (dtype == TTDFT_F64)
        hipLaunchKernelGGL((cm_to_bm_kernel<double>), grid, dim3(256), 0, ctx->stream, src, lda, n,
                           k, (double *)dst, ldd);
    else
        hipLaunchKernelGGL((cm_to_bm_kernel<float>), grid, dim3(256), 0, ctx->stream, src, lda, n,
                           k, (float *)dst, ldd);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int launch_bm_to_cm(ttdft_ctx ctx, const void *src, int64_t lds, int dtype, int64_t n,
                    int32_t k, double *dst, int64_t lda) {
    const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((k + 31) / 32));
    if (dtype == TTDFT_F64)
        hipLaunchKernelGGL((bm_to_cm_kernel<double>), grid, dim3(256), 0, ctx->stream,
                           (const double *)src, lds, n, k, dst, lda);
    else
        hipLaunchKernelGGL((bm_to_cm_kernel<float>), grid, dim3(256), 0, ctx->stream,
                           (const float *)src, lds, n, k, dst, lda);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int launch_lower_transpose(ttdft_ctx ctx, const double *L, double *LT, int32_t k) {
    const int64_t kk = (int64_t)k * k;
    hipLaunchKernelGGL(lower_transpose_kernel, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0,
                       ctx->stream, L, LT, k);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // namespace ttdft
