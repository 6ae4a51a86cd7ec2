// Electron density at tensor-grid points,
//   rho(qx,qy,qz) = sum_i w_i psi_i(qx,qy,qz)^2,   w_i = 2 f_i,
//   psi_i = sum_abc C_i[a,b,c] Bx[qx,a] By[qy,b] Bz[qz,c],
// computed for batches of S states:
//   1. gather the batch's coefficients state-major (Cb[s][c][b][a]),
//   2. T1[s][qz][b][a]  = sum_c Bz[qz,c] Cb[s][c][b][a]      (batched GEMM),
//   3. T2[s][qz][b][qx] = sum_a T1[s][qz][b][a] Bx[qx,a]     (one GEMM),
//   4. fused last contraction + square + weighted sum over the batch, with
//      psi never written to memory: rho tile += sum_s w_s (By T2_s)^2.
// The contraction order (z, then x, then y) is that of the reference's TTM
// chain (dft/ProjectHamiltonianSparse.cc:1052-1067).
#include "internal.h"

namespace ttdft {
namespace {

template <typename T>
__global__ void gather_states_kernel(const T *__restrict__ X, int64_t ldx, int32_t s0,
                                     int32_t S, int64_t N, double *__restrict__ Cb) {
    const int64_t total = N * S;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t I = idx / S;
        const int s = (int)(idx - I * S);
        Cb[(int64_t)s * N + I] = (double)X[I * ldx + s0 + s];
    }
}

// C[b](i,j) = sum_kk A[b](i,kk) B[b](kk,j), all FP64, generic strides.
// 64x64 tile, BK = 16, 256 threads, 4x4 outputs per thread.
__global__ __launch_bounds__(256) void gemm_batched_f64_kernel(
    int64_t M, int64_t N, int64_t K, const double *__restrict__ A, int64_t sab, int64_t sar,
    int64_t sac, const double *__restrict__ B, int64_t sbb, int64_t sbr, int64_t sbc,
    double *__restrict__ C, int64_t scb, int64_t ldc) {
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
    const int64_t bz = blockIdx.z;
    A += bz * sab;
    B += bz * sbb;
    C += bz * scb;
    double acc[4][4] = {};
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = tid + 256 * t;        // 0..1023
            const int ii = e & 63, kk = e >> 6; // A tile: i fastest
            const int64_t gi = i0 + ii, gk = k0 + kk;
            As[kk][ii] = (gi < M && gk < K) ? A[gi * sar + gk * sac] : 0.0;
            const int jj = e & 63;
            const int64_t gj = j0 + jj;
            Bs[kk][jj] = (gj < N && gk < K) ? B[gk * sbr + gj * sbc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = As[kk][ty + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u) b[u] = Bs[kk][tx + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[v][u] += a[v] * b[u];
        }
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t gi = i0 + ty + 16 * v;
        if (gi >= M) continue;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t gj = j0 + tx + 16 * u;
            if (gj < N) C[gi * ldc + gj] = acc[v][u];
        }
    }
}

// rho[qz][qy][qx] (+)= sum_s w[s] (sum_b By[qy + py b] T2[s][qz][b][qx])^2
// Block: 64 qx x 64 qy at one qz.  Dynamic LDS: By tile [ry][64], T2 tile [ry][64].
__global__ __launch_bounds__(256) void density_square_kernel(
    const double *__restrict__ T2, const double *__restrict__ By, const double *__restrict__ w,
    int32_t S, int32_t px, int32_t py, int32_t pz, int32_t ry, double *__restrict__ rho,
    int accumulate) {
    extern __shared__ double lds[];
    double *Bys = lds;            // [ry][64]
    double *Ts = lds + ry * 64;   // [ry][64]
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int qx0 = blockIdx.x * 64, qy0 = blockIdx.y * 64, qz = blockIdx.z;
    for (int e = tid; e < ry * 64; e += 256) {
        const int q = e & 63, b = e >> 6;
        const int qy = qy0 + q;
        Bys[b * 64 + q] = (qy < py) ? By[qy + (int64_t)py * b] : 0.0;
    }
    double acc[4][4] = {};
    for (int s = 0; s < S; ++s) {
        __syncthreads();
        const double *T2s = T2 + (((int64_t)s * pz + qz) * ry) * px;
        for (int e = tid; e < ry * 64; e += 256) {
            const int q = e & 63, b = e >> 6;
            const int qx = qx0 + q;
            Ts[b * 64 + q] = (qx < px) ? T2s[(int64_t)b * px + qx] : 0.0;
        }
        __syncthreads();
        double psi[4][4] = {};
        for (int b = 0; b < ry; ++b) {
            double yv[4], tv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) yv[v] = Bys[b * 64 + ty + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u) tv[u] = Ts[b * 64 + tx + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int u = 0; u < 4; ++u) psi[v][u] += yv[v] * tv[u];
        }
        const double ws = w[s];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[v][u] += ws * psi[v][u] * psi[v][u];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int qy = qy0 + ty + 16 * v;
        if (qy >= py) continue;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int qx = qx0 + tx + 16 * u;
            if (qx >= px) continue;
            const int64_t o = qx + (int64_t)px * qy + (int64_t)px * py * qz;
            rho[o] = accumulate ? rho[o] + acc[v][u] : acc[v][u];
        }
    }
}

int gemm_batched(ttdft_ctx ctx, int64_t M, int64_t N, int64_t K, int64_t batch,
                 const double *A, int64_t sab, int64_t sar, int64_t sac, const double *B,
                 int64_t sbb, int64_t sbr, int64_t sbc, double *C, int64_t scb, int64_t ldc) {
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)batch);
    TT_CHECK((M + 63) / 64 < 65536 && batch < 65536, TTDFT_EBADARG, "gemm grid too large");
    hipLaunchKernelGGL(gemm_batched_f64_kernel, grid, dim3(256), 0, ctx->stream, M, N, K, A, sab,
                       sar, sac, B, sbb, sbr, sbc, C, scb, ldc);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // namespace

int density_compute(ttdft_ctx ctx, const void *X, int64_t ldx, int dtype, int32_t k,
                    const double *occ, int32_t rx, int32_t ry, int32_t rz, const double *Bx,
                    const double *By, const double *Bz, int32_t px, int32_t py, int32_t pz,
                    double *rho) {
    TT_CHECK(ry <= 128, TTDFT_EBADARG, "density: ry=%d > 128 not supported", ry);
    const int64_t N = (int64_t)rx * ry * rz;
    // states with a nonzero weight
    std::vector<int32_t> idx;
    std::vector<double> wts;
    for (int i = 0; i < k; ++i) {
        const double w = 2.0 * occ[i];
        if (w != 0.0) {
            idx.push_back(i);
            wts.push_back(w);
        }
    }
    const int64_t t2_per_state = (int64_t)pz * ry * px;
    const int64_t t1_per_state = (int64_t)pz * rx * ry;
    int32_t S = 16;
    const int64_t budget = (int64_t)4 << 30;  // bytes for T2 of one batch
    while (S > 1 && (int64_t)S * t2_per_state * 8 > budget) S /= 2;
    TT_TRY(ctx->scratch[5].ensure((size_t)S * (N + t1_per_state + t2_per_state) * 8 + 8 * S));
    double *Cb = (double *)ctx->scratch[5].ptr;
    double *T1 = Cb + (int64_t)S * N;
    double *T2 = T1 + (int64_t)S * t1_per_state;
    double *wdev = T2 + (int64_t)S * t2_per_state;
    if (idx.empty()) {
        TT_HIP(hipMemsetAsync(rho, 0, sizeof(double) * (size_t)px * py * pz, ctx->stream));
        return TTDFT_OK;
    }
    PendingTiming t;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_DENSITY, &t));
    int first = 1;
    for (size_t b0 = 0; b0 < idx.size();) {
        // states of this batch must be consecutive in X for the gather; split
        // runs of consecutive indices
        int32_t s0 = idx[b0];
        int32_t cnt = 1;
        while ((size_t)(b0 + cnt) < idx.size() && cnt < S && idx[b0 + cnt] == s0 + cnt) ++cnt;
        TT_HIP(hipMemcpyAsync(wdev, wts.data() + b0, sizeof(double) * cnt, hipMemcpyHostToDevice,
                              ctx->stream));
        {
            int64_t blocks = (N * cnt + 255) / 256;
            if (blocks > 65536) blocks = 65536;
            if (dtype == TTDFT_F64)
                hipLaunchKernelGGL((gather_states_kernel<double>), dim3((unsigned)blocks),
                                   dim3(256), 0, ctx->stream, (const double *)X, ldx, s0, cnt, N,
                                   Cb);
            else
                hipLaunchKernelGGL((gather_states_kernel<float>), dim3((unsigned)blocks),
                                   dim3(256), 0, ctx->stream, (const float *)X, ldx, s0, cnt, N,
                                   Cb);
            TT_LAUNCH_CHECK();
        }
        // T1[s](qz, ab) = sum_c Bz(qz, c) Cb[s](c, ab)
        const int64_t rxy = (int64_t)rx * ry;
        TT_TRY(gemm_batched(ctx, pz, rxy, rz, cnt, Bz, 0, 1, pz, Cb, N, rxy, 1, T1,
                            t1_per_state, rxy));
        // T2(row, qx) = sum_a T1(row, a) Bx(qx, a),  row = (s, qz, b)
        TT_TRY(gemm_batched(ctx, (int64_t)cnt * pz * ry, px, rx, 1, T1, 0, rx, 1, Bx, 0, px, 1,
                            T2, 0, px));
        const dim3 grid((unsigned)((px + 63) / 64), (unsigned)((py + 63) / 64), (unsigned)pz);
        const size_t lds = (size_t)2 * ry * 64 * sizeof(double);
        hipLaunchKernelGGL(density_square_kernel, grid, dim3(256), lds, ctx->stream, T2, By, wdev,
                           cnt, px, py, pz, ry, rho, first ? 0 : 1);
        TT_LAUNCH_CHECK();
        first = 0;
        b0 += cnt;
    }
    const double nst = (double)idx.size();
    const double flops = 2.0 * nst *
                         ((double)rx * ry * rz * pz + (double)pz * rx * ry * px +
                          (double)px * py * pz * ry) +
                         3.0 * nst * px * py * pz;
    const double bytes = (double)N * k * dtype_size(dtype) + 8.0 * px * py * pz;
    TT_TRY(timing_end(ctx, &t, bytes, flops));
    return TTDFT_OK;
}

}  // namespace ttdft
