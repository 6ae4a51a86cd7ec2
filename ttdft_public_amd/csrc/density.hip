// Electron density at tensor-grid points,
//   rho(qx,qy,qz) = sum_i w_i psi_i(qx,qy,qz)^2,   w_i = 2 f_i,
//   psi_i = sum_abc C_i[a,b,c] Bx[qx,a] By[qy,b] Bz[qz,c],
// computed for batches of S states:
//   1. gather the batch's coefficients state-major (Cb[s][c][b][a]),
//   2. T1[s][qz][b][a]  = sum_c Bz[qz,c] Cb[s][c][b][a]      (batched GEMM),
//   3. T2[s][qz][b][qx] = sum_a T1[s][qz][b][a] Bx[qx,a]     (one GEMM),
//   4. fused last contraction on FP64 MFMA + square + weighted sum over the
//      batch, with psi never written to memory: rho tile += sum_s w_s (By T2_s)^2.
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

// rho[qz][qy][qx] (+)= sum_s w[s] (sum_b By[qy + py b] T2[s][qz][b][qx])^2 on
// FP64 MFMA (16x16x4).  Block = 8 waves = 128 qy x 64 qx at one qz; wave =
// 32 qy x 32 qx (2x2 MFMA tiles).  A[i][kk] = By[qy0+i][b+kk] is the same for
// every state and stays in registers; B[kk][j] = T2_s[b+kk][qx0+j] is staged
// per state in LDS ([ry][64], shared by the four qy-groups).  psi never leaves
// the accumulators: it is squared, weighted and summed over the batch there.
// f64 MFMA layouts: A[i][kk] in lane i + 16 kk, B[kk][j] in lane j + 16 kk,
// D[i][j] in lane j + 16 (i%4), element i/4.
typedef double d4 __attribute__((ext_vector_type(4)));

// KS = number of 4-deep K steps (ceil(ry/4) rounded up to a supported size).
// The next state's T2 tile is prefetched into registers while the current
// state's MFMAs run, then written to the other LDS buffer (one barrier per
// state).
template <int KS>
__global__ __launch_bounds__(512) void density_square_mfma_kernel(
    const double *__restrict__ T2, const double *__restrict__ By, const double *__restrict__ w,
    int32_t S, int32_t px, int32_t py, int32_t pz, int32_t ry, double *__restrict__ rho,
    int accumulate) {
    constexpr int TILE = KS * 4 * 64;          // doubles per state tile
    constexpr int PER_THREAD = (TILE + 511) / 512;
    extern __shared__ double Ts[];             // [2][TILE]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int qx0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int qy0 = blockIdx.y * 128 + (wave >> 1) * 32;
    const int qz = blockIdx.z;
    const int li = lane & 15, lk = lane >> 4;
    double a0[KS], a1[KS];
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        const int b = 4 * t + lk;
        const int qa = qy0 + li, qb = qy0 + 16 + li;
        a0[t] = (b < ry && qa < py) ? By[qa + (int64_t)py * b] : 0.0;
        a1[t] = (b < ry && qb < py) ? By[qb + (int64_t)py * b] : 0.0;
    }
    auto fetch = [&](int s, double *regs) {
        const double *T2s = T2 + (((int64_t)s * pz + qz) * ry) * px;
#pragma unroll
        for (int u = 0; u < PER_THREAD; ++u) {
            const int e = tid + 512 * u;
            const int q = e & 63, b = e >> 6;
            const int qx = blockIdx.x * 64 + q;
            regs[u] = (e < TILE && b < ry && qx < px) ? T2s[(int64_t)b * px + qx] : 0.0;
        }
    };
    auto store = [&](int buf, const double *regs) {
#pragma unroll
        for (int u = 0; u < PER_THREAD; ++u) {
            const int e = tid + 512 * u;
            if (e < TILE) Ts[buf * TILE + e] = regs[u];
        }
    };
    double pre[PER_THREAD];
    fetch(0, pre);
    store(0, pre);
    __syncthreads();
    d4 r00 = {0, 0, 0, 0}, r01 = r00, r10 = r00, r11 = r00;
    const int lx = (wave & 1) * 32;
    for (int s = 0; s < S; ++s) {
        const int buf = s & 1;
        if (s + 1 < S) fetch(s + 1, pre);
        const double *Tb = Ts + buf * TILE;
        d4 p00 = {0, 0, 0, 0}, p01 = p00, p10 = p00, p11 = p00;
#pragma unroll
        for (int t = 0; t < KS; ++t) {
            const double b0 = Tb[(4 * t + lk) * 64 + lx + li];
            const double b1 = Tb[(4 * t + lk) * 64 + lx + 16 + li];
            p00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b0, p00, 0, 0, 0);
            p01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b1, p01, 0, 0, 0);
            p10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b0, p10, 0, 0, 0);
            p11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b1, p11, 0, 0, 0);
        }
        const double ws = w[s];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            r00[v] += ws * p00[v] * p00[v];
            r01[v] += ws * p01[v] * p01[v];
            r10[v] += ws * p10[v] * p10[v];
            r11[v] += ws * p11[v] * p11[v];
        }
        if (s + 1 < S) store(buf ^ 1, pre);
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = lk + 4 * v;
        const int qyA = qy0 + i, qyB = qy0 + 16 + i;
        const int qxA = qx0 + li, qxB = qx0 + 16 + li;
        const int64_t base = (int64_t)px * py * qz;
        if (qyA < py && qxA < px) {
            const int64_t o = base + qxA + (int64_t)px * qyA;
            rho[o] = accumulate ? rho[o] + r00[v] : r00[v];
        }
        if (qyA < py && qxB < px) {
            const int64_t o = base + qxB + (int64_t)px * qyA;
            rho[o] = accumulate ? rho[o] + r01[v] : r01[v];
        }
        if (qyB < py && qxA < px) {
            const int64_t o = base + qxA + (int64_t)px * qyB;
            rho[o] = accumulate ? rho[o] + r10[v] : r10[v];
        }
        if (qyB < py && qxB < px) {
            const int64_t o = base + qxB + (int64_t)px * qyB;
            rho[o] = accumulate ? rho[o] + r11[v] : r11[v];
        }
    }
}

template <int KS>
int launch_square(ttdft_ctx ctx, const double *T2, const double *By, const double *w, int32_t S,
                  int32_t px, int32_t py, int32_t pz, int32_t ry, double *rho, int accumulate) {
    const dim3 grid((unsigned)((px + 63) / 64), (unsigned)((py + 127) / 128), (unsigned)pz);
    const size_t lds = (size_t)2 * KS * 4 * 64 * sizeof(double);
    hipLaunchKernelGGL((density_square_mfma_kernel<KS>), grid, dim3(512), lds, ctx->stream, T2,
                       By, w, S, px, py, pz, ry, rho, accumulate);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int density_square(ttdft_ctx ctx, const double *T2, const double *By, const double *w, int32_t S,
                   int32_t px, int32_t py, int32_t pz, int32_t ry, double *rho, int accumulate) {
    const int ks = (ry + 3) / 4;
    if (ks <= 4) return launch_square<4>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
    if (ks <= 8) return launch_square<8>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
    if (ks <= 13) return launch_square<13>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
    if (ks <= 18) return launch_square<18>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
    if (ks <= 24) return launch_square<24>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
    return launch_square<32>(ctx, T2, By, w, S, px, py, pz, ry, rho, accumulate);
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
        TT_TRY(density_square(ctx, T2, By, wdev, cnt, px, py, pz, ry, rho, first ? 0 : 1));
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
