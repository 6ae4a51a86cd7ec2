// Layers of the 3-D convolutional network that predicts a correction to the
// atomic-superposition initial density (dft/InitialGuess.cc:573-762 is the
// start it corrects; the network has no reference counterpart).
//
// Tensors are one sample's [C][D][H][W] FP32 arrays (x fastest); a channel
// concatenation is two adjacent channel blocks, so it costs nothing.
//
// Convolutions (kernel ks = 1 or 3, stride 1, zero padding ks / 2) run as
// implicit GEMMs on the vector ALUs: Y[co][p] = sum_k W[co][k] X_k[p] with
// k = (ci, tap) and X_k the input shifted by the tap, gathered into LDS with
// the padding applied on the fly (no im2col buffer).  A 256-thread workgroup
// computes a 32 x 128 tile of Y in 16-deep K steps; each thread holds 4 x 4
// accumulators, its 4 positions strided by 32 so stores coalesce.
//   backward data:   dX = conv(dY, W') with W'[ci][co][ks^3-1-t] = W[co][ci][t]
//   backward weight: dW[co][k] = sum_p dY[co][p] X_k[p], split over position
//                    ranges with the partial sums reduced in a fixed order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

constexpr int kTM = 32, kTN = 128, kTK = 16;

struct Geo {
    int32_t D, H, W;
    int64_t P;
};

// X_k at the position (x, y, z), k = (c, tap); zero outside the grid
template <int KS>
__device__ __forceinline__ float gather(const float *__restrict__ X, const Geo g, int x, int y,
                                        int z, int k) {
    constexpr int KT = KS * KS * KS, PAD = KS / 2;
    const int c = k / KT, tap = k - c * KT;
    const int xx = x + tap % KS - PAD, yy = y + (tap / KS) % KS - PAD, zz = z + tap / (KS * KS) - PAD;
    if (xx < 0 || xx >= g.W || yy < 0 || yy >= g.H || zz < 0 || zz >= g.D) return 0.0f;
    return X[(int64_t)c * g.P + ((int64_t)zz * g.H + yy) * g.W + xx];
}

__device__ __forceinline__ void decompose(const Geo g, int64_t p, int &x, int &y, int &z) {
    x = (int)(p % g.W);
    const int64_t r = p / g.W;
    y = (int)(r % g.H);
    z = (int)(r / g.H);
}

// Y[co][p] (=|+=) b[co] + sum_k W[co][k] X_k[p],  K = ci ks^3
template <int KS>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const float *__restrict__ X, int ci,
                                                       Geo g, const float *__restrict__ Wt,
                                                       const float *__restrict__ bias, int co,
                                                       float *__restrict__ Y, int accumulate) {
    __shared__ float As[kTK][kTM + 1];
    __shared__ float Bs[kTK][kTN];
    const int t = threadIdx.x;
    const int tm = t >> 5, tn = t & 31;
    const int m0 = blockIdx.y * kTM;
    const int64_t n0 = (int64_t)blockIdx.x * kTN;
    constexpr int KT = KS * KS * KS;
    const int K = ci * KT;
    // every B element this thread loads sits at the same position
    const int64_t pb = n0 + (t & 127);
    int bx = 0, by = 0, bz = 0;
    decompose(g, pb < g.P ? pb : 0, bx, by, bz);
    float acc[4][4] = {};
    for (int kc = 0; kc < K; kc += kTK) {
        for (int r = 0; r < 2; ++r) {  // A: 32 x 16
            const int e = t + 256 * r;
            const int m = e >> 4, kk = e & 15;
            const int k = kc + kk;
            As[kk][m] = (m0 + m < co && k < K) ? Wt[(int64_t)(m0 + m) * K + k] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {  // B: 16 x 128, gathered
            const int kk = (t >> 7) + 2 * r;
            const int k = kc + kk;
            Bs[kk][t & 127] = (k < K && pb < g.P) ? gather<KS>(X, g, bx, by, bz, k) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tm * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tn + 32 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + tm * 4 + i;
        if (m >= co) continue;
        const float bv = bias ? bias[m] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t p = n0 + tn + 32 * j;
            if (p >= g.P) continue;
            const float v = acc[i][j] + bv;
            float *dst = Y + (int64_t)m * g.P + p;
            *dst = accumulate ? *dst + v : v;
        }
    }
}

// ---- BF16 MFMA forms -----------------------------------------------------------
// The same products on v_mfma_f32_16x16x16_bf16: operands rounded to BF16
// (nearest even) as they are staged into LDS, FP32 accumulation.  MFMA
// layouts: A[i][kk] in lane i + 16 (kk/4), element kk%4; B[kk][j] in lane
// j + 16 (kk/4), element kk%4; element v of lane l holds D[4 (l/16) + v][l%16].
typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float x) {
    unsigned u = __builtin_bit_cast(unsigned, x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Forward: a workgroup computes 32 (co) x 128 (positions) of Y; wave w owns
// positions 32w .. 32w+31 (two 16-column MFMA tiles) and both 16-row tiles.
template <int KS>
__global__ __launch_bounds__(256) void conv_fwd_bf16_kernel(const float *__restrict__ X, int ci,
                                                            Geo g, const float *__restrict__ Wt,
                                                            const float *__restrict__ bias,
                                                            int co, float *__restrict__ Y,
                                                            int accumulate) {
    __shared__ unsigned short As[kTM][kTK + 4];   // [m][k]
    __shared__ unsigned short Bs[kTN][kTK + 4];   // [position][k]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int m0 = blockIdx.y * kTM;
    const int64_t n0 = (int64_t)blockIdx.x * kTN;
    constexpr int KT = KS * KS * KS;
    const int K = ci * KT;
    const int64_t pb = n0 + (t & 127);
    int bx = 0, by = 0, bz = 0;
    decompose(g, pb < g.P ? pb : 0, bx, by, bz);
    f4v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    const int li = lane & 15, lk = lane >> 4;
    for (int kc = 0; kc < K; kc += kTK) {
        for (int r = 0; r < 2; ++r) {  // A: 32 x 16
            const int e = t + 256 * r;
            const int m = e >> 4, kk = e & 15;
            const int k = kc + kk;
            As[m][kk] = to_bf16((m0 + m < co && k < K) ? Wt[(int64_t)(m0 + m) * K + k] : 0.0f);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {  // B: 16 x 128, gathered
            const int kk = (t >> 7) + 2 * r;
            const int k = kc + kk;
            Bs[t & 127][kk] =
                to_bf16((k < K && pb < g.P) ? gather<KS>(X, g, bx, by, bz, k) : 0.0f);
        }
        __syncthreads();
        s4v a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[i][v] = (short)As[16 * i + li][4 * lk + v];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) b[j][v] = (short)Bs[32 * wave + 16 * j + li][4 * lk + v];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = m0 + 16 * i + 4 * lk + v;
            if (m >= co) continue;
            const float bv = bias ? bias[m] : 0.0f;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int64_t p = n0 + 32 * wave + 16 * j + li;
                if (p >= g.P) continue;
                const float val = acc[i][j][v] + bv;
                float *dst = Y + (int64_t)m * g.P + p;
                *dst = accumulate ? *dst + val : val;
            }
        }
}

// Weight gradient: a workgroup computes partial[s][m][k] for 32 (co) x 64 (k)
// over one position range; the MFMA reduction dimension is 16 positions per
// step; wave w owns k columns 16w .. 16w+15 and both 16-row tiles.
template <int KS>
__global__ __launch_bounds__(256) void conv_wgrad_bf16_kernel(const float *__restrict__ X, int ci,
                                                              Geo g, const float *__restrict__ dY,
                                                              int co, int64_t span,
                                                              float *__restrict__ partial) {
    __shared__ unsigned short Ds[kTM][16 + 4];  // [m][position]
    __shared__ unsigned short Cs[64][16 + 4];   // [k][position]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int m0 = blockIdx.y * kTM;
    const int k0 = blockIdx.x * 64;
    constexpr int KT = KS * KS * KS;
    const int K = ci * KT;
    const int64_t pb = (int64_t)blockIdx.z * span;
    const int64_t pe = min(pb + span, g.P);
    const int li = lane & 15, lk = lane >> 4;
    f4v acc[2] = {f4v{0.0f, 0.0f, 0.0f, 0.0f}, f4v{0.0f, 0.0f, 0.0f, 0.0f}};
    for (int64_t pc = pb; pc < pe; pc += 16) {
        for (int r = 0; r < 2; ++r) {  // dY: 32 x 16
            const int e = t + 256 * r;
            const int pp = e & 15, m = e >> 4;
            const int64_t p = pc + pp;
            Ds[m][pp] = to_bf16((m0 + m < co && p < pe) ? dY[(int64_t)(m0 + m) * g.P + p] : 0.0f);
        }
        const int64_t p = pc + (t & 15);
        int x = 0, y = 0, z = 0;
        decompose(g, p < pe ? p : pb, x, y, z);
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // X_k: 64 x 16, gathered
            const int kk = (t >> 4) + 16 * r;
            const int k = k0 + kk;
            Cs[kk][t & 15] = to_bf16((k < K && p < pe) ? gather<KS>(X, g, x, y, z, k) : 0.0f);
        }
        __syncthreads();
        s4v a[2], b;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[i][v] = (short)Ds[16 * i + li][4 * lk + v];
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v] = (short)Cs[16 * wave + li][4 * lk + v];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b, acc[i], 0, 0, 0);
        __syncthreads();
    }
    float *out = partial + (int64_t)blockIdx.z * co * K;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = m0 + 16 * i + 4 * lk + v;
            const int k = k0 + 16 * wave + li;
            if (m < co && k < K) out[(int64_t)m * K + k] = acc[i][v];
        }
}

// W'[c][m ks^3 + (ks^3 - 1 - tap)] = W[m][c ks^3 + tap]
__global__ __launch_bounds__(256) void flip_weights_kernel(const float *__restrict__ Wt, int co,
                                                           int ci, int KT,
                                                           float *__restrict__ Wf) {
    const int64_t total = (int64_t)co * ci * KT;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int tap = (int)(i % KT);
        const int64_t r = i / KT;
        const int c = (int)(r % ci);
        const int m = (int)(r / ci);
        Wf[((int64_t)c * co + m) * KT + (KT - 1 - tap)] = Wt[i];
    }
}

// partial[s][m][k] = sum_{p in range s} dY[m][p] X_k[p]
template <int KS>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(const float *__restrict__ X, int ci,
                                                         Geo g, const float *__restrict__ dY,
                                                         int co, int64_t span,
                                                         float *__restrict__ partial) {
    __shared__ float Ds[16][kTM + 1];
    __shared__ float Cs[16][64 + 1];
    const int t = threadIdx.x;
    const int tm = t >> 4, tk = t & 15;
    const int m0 = blockIdx.y * kTM;
    const int k0 = blockIdx.x * 64;
    constexpr int KT = KS * KS * KS;
    const int K = ci * KT;
    const int64_t pb = (int64_t)blockIdx.z * span;
    const int64_t pe = min(pb + span, g.P);
    float acc[2][4] = {};
    for (int64_t pc = pb; pc < pe; pc += 16) {
        for (int r = 0; r < 2; ++r) {  // dY: 32 x 16
            const int e = t + 256 * r;
            const int pp = e & 15, m = e >> 4;
            const int64_t p = pc + pp;
            Ds[pp][m] = (m0 + m < co && p < pe) ? dY[(int64_t)(m0 + m) * g.P + p] : 0.0f;
        }
        // every X_k element this thread loads sits at the same position
        const int64_t p = pc + (t & 15);
        int x = 0, y = 0, z = 0;
        decompose(g, p < pe ? p : pb, x, y, z);
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // X_k: 64 x 16, gathered
            const int kk = (t >> 4) + 16 * r;
            const int k = k0 + kk;
            Cs[t & 15][kk] = (k < K && p < pe) ? gather<KS>(X, g, x, y, z, k) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < 16; ++pp) {
            float d[2], c[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) d[i] = Ds[pp][tm * 2 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) c[j] = Cs[pp][tk + 16 * j];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += d[i] * c[j];
        }
        __syncthreads();
    }
    float *out = partial + (int64_t)blockIdx.z * co * K;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = m0 + tm * 2 + i;
        if (m >= co) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + tk + 16 * j;
            if (k < K) out[(int64_t)m * K + k] = acc[i][j];
        }
    }
}

// dW[i] (=|+=) sum_s partial[s][i], s in order
__global__ __launch_bounds__(256) void split_reduce_kernel(const float *__restrict__ partial,
                                                           int nsplit, int64_t n,
                                                           float *__restrict__ dW,
                                                           int accumulate) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int k = 0; k < nsplit; ++k) s += partial[(int64_t)k * n + i];
        dW[i] = accumulate ? dW[i] + s : s;
    }
}

// db[m] (=|+=) sum_p dY[m][p]: one workgroup per channel, fixed tree
__global__ __launch_bounds__(256) void bias_grad_kernel(const float *__restrict__ dY, int64_t P,
                                                        float *__restrict__ db, int accumulate) {
    __shared__ float red[256];
    const float *row = dY + (int64_t)blockIdx.x * P;
    float s = 0.0f;
    for (int64_t p = threadIdx.x; p < P; p += 256) s += row[p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[blockIdx.x] = accumulate ? db[blockIdx.x] + red[0] : red[0];
}

__global__ __launch_bounds__(256) void leaky_fwd_kernel(float *__restrict__ y, int64_t n,
                                                        float slope) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = y[i];
        y[i] = v > 0.0f ? v : slope * v;
    }
}

// dy *= (y > 0 ? 1 : slope), y the activation's output
__global__ __launch_bounds__(256) void leaky_bwd_kernel(const float *__restrict__ y,
                                                        float *__restrict__ dy, int64_t n,
                                                        float slope) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dy[i] = y[i] > 0.0f ? dy[i] : slope * dy[i];
}

// coarse[c][z][y][x] = mean of the 2x2x2 fine children (fine dims 2D x 2H x 2W)
__global__ __launch_bounds__(256) void pool_fwd_kernel(const float *__restrict__ fine, int c,
                                                       Geo cg, float *__restrict__ coarse) {
    const int64_t total = (int64_t)c * cg.P;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i / cg.P, p = i % cg.P;
        const int x = (int)(p % cg.W);
        const int64_t r = p / cg.W;
        const int y = (int)(r % cg.H), z = (int)(r / cg.H);
        const int W2 = 2 * cg.W, H2 = 2 * cg.H;
        const float *f = fine + ch * 8 * cg.P;
        float s = 0.0f;
        for (int q = 0; q < 8; ++q)
            s += f[((int64_t)(2 * z + (q >> 2)) * H2 + 2 * y + ((q >> 1) & 1)) * W2 + 2 * x +
                   (q & 1)];
        coarse[i] = 0.125f * s;
    }
}

// fine (=|+=) coarse / 8 (pool backward) or coarse (upsample forward), by fine index
__global__ __launch_bounds__(256) void spread_kernel(const float *__restrict__ coarse, int c,
                                                     Geo fg, float scale,
                                                     float *__restrict__ fine, int accumulate) {
    const int64_t total = (int64_t)c * fg.P;
    const int Wc = fg.W / 2, Hc = fg.H / 2;
    const int64_t Pc = fg.P / 8;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i / fg.P, p = i % fg.P;
        const int x = (int)(p % fg.W);
        const int64_t r = p / fg.W;
        const int y = (int)(r % fg.H), z = (int)(r / fg.H);
        const float v =
            scale * coarse[ch * Pc + ((int64_t)(z >> 1) * Hc + (y >> 1)) * Wc + (x >> 1)];
        fine[i] = accumulate ? fine[i] + v : v;
    }
}

// coarse (=|+=) sum of the 8 fine children (upsample backward)
__global__ __launch_bounds__(256) void gather8_kernel(const float *__restrict__ fine, int c,
                                                      Geo cg, float *__restrict__ coarse,
                                                      int accumulate) {
    const int64_t total = (int64_t)c * cg.P;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i / cg.P, p = i % cg.P;
        const int x = (int)(p % cg.W);
        const int64_t r = p / cg.W;
        const int y = (int)(r % cg.H), z = (int)(r / cg.H);
        const int W2 = 2 * cg.W, H2 = 2 * cg.H;
        const float *f = fine + ch * 8 * cg.P;
        float s = 0.0f;
        for (int q = 0; q < 8; ++q)
            s += f[((int64_t)(2 * z + (q >> 2)) * H2 + 2 * y + ((q >> 1) & 1)) * W2 + 2 * x +
                   (q & 1)];
        coarse[i] = accumulate ? coarse[i] + s : s;
    }
}

// per-block sums of (y - t)^2; dy = 2 (y - t) / n
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ y,
                                                  const float *__restrict__ tg, int64_t n,
                                                  float *__restrict__ dy,
                                                  double *__restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    const float sc = 2.0f / (float)n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float d = y[i] - tg[i];
        s += (double)d * d;
        if (dy) dy[i] = sc * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p,
                                                   const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v,
                                                   int64_t n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

__global__ __launch_bounds__(256) void f64_to_f32_scaled(const double *__restrict__ s,
                                                         float *__restrict__ d, int64_t n,
                                                         double scale) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (float)(s[i] * scale);
}

__global__ __launch_bounds__(256) void f32_to_f64_scaled(const float *__restrict__ s,
                                                         double *__restrict__ d, int64_t n,
                                                         double scale) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (double)s[i] * scale;
}

unsigned nblk(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

int make_geo(const int32_t dims[3], Geo &g) {
    TT_CHECK(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, TTDFT_EBADARG,
             "nn: bad spatial dims");
    g.D = dims[0];
    g.H = dims[1];
    g.W = dims[2];
    g.P = (int64_t)g.D * g.H * g.W;
    return TTDFT_OK;
}

int conv_forward(ttdft_ctx ctx, const float *x, int ci, const Geo &g, const float *w,
                 const float *b, int co, int ks, float *y, int accumulate) {
    const dim3 grid((unsigned)((g.P + kTN - 1) / kTN), (unsigned)((co + kTM - 1) / kTM));
    TT_CHECK(grid.x <= 0x7fffffffu && grid.y <= 65535u, TTDFT_EBADARG, "conv3d: grid too large");
    if (ctx->nn_math == TTDFT_NN_BF16) {
        if (ks == 3)
            hipLaunchKernelGGL(conv_fwd_bf16_kernel<3>, grid, dim3(256), 0, ctx->stream, x, ci, g,
                               w, b, co, y, accumulate);
        else
            hipLaunchKernelGGL(conv_fwd_bf16_kernel<1>, grid, dim3(256), 0, ctx->stream, x, ci, g,
                               w, b, co, y, accumulate);
        TT_LAUNCH_CHECK();
        return TTDFT_OK;
    }
    if (ks == 3)
        hipLaunchKernelGGL(conv_fwd_kernel<3>, grid, dim3(256), 0, ctx->stream, x, ci, g, w, b, co,
                           y, accumulate);
    else
        hipLaunchKernelGGL(conv_fwd_kernel<1>, grid, dim3(256), 0, ctx->stream, x, ci, g, w, b, co,
                           y, accumulate);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int check_conv(ttdft_ctx ctx, int ci, int co, int ks) {
    TT_CHECK(ctx, TTDFT_EBADARG, "nn: NULL context");
    TT_CHECK(ci >= 1 && co >= 1 && (ks == 1 || ks == 3), TTDFT_EBADARG,
             "conv3d: channels %d -> %d, kernel %d (1 or 3)", ci, co, ks);
    TT_CHECK((int64_t)ci * ks * ks * ks < (1 << 30), TTDFT_EBADARG, "conv3d: too many inputs");
    return TTDFT_OK;
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" {

int ttdft_nn_set_math(ttdft_ctx ctx, int math) {
    clear_error();
    TT_CHECK(ctx && (math == TTDFT_NN_FP32 || math == TTDFT_NN_BF16), TTDFT_EBADARG,
             "nn_set_math: unknown mode %d", math);
    ctx->nn_math = math;
    return TTDFT_OK;
}

int ttdft_nn_conv3d_forward(ttdft_ctx ctx, const float *d_x, int32_t ci, const int32_t dims[3],
                            const float *d_w, const float *d_b, int32_t co, int32_t ks,
                            float *d_y, int accumulate) {
    clear_error();
    TT_TRY(check_conv(ctx, ci, co, ks));
    Geo g;
    TT_TRY(make_geo(dims, g));
    TT_CHECK(d_x && d_w && d_y, TTDFT_EBADARG, "conv3d_forward: NULL tensor");
    TT_HIP(hipSetDevice(ctx->device));
    return conv_forward(ctx, d_x, ci, g, d_w, d_b, co, ks, d_y, accumulate);
}

int ttdft_nn_conv3d_backward_data(ttdft_ctx ctx, const float *d_dy, int32_t co,
                                  const int32_t dims[3], const float *d_w, int32_t ci,
                                  int32_t ks, float *d_dx, int accumulate) {
    clear_error();
    TT_TRY(check_conv(ctx, ci, co, ks));
    Geo g;
    TT_TRY(make_geo(dims, g));
    TT_CHECK(d_dy && d_w && d_dx, TTDFT_EBADARG, "conv3d_backward_data: NULL tensor");
    TT_HIP(hipSetDevice(ctx->device));
    const int KT = ks * ks * ks;
    const int64_t nw = (int64_t)co * ci * KT;
    TT_TRY(ctx->nn_scratch[0].ensure(sizeof(float) * nw));
    float *wf = (float *)ctx->nn_scratch[0].ptr;
    hipLaunchKernelGGL(flip_weights_kernel, dim3(nblk(nw)), dim3(256), 0, ctx->stream, d_w, co,
                       ci, KT, wf);
    TT_LAUNCH_CHECK();
    return conv_forward(ctx, d_dy, co, g, wf, nullptr, ci, ks, d_dx, accumulate);
}

int ttdft_nn_conv3d_backward_weight(ttdft_ctx ctx, const float *d_x, int32_t ci,
                                    const int32_t dims[3], const float *d_dy, int32_t co,
                                    int32_t ks, float *d_dw, float *d_db, int accumulate) {
    clear_error();
    TT_TRY(check_conv(ctx, ci, co, ks));
    Geo g;
    TT_TRY(make_geo(dims, g));
    TT_CHECK(d_x && d_dy && d_dw, TTDFT_EBADARG, "conv3d_backward_weight: NULL tensor");
    TT_HIP(hipSetDevice(ctx->device));
    const int K = ci * ks * ks * ks;
    const int gx = (K + 63) / 64, gy = (co + kTM - 1) / kTM;
    // position ranges: about 1024 workgroups in all, each range a multiple of 16
    int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(1024 / (gx * gy), (g.P + 255) / 256));
    int64_t span = ((g.P + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = (g.P + span - 1) / span;
    const int64_t nw = (int64_t)co * K;
    TT_TRY(ctx->nn_scratch[1].ensure(sizeof(float) * nw * nsplit));
    float *partial = (float *)ctx->nn_scratch[1].ptr;
    const dim3 wg(gx, gy, (unsigned)nsplit);
    const bool bf = ctx->nn_math == TTDFT_NN_BF16;
    if (ks == 3 && bf)
        hipLaunchKernelGGL(conv_wgrad_bf16_kernel<3>, wg, dim3(256), 0, ctx->stream, d_x, ci, g,
                           d_dy, co, span, partial);
    else if (bf)
        hipLaunchKernelGGL(conv_wgrad_bf16_kernel<1>, wg, dim3(256), 0, ctx->stream, d_x, ci, g,
                           d_dy, co, span, partial);
    else if (ks == 3)
        hipLaunchKernelGGL(conv_wgrad_kernel<3>, wg, dim3(256), 0, ctx->stream, d_x, ci, g, d_dy,
                           co, span, partial);
    else
        hipLaunchKernelGGL(conv_wgrad_kernel<1>, wg, dim3(256), 0, ctx->stream, d_x, ci, g, d_dy,
                           co, span, partial);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(split_reduce_kernel, dim3(nblk(nw)), dim3(256), 0, ctx->stream, partial,
                       (int)nsplit, nw, d_dw, accumulate);
    TT_LAUNCH_CHECK();
    if (d_db) {
        hipLaunchKernelGGL(bias_grad_kernel, dim3(co), dim3(256), 0, ctx->stream, d_dy, g.P, d_db,
                           accumulate);
        TT_LAUNCH_CHECK();
    }
    return TTDFT_OK;
}

int ttdft_nn_leaky_relu(ttdft_ctx ctx, float *d_y, int64_t n, float slope) {
    clear_error();
    TT_CHECK(ctx && d_y && n >= 0, TTDFT_EBADARG, "leaky_relu: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(leaky_fwd_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_y, n, slope);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_leaky_relu_backward(ttdft_ctx ctx, const float *d_y, float *d_dy, int64_t n,
                                 float slope) {
    clear_error();
    TT_CHECK(ctx && d_y && d_dy && n >= 0, TTDFT_EBADARG, "leaky_relu_backward: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(leaky_bwd_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_y, d_dy, n,
                       slope);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_avgpool2(ttdft_ctx ctx, const float *d_fine, int32_t c, const int32_t fine_dims[3],
                      float *d_coarse) {
    clear_error();
    Geo f;
    TT_TRY(make_geo(fine_dims, f));
    TT_CHECK(ctx && d_fine && d_coarse && c >= 1 && f.D % 2 == 0 && f.H % 2 == 0 && f.W % 2 == 0,
             TTDFT_EBADARG, "avgpool2: bad arguments (even dims needed)");
    TT_HIP(hipSetDevice(ctx->device));
    const int32_t cd[3] = {f.D / 2, f.H / 2, f.W / 2};
    Geo cg;
    TT_TRY(make_geo(cd, cg));
    hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblk(c * cg.P)), dim3(256), 0, ctx->stream, d_fine,
                       c, cg, d_coarse);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_avgpool2_backward(ttdft_ctx ctx, const float *d_dcoarse, int32_t c,
                               const int32_t fine_dims[3], float *d_dfine, int accumulate) {
    clear_error();
    Geo f;
    TT_TRY(make_geo(fine_dims, f));
    TT_CHECK(ctx && d_dcoarse && d_dfine && c >= 1 && f.D % 2 == 0 && f.H % 2 == 0 &&
                 f.W % 2 == 0,
             TTDFT_EBADARG, "avgpool2_backward: bad arguments (even dims needed)");
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(spread_kernel, dim3(nblk(c * f.P)), dim3(256), 0, ctx->stream, d_dcoarse,
                       c, f, 0.125f, d_dfine, accumulate);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_upsample2(ttdft_ctx ctx, const float *d_coarse, int32_t c,
                       const int32_t fine_dims[3], float *d_fine) {
    clear_error();
    Geo f;
    TT_TRY(make_geo(fine_dims, f));
    TT_CHECK(ctx && d_coarse && d_fine && c >= 1 && f.D % 2 == 0 && f.H % 2 == 0 && f.W % 2 == 0,
             TTDFT_EBADARG, "upsample2: bad arguments (even dims needed)");
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(spread_kernel, dim3(nblk(c * f.P)), dim3(256), 0, ctx->stream, d_coarse, c,
                       f, 1.0f, d_fine, 0);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_upsample2_backward(ttdft_ctx ctx, const float *d_dfine, int32_t c,
                                const int32_t fine_dims[3], float *d_dcoarse, int accumulate) {
    clear_error();
    Geo f;
    TT_TRY(make_geo(fine_dims, f));
    TT_CHECK(ctx && d_dfine && d_dcoarse && c >= 1 && f.D % 2 == 0 && f.H % 2 == 0 &&
                 f.W % 2 == 0,
             TTDFT_EBADARG, "upsample2_backward: bad arguments (even dims needed)");
    TT_HIP(hipSetDevice(ctx->device));
    const int32_t cd[3] = {f.D / 2, f.H / 2, f.W / 2};
    Geo cg;
    TT_TRY(make_geo(cd, cg));
    hipLaunchKernelGGL(gather8_kernel, dim3(nblk(c * cg.P)), dim3(256), 0, ctx->stream, d_dfine,
                       c, cg, d_dcoarse, accumulate);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_mse(ttdft_ctx ctx, const float *d_y, const float *d_t, int64_t n, float *d_dy,
                 double *loss) {
    clear_error();
    TT_CHECK(ctx && d_y && d_t && n > 0 && loss, TTDFT_EBADARG, "mse: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    constexpr int kBlocks = 512;
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * kBlocks));
    double *partial = (double *)ctx->scratch[6].ptr;
    hipLaunchKernelGGL(mse_kernel, dim3(kBlocks), dim3(256), 0, ctx->stream, d_y, d_t, n, d_dy,
                       partial);
    TT_LAUNCH_CHECK();
    std::vector<double> h(kBlocks);
    TT_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * kBlocks, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / (double)n;
    return TTDFT_OK;
}

int ttdft_nn_adam(ttdft_ctx ctx, float *d_p, const float *d_g, float *d_m, float *d_v, int64_t n,
                  float lr, float beta1, float beta2, float eps, int32_t step) {
    clear_error();
    TT_CHECK(ctx && d_p && d_g && d_m && d_v && n >= 0 && step >= 1, TTDFT_EBADARG,
             "adam: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, step)));
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_p, d_g, d_m, d_v,
                       n, lr, beta1, beta2, eps, c1, c2);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_to_f32(ttdft_ctx ctx, const double *d_src, float *d_dst, int64_t n, double scale) {
    clear_error();
    TT_CHECK(ctx && d_src && d_dst && n >= 0, TTDFT_EBADARG, "to_f32: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(f64_to_f32_scaled, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_src, d_dst,
                       n, scale);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_nn_to_f64(ttdft_ctx ctx, const float *d_src, double *d_dst, int64_t n, double scale) {
    clear_error();
    TT_CHECK(ctx && d_src && d_dst && n >= 0, TTDFT_EBADARG, "to_f64: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(f32_to_f64_scaled, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_src, d_dst,
                       n, scale);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // extern "C"
