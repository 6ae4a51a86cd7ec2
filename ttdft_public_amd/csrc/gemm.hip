// Tall-skinny dense contractions of the subspace iteration on MFMA:
//   Gram      S = X^T Y   (n x k)^T (n x k) -> k x k     (MultTM::mult)
//   rotation  Z = X Q     (n x k)(k x k)    -> n x k     (orth, X <- X Q)
// MFMA output layouts differ by type: for f64 16x16x4, element v of lane l
// holds D[l/16 + 4v][l%16]; for f32 16x16x16 bf16 it holds D[4(l/16) + v][l%16].
// Two arithmetic modes:
//   F64    : FP64 blocks, v_mfma_f64_16x16x4_f64, FP64 accumulation.
//   BF16X3 : FP32 blocks, each operand split as x = hi + lo (both bf16), the
//            product taken as hi*hi' + hi*lo' + lo*hi' on
//            v_mfma_f32_16x16x16_bf16 with FP32 accumulation.
// The Gram is split over row chunks; each chunk's k x k partial is reduced
// in FP64 by a second kernel in a fixed order (deterministic).
#include "internal.h"

namespace ttdft {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma_bf16(s4 a, s4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ short f2bf(float x) {
    unsigned u = __builtin_bit_cast(unsigned, x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) {
    return __builtin_bit_cast(float, ((unsigned)(unsigned short)h) << 16);
}
__device__ __forceinline__ void split_bf16(float x, short &hi, short &lo) {
    hi = f2bf(x);
    lo = f2bf(x - bf2f(hi));
}

// One lane's split bf16 operand of a 16x16x16 MFMA: four values as hi and lo
// halves (16 bytes).
struct QFrag {
    s4 hi, lo;
};

// ---------------------------------------------------------------------------
// Gram, split-BF16, staged through LDS.  A workgroup (4 waves) owns a block of
// 160 x 160 of S (10 x 10 tiles of 16) and a chunk of rows; a wave owns 5 x 5
// tiles.  Per 16-row step the workgroup loads the block's 160 columns of X
// and of Y once (float4 along a row when aligned), splits every value once
// into bf16 hi + lo and stores it in LDS in fragment order: column c holds,
// for each row quad q, {hi[4], lo[4]} of rows 4q..4q+3 as 16 bytes at
// c * 80 + 16 q (the 80-byte column pitch keeps a quarter-wave's 16-byte
// reads on distinct banks).  A fragment (lane i + 16 q) is then one 16-byte
// LDS read; each A fragment serves 5 B fragments and each B fragment 5 A
// fragments (3 MFMAs per pair).  The next step's global loads are issued
// before the current step's MFMAs and two LDS stages alternate, so one
// barrier per step suffices.
//
// With sym (X == Y) only blocks bi <= bj are launched and, within diagonal
// blocks, only tiles ta <= tb are computed; the reduction reads S[a][b] from
// the tile of (min(a,b), max(a,b)), so S comes out exactly symmetric.
// Partials are written per chunk in tile order: tile (ta, tb) of chunk c at
// P + (c * nt * nt + tb * nt + ta) * 256, lane l's accumulator as a float4 at
// 4 l (element v = D[4 (l/16) + v][l % 16]).
constexpr int kGramTiles = 5;                 // tiles per wave per dimension
constexpr int kGramBlock = 32 * kGramTiles;   // block edge: 2 waves x 5 tiles x 16
constexpr int kGramPitch = 80;                // LDS bytes per staged column
constexpr int kGramItems = 2 * (kGramBlock / 4) * 4;  // (matrix, column quad, row quad)

struct StageRegs {
    f4 v[2][4];  // up to two items of 4 rows x 4 columns
};

__device__ __forceinline__ void gram_stage_load(const float *__restrict__ X, int64_t ldx,
                                                const float *__restrict__ Y, int64_t ldy,
                                                int64_t r, int64_t r1, int32_t k, int a0, int b0,
                                                int vec, StageRegs &s) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int item = threadIdx.x + it * 256;
        if (item >= kGramItems) break;
        const int m = item / (kGramItems / 2);
        const int rem = item % (kGramItems / 2);
        const int cq = rem >> 2, q = rem & 3;
        const float *M = m ? Y : X;
        const int64_t ld = m ? ldy : ldx;
        const int c = (m ? b0 : a0) + 4 * cq;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = r + 4 * q + v;
            f4 x = {0.f, 0.f, 0.f, 0.f};
            if (row < r1) {
                const float *p = M + row * ld + c;
                if (vec && c + 3 < k) {
                    x = *reinterpret_cast<const f4 *>(p);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) x[j] = c + j < k ? p[j] : 0.f;
                }
            }
            s.v[it][v] = x;
        }
    }
}

__device__ __forceinline__ void gram_stage_store(unsigned char *lds, const StageRegs &s) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int item = threadIdx.x + it * 256;
        if (item >= kGramItems) break;
        const int m = item / (kGramItems / 2);
        const int rem = item % (kGramItems / 2);
        const int cq = rem >> 2, q = rem & 3;
        unsigned char *base = lds + m * (kGramBlock * kGramPitch);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s4 hi, lo;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                short h, l;
                split_bf16(s.v[it][v][j], h, l);
                hi[v] = h;
                lo[v] = l;
            }
            QFrag f;
            f.hi = hi;
            f.lo = lo;
            *reinterpret_cast<QFrag *>(base + (4 * cq + j) * kGramPitch + 16 * q) = f;
        }
    }
}

__global__ __launch_bounds__(256, 2) void gram_bf16x3_lds_kernel(
    const float *__restrict__ X, int64_t ldx, const float *__restrict__ Y, int64_t ldy,
    int64_t n, int32_t k, int32_t nb, int32_t nt, int sym, int64_t rows_per_chunk, int vec,
    float *__restrict__ P) {
    __shared__ __align__(16) unsigned char lds[2][2 * kGramBlock * kGramPitch];
    int bi, bj;
    if (sym) {  // upper triangle of blocks, row by row
        int p = blockIdx.x, len = nb;
        bi = 0;
        while (p >= len) {
            p -= len;
            ++bi;
            --len;
        }
        bj = bi + p;
    } else {
        bi = blockIdx.x % nb;
        bj = blockIdx.x / nb;
    }
    const int a0 = bi * kGramBlock, b0 = bj * kGramBlock;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wa = (wave & 1) * 16 * kGramTiles, wb = (wave >> 1) * 16 * kGramTiles;
    const int ta0 = (a0 + wa) / 16, tb0 = (b0 + wb) / 16;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t r1 = (r0 + rows_per_chunk) < n ? (r0 + rows_per_chunk) : n;
    // tile (ta0 + i, tb0 + j) is computed when inside S and, with sym, on or
    // above the diagonal (wave-uniform)
    bool on[kGramTiles][kGramTiles];
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j)
            on[i][j] = ta0 + i < nt && tb0 + j < nt && (!sym || ta0 + i <= tb0 + j);
    f4 acc[kGramTiles][kGramTiles];
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    const int li = lane & 15, lq = lane >> 4;
    StageRegs sr;
    gram_stage_load(X, ldx, Y, ldy, r0, r1, k, a0, b0, vec, sr);
    int buf = 0;
    for (int64_t r = r0; r < r1; r += 16) {
        gram_stage_store(lds[buf], sr);
        __syncthreads();
        // the next step's loads fly while this step's MFMAs run
        if (r + 16 < r1) gram_stage_load(X, ldx, Y, ldy, r + 16, r1, k, a0, b0, vec, sr);
        const unsigned char *lx = lds[buf] + (wa + li) * kGramPitch + 16 * lq;
        const unsigned char *ly = lds[buf] + kGramBlock * kGramPitch + (wb + li) * kGramPitch +
                                  16 * lq;
        QFrag fa[kGramTiles];
#pragma unroll
        for (int i = 0; i < kGramTiles; ++i)
            fa[i] = *reinterpret_cast<const QFrag *>(lx + i * 16 * kGramPitch);
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j) {
            const QFrag fb = *reinterpret_cast<const QFrag *>(ly + j * 16 * kGramPitch);
#pragma unroll
            for (int i = 0; i < kGramTiles; ++i) {
                if (!on[i][j]) continue;
                acc[i][j] = mfma_bf16(fa[i].lo, fb.hi, acc[i][j]);
                acc[i][j] = mfma_bf16(fa[i].hi, fb.lo, acc[i][j]);
                acc[i][j] = mfma_bf16(fa[i].hi, fb.hi, acc[i][j]);
            }
        }
        buf ^= 1;
    }
    float *Pc = P + (int64_t)blockIdx.y * nt * nt * 256;
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j)
            if (on[i][j])
                *reinterpret_cast<f4 *>(Pc + ((int64_t)(tb0 + j) * nt + ta0 + i) * 256 +
                                        4 * lane) = acc[i][j];
}

// ---------------------------------------------------------------------------
// Gram, FP64, staged through LDS.  The blocking of the split-BF16 kernel
// above: a workgroup (4 waves) owns a 160 x 160 block of S and a chunk of
// rows, a wave 5 x 5 tiles, 16-row steps, only blocks bi <= bj and tiles
// ta <= tb for X^T X.  The step's rows are staged row-major (a row's 160
// block columns contiguous, 32-byte stores from coalesced global loads).  In
// a K-step lane i + 16 kk reads row 4 ks + kk, column i; the row pitch of
// 1408 bytes (352 dwords, 32 mod 64) puts rows kk and kk + 1 on opposite bank
// halves, so each half-wave's 32 reads hit 64 distinct banks.  Per 16-row
// step a wave does 4 K-steps x (5 + 5) 8-byte LDS reads
// and 100 MFMAs.  Partials: tile (ta, tb) of chunk c at
// P + (c nt nt + tb nt + ta) 256, lane l's d4 at 4 l (D[l/16 + 4v][l%16]).
constexpr int kG64Row = 176;  // doubles per staged row (160 + 16 padding)
constexpr int kG64Items = 2 * 16 * (kGramBlock / 4);  // (matrix, row, column quad)

struct Stage64 {
    d4 v[5];  // up to five items of 4 consecutive columns of one row
};

__device__ __forceinline__ void gram64_stage_load(const double *__restrict__ X, int64_t ldx,
                                                  const double *__restrict__ Y, int64_t ldy,
                                                  int64_t r, int64_t r1, int32_t k, int a0,
                                                  int b0, int vec, Stage64 &s) {
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const int item = threadIdx.x + it * 256;
        if (item >= kG64Items) break;
        const int m = item / (kG64Items / 2);
        const int rem = item % (kG64Items / 2);
        const int row = rem / (kGramBlock / 4), cq = rem % (kGramBlock / 4);
        const double *M = m ? Y : X;
        const int64_t ld = m ? ldy : ldx;
        const int c = (m ? b0 : a0) + 4 * cq;
        const int64_t gr = r + row;
        d4 x = {0.0, 0.0, 0.0, 0.0};
        if (gr < r1) {
            const double *p = M + gr * ld + c;
            if (vec && c + 3 < k) {
                typedef double d2 __attribute__((ext_vector_type(2)));
                const d2 lo = *reinterpret_cast<const d2 *>(p);
                const d2 hi = *reinterpret_cast<const d2 *>(p + 2);
                x = d4{lo[0], lo[1], hi[0], hi[1]};
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) x[j] = c + j < k ? p[j] : 0.0;
            }
        }
        s.v[it] = x;
    }
}

__device__ __forceinline__ void gram64_stage_store(double *lds, const Stage64 &s) {
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const int item = threadIdx.x + it * 256;
        if (item >= kG64Items) break;
        const int m = item / (kG64Items / 2);
        const int rem = item % (kG64Items / 2);
        const int row = rem / (kGramBlock / 4), cq = rem % (kGramBlock / 4);
        *reinterpret_cast<d4 *>(lds + (m * 16 + row) * kG64Row + 4 * cq) = s.v[it];
    }
}

__global__ __launch_bounds__(256, 1) void gram_f64_lds_kernel(
    const double *__restrict__ X, int64_t ldx, const double *__restrict__ Y, int64_t ldy,
    int64_t n, int32_t k, int32_t nb, int32_t nt, int sym, int64_t rows_per_chunk, int vec,
    double *__restrict__ P) {
    __shared__ __align__(32) double lds[2][2 * 16 * kG64Row];
    int bi, bj;
    if (sym) {
        int p = blockIdx.x, len = nb;
        bi = 0;
        while (p >= len) {
            p -= len;
            ++bi;
            --len;
        }
        bj = bi + p;
    } else {
        bi = blockIdx.x % nb;
        bj = blockIdx.x / nb;
    }
    const int a0 = bi * kGramBlock, b0 = bj * kGramBlock;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wa = (wave & 1) * 16 * kGramTiles, wb = (wave >> 1) * 16 * kGramTiles;
    const int ta0 = (a0 + wa) / 16, tb0 = (b0 + wb) / 16;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t r1 = (r0 + rows_per_chunk) < n ? (r0 + rows_per_chunk) : n;
    bool on[kGramTiles][kGramTiles];
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j)
            on[i][j] = ta0 + i < nt && tb0 + j < nt && (!sym || ta0 + i <= tb0 + j);
    d4 acc[kGramTiles][kGramTiles];
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    const int li = lane & 15, lq = lane >> 4;
    Stage64 sr;
    gram64_stage_load(X, ldx, Y, ldy, r0, r1, k, a0, b0, vec, sr);
    int buf = 0;
    for (int64_t r = r0; r < r1; r += 16) {
        gram64_stage_store(lds[buf], sr);
        __syncthreads();
        if (r + 16 < r1) gram64_stage_load(X, ldx, Y, ldy, r + 16, r1, k, a0, b0, vec, sr);
        const double *lx = lds[buf] + lq * kG64Row + wa + li;
        const double *ly = lds[buf] + (16 + lq) * kG64Row + wb + li;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            double fa[kGramTiles];
#pragma unroll
            for (int i = 0; i < kGramTiles; ++i) fa[i] = lx[4 * ks * kG64Row + 16 * i];
#pragma unroll
            for (int j = 0; j < kGramTiles; ++j) {
                const double fb = ly[4 * ks * kG64Row + 16 * j];
#pragma unroll
                for (int i = 0; i < kGramTiles; ++i)
                    if (on[i][j]) acc[i][j] = mfma_f64(fa[i], fb, acc[i][j]);
            }
        }
        buf ^= 1;
    }
    double *Pc = P + (int64_t)blockIdx.y * nt * nt * 256;
#pragma unroll
    for (int i = 0; i < kGramTiles; ++i)
#pragma unroll
        for (int j = 0; j < kGramTiles; ++j)
            if (on[i][j])
                *reinterpret_cast<d4 *>(Pc + ((int64_t)(tb0 + j) * nt + ta0 + i) * 256 +
                                        4 * lane) = acc[i][j];
}

// S[a + k b] = sum over chunks (FP64, fixed order) of the tile partials; with
// sym the pair is read as (min(a,b), max(a,b)).  Within a tile, D[ai][bj]
// is element v of lane l: for the bf16 MFMA (P float) l = bj + 16 (ai/4),
// v = ai%4; for the f64 MFMA (P double) l = bj + 16 (ai%4), v = ai/4.
template <typename P_t>
__global__ void reduce_tile_partials_kernel(const P_t *__restrict__ P, int nchunks, int32_t k,
                                            int32_t nt, int sym, double *__restrict__ S) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)k * k) return;
    int a = (int)(idx % k), b = (int)(idx / k);
    if (sym && a > b) {
        const int t = a;
        a = b;
        b = t;
    }
    const int ta = a >> 4, tb = b >> 4, ai = a & 15, bj = b & 15;
    const int lv = sizeof(P_t) == 4 ? 4 * (bj + 16 * (ai >> 2)) + (ai & 3)
                                    : 4 * (bj + 16 * (ai & 3)) + (ai >> 2);
    const int64_t off = ((int64_t)tb * nt + ta) * 256 + lv;
    const int64_t stride = (int64_t)nt * nt * 256;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += (double)P[c * stride + off];
    S[idx] = s;
}

// ---------------------------------------------------------------------------
// Q (k x k column-major FP64) split once into bf16 hi/lo B fragments of the
// 16x16x16 MFMA: fragment (step s, tile t) holds Q[16s + kk][16t + j] in lane
// j + 16 (kk/4), element kk%4, as 8 shorts per lane (hi[4], lo[4]), zero past
// k.  Qf[((s * ntile + t) * 64 + lane)].
__global__ void split_q_frag_kernel(const double *__restrict__ Q, int32_t k, int32_t nstep,
                                    int32_t ntile, QFrag *__restrict__ Qf) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nstep * ntile * 64) return;
    const int lane = (int)(idx & 63);
    const int64_t st = idx >> 6;
    const int t = (int)(st % ntile), s = (int)(st / ntile);
    const int b = 16 * t + (lane & 15);
    QFrag f;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int a = 16 * s + 4 * (lane >> 4) + v;
        short h = 0, l = 0;
        if (a < k && b < k) split_bf16((float)Q[a + (int64_t)k * b], h, l);
        f.hi[v] = h;
        f.lo[v] = l;
    }
    Qf[idx] = f;
}

// Rotation, split-BF16: Z = X Q for FP32 X.  A wave computes 32 rows (two
// 16-row MFMA tiles) by RT16 column tiles: per 16-deep step it loads its two
// rows' 4 values each as one 16-byte vector (split once into hi/lo), then one
// 16-byte pre-split Q fragment per column tile, shared by both row tiles
// (three MFMAs each).  A[i][kk] = X[row i][a + kk] in lane i + 16 (kk/4),
// element kk%4.
template <int RT16>
__global__ __launch_bounds__(256) void rotate_bf16x3_frag_kernel(
    const float *__restrict__ X, int64_t ldx, const QFrag *__restrict__ Qf, int32_t nstep,
    int32_t ntile, float *__restrict__ Z, int64_t ldz, int64_t n, int32_t k, int vec) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n0 = (int64_t)blockIdx.x * 128 + wave * 32;
    const int t0 = blockIdx.y * RT16;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t row0 = n0 + li, row1 = n0 + 16 + li;
    const float *x0 = X + (row0 < n ? row0 : n - 1) * ldx;
    const float *x1 = X + (row1 < n ? row1 : n - 1) * ldx;
    f4 acc[2][RT16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int t = 0; t < RT16; ++t) acc[r][t] = f4{0, 0, 0, 0};
    for (int s = 0; s < nstep; ++s) {
        const int a = 16 * s + 4 * lk;
        f4 v0, v1;
        if (vec && a + 3 < k) {  // rows 16-byte aligned (checked on the host)
            v0 = *reinterpret_cast<const f4 *>(x0 + a);
            v1 = *reinterpret_cast<const f4 *>(x1 + a);
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                v0[v] = a + v < k ? x0[a + v] : 0.0f;
                v1[v] = a + v < k ? x1[a + v] : 0.0f;
            }
        }
        s4 h0, l0, h1, l1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            short h, l;
            split_bf16(v0[v], h, l);
            h0[v] = h;
            l0[v] = l;
            split_bf16(v1[v], h, l);
            h1[v] = h;
            l1[v] = l;
        }
        const QFrag *qs = Qf + ((int64_t)s * ntile + t0) * 64 + lane;
#pragma unroll
        for (int t = 0; t < RT16; ++t) {
            if (t0 + t >= ntile) break;
            const QFrag q = qs[t * 64];
            acc[0][t] = mfma_bf16(l0, q.hi, acc[0][t]);
            acc[0][t] = mfma_bf16(h0, q.lo, acc[0][t]);
            acc[0][t] = mfma_bf16(h0, q.hi, acc[0][t]);
            acc[1][t] = mfma_bf16(l1, q.hi, acc[1][t]);
            acc[1][t] = mfma_bf16(h1, q.lo, acc[1][t]);
            acc[1][t] = mfma_bf16(h1, q.hi, acc[1][t]);
        }
    }
    // D[4(l/16) + v][l%16]: rows n0 + 16 r + 4 lk + v, column 16 (t0 + t) + li
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = n0 + 16 * r + 4 * lk + v;
            if (row >= n) continue;
#pragma unroll
            for (int t = 0; t < RT16; ++t) {
                const int b = 16 * (t0 + t) + li;
                if (t0 + t < ntile && b < k) Z[row * ldz + b] = acc[r][t][v];
            }
        }
}

// Q (k x k column-major FP64) in B-fragment order of the f64 16x16x4 MFMA,
// the K index of each 16-row chunk permuted so that an A-operand lane reads
// four consecutive columns of X: chunk s, K-step t, column tile c, lane
// j + 16 kk holds Q[16 s + 4 kk + t][16 c + j] (zero past k) at
// Qd[((s * 4 + t) * ntile + c) * 64 + lane].
__global__ void q_frag64_kernel(const double *__restrict__ Q, int32_t k, int32_t nstep,
                                int32_t ntile, double *__restrict__ Qd) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nstep * 4 * ntile * 64) return;
    const int lane = (int)(idx & 63);
    int64_t rest = idx >> 6;
    const int c = (int)(rest % ntile);
    rest /= ntile;
    const int t = (int)(rest & 3), s = (int)(rest >> 2);
    const int a = 16 * s + 4 * (lane >> 4) + t, b = 16 * c + (lane & 15);
    Qd[idx] = (a < k && b < k) ? Q[a + (int64_t)k * b] : 0.0;
}

// Rotation, FP64: Z = X Q.  A wave computes 32 rows (two 16-row tiles) by
// RT16 column tiles, a workgroup 128 rows.  Per 16-column chunk s of X a
// lane loads its two rows' columns 16 s + 4 kk .. + 3 once (two 16-byte
// loads each when aligned); K-step t of the chunk takes column 16 s + 4 kk + t,
// so the four values are the A operands of the chunk's four K-steps.  Each
// pre-arranged Q fragment (one 8-byte load per lane, from L2) serves both
// row tiles.
template <int RT16>
__global__ __launch_bounds__(256) void rotate_f64_frag_kernel(
    const double *__restrict__ X, int64_t ldx, const double *__restrict__ Qd, int32_t nstep,
    int32_t ntile, double *__restrict__ Z, int64_t ldz, int64_t n, int32_t k, int vec) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n0 = (int64_t)blockIdx.x * 128 + wave * 32;
    const int t0 = blockIdx.y * RT16;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t row0 = n0 + li, row1 = n0 + 16 + li;
    // rows past n read row n - 1; their products are never stored
    const double *x0 = X + (row0 < n ? row0 : n - 1) * ldx;
    const double *x1 = X + (row1 < n ? row1 : n - 1) * ldx;
    d4 acc[2][RT16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < RT16; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < nstep; ++s) {
        const int a = 16 * s + 4 * lk;
        d4 v0, v1;
        if (vec && a + 3 < k) {  // rows 16-byte aligned (checked on the host)
            const d2 p0 = *reinterpret_cast<const d2 *>(x0 + a);
            const d2 p1 = *reinterpret_cast<const d2 *>(x0 + a + 2);
            const d2 q0 = *reinterpret_cast<const d2 *>(x1 + a);
            const d2 q1 = *reinterpret_cast<const d2 *>(x1 + a + 2);
            v0 = d4{p0[0], p0[1], p1[0], p1[1]};
            v1 = d4{q0[0], q0[1], q1[0], q1[1]};
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v0[j] = a + j < k ? x0[a + j] : 0.0;
                v1[j] = a + j < k ? x1[a + j] : 0.0;
            }
        }
        const double *qs = Qd + ((int64_t)s * 4 * ntile + t0) * 64 + lane;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < RT16; ++c) {
                if (t0 + c >= ntile) break;
                const double q = qs[((int64_t)t * ntile + c) * 64];
                acc[0][c] = mfma_f64(v0[t], q, acc[0][c]);
                acc[1][c] = mfma_f64(v1[t], q, acc[1][c]);
            }
    }
    // D[l/16 + 4v][l%16]: rows n0 + 16 r + lk + 4 v, column 16 (t0 + c) + li
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = n0 + 16 * r + lk + 4 * v;
            if (row >= n) continue;
#pragma unroll
            for (int c = 0; c < RT16; ++c) {
                const int b = 16 * (t0 + c) + li;
                if (t0 + c < ntile && b < k) Z[row * ldz + b] = acc[r][c][v];
            }
        }
}

__global__ void transpose_small_kernel(const double *__restrict__ A, double *__restrict__ AT,
                                       int32_t k) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)k * k) return;
    const int i = (int)(idx % k), j = (int)(idx / k);
    AT[j + (int64_t)k * i] = A[i + (int64_t)k * j];
}

template <typename S, typename D>
__global__ void copy_block_kernel(const S *__restrict__ src, int64_t lds, D *__restrict__ dst,
                                  int64_t ldd, int64_t n, int32_t k) {
    const int64_t total = n * (int64_t)k;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx / k;
        const int j = (int)(idx - i * k);
        dst[i * ldd + j] = (D)src[i * lds + j];
    }
}

}  // namespace

int launch_gram(ttdft_ctx ctx, const void *X, int64_t ldx, const void *Y, int64_t ldy,
                int64_t n, int32_t k, int dtype, int mode, double *dS) {
    TT_CHECK(k >= 1 && n >= 1, TTDFT_EBADARG, "gram: n=%lld k=%d", (long long)n, k);
    TT_CHECK((mode == TTDFT_GEMM_F64 && dtype == TTDFT_F64) ||
                 (mode == TTDFT_GEMM_BF16X3 && dtype == TTDFT_F32),
             TTDFT_EBADARG, "gram: mode %d does not match dtype %d", mode, dtype);
    const int64_t kk = (int64_t)k * k;
    PendingTiming t;
    // X^T X is computed as the upper triangle of blocks and mirrored
    const int sym = (X == Y && ldx == ldy) ? 1 : 0;
    const int nb = (k + kGramBlock - 1) / kGramBlock;
    const int nt = (k + 15) / 16;
    const int blocks = sym ? nb * (nb + 1) / 2 : nb * nb;
    // about two workgroups per CU in all, each chunk at least 8 steps
    int64_t nchunks = (512 + blocks - 1) / blocks;
    const int64_t max_chunks = (n + 127) / 128;
    if (nchunks > max_chunks) nchunks = max_chunks;
    if (nchunks < 1) nchunks = 1;
    int64_t rows = (n + nchunks - 1) / nchunks;
    rows = (rows + 15) / 16 * 16;
    nchunks = (n + rows - 1) / rows;
    const size_t es_p = mode == TTDFT_GEMM_F64 ? sizeof(double) : sizeof(float);
    TT_TRY(ctx->scratch[2].ensure((size_t)nchunks * nt * nt * 256 * es_p));
    const dim3 grid((unsigned)blocks, (unsigned)nchunks), red((unsigned)((kk + 255) / 256));
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_GRAM, &t));
    if (mode == TTDFT_GEMM_F64) {
        double *P = (double *)ctx->scratch[2].ptr;
        const int vec = (ldx % 2 == 0) && (ldy % 2 == 0) &&
                        (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                        (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
        hipLaunchKernelGGL(gram_f64_lds_kernel, grid, dim3(256), 0, ctx->stream,
                           (const double *)X, ldx, (const double *)Y, ldy, n, k, nb, nt, sym,
                           rows, vec, P);
        TT_LAUNCH_CHECK();
        hipLaunchKernelGGL(reduce_tile_partials_kernel<double>, red, dim3(256), 0, ctx->stream, P,
                           (int)nchunks, k, nt, sym, dS);
        TT_LAUNCH_CHECK();
    } else {
        float *P = (float *)ctx->scratch[2].ptr;
        const int vec = (ldx % 4 == 0) && (ldy % 4 == 0) &&
                        (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                        (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
        hipLaunchKernelGGL(gram_bf16x3_lds_kernel, grid, dim3(256), 0, ctx->stream,
                           (const float *)X, ldx, (const float *)Y, ldy, n, k, nb, nt, sym, rows,
                           vec, P);
        TT_LAUNCH_CHECK();
        hipLaunchKernelGGL(reduce_tile_partials_kernel<float>, red, dim3(256), 0, ctx->stream, P,
                           (int)nchunks, k, nt, sym, dS);
        TT_LAUNCH_CHECK();
    }
    const double es = (double)dtype_size(dtype);
    TT_TRY(timing_end(ctx, &t, 2.0 * n * k * es + kk * 8.0,
                      2.0 * n * kk * (mode == TTDFT_GEMM_F64 ? 1.0 : 3.0)));
    return TTDFT_OK;
}

int launch_rotate(ttdft_ctx ctx, const void *X, int64_t ldx, void *Z, int64_t ldz, int64_t n,
                  int32_t k, const double *dQ, int dtype, int mode) {
    TT_CHECK(k >= 1 && n >= 1, TTDFT_EBADARG, "rotate: n=%lld k=%d", (long long)n, k);
    TT_CHECK((mode == TTDFT_GEMM_F64 && dtype == TTDFT_F64) ||
                 (mode == TTDFT_GEMM_BF16X3 && dtype == TTDFT_F32),
             TTDFT_EBADARG, "rotate: mode %d does not match dtype %d", mode, dtype);
    const int64_t kk = (int64_t)k * k;
    PendingTiming t;
    if (mode == TTDFT_GEMM_F64) {
        constexpr int RT16 = 8;  // column tiles per wave
        const int nstep = (k + 15) / 16, ntile = (k + 15) / 16;
        const int64_t nq = (int64_t)nstep * 4 * ntile * 64;
        TT_TRY(ctx->scratch[4].ensure((size_t)nq * sizeof(double)));
        double *Qd = (double *)ctx->scratch[4].ptr;
        hipLaunchKernelGGL(q_frag64_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                           ctx->stream, dQ, k, nstep, ntile, Qd);
        TT_LAUNCH_CHECK();
        const int vec = (ldx % 2 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
        const dim3 g((unsigned)((n + 127) / 128), (unsigned)((ntile + RT16 - 1) / RT16));
        TT_TRY(timing_begin(ctx, TTDFT_KSTAT_ROTATE, &t));
        hipLaunchKernelGGL(rotate_f64_frag_kernel<RT16>, g, dim3(256), 0, ctx->stream,
                           (const double *)X, ldx, Qd, nstep, ntile, (double *)Z, ldz, n, k, vec);
        TT_LAUNCH_CHECK();
        TT_TRY(timing_end(ctx, &t, 2.0 * n * k * 8.0 + kk * 8.0, 2.0 * n * kk));
    } else {
        constexpr int RT16 = 8;  // column tiles per wave
        const int nstep = (k + 15) / 16, ntile = (k + 15) / 16;
        const int64_t nfrag = (int64_t)nstep * ntile * 64;
        TT_TRY(ctx->scratch[4].ensure((size_t)nfrag * sizeof(QFrag)));
        QFrag *Qf = (QFrag *)ctx->scratch[4].ptr;
        hipLaunchKernelGGL(split_q_frag_kernel, dim3((unsigned)((nfrag + 255) / 256)), dim3(256),
                           0, ctx->stream, dQ, k, nstep, ntile, Qf);
        TT_LAUNCH_CHECK();
        const int vec = (ldx % 4 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
        const dim3 g((unsigned)((n + 127) / 128), (unsigned)((ntile + RT16 - 1) / RT16));
        TT_TRY(timing_begin(ctx, TTDFT_KSTAT_ROTATE, &t));
        hipLaunchKernelGGL(rotate_bf16x3_frag_kernel<RT16>, g, dim3(256), 0, ctx->stream,
                           (const float *)X, ldx, Qf, nstep, ntile, (float *)Z, ldz, n, k, vec);
        TT_LAUNCH_CHECK();
        TT_TRY(timing_end(ctx, &t, 2.0 * n * k * 4.0 + kk * 4.0, 3.0 * 2.0 * n * kk));
    }
    return TTDFT_OK;
}

int launch_transpose_small(ttdft_ctx ctx, const double *A, double *AT, int32_t k) {
    const int64_t kk = (int64_t)k * k;
    hipLaunchKernelGGL(transpose_small_kernel, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0,
                       ctx->stream, A, AT, k);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int launch_copy_block(ttdft_ctx ctx, const void *src, int64_t lds, int src_dtype, void *dst,
                      int64_t ldd, int dst_dtype, int64_t n, int32_t k) {
    const int64_t total = n * (int64_t)k;
    if (total == 0) return TTDFT_OK;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const dim3 g((unsigned)blocks), b(256);
    if (src_dtype == TTDFT_F64 && dst_dtype == TTDFT_F64)
        hipLaunchKernelGGL((copy_block_kernel<double, double>), g, b, 0, ctx->stream,
                           (const double *)src, lds, (double *)dst, ldd, n, k);
    else if (src_dtype == TTDFT_F64 && dst_dtype == TTDFT_F32)
        hipLaunchKernelGGL((copy_block_kernel<double, float>), g, b, 0, ctx->stream,
                           (const double *)src, lds, (float *)dst, ldd, n, k);
    else if (src_dtype == TTDFT_F32 && dst_dtype == TTDFT_F64)
        hipLaunchKernelGGL((copy_block_kernel<float, double>), g, b, 0, ctx->stream,
                           (const float *)src, lds, (double *)dst, ldd, n, k);
    else
        hipLaunchKernelGGL((copy_block_kernel<float, float>), g, b, 0, ctx->stream,
                           (const float *)src, lds, (float *)dst, ldd, n, k);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // namespace ttdft
