// Electron density at tensor-grid points,
//   rho(qx,qy,qz) = sum_i w_i psi_i(qx,qy,qz)^2,   w_i = 2 f_i,
//   psi_i = sum_abc C_i[a,b,c] Bx[qx,a] By[qy,b] Bz[qz,c].
// The contraction order (z, then x, then y) is that of the reference's TTM
// chain (dft/ProjectHamiltonianSparse.cc:1052-1067).  The y step is regrouped
// through the y-factor density matrix of every (qz,qx) column,
//   M[qz,qx](b,b') = sum_i w_i T2_i[qz,qx](b) T2_i[qz,qx](b'),
//   rho(qx,qy,qz)  = sum_{b<=b'} (2 - [b=b']) By[qy,b] By[qy,b'] M[qz,qx](b,b'),
// which needs 2 R^2 Q^2 k + Q^3 R(R+1) flops instead of the 2 R Q^3 k of
// squaring every psi_i (R ranks, Q points per axis, k states): at R = 50,
// Q = 600, k = 300 about 4x fewer, and all states are summed in one pass.
// Every step is a dense matrix product on the matrix cores:
//   1. Cg[c][a][s][b]  = sqrt(w_s) X[a + rx b + rx ry c][idx_s]   (gather)
//   2. T1[qz][a][s][b] = sum_c Bz[qz,c] Cg[c][a][s][b]   (rocBLAS GEMM)
//   3. T2[qz][qx][s][b]= sum_a Bx[qx,a] T1[qz][a][s][b]
//   4. m[qz][qx][pair] = upper triangle of T2[qz][qx] T2[qz][qx]^T, packed
//      3 + 4 fused in xsyrk_kernel (rx, ry <= 64 in FP64, <= 80 in FP32): T2
//      lives in LDS only.  Above: a rocBLAS GEMM per plane writes T2,
//      syrk_pack_kernel reads it.
//   5. rho[qz]^T       = m[qz] E^T,  E[qy][pair] = (2 - [b=b']) By[qy,b] By[qy,b']
//                        (one rocBLAS GEMM over all columns of a chunk, then
//                         widen_planes_kernel into the planes' layout)
// Steps 2-5 run over chunks of qz planes to bound the workspace.  All steps
// run in FP64 for an FP64 block and in FP32 (FP32 accumulation) for an FP32
// block, which runs the products at the FP32 rate; an FP32 plane's rho is
// widened into the FP64 output.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

// Cg[((c rx + a) K + s) ry + b] = sqrt(w_s) X[(a + rx b + rx ry c) ldx + idx_s]
// (states before b: one state's y-factor coefficients are contiguous all the
// way to the T2 columns read by syrk_pack_kernel)
template <typename T, typename Tm>
__global__ void gather_scaled_kernel(const T *__restrict__ X, int64_t ldx,
                                     const int32_t *__restrict__ idx,
                                     const double *__restrict__ sw, int32_t K, int32_t rx,
                                     int32_t ry, int64_t N, Tm *__restrict__ Cg) {
    const int64_t total = N * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(e % ry);
        const int64_t J = e / ry;  // (c rx + a) K + s
        const int s = (int)(J % K);
        const int64_t ca = J / K;
        const int a = (int)(ca % rx);
        const int64_t c = ca / rx;
        const int64_t I = a + (int64_t)rx * (b + (int64_t)ry * c);
        Cg[e] = (Tm)(sw[s] * (double)X[I * ldx + idx[s]]);
    }
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int tri_col(int p) {  // column b' of upper-triangle pair p (column-major order)
    int bp = 0;
    while ((bp + 1) * (bp + 2) / 2 <= p) ++bp;
    return bp;
}

// Packed upper triangle of M = V V^T for one column t per wave:
//   V[s][b] = T2[t R K + s R + b] (R <= 16 RT, K states),
//   m[t np + b'(b'+1)/2 + b] = M[b][b'],  b <= b'.
// The 16x16 output tiles (bi, bj), bi <= bj, are numbered column-major; the
// wave computes tiles [P0, P0 + NP).  Per 16-state chunk, lane (i, kq) holds
// V[c + 4 kq + e][16 bi + i] for e = 0..3; MFMA step e contracts the four
// states {c + 4 kq + e}: both operands come from the same registers
// (A[i][kq] and B[kq][j] have the same lane layout), and the order in which
// states are summed does not matter.  For a fixed state the 16 lanes of a row
// block read 16 consecutive b.
// The kernel streams T2 once from HBM; to keep enough bytes in flight, two
// chunks are loaded ahead of the one being multiplied (three register sets,
// loop unrolled by 3).  Lane offsets inside a chunk are fixed and computed
// once; rows past R load a valid row whose products land only in tiles that
// are never stored, and states past K (last chunk only) are zeroed, reading
// into the 16 R doubles of slack that follow the T2 buffer.
// Tm = float (the density of an FP32 block): T2 in FP32, FP32 16x16x4 MFMAs
// with FP32 accumulation (operands in the same lane layout; element v of
// lane l then holds D[4(l/16) + v][l%16]), M written in FP64.
template <typename Tm, int RT, int P0, int NP>
__global__ __launch_bounds__(256) void syrk_pack_kernel(const Tm *__restrict__ T2, int32_t R,
                                                         int32_t K, int64_t ncols,
                                                         Tm *__restrict__ mp) {
    typedef typename std::conditional<sizeof(Tm) == 8, d4_t, f4_t>::type acc_t;
    constexpr int NBLK = tri_col(P0 + NP - 1) + 1;  // row blocks the tiles use
    static_assert(NBLK <= RT, "tile range");
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ncols) return;  // no barriers: waves are independent
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
    const Tm *V = T2 + t * R * K;
    int32_t off[NBLK][4];
#pragma unroll
    for (int bi = 0; bi < NBLK; ++bi) {
        const int row = 16 * bi + i;
        const int rc = row < R ? row : R - 1;
#pragma unroll
        for (int e = 0; e < 4; ++e) off[bi][e] = (4 * kq + e) * R + rc;
    }
    acc_t acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = acc_t{0, 0, 0, 0};
    Tm b0[NBLK][4], b1[NBLK][4], b2[NBLK][4];
    const int nchunk = (K + 15) / 16;
    auto load = [&](int ch, Tm (*v)[4]) {
        const Tm *Vc = V + (int64_t)(ch < nchunk ? ch : nchunk - 1) * 16 * R;
#pragma unroll
        for (int bi = 0; bi < NBLK; ++bi)
#pragma unroll
            for (int e = 0; e < 4; ++e) v[bi][e] = Vc[off[bi][e]];
    };
    auto mul = [&](int ch, Tm (*v)[4]) {
        const int c = ch * 16;
        if (c + 16 > K) {  // last chunk: states past K contribute nothing
#pragma unroll
            for (int bi = 0; bi < NBLK; ++bi)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (c + 4 * kq + e >= K) v[bi][e] = Tm(0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int p = P0 + q;
                const int bj = tri_col(p), bi = p - bj * (bj + 1) / 2;
                if constexpr (sizeof(Tm) == 8)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[bi][e], v[bj][e], acc[q], 0, 0,
                                                                  0);
                else
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[bi][e], v[bj][e], acc[q], 0, 0,
                                                                  0);
            }
    };
    load(0, b0);
    load(1, b1);
    for (int ch = 0; ch < nchunk; ch += 3) {
        load(ch + 2, b2);
        mul(ch, b0);
        if (ch + 1 >= nchunk) break;
        load(ch + 3, b0);
        mul(ch + 1, b1);
        if (ch + 2 >= nchunk) break;
        load(ch + 4, b1);
        mul(ch + 2, b2);
    }
    // 16x16x4 D layout: element v of lane l is D[l/16 + 4v][l%16] (FP64),
    // D[4(l/16) + v][l%16] (FP32)
    const int64_t np = (int64_t)R * (R + 1) / 2;
    Tm *out = mp + t * np;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int p = P0 + q;
        const int bj = tri_col(p), bi = p - bj * (bj + 1) / 2;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int b = 16 * bi + (sizeof(Tm) == 8 ? kq + 4 * v : 4 * kq + v), bp = 16 * bj + i;
            if (b <= bp && bp < R) out[(int64_t)bp * (bp + 1) / 2 + b] = acc[q][v];
        }
    }
}

template <typename Tm, int RT, int P0, int NP>
int launch_syrk_group(ttdft_ctx ctx, const Tm *T2, int32_t R, int32_t K, int64_t ncols,
                      Tm *mp) {
    const int64_t blocks = (ncols + 3) / 4;
    hipLaunchKernelGGL((syrk_pack_kernel<Tm, RT, P0, NP>), dim3((unsigned)blocks), dim3(256), 0,
                       ctx->stream, T2, R, K, ncols, mp);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

// all upper-triangle tiles of an R x R result, at most 11 tiles per wave
template <typename Tm>
int launch_syrk_pack(ttdft_ctx ctx, const Tm *T2, int32_t R, int32_t K, int64_t ncols,
                     Tm *mp) {
    switch ((R + 15) / 16) {
    case 1:
        return launch_syrk_group<Tm, 1, 0, 1>(ctx, T2, R, K, ncols, mp);
    case 2:
        return launch_syrk_group<Tm, 2, 0, 3>(ctx, T2, R, K, ncols, mp);
    case 3:
        return launch_syrk_group<Tm, 3, 0, 6>(ctx, T2, R, K, ncols, mp);
    case 4:
        return launch_syrk_group<Tm, 4, 0, 10>(ctx, T2, R, K, ncols, mp);
    case 5:
        TT_TRY((launch_syrk_group<Tm, 5, 0, 8>(ctx, T2, R, K, ncols, mp)));
        return launch_syrk_group<Tm, 5, 8, 7>(ctx, T2, R, K, ncols, mp);
    case 6:
        TT_TRY((launch_syrk_group<Tm, 6, 0, 11>(ctx, T2, R, K, ncols, mp)));
        return launch_syrk_group<Tm, 6, 11, 10>(ctx, T2, R, K, ncols, mp);
    case 7:
        TT_TRY((launch_syrk_group<Tm, 7, 0, 10>(ctx, T2, R, K, ncols, mp)));
        TT_TRY((launch_syrk_group<Tm, 7, 10, 9>(ctx, T2, R, K, ncols, mp)));
        return launch_syrk_group<Tm, 7, 19, 9>(ctx, T2, R, K, ncols, mp);
    case 8:
        TT_TRY((launch_syrk_group<Tm, 8, 0, 9>(ctx, T2, R, K, ncols, mp)));
        TT_TRY((launch_syrk_group<Tm, 8, 9, 9>(ctx, T2, R, K, ncols, mp)));
        TT_TRY((launch_syrk_group<Tm, 8, 18, 9>(ctx, T2, R, K, ncols, mp)));
        return launch_syrk_group<Tm, 8, 27, 9>(ctx, T2, R, K, ncols, mp);
    default:
        break;
    }
    set_error("density: ry=%d outside 1..128", R);
    return TTDFT_EBADARG;
}

// ---- fused x contraction + y-factor density matrices (steps 3 and 4) ----
// For one qz plane of T1 and 16 consecutive qx points the kernel forms
//   V[qx][s][b] = sum_a Bx[qx,a] T1[qz][a][s][b]
// a chunk of states at a time on the matrix cores, hands it to the y step through
// LDS and accumulates the packed upper triangle of M[qx] = V[qx]^T V[qx] over
// all states in registers, so the x-contracted block never reaches memory.
//
// Work item = (plane, 16-qx tile); items are renumbered so that each XCD
// takes a contiguous range (dispatch deals workgroups round-robin over the 8
// XCDs): the 16-qx tiles of one plane then read that plane's T1 from one L2.
// x step (per chunk of SC states: 16 in FP32 up to rank 64, else 8): the
//   chunk's (s, b) pairs are flattened, n = s_local R + b, and cut into
//   16-wide tiles; wave w takes tiles w, w + W, ... (W waves).  MFMA D[qx][n]
//   += Bx[qx][a] T1[a][n] over a in steps of 4, with the Bx fragments read
//   from LDS (written once per workgroup) and the T1 operand 16 consecutive
//   values of one row a per lane group (one coalesced load).  Operand k of
//   the wave's next tile, across chunk boundaries too, is loaded right after
//   MFMA k of the current one.  Each finished tile is stored to LDS as
//   V[s_local][qx][b]; b runs to the pitch (64, or 80 at five row blocks)
//   with zeros past R.
// y step: wave w owns qx = CW w .. CW w + CW - 1 (CW = 16 / W: two qx points
//   with 8 waves, one with 16).  Per 4 states lane (i, kq) reads
//   V[4 ks + kq][q][16 bi + i] and accumulates the M tiles (bi <= bj) as in
//   syrk_pack_kernel (both MFMA operands from the same registers).
// Two LDS buffers: the x step of chunk c writes buffer c & 1, so one barrier
// per chunk orders it (every wave finished the y step of chunk c - 1, which
// read the other buffer, before it reached the barrier of chunk c).
// KA = a steps of 4 (rx <= 4 KA), NB = 16-row blocks of b (R <= 16 NB).
constexpr int kXsQ = 16;   // qx points per workgroup
// states per chunk: 8, or 16 in FP32 up to 4 row blocks (half the barriers,
// the same LDS bytes)
template <typename Tm, int NB>
constexpr int xs_states() { return sizeof(Tm) == 4 && NB <= 4 ? 16 : 8; }
// LDS pitch of b: 64, or 80 for 5 row blocks (FP32 only)
template <int NB>
constexpr int xs_pitch() { return NB <= 4 ? 64 : 16 * NB; }
// waves per workgroup: 8 (two per SIMD, M tiles of two qx points per wave:
// FP64, and FP32 at 5 row blocks) or 16 (FP32 up to 4 row blocks: four per
// SIMD, one qx point per wave)
template <typename Tm, int NB>
constexpr int xs_waves() { return sizeof(Tm) == 8 || NB > 4 ? 8 : 16; }

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// row of the 16x16 D tile held in element v of lane l
template <typename Tm>
__device__ __forceinline__ int drow(int l, int v) {
    return sizeof(Tm) == 8 ? (l >> 4) + 4 * v : 4 * (l >> 4) + v;
}

template <typename Tm, int KA, int NB>
__global__ __launch_bounds__((64 * xs_waves<Tm, NB>()), (xs_waves<Tm, NB>() == 16 ? 2 : 1)) void xsyrk_kernel(
    const Tm *__restrict__ T1, const Tm *__restrict__ Bx, int32_t px, int32_t rx, int32_t R,
    int32_t K, int32_t nplanes, int32_t ntx, int32_t per_xcd, Tm *__restrict__ mp) {
    typedef typename std::conditional<sizeof(Tm) == 8, d4_t, f4_t>::type acc_t;
    constexpr int NP = NB * (NB + 1) / 2;
    constexpr int kXsW = xs_waves<Tm, NB>(), kXsCW = kXsQ / kXsW;  // qx points per wave, y step
    constexpr int kXsSC = xs_states<Tm, NB>(), kXsPB = xs_pitch<NB>();
    __shared__ __align__(16) Tm V[2][kXsSC][kXsQ][kXsPB];
    __shared__ __align__(16) Tm Af[KA][64];
    const int64_t item = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (item >= (int64_t)nplanes * ntx) return;  // whole workgroup, before any barrier
    const int plane = (int)(item / ntx), qx0 = (int)(item % ntx) * kXsQ;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, kq = lane >> 4;
    {
        Tm *v = &V[0][0][0][0];
        for (int e = tid; e < 2 * kXsSC * kXsQ * kXsPB; e += 64 * kXsW) v[e] = Tm(0);
    }
    // A operand of the x step, lane-major: Af[k][lane] = Bx[qx0 + i][4 k + kq]
    for (int e = tid; e < KA * 64; e += 64 * kXsW) {
        const int l = e & 63, a = 4 * (e >> 6) + (l >> 4), q = qx0 + (l & 15);
        Af[e >> 6][l] = (a < rx && q < px) ? Bx[q + (int64_t)px * a] : Tm(0);
    }
    // T1 operand offsets in elements within the plane (< 2^31: checked by
    // the launcher), from one uniform base so the loads take a 32-bit offset
    const int32_t KR = K * R;
    const char *T1p = reinterpret_cast<const char *>(T1 + (int64_t)plane * rx * KR);
    const int nch = (K + kXsSC - 1) / kXsSC;
    auto nnz = [&](int ch) { return min(kXsSC, K - ch * kXsSC) * R; };  // (s, b) pairs
    // operands of tile t of chunk ch; rows past rx and columns past the chunk
    // load a valid element: the rows meet zero Bx fragments, the columns are
    // never stored
    auto column = [&](int ch, int t) {
        const int n = 16 * t + li, nn = nnz(ch);
        return (uint32_t)(ch * kXsSC * R + (n < nn ? n : nn - 1));
    };
    auto load1 = [&](uint32_t col, int k) {
        const int a = min(4 * k + kq, rx - 1);
        return *reinterpret_cast<const Tm *>(
            T1p + (uint32_t)((col + (uint32_t)a * (uint32_t)KR) * sizeof(Tm)));
    };
    // this wave's tiles in order: t = w, w + W, ... of every chunk; operand k
    // of the next tile is loaded into the register that MFMA k of the current
    // tile has just read
    auto next = [&](int &c, int &t) {
        t += kXsW;
        while (c < nch && t >= (nnz(c) + 15) >> 4) {
            ++c;
            t = w;
        }
    };
    int lc = 0, lt = w - kXsW;
    next(lc, lt);
    Tm bc[KA];
    if (lc < nch) {
        const uint32_t col = column(lc, lt);
#pragma unroll
        for (int k = 0; k < KA; ++k) bc[k] = load1(col, k);
    }
    acc_t acc[kXsCW][NP];
#pragma unroll
    for (int c = 0; c < kXsCW; ++c)
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[c][p] = acc_t{0, 0, 0, 0};
    __syncthreads();  // LDS zeroed, Af written

    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1, nv = min(kXsSC, K - ch * kXsSC), nn = nv * R;
        while (lc == ch) {
            int nc = lc, nt = lt;
            next(nc, nt);
            // unconditional (past the last tile: any valid column), so the
            // wait before MFMA k counts exactly the KA - 1 younger loads
            const uint32_t ncol = nc < nch ? column(nc, nt) : column(lc, lt);
            acc_t d = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < KA; ++k) {
                d = mfma16(Af[k][lane], bc[k], d);
                bc[k] = load1(ncol, k);
            }
            // keep the order MFMA k, load k (each load one tile ahead of its use)
#pragma unroll
            for (int k = 0; k < KA; ++k) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one global load
            }
            const int n = 16 * lt + li;
            if (n < nn) {
                const int s = n / R, b = n - s * R;
#pragma unroll
                for (int v = 0; v < 4; ++v) V[buf][s][drow<Tm>(lane, v)][b] = d[v];
            }
            lc = nc;
            lt = nt;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kXsSC / 4; ++ks) {
            if (4 * ks >= nv) break;
            const int s = 4 * ks + kq;
#pragma unroll
            for (int c = 0; c < kXsCW; ++c) {
                Tm op[NB];
#pragma unroll
                for (int bi = 0; bi < NB; ++bi) {
                    const Tm x = V[buf][s < nv ? s : 0][kXsCW * w + c][16 * bi + li];
                    op[bi] = s < nv ? x : Tm(0);
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    const int bj = tri_col(p), bi = p - bj * (bj + 1) / 2;
                    acc[c][p] = mfma16(op[bi], op[bj], acc[c][p]);
                }
            }
        }
    }
    const int64_t np = (int64_t)R * (R + 1) / 2;
#pragma unroll
    for (int c = 0; c < kXsCW; ++c) {
        const int q = qx0 + kXsCW * w + c;
        if (q >= px) continue;
        Tm *out = mp + ((int64_t)plane * px + q) * np;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int bj = tri_col(p), bi = p - bj * (bj + 1) / 2;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int b = 16 * bi + drow<Tm>(lane, v), bp = 16 * bj + li;
                if (b <= bp && bp < R) out[(int64_t)bp * (bp + 1) / 2 + b] = acc[c][p][v];
            }
        }
    }
}

// the fused path covers ranks up to 64 (FP64) or 80 (FP32: the M tiles of
// two qx points at 5 row blocks fit half a SIMD's registers in FP32 only)
bool xsyrk_fits(int32_t rx, int32_t ry, bool f32) {
    const int m = f32 ? 80 : 64;
    return rx >= 1 && rx <= m && ry >= 1 && ry <= m;
}

template <typename Tm, int KA>
int launch_xsyrk_ka(ttdft_ctx ctx, const Tm *T1, const Tm *Bx, int32_t px, int32_t rx,
                    int32_t R, int32_t K, int32_t nplanes, Tm *mp) {
    TT_CHECK((int64_t)rx * K * R * (int64_t)sizeof(Tm) < ((int64_t)1 << 32), TTDFT_EBADARG,
             "density: one plane of %d x %d x %d exceeds 4 GiB", rx, K, R);
    const int32_t ntx = (px + kXsQ - 1) / kXsQ;
    const int64_t items = (int64_t)nplanes * ntx;
    const int32_t per_xcd = (int32_t)((items + 7) / 8);
    const dim3 grid((unsigned)(8 * (int64_t)per_xcd));
    const dim3 block(64 * (R > 64 ? xs_waves<Tm, 5>() : xs_waves<Tm, 4>()));
    switch ((R + 15) / 16) {
    case 1:
        hipLaunchKernelGGL((xsyrk_kernel<Tm, KA, 1>), grid, block, 0, ctx->stream, T1, Bx, px, rx,
                           R, K, nplanes, ntx, per_xcd, mp);
        break;
    case 2:
        hipLaunchKernelGGL((xsyrk_kernel<Tm, KA, 2>), grid, block, 0, ctx->stream, T1, Bx, px, rx,
                           R, K, nplanes, ntx, per_xcd, mp);
        break;
    case 3:
        hipLaunchKernelGGL((xsyrk_kernel<Tm, KA, 3>), grid, block, 0, ctx->stream, T1, Bx, px, rx,
                           R, K, nplanes, ntx, per_xcd, mp);
        break;
    case 4:
        hipLaunchKernelGGL((xsyrk_kernel<Tm, KA, 4>), grid, block, 0, ctx->stream, T1, Bx, px, rx,
                           R, K, nplanes, ntx, per_xcd, mp);
        break;
    default:
        if constexpr (sizeof(Tm) == 4) {
            hipLaunchKernelGGL((xsyrk_kernel<Tm, KA, 5>), grid, block, 0, ctx->stream, T1, Bx, px,
                               rx, R, K, nplanes, ntx, per_xcd, mp);
            break;
        }
        set_error("density: fused y step with %d y ranks", R);
        return TTDFT_EBADARG;
    }
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

// a steps rounded up to 4, 8, 13, 16 or 20 (rx <= 16, 32, 52, 64, 80)
template <typename Tm>
int launch_xsyrk(ttdft_ctx ctx, const Tm *T1, const Tm *Bx, int32_t px, int32_t rx, int32_t R,
                 int32_t K, int32_t nplanes, Tm *mp) {
    if (rx <= 16) return launch_xsyrk_ka<Tm, 4>(ctx, T1, Bx, px, rx, R, K, nplanes, mp);
    if (rx <= 32) return launch_xsyrk_ka<Tm, 8>(ctx, T1, Bx, px, rx, R, K, nplanes, mp);
    if (rx <= 52) return launch_xsyrk_ka<Tm, 13>(ctx, T1, Bx, px, rx, R, K, nplanes, mp);
    if (rx <= 64) return launch_xsyrk_ka<Tm, 16>(ctx, T1, Bx, px, rx, R, K, nplanes, mp);
    if constexpr (sizeof(Tm) == 4)
        return launch_xsyrk_ka<Tm, 20>(ctx, T1, Bx, px, rx, R, K, nplanes, mp);
    set_error("density: fused x step with %d x ranks", rx);
    return TTDFT_EBADARG;
}

// E[qy][pair(b,b')] = (b == b' ? 1 : 2) By[qy,b] By[qy,b'];  By column-major py x R
template <typename Tm>
__global__ void pair_table_kernel(const double *__restrict__ By, int32_t py, int32_t R,
                                  Tm *__restrict__ E) {
    const int64_t np = (int64_t)R * (R + 1) / 2;
    const int64_t total = (int64_t)py * np;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int qy = (int)(e / np);
        const int pr = (int)(e - (int64_t)qy * np);
        int bp = (int)((sqrt(8.0 * pr + 1.0) - 1.0) * 0.5);
        while ((bp + 1) * (bp + 2) / 2 <= pr) ++bp;
        while (bp * (bp + 1) / 2 > pr) --bp;
        const int b = pr - bp * (bp + 1) / 2;
        const double v = By[qy + (int64_t)py * b] * By[qy + (int64_t)py * bp];
        E[e] = (Tm)(b == bp ? v : 2.0 * v);
    }
}

unsigned grid_for(int64_t total) {
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

#define TT_BLAS(call)                                                                     \
    do {                                                                                  \
        const rocblas_status st_ = (call);                                                \
        TT_CHECK(st_ == rocblas_status_success, TTDFT_ELIB, "%s: rocBLAS status %d", #call, \
                 (int)st_);                                                               \
    } while (0)

// Level-3 BLAS of the intermediate type (FP64 for FP64 blocks, FP32 for FP32)
rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, rocblas_int m,
                    rocblas_int n, rocblas_int k, const double *A, rocblas_int lda,
                    const double *B, rocblas_int ldb, double *C, rocblas_int ldc) {
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}
rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, rocblas_int m,
                    rocblas_int n, rocblas_int k, const float *A, rocblas_int lda,
                    const float *B, rocblas_int ldb, float *C, rocblas_int ldc) {
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}
rocblas_status gemm_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                            rocblas_int m, rocblas_int n, rocblas_int k, const double *A,
                            rocblas_int lda, int64_t sa, const double *B, rocblas_int ldb,
                            int64_t sb, double *C, rocblas_int ldc, int64_t sc, rocblas_int cnt) {
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                         C, ldc, sc, cnt);
}
rocblas_status gemm_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                            rocblas_int m, rocblas_int n, rocblas_int k, const float *A,
                            rocblas_int lda, int64_t sa, const float *B, rocblas_int ldb,
                            int64_t sb, float *C, rocblas_int ldc, int64_t sc, rocblas_int cnt) {
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                         C, ldc, sc, cnt);
}

__global__ void to_float_kernel(const double *__restrict__ a, float *__restrict__ b,
                                int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) b[i] = (float)a[i];
}
// rho[qzl][qy][qx] = a[(qzl px + qx) + qc px qy]: the FP32 expansion onto the
// y points runs as one product over all (qz, qx) columns of a chunk, and this
// pass widens its column-major result into the planes' layout (reads and
// writes run along qx)
template <typename Tm>
__global__ void widen_planes_kernel(const Tm *__restrict__ a, double *__restrict__ rho,
                                    int32_t qc, int32_t px, int32_t py) {
    const int64_t count = (int64_t)qc * py * px;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(i % px);
        const int64_t r = i / px;
        const int qy = (int)(r % py), qz = (int)(r / py);
        rho[i] = (double)a[(int64_t)qz * px + qx + (int64_t)qc * px * qy];
    }
}

// Steps 1-5 with the T1 / T2 intermediates and their contractions in Tm.
// Workspace (scratch 5, bytes): Cg (N K Tm), E (py np double), sqrt(w) (K
// double), state indices (K int32, padded to K doubles), Bx / Bz in Tm (FP32
// only), then per qz plane T1 (rx ry K Tm), T2 (px ry K Tm), the packed
// M (px np Tm) and, for FP32, the plane's rho (px py float) before it is
// widened into the FP64 output; 16 ry Tm of slack follow T2 (read, never
// used, by the last state chunk of the y contraction).
template <typename Tm>
int density_run(ttdft_ctx ctx, const void *X, int64_t ldx, int dtype, const std::vector<int32_t> &idx,
                const std::vector<double> &sw, int32_t rx, int32_t ry, int32_t rz,
                const double *Bx, const double *By, const double *Bz, int32_t px, int32_t py,
                int32_t pz, double *rho) {
    const int64_t N = (int64_t)rx * ry * rz;
    const int32_t K = (int32_t)idx.size();
    const int64_t rxy = (int64_t)rx * ry;
    const int64_t np = (int64_t)ry * (ry + 1) / 2;
    constexpr int64_t ES = sizeof(Tm);
    const bool f32 = ES == 4;
    // steps 3 and 4 fused (T2 never stored) unless a rank exceeds 64 (FP64) or 80;
    // TTDFT_DENSITY_FUSED=0 selects the two-step form
    bool fused = xsyrk_fits(rx, ry, f32);
    if (const char *env = std::getenv("TTDFT_DENSITY_FUSED"))
        if (env[0] == '0') fused = false;
    const int64_t t2_plane = fused ? 0 : (int64_t)px * ry * K;
    // the expansion onto the y points as one product over all (qz, qx) columns
    // of a chunk, through a column-major buffer that widen_planes_kernel turns
    // into the planes: rocBLAS runs the one product with a wider tile than the
    // per-plane batch (600^3, k = 300: FP64 35.9 -> 35.3 ms, FP32 -2.8 ms).
    // TTDFT_DENSITY_PAIR=0 selects the batch for FP64.
    bool pair_one = true;
    if (const char *env = std::getenv("TTDFT_DENSITY_PAIR"))
        if (env[0] == '0' && !f32) pair_one = false;
    const int64_t per_plane = (rxy * K + t2_plane + (int64_t)px * np) * ES +
                              (pair_one ? (int64_t)px * py * ES : 0);
    const int64_t budget = (int64_t)6 << 30;  // bytes for the chunk buffers
    int64_t QC = budget / per_plane;
    if (QC < 1) QC = 1;
    if (QC > pz) QC = pz;
    // TTDFT_DENSITY_PLANES caps the planes per chunk (smaller workspace)
    if (const char *env = std::getenv("TTDFT_DENSITY_PLANES")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v > 0 && v < QC) QC = v;
    }
    {  // equal chunks: the last one is not a small tail
        const int64_t nchunk = (pz + QC - 1) / QC;
        QC = (pz + nchunk - 1) / nchunk;
    }
    auto align = [](int64_t b) { return (b + 255) / 256 * 256; };
    const int64_t o_cg = 0, o_e = align(N * K * ES), o_sw = o_e + align(py * np * ES),
                  o_idx = o_sw + align(K * 8), o_bx = o_idx + align(K * 8),
                  o_bz = o_bx + (f32 ? align((int64_t)px * rx * ES) : 0),
                  o_t1 = o_bz + (f32 ? align((int64_t)pz * rz * ES) : 0),
                  o_t2 = o_t1 + align(QC * rxy * K * ES),
                  o_mp = o_t2 + (fused ? 0 : align(QC * px * ry * K * ES + 16 * ry * ES)),
                  o_rf = o_mp + align(QC * px * np * ES),
                  total = o_rf + (pair_one ? align(QC * px * py * ES) : 0);
    TT_TRY(ctx->scratch[5].ensure((size_t)total));
    char *base = (char *)ctx->scratch[5].ptr;
    Tm *Cg = (Tm *)(base + o_cg);
    Tm *E = (Tm *)(base + o_e);
    double *swd = (double *)(base + o_sw);
    int32_t *idxd = (int32_t *)(base + o_idx);
    const Tm *Bxm = (const Tm *)Bx, *Bzm = (const Tm *)Bz;
    if (f32) {
        Bxm = (const Tm *)(base + o_bx);
        Bzm = (const Tm *)(base + o_bz);
        hipLaunchKernelGGL(to_float_kernel, dim3(grid_for((int64_t)px * rx)), dim3(256), 0,
                           ctx->stream, Bx, (float *)(base + o_bx), (int64_t)px * rx);
        TT_LAUNCH_CHECK();
        hipLaunchKernelGGL(to_float_kernel, dim3(grid_for((int64_t)pz * rz)), dim3(256), 0,
                           ctx->stream, Bz, (float *)(base + o_bz), (int64_t)pz * rz);
        TT_LAUNCH_CHECK();
    }
    Tm *T1 = (Tm *)(base + o_t1);
    Tm *T2 = (Tm *)(base + o_t2);
    Tm *mp = (Tm *)(base + o_mp);
    Tm *rho_c = (Tm *)(base + o_rf);  // one-product expansion only

    TT_HIP(hipMemcpyAsync(swd, sw.data(), sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
    TT_HIP(hipMemcpyAsync(idxd, idx.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice,
                          ctx->stream));
    if (dtype == TTDFT_F64)
        hipLaunchKernelGGL((gather_scaled_kernel<double, Tm>), dim3(grid_for(N * K)), dim3(256), 0,
                           ctx->stream, (const double *)X, ldx, idxd, swd, K, rx, ry, N, Cg);
    else
        hipLaunchKernelGGL((gather_scaled_kernel<float, Tm>), dim3(grid_for(N * K)), dim3(256), 0,
                           ctx->stream, (const float *)X, ldx, idxd, swd, K, rx, ry, N, Cg);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(pair_table_kernel<Tm>, dim3(grid_for((int64_t)py * np)), dim3(256), 0,
                       ctx->stream, By, py, ry, E);
    TT_LAUNCH_CHECK();

    for (int64_t q0 = 0; q0 < pz; q0 += QC) {
        const int qc = (int)std::min<int64_t>(QC, pz - q0);
        // 2. T1[qzl][a][s][b] = sum_c Cg[c][a][s][b] Bz[q0 + qzl, c]
        TT_BLAS(gemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose,
                     (rocblas_int)(rxy * K), qc, rz, Cg, (rocblas_int)(rxy * K), Bzm + q0, pz, T1,
                     (rocblas_int)(rxy * K)));
        if (fused) {
            // 3 + 4. packed M[qzl][qx] with T2 formed on chip
            TT_TRY(launch_xsyrk<Tm>(ctx, T1, Bxm, px, rx, ry, K, qc, mp));
        } else {
            // 3. T2[qzl][qx][(s, b)] = sum_a T1[qzl][a][(s, b)] Bx[qx, a], one GEMM per plane
            TT_BLAS(gemm_batched(ctx->blas, rocblas_operation_none, rocblas_operation_transpose,
                                 (rocblas_int)((int64_t)ry * K), px, rx, T1,
                                 (rocblas_int)((int64_t)ry * K), rxy * K, Bxm, px, 0, T2,
                                 (rocblas_int)((int64_t)ry * K), (int64_t)px * ry * K, qc));
            // 4. packed upper triangles of T2[qzl][qx] T2[qzl][qx]^T
            TT_TRY(launch_syrk_pack<Tm>(ctx, T2, ry, K, (int64_t)qc * px, mp));
        }
        // 5. rho[q0 + qzl](qx, qy) = sum_pair m[qzl][qx][pair] E[qy][pair]
        if constexpr (sizeof(Tm) == 8) {
            if (!pair_one)
                TT_BLAS(gemm_batched(ctx->blas, rocblas_operation_transpose,
                                     rocblas_operation_none, px, py, (rocblas_int)np, mp,
                                     (rocblas_int)np, (int64_t)px * np, E, (rocblas_int)np, 0,
                                     rho + q0 * (int64_t)px * py, px, (int64_t)px * py, qc));
        }
        if (pair_one) {
            TT_BLAS(gemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                         (rocblas_int)((int64_t)qc * px), py, (rocblas_int)np, mp,
                         (rocblas_int)np, E, (rocblas_int)np, rho_c,
                         (rocblas_int)((int64_t)qc * px)));
            const int64_t cnt = (int64_t)qc * px * py;
            hipLaunchKernelGGL(widen_planes_kernel<Tm>, dim3(grid_for(cnt)), dim3(256), 0,
                               ctx->stream, rho_c, rho + q0 * (int64_t)px * py, qc, px, py);
            TT_LAUNCH_CHECK();
        }
    }
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
    std::vector<double> sw;
    for (int i = 0; i < k; ++i) {
        const double w = 2.0 * occ[i];
        TT_CHECK(w >= 0.0, TTDFT_EBADARG, "density: negative occupancy %g of state %d", occ[i], i);
        if (w != 0.0) {
            idx.push_back(i);
            sw.push_back(std::sqrt(w));
        }
    }
    if (idx.empty()) {
        TT_HIP(hipMemsetAsync(rho, 0, sizeof(double) * (size_t)px * py * pz, ctx->stream));
        return TTDFT_OK;
    }
    const int32_t K = (int32_t)idx.size();
    const int64_t rxy = (int64_t)rx * ry;
    const int64_t np = (int64_t)ry * (ry + 1) / 2;
    TT_CHECK(rxy * K < ((int64_t)1 << 31), TTDFT_EBADARG,
             "density: rx*ry*states = %lld exceeds the 32-bit GEMM dimension",
             (long long)(rxy * K));
    PendingTiming t;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_DENSITY, &t));
    // An FP32 block takes FP32 intermediates (T1, T2, the state sums of M);
    // M and the expansion onto the y points stay FP64.
    if (dtype == TTDFT_F32)
        TT_TRY(density_run<float>(ctx, X, ldx, dtype, idx, sw, rx, ry, rz, Bx, By, Bz, px, py,
                                  pz, rho));
    else
        TT_TRY(density_run<double>(ctx, X, ldx, dtype, idx, sw, rx, ry, rz, Bx, By, Bz, px, py,
                                   pz, rho));
    const double nst = (double)K;
    const double flops = 2.0 * nst * ((double)N * pz + (double)pz * rx * ry * px) +
                         nst * (double)ry * (ry + 1) * px * pz +
                         2.0 * (double)np * px * py * pz;
    const double bytes = (double)N * k * dtype_size(dtype) + 8.0 * px * py * pz;
    TT_TRY(timing_end(ctx, &t, bytes, flops));
    return TTDFT_OK;
}

}  // namespace ttdft
