// Matrix-core SpMM for H_loc with Kronecker band structure.
//
// In a Tucker product basis the projected Hamiltonian couples (x,y,z) with
// (x',y',z') only for |x-x'| <= bx, |y-y'| <= by, |z-z'| <= bz
// (ProjectHamiltonianSparse.cc:464-548 builds each row as a Kronecker dyad and
// truncates it).  The rows are grouped into 4x4 "bricks" (4 consecutive x, 4
// consecutive y, one z).  For one brick and one neighbouring x-line (y',z'),
// the coupling is a dense 16 x (4 + 2bx) block: 16 brick rows against the x'
// window of that line.  It is padded to 16 x 4KS and stored pre-swizzled in
// the A-operand layout of the 16x16x4 MFMA.  HX for a brick and a 16-state
// tile is then a sum of KS MFMAs per neighbouring line, so the flops of the
// band run on the matrix cores in FP64 or FP32 instead of the VALU.
//
// A workgroup owns one y-brick row (4 y), one z, a chunk of up to 16 x-bricks
// (one wavefront each) and 16*NST states.  For every neighbouring line it
// stages the line's X window (all x-bricks of the chunk plus the 2bx halo,
// 16*NST states) in LDS once; every wave reads its B operands from there.
// The stage is double-buffered in LDS; the global loads of X and of the A
// fragments run two lines ahead through register buffers, so HBM/L2 latency
// hides behind MFMA work.  The nonlocal term C^T W and the Chebyshev three-term
// recurrence are applied in the epilogue, as in the CSR kernel.  Workgroups are
// renumbered so consecutive logical workgroups land on the same XCD (dispatch
// is round-robin over the 8 XCDs), keeping neighbouring lines' X in one L2.
#include <algorithm>
#include <thread>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16x16x4(float a, float b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
struct Acc4;
template <>
struct Acc4<double> {
    typedef d4_t type;
};
template <>
struct Acc4<float> {
    typedef f4_t type;
};

// states per workgroup = 16 * kStateTiles<T>
template <typename T>
constexpr int kStateTiles() { return sizeof(T) == 8 ? 4 : 5; }

template <typename T>
struct BrickArgs {
    const T *mf;
    const T *X;
    int64_t ldx;
    const int64_t *ct_rowptr;
    const int32_t *ct_col;
    const T *ct_val;
    const T *Wnl;
    int32_t ldw;
    T *Y;
    int64_t ldy;
    const T *Z;
    int64_t ldz;
    int32_t k, rx, ry, rz, bx, by, bz;
    int32_t gxn, hyn;  // x-bricks, y-bricks
    int32_t gxb, nxc;  // x-bricks per workgroup, x-chunks
    int32_t nsplit;    // trailing x-bricks whose state tiles are split over two waves
    int32_t nsg;       // state groups
    int64_t nblocks;
    T alpha, beta, delta;
    int use_beta, use_delta;
};

// KS: 4-column MFMA steps covering the x' window (4 KS >= 4 + 2 bx).
// NST: 16-state tiles per workgroup.  PF: staged X elements per thread.
template <typename T, int KS, int NST, int PF>
__global__ __launch_bounds__(1024) void brick_spmm_kernel(BrickArgs<T> a) {
    typedef typename Acc4<T>::type acc_t;
    constexpr int S = 16 * NST;
    extern __shared__ __align__(16) unsigned char brick_lds[];
    T *Xs = reinterpret_cast<T *>(brick_lds);
    const int nthr = (a.gxb + a.nsplit) * 64;
    const int TR = 4 * (a.gxb - 1) + 4 * KS;  // staged x' rows
    const int TILE = TR * S;
    const int lane = threadIdx.x & 63;
    const int wl = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;

    const unsigned nb = gridDim.x;  // a multiple of 8
    const unsigned bid = blockIdx.x;
    const unsigned L = (bid & 7u) * (nb >> 3) + (bid >> 3);
    if ((int64_t)L >= a.nblocks) return;
    int64_t t = L;
    const int sg = (int)(t % a.nsg);
    t /= a.nsg;
    const int xc = (int)(t % a.nxc);
    t /= a.nxc;
    const int h = (int)(t % a.hyn);
    const int z = (int)(t / a.hyn);
    const int s0 = sg * S;
    // Waves are spread round-robin over the 4 SIMDs and the MFMA work of a
    // line ends at a barrier, so the SIMD with the most tile work sets the
    // pace.  With gxb = 1 or 2 (mod 4) the state tiles of the last one or two
    // x-bricks are split over two waves each (mode 1: lower tiles, mode 2:
    // upper tiles), which evens out the MFMA work per SIMD.
    constexpr int NH = NST - NST / 2;  // tiles of the lower half
    int bl = wl, mode = 0;
    if (wl >= a.gxb - a.nsplit) {
        const int q = wl - (a.gxb - a.nsplit);
        bl = a.gxb - a.nsplit + (q >> 1);
        mode = 1 + (q & 1);
    }
    const int g0 = xc * a.gxb;
    const int g = g0 + bl;
    const bool active = g < a.gxn;
    const int xw0 = 4 * g0 - a.bx;  // x' of staged row 0
    const int NPY = 2 * a.by + 4;
    const int NP = NPY * (2 * a.bz + 1);

    acc_t acc[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) acc[st] = acc_t{0, 0, 0, 0};

    const int y0 = 4 * h;
    const int zlo = z - a.bz < 0 ? 0 : z - a.bz;
    const int zhi = z + a.bz > a.rz - 1 ? a.rz - 1 : z + a.bz;
    const int ylo = y0 - a.by < 0 ? 0 : y0 - a.by;
    const int yhi = y0 + 3 + a.by > a.ry - 1 ? a.ry - 1 : y0 + 3 + a.by;
    const int ny = yhi - ylo + 1;
    const int nlines = ny * (zhi - zlo + 1);

    T ra[PF], rb[PF];
    // Loads are unconditional from clamped, always-valid addresses, and the
    // out-of-range elements are zeroed only when the line is staged: a load
    // under a branch, or a select right after it, would make the compiler wait
    // for the load there instead of two lines later.
    auto elem_ok = [&](int li_, int u) {
        const int e = threadIdx.x + nthr * u;
        const int row = e / S, s = e - row * S;
        const int xp = xw0 + row;
        return li_ < nlines && e < TILE && xp >= 0 && xp < a.rx && s0 + s < a.k;
    };
    auto fetch = [&](int li_, T *r) {
        const int lc = li_ < nlines ? li_ : nlines - 1;
        const int yy = ylo + lc % ny, zz = zlo + lc / ny;
        const T *xb = a.X + ((int64_t)(zz * a.ry + yy) * a.rx) * a.ldx;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + nthr * u;
            const int row = e / S, s = e - row * S;
            const int xp = xw0 + row, st = s0 + s;
            const int xpc = xp < 0 ? 0 : (xp >= a.rx ? a.rx - 1 : xp);
            const int stc = st < a.k ? st : a.k - 1;
            r[u] = xb[(int64_t)xpc * a.ldx + stc];
        }
    };
    auto stage = [&](int li_, const T *r) {
        T *dst = Xs + (li_ & 1) * TILE;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + nthr * u;
            if (e < TILE) dst[e] = elem_ok(li_, u) ? r[u] : T(0);
        }
    };
    // A fragments of line li_ (wave-private, from HBM/L2): loaded one line ahead
    auto load_a = [&](int li_, T *av) {
        const int lc = li_ < nlines ? li_ : nlines - 1;
        const int gc = active ? g : a.gxn - 1;
        const int yy = ylo + lc % ny, zz = zlo + lc / ny;
        const int tt = yy - y0;  // in [-by, 3 + by] by construction
        const int p = (tt + a.by) + NPY * (zz - z + a.bz);
        const T *Ab = a.mf + ((((int64_t)(z * a.hyn + h) * a.gxn + gc) * NP + p) * KS) * 64 + lane;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) av[ks] = Ab[ks * 64];
    };
    auto compute = [&](int li_, const T *av) {
        const T *Xb = Xs + (li_ & 1) * TILE + (4 * bl + lk) * S + li;
        if (mode == 0) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int st = 0; st < NST; ++st)
                    acc[st] = mfma16x16x4(av[ks], Xb[4 * ks * S + 16 * st], acc[st]);
        } else if (mode == 1) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int st = 0; st < NH; ++st)
                    acc[st] = mfma16x16x4(av[ks], Xb[4 * ks * S + 16 * st], acc[st]);
        } else {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int st = NH; st < NST; ++st)
                    acc[st] = mfma16x16x4(av[ks], Xb[4 * ks * S + 16 * st], acc[st]);
        }
    };

    // Line i is staged in LDS buffer i&1.  While it is multiplied, the X
    // windows of lines i+1 and i+2 (register sets xr[(i+1)&1], xr[i&1]) and the
    // A fragments of lines i+1 and i+2 (sets av[(i+1)%3], av[(i+2)%3]) are in
    // flight, so both loads have two lines of MFMA work to arrive.  The loop
    // is unrolled by 6 so every register set is addressed statically.
    T xr[2][PF];
    T av[3][KS];
    fetch(0, xr[0]);
    stage(0, xr[0]);
    fetch(1, xr[1]);
    load_a(0, av[0]);
    load_a(1, av[1]);
    __syncthreads();
    for (int base = 0; base < nlines; base += 6) {
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int li_ = base + j;
            if (li_ < nlines) {
                fetch(li_ + 2, xr[j & 1]);
                load_a(li_ + 2, av[(j + 2) % 3]);
                if (active) compute(li_, av[j % 3]);
                if (li_ + 1 < nlines) stage(li_ + 1, xr[(j + 1) & 1]);
                __syncthreads();
            }
        }
    }
    if (!active) return;

    // D layout of the 16x16x4 MFMA: FP64 element v of lane l is D[l/16 + 4v][l%16],
    // FP32 element v is D[4(l/16) + v][l%16].  D rows are brick rows i =
    // (x - 4g) + 4 (y - 4h), D columns are states.
#pragma unroll
    for (int st = 0; st < NST; ++st) {
        if ((mode == 1 && st >= NH) || (mode == 2 && st < NH)) continue;
        const int s = s0 + 16 * st + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = sizeof(T) == 8 ? (lk + 4 * v) : (4 * lk + v);
            const int x = 4 * g + (i & 3), y = y0 + (i >> 2);
            if (x < a.rx && y < a.ry && s < a.k) {
                const int64_t row = x + (int64_t)a.rx * (y + (int64_t)a.ry * z);
                T val = acc[st][v];
                if (a.ct_rowptr) {
                    const int64_t f0 = a.ct_rowptr[row], f1 = a.ct_rowptr[row + 1];
                    for (int64_t e = f0; e < f1; ++e)
                        val += a.ct_val[e] * a.Wnl[(int64_t)a.ct_col[e] * a.ldw + s];
                }
                T out = val * a.alpha;
                if (a.use_beta) out = out + a.beta * a.X[row * a.ldx + s];
                if (a.use_delta) out = out + a.delta * a.Z[row * a.ldz + s];
                a.Y[row * a.ldy + s] = out;
            }
        }
    }
}

// staged elements per thread: TR * S / (64 gxb) <= NST (gxb - 1 + KS) / gxb
constexpr int pf_for(int nst, int ks, int gxb_min) {
    return (nst * (gxb_min - 1 + ks) + gxb_min - 1) / gxb_min;
}

template <typename T, int KS>
int launch_ks(ttdft_ctx ctx, const BrickArgs<T> &a) {
    constexpr int NST = kStateTiles<T>();
    const int S = 16 * NST;
    const int TR = 4 * (a.gxb - 1) + 4 * KS;
    const size_t lds = (size_t)2 * TR * S * sizeof(T);
    const int nthr = 64 * (a.gxb + a.nsplit);
    const int pf = (TR * S + nthr - 1) / nthr;
    int64_t blocks = (a.nblocks + 7) / 8 * 8;
    TT_CHECK(lds <= 160 * 1024, TTDFT_EBADARG, "band SpMM: LDS stage too large");
    // two variants: wide chunks (>= 12 x-bricks) and narrow ones (>= 4)
    constexpr int PF_WIDE = pf_for(NST, KS, 12), PF_NARROW = pf_for(NST, KS, 4);
    if (a.gxb >= 12 && pf <= PF_WIDE) {
        auto kf = brick_spmm_kernel<T, KS, NST, PF_WIDE>;
        TT_HIP(hipFuncSetAttribute((const void *)kf, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(kf, dim3((unsigned)blocks), dim3(nthr), lds, ctx->stream, a);
    } else {
        TT_CHECK(pf <= PF_NARROW, TTDFT_EBADARG, "band SpMM: staging exceeds the register buffer");
        auto kf = brick_spmm_kernel<T, KS, NST, PF_NARROW>;
        TT_HIP(hipFuncSetAttribute((const void *)kf, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(kf, dim3((unsigned)blocks), dim3(nthr), lds, ctx->stream, a);
    }
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

template <typename T>
int launch_brick(ttdft_ctx ctx, BrickArgs<T> a, int ks) {
    a.nsg = (a.k + 16 * kStateTiles<T>() - 1) / (16 * kStateTiles<T>());
    a.nsplit = a.gxb % 4 == 1 ? 1 : (a.gxb % 4 == 2 ? 2 : 0);
    a.nblocks = (int64_t)a.nsg * a.nxc * a.hyn * a.rz;
    switch (ks) {
    case 2:
        return launch_ks<T, 2>(ctx, a);
    case 3:
        return launch_ks<T, 3>(ctx, a);
    case 4:
        return launch_ks<T, 4>(ctx, a);
    case 5:
        return launch_ks<T, 5>(ctx, a);
    default:
        break;
    }
    set_error("band SpMM: unsupported x half-bandwidth %d", a.bx);
    return TTDFT_EBADARG;
}

}  // namespace

int launch_band_spmm(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                     const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                     int32_t k, int dtype, double alpha, double beta, double delta,
                     int use_beta, int use_delta) {
    if (dtype == TTDFT_F64) {
        TT_CHECK(H->mf64, TTDFT_EBADARG, "no FP64 band copy of H_loc uploaded");
        BrickArgs<double> a{H->mf64, (const double *)X, ldx,
                            H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val64,
                            (const double *)W, k, (double *)Y, ldy, (const double *)Z, ldz,
                            k, H->rx, H->ry, H->rz, H->bx, H->by, H->bz, H->gxn, H->hyn,
                            H->gxb, H->nxc, 0, 0, 0, alpha, beta, delta, use_beta, use_delta};
        return launch_brick<double>(ctx, a, H->ks);
    }
    TT_CHECK(H->mf32, TTDFT_EBADARG, "no FP32 band copy of H_loc uploaded");
    BrickArgs<float> a{H->mf32, (const float *)X, ldx,
                       H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val32,
                       (const float *)W, k, (float *)Y, ldy, (const float *)Z, ldz,
                       k, H->rx, H->ry, H->rz, H->bx, H->by, H->bz, H->gxn, H->hyn,
                       H->gxb, H->nxc, 0, 0, 0, (float)alpha, (float)beta, (float)delta,
                       use_beta, use_delta};
    return launch_brick<float>(ctx, a, H->ks);
}

// ---------------------------------------------------------------------------
// Host: detect the band and scatter the CSR values into MFMA A fragments.
//
// Fragment layout: mf[((((z * hyn + h) * gxn + g) * NP + p) * KS + ks) * 64 + lane]
// holds A[i][kk] of the 16x16x4 MFMA (lane = i + 16 kk) for brick (g, h, z),
// neighbouring line p = (y' - 4h + by) + (2by + 4)(z' - z + bz) and x' =
// 4g - bx + 4ks + kk; brick row i = (x - 4g) + 4 (y - 4h).  Zero where H_loc
// has no entry.
int build_band(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz,
               const int64_t *rowptr, const int32_t *colind, const double *val, int copies) {
    const int64_t n = (int64_t)rx * ry * rz;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 32) nt = 32;
    std::vector<int> mx(nt, 0), my(nt, 0), mz(nt, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                int bx = 0, by = 0, bz = 0;
                for (int64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) {
                    const int x = (int)(i % rx), y = (int)((i / rx) % ry),
                              z = (int)(i / ((int64_t)rx * ry));
                    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
                        const int64_t j = colind[e];
                        const int xj = (int)(j % rx), yj = (int)((j / rx) % ry),
                                  zj = (int)(j / ((int64_t)rx * ry));
                        bx = std::max(bx, std::abs(xj - x));
                        by = std::max(by, std::abs(yj - y));
                        bz = std::max(bz, std::abs(zj - z));
                    }
                }
                mx[t] = bx;
                my[t] = by;
                mz[t] = bz;
            });
        for (auto &x : th) x.join();
    }
    const int bx = *std::max_element(mx.begin(), mx.end());
    const int by = *std::max_element(my.begin(), my.end());
    const int bz = *std::max_element(mz.begin(), mz.end());
    const int64_t band_slots = n * (int64_t)(2 * bx + 1) * (2 * by + 1) * (2 * bz + 1);
    const double fill = band_slots > 0 ? (double)rowptr[n] / (double)band_slots : 0.0;
    if (bx < 1 || bx > 7 || fill < 0.15 || rx < 4) return TTDFT_OK;  // keep CSR

    const int KS = (4 + 2 * bx + 3) / 4;
    const int gxn = (rx + 3) / 4, hyn = (ry + 3) / 4;
    const int NPY = 2 * by + 4, NP = NPY * (2 * bz + 1);
    const int64_t slots = (int64_t)rz * hyn * gxn * NP * KS * 64;
    std::vector<double> mf((size_t)slots, 0.0);
    {
        // rows of one thread fill distinct brick rows: no two threads write a slot
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (int64_t r = n * t / nt; r < n * (t + 1) / nt; ++r) {
                    const int x = (int)(r % rx), y = (int)((r / rx) % ry),
                              z = (int)(r / ((int64_t)rx * ry));
                    const int g = x >> 2, h = y >> 2;
                    const int i16 = (x & 3) + 4 * (y & 3);
                    const int64_t base = (((int64_t)z * hyn + h) * gxn + g) * NP;
                    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
                        const int64_t j = colind[e];
                        const int xj = (int)(j % rx), yj = (int)((j / rx) % ry),
                                  zj = (int)(j / ((int64_t)rx * ry));
                        const int p = (yj - 4 * h + by) + NPY * (zj - z + bz);
                        const int c = xj - 4 * g + bx;  // in [0, 4 KS)
                        mf[(size_t)(((base + p) * KS + (c >> 2)) * 64 + i16 + 16 * (c & 3))] +=
                            val[e];
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    if (copies & TTDFT_HAM_F64) {
        TT_HIP(hipMalloc(&H->mf64, sizeof(double) * mf.size()));
        TT_HIP(hipMemcpyAsync(H->mf64, mf.data(), sizeof(double) * mf.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (copies & TTDFT_HAM_F32) {
        std::vector<float> m32(mf.size());
        for (size_t q = 0; q < mf.size(); ++q) m32[q] = (float)mf[q];
        TT_HIP(hipMalloc(&H->mf32, sizeof(float) * m32.size()));
        TT_HIP(hipMemcpyAsync(H->mf32, m32.data(), sizeof(float) * m32.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    const int nxc = (gxn + 15) / 16;
    H->format = 1;
    H->rx = rx;
    H->ry = ry;
    H->rz = rz;
    H->bx = bx;
    H->by = by;
    H->bz = bz;
    H->ks = KS;
    H->gxn = gxn;
    H->hyn = hyn;
    H->nxc = nxc;
    H->gxb = std::max(4, (gxn + nxc - 1) / nxc);
    return TTDFT_OK;
}

}  // namespace ttdft
