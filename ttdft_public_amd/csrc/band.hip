// Register-blocked SpMM for H_loc with Kronecker band structure.
//
// In a Tucker product basis the projected Hamiltonian couples
// (x,y,z) with (x',y',z') only for |x-x'| <= bx, |y-y'| <= by, |z-z'| <= bz
// (ProjectHamiltonianSparse.cc:464-548 builds each row as a Kronecker dyad and
// truncates it).  Such an H_loc is stored index-free: for every x-line
// (y,z) and every (dy,dz) pair, an rx x (2bx+1) block of band values
// (zeros where the CSR has no entry).
//
// One wavefront owns RW consecutive x-rows of LPW consecutive y-lines (same
// z) and a 64-state panel (lane = state); small tiles (RW = 5, LPW = 2 in
// FP64, 4 in FP32) keep ~100 VGPRs per wave so enough waves are resident to
// hide the scalar-cache latency of the band values.  For every neighbouring line
// (y',z') it loads the RW + 2bx X rows of that line's x-window into registers
// once (coalesced 64-state rows) and feeds up to LPW * RW * (2bx+1) FMAs from
// them; the band values are wavefront-uniform and arrive through the scalar
// cache.  No LDS and no barriers: waves are independent.  The nonlocal term
// C^T W and the Chebyshev three-term recurrence are applied in the epilogue,
// as in the CSR kernel.  Blocks are renumbered so consecutive logical blocks
// land on the same XCD (blocks are dispatched round-robin over the 8 XCDs),
// keeping neighbouring lines' X rows in one L2.
#include <algorithm>
#include <thread>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

template <typename T>
struct BandArgs {
    const T *band;
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
    int32_t k, rx, ry, rz, by, bz, rxp, ngx, nly, npanel;
    int64_t nwaves;
    T alpha, beta, delta;
    int use_beta, use_delta;
};

template <typename T, int W, int RW, int kLPW>
__global__ __launch_bounds__(256) void band_spmm_kernel(BandArgs<T> a) {
    constexpr int BX = (W - 1) / 2;
    constexpr int NWIN = RW + W - 1;
    const unsigned nb = gridDim.x;  // a multiple of 8
    const unsigned bid = blockIdx.x;
    const unsigned L = (bid & 7u) * (nb >> 3) + (bid >> 3);
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)L * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= a.nwaves) return;
    int64_t t = wid;
    const int panel = (int)(t % a.npanel);
    t /= a.npanel;
    const int gx = (int)(t % a.ngx);
    t /= a.ngx;
    const int ly = (int)(t % a.nly);
    const int z = (int)(t / a.nly);
    const int xr0 = gx * RW, y0 = ly * kLPW;
    const int c = panel * 64 + lane;
    const bool cok = c < a.k;
    const int wy = 2 * a.by + 1;
    const int P = wy * (2 * a.bz + 1);

    T acc[kLPW][RW];
#pragma unroll
    for (int l = 0; l < kLPW; ++l)
#pragma unroll
        for (int r = 0; r < RW; ++r) acc[l][r] = T(0);

    const int zlo = z - a.bz < 0 ? 0 : z - a.bz;
    const int zhi = z + a.bz > a.rz - 1 ? a.rz - 1 : z + a.bz;
    const int ylo = y0 - a.by < 0 ? 0 : y0 - a.by;
    const int yhi = y0 + kLPW - 1 + a.by > a.ry - 1 ? a.ry - 1 : y0 + kLPW - 1 + a.by;
    for (int zz = zlo; zz <= zhi; ++zz) {
        for (int yy = ylo; yy <= yhi; ++yy) {
            const T *xb = a.X + ((int64_t)(zz * a.ry + yy) * a.rx) * a.ldx + c;
            T xw[NWIN];
#pragma unroll
            for (int j = 0; j < NWIN; ++j) {
                const int xp = xr0 - BX + j;
                xw[j] = (cok && xp >= 0 && xp < a.rx) ? xb[(int64_t)xp * a.ldx] : T(0);
            }
#pragma unroll
            for (int l = 0; l < kLPW; ++l) {
                const int y = y0 + l;
                const int dy = yy - y;
                if (y < a.ry && dy >= -a.by && dy <= a.by) {
                    const int pair = (dy + a.by) + wy * (zz - z + a.bz);
                    const T *Al =
                        a.band + (((int64_t)(y + a.ry * z) * P + pair) * a.rxp + xr0) * W;
#pragma unroll
                    for (int r = 0; r < RW; ++r)
#pragma unroll
                        for (int d = 0; d < W; ++d) acc[l][r] += Al[r * W + d] * xw[r + d];
                }
            }
        }
    }
    if (!cok) return;
#pragma unroll
    for (int l = 0; l < kLPW; ++l) {
        const int y = y0 + l;
        if (y >= a.ry) continue;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int x = xr0 + r;
            if (x >= a.rx) continue;
            const int64_t row = x + (int64_t)a.rx * (y + (int64_t)a.ry * z);
            T v = acc[l][r];
            if (a.ct_rowptr) {
                const int64_t f0 = a.ct_rowptr[row], f1 = a.ct_rowptr[row + 1];
                for (int64_t e = f0; e < f1; ++e)
                    v += a.ct_val[e] * a.Wnl[(int64_t)a.ct_col[e] * a.ldw + c];
            }
            T out = v * a.alpha;
            if (a.use_beta) out = out + a.beta * a.X[row * a.ldx + c];
            if (a.use_delta) out = out + a.delta * a.Z[row * a.ldz + c];
            a.Y[row * a.ldy + c] = out;
        }
    }
}

template <typename T>
constexpr int lines_per_wave() { return sizeof(T) == 8 ? 2 : 4; }

template <typename T, int W, int RW>
int launch_tmpl(ttdft_ctx ctx, const BandArgs<T> &a) {
    int64_t blocks = (a.nwaves + 3) / 4;
    blocks = (blocks + 7) / 8 * 8;
    hipLaunchKernelGGL((band_spmm_kernel<T, W, RW, lines_per_wave<T>()>), dim3((unsigned)blocks),
                       dim3(256), 0, ctx->stream, a);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

template <typename T>
int dispatch(ttdft_ctx ctx, const BandArgs<T> &a, int W, int RW) {
#define TT_BAND_CASE(WV)                                                              \
    case WV:                                                                          \
        return RW == 5 ? launch_tmpl<T, WV, 5>(ctx, a)                                \
                       : (RW == 8 ? launch_tmpl<T, WV, 8>(ctx, a) : launch_tmpl<T, WV, 10>(ctx, a));
    switch (W) {
        TT_BAND_CASE(5)
        TT_BAND_CASE(7)
        TT_BAND_CASE(9)
        TT_BAND_CASE(11)
        TT_BAND_CASE(13)
        TT_BAND_CASE(15)
    default:
        break;
    }
#undef TT_BAND_CASE
    set_error("band SpMM: unsupported band width %d", W);
    return TTDFT_EBADARG;
}

}  // namespace

bool band_supported(int32_t W) { return W >= 5 && W <= 15 && (W & 1); }

int launch_band_spmm(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                     const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                     int32_t k, int dtype, double alpha, double beta, double delta,
                     int use_beta, int use_delta) {
    const int Wd = 2 * H->bx + 1;
    const int npanel = (k + 63) / 64;
    const int ngx = H->rxp / H->rw;
    const int lpw = dtype == TTDFT_F64 ? lines_per_wave<double>() : lines_per_wave<float>();
    const int nly = (H->ry + lpw - 1) / lpw;
    const int64_t nwaves = (int64_t)npanel * ngx * nly * H->rz;
    if (dtype == TTDFT_F64) {
        TT_CHECK(H->band64, TTDFT_EBADARG, "no FP64 band copy of H_loc uploaded");
        BandArgs<double> a{H->band64, (const double *)X, ldx,
                           H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val64,
                           (const double *)W, k, (double *)Y, ldy, (const double *)Z, ldz,
                           k, H->rx, H->ry, H->rz, H->by, H->bz, H->rxp, ngx, nly, npanel,
                           nwaves, alpha, beta, delta, use_beta, use_delta};
        return dispatch<double>(ctx, a, Wd, H->rw);
    }
    TT_CHECK(H->band32, TTDFT_EBADARG, "no FP32 band copy of H_loc uploaded");
    BandArgs<float> a{H->band32, (const float *)X, ldx,
                      H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val32,
                      (const float *)W, k, (float *)Y, ldy, (const float *)Z, ldz,
                      k, H->rx, H->ry, H->rz, H->by, H->bz, H->rxp, ngx, nly, npanel,
                      nwaves, (float)alpha, (float)beta, (float)delta, use_beta, use_delta};
    return dispatch<float>(ctx, a, Wd, H->rw);
}

// ---------------------------------------------------------------------------
// Host: detect the band and build the band blocks from CSR.
int build_band(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz,
               const int64_t *rowptr, const int32_t *colind, const double *val, int copies) {
    const int64_t n = (int64_t)rx * ry * rz;
    // half-bandwidths (parallel max over rows)
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
                    const int x = (int)(i % rx), y = (int)((i / rx) % ry), z = (int)(i / ((int64_t)rx * ry));
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
    const int W = 2 * bx + 1;
    const int64_t P = (int64_t)(2 * by + 1) * (2 * bz + 1);
    // rows per wavefront: the register block with the least padding of rx
    // (ties go to the smaller block, which keeps more waves resident)
    const int pad5 = (5 - rx % 5) % 5, pad8 = (8 - rx % 8) % 8, pad10 = (10 - rx % 10) % 10;
    const int rw = (pad5 <= pad8 && pad5 <= pad10) ? 5 : (pad8 <= pad10 ? 8 : 10);
    const int rxp = (rx + rw - 1) / rw * rw;
    const int64_t slots = (int64_t)ry * rz * P * rxp * W;
    const int64_t nnz = rowptr[n];
    const double fill = slots > 0 ? (double)nnz / (double)slots : 0.0;
    if (!band_supported(W) || fill < 0.15 || rx < 4) return TTDFT_OK;  // keep CSR
    std::vector<double> band((size_t)slots, 0.0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (int64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) {
                    const int x = (int)(i % rx), y = (int)((i / rx) % ry), z = (int)(i / ((int64_t)rx * ry));
                    const int64_t line = y + (int64_t)ry * z;
                    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
                        const int64_t j = colind[e];
                        const int dx = (int)(j % rx) - x, dy = (int)((j / rx) % ry) - y,
                                  dz = (int)(j / ((int64_t)rx * ry)) - z;
                        const int64_t pair = (dy + by) + (int64_t)(2 * by + 1) * (dz + bz);
                        band[((line * P + pair) * rxp + x) * W + dx + bx] += val[e];
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    if (copies & TTDFT_HAM_F64) {
        TT_HIP(hipMalloc(&H->band64, sizeof(double) * band.size()));
        TT_HIP(hipMemcpyAsync(H->band64, band.data(), sizeof(double) * band.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (copies & TTDFT_HAM_F32) {
        std::vector<float> b32(band.size());
        for (size_t q = 0; q < band.size(); ++q) b32[q] = (float)band[q];
        TT_HIP(hipMalloc(&H->band32, sizeof(float) * b32.size()));
        TT_HIP(hipMemcpyAsync(H->band32, b32.data(), sizeof(float) * b32.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    H->format = 1;
    H->rx = rx;
    H->ry = ry;
    H->rz = rz;
    H->bx = bx;
    H->by = by;
    H->bz = bz;
    H->rw = rw;
    H->rxp = rxp;
    return TTDFT_OK;
}

}  // namespace ttdft
