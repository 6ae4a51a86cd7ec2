// Anderson density mixing with the history held in device memory
// (dft/AndersonMixing.cc).
//
// The reference keeps up to maxIterHistory (rho_in, rho_out) pairs of the
// nodal and quadrature-grid densities, writes every entry to disk, and for
// the mixing constants rebuilds each difference d_m = F_n - F_{h-2-m}
// (F = out - in) for every (m, k) pair: O(h^2) passes over the grid.  Here
// the history stays in HBM (a 600^3 grid entry is 1.7 GB; 288 GB holds tens
// of them) and one pass over it gives the whole system:
//   G[m][k] = sum_i d_m[i] d_k[i] jw[i],  G[m][h-1] = sum_i d_m[i] F_n[i] jw[i]
// (jw = jacobian x quadrature weight), as Dm^T Dw on the FP64 16x16x4 MFMA
// with Dm[m][p] = d_m[p] and Dw[k][p] = d_k[p] jw[p] (row h-1: F_n jw) staged
// in LDS per 64 points.  Per-workgroup partial sums are added in a fixed
// order, the (h-1) x (h-1) system is solved on the host by LU with partial
// pivoting, and a second pass forms the mixed density.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "internal.h"

constexpr int kMixMaxHistory = 64;

struct ttdft_mixer_s {
    struct Entry {
        double *grid_in = nullptr, *grid_out = nullptr, *nodal_in = nullptr, *nodal_out = nullptr;
    };
    ttdft_ctx ctx = nullptr;
    int64_t n_grid = 0, n_nodal = 0;
    double alpha = 0.0;
    int32_t max_history = 0;
    int32_t history_offset = 0;
    double *jw = nullptr;                 // jacobian x weight on the grid
    std::deque<Entry> hist;               // oldest first
    std::vector<Entry> spare;             // buffers of dropped entries, reused
    std::vector<double> constants;        // of the last Anderson step
    double *partial = nullptr;            // per-workgroup Gram partials
    size_t partial_bytes = 0;
    double *gram = nullptr;               // reduced system (64 x 64 + rhs)
};

namespace ttdft {
namespace {

struct FieldTable {
    const double *in[kMixMaxHistory];
    const double *out[kMixMaxHistory];
};

typedef double md4_t __attribute__((ext_vector_type(4)));

constexpr int kChunk = 64;          // points per LDS stage
constexpr int kLdsStride = kChunk + 1;
constexpr int kGramBlocks = 2048;   // workgroups of the Gram pass (8 per CU)

// One pass over the grid.  Output: partial[(b * 4 + w)][m][k] (mp x kp,
// row-major) per workgroup b and wave w.  LDS holds mp + kp rows of one
// 64-point stage (sized to the history, so short histories leave room for
// many workgroups and loads in flight).  Wave w stages rows w, w + 4, ...
// (RPW of them, loads unrolled and clamped so they issue back to back).  Up to
// 4 Gram tiles: every wave multiplies all tiles over its 16 of the 64 points;
// more tiles: wave w owns tiles w, w + 4, ... over all 64 points.
template <int RPW>
__global__ __launch_bounds__(256) void mix_gram_kernel(FieldTable t, int h, int mp, int kp,
                                                       const double *__restrict__ jw, int64_t n,
                                                       double *__restrict__ partial) {
    extern __shared__ double mix_lds[];
    double *Dm = mix_lds;
    double *Dw = mix_lds + mp * kLdsStride;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m1 = h - 1;
    // padding rows stay zero for the whole pass
    for (int e = threadIdx.x; e < mp * kLdsStride; e += 256)
        if (e / kLdsStride >= m1) Dm[e] = 0.0;
    for (int e = threadIdx.x; e < kp * kLdsStride; e += 256)
        if (e / kLdsStride > m1) Dw[e] = 0.0;
    const int ntm = mp / 16, ntk = kp / 16, ntiles = ntm * ntk;
    const bool split_points = ntiles <= 4;
    md4_t acc[4];
    for (int q = 0; q < 4; ++q) acc[q] = md4_t{0.0, 0.0, 0.0, 0.0};
    const double *inn = t.in[h - 1], *outn = t.out[h - 1];
    // this wave's history entries (clamped; rows >= m1 are not stored)
    const double *pin[RPW], *pout[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int m = min(w + 4 * r, m1 - 1);
        pin[r] = t.in[h - 2 - m];
        pout[r] = t.out[h - 2 - m];
    }
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        __syncthreads();  // the previous chunk's MFMAs have read the LDS stage
        const int64_t i = c * kChunk + lane;
        const bool ok = i < n;
        const int64_t ii = ok ? i : n - 1;
        const double vin = inn[ii], vout = outn[ii], vjw = jw[ii];
        double fin[RPW], fout[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            fin[r] = pin[r][ii];
            fout[r] = pout[r][ii];
        }
        const double fn = ok ? vout - vin : 0.0;
        const double wv = ok ? vjw : 0.0;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int m = w + 4 * r;
            if (m < m1) {
                const double d = ok ? fn - (fout[r] - fin[r]) : 0.0;
                Dm[m * kLdsStride + lane] = d;
                Dw[m * kLdsStride + lane] = d * wv;
            }
        }
        if (w == 0) Dw[m1 * kLdsStride + lane] = fn * wv;
        __syncthreads();
        if (split_points) {
            for (int q = 0; q < ntiles; ++q) {
                const int mt = q / ntk, kt = q - mt * ntk;
                const double *ar = Dm + (16 * mt + (lane & 15)) * kLdsStride + (lane >> 4) + 16 * w;
                const double *br = Dw + (16 * kt + (lane & 15)) * kLdsStride + (lane >> 4) + 16 * w;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[4 * s], br[4 * s], acc[q],
                                                                  0, 0, 0);
            }
        } else {
            for (int q = 0; q < 4; ++q) {
                const int tile = w + 4 * q;
                if (tile >= ntiles) break;
                const int mt = tile / ntk, kt = tile - mt * ntk;
                const double *ar = Dm + (16 * mt + (lane & 15)) * kLdsStride + (lane >> 4);
                const double *br = Dw + (16 * kt + (lane & 15)) * kLdsStride + (lane >> 4);
                for (int s = 0; s < kChunk / 4; ++s)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[4 * s], br[4 * s], acc[q],
                                                                  0, 0, 0);
            }
        }
    }
    // D layout: element v of lane l is [l/16 + 4v][l%16] of the tile; every
    // tile of this wave's slot is written (zero where another wave owns it)
    double *pb = partial + ((int64_t)blockIdx.x * 4 + w) * mp * kp;
    for (int tile = 0; tile < ntiles; ++tile) {
        const int mt = tile / ntk, kt = tile - mt * ntk;
        md4_t v4 = md4_t{0.0, 0.0, 0.0, 0.0};
        if (split_points)
            v4 = acc[tile];
        else if ((tile & 3) == w)
            v4 = acc[tile >> 2];
        for (int v = 0; v < 4; ++v)
            pb[(16 * mt + (lane >> 4) + 4 * v) * kp + 16 * kt + (lane & 15)] = v4[v];
    }
}

// gram[e] = sum over slots b of partial[b][e]: one workgroup per entry, a
// strided sum per thread, then a fixed tree (the same order on every run).
__global__ __launch_bounds__(256) void mix_reduce_kernel(const double *__restrict__ partial,
                                                         int nslots, int count,
                                                         double *__restrict__ gram) {
    __shared__ double red[256];
    const int e = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nslots; b += 256) s += partial[(int64_t)b * count + e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) gram[e] = red[0];
}

struct MixCoef {
    double c[kMixMaxHistory];
};

// MODE 0: rho = |rho| (first iteration after a reset);
// MODE 1: rho = alpha out_last + (1 - alpha) in_last (second iteration);
// MODE 2: Anderson: in_mix = sum_k c_k (in_{h-2-k} - in_n) + in_n, out_mix
//         likewise, rho = |alpha out_mix + (1 - alpha) in_mix|   (:264-323)
template <int MODE>
__global__ __launch_bounds__(256) void mix_combine_kernel(FieldTable t, int h, MixCoef cf,
                                                          double alpha, int64_t n,
                                                          double *__restrict__ rho) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (MODE == 0) {
            rho[i] = fabs(rho[i]);
        } else if (MODE == 1) {
            rho[i] = alpha * t.out[h - 1][i] + (1.0 - alpha) * t.in[h - 1][i];
        } else {
            const double inn = t.in[h - 1][i], outn = t.out[h - 1][i];
            double a = 0.0, b = 0.0;
#pragma unroll 4
            for (int k = 0; k < h - 1; ++k) {
                a += cf.c[k] * (t.in[h - 2 - k][i] - inn);
                b += cf.c[k] * (t.out[h - 2 - k][i] - outn);
            }
            a += inn;
            b += outn;
            rho[i] = fabs(alpha * b + (1.0 - alpha) * a);
        }
    }
}

__global__ __launch_bounds__(256) void mix_weights_kernel(const double *__restrict__ jacob,
                                                          const double *__restrict__ weight,
                                                          int64_t n, double *__restrict__ jw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) jw[i] = jacob[i] * weight[i];
}

// A x = b, A n x n column-major (overwritten), b overwritten by x: LU with
// partial pivoting, the factorisation of LAPACK dgesv.  0 or pivot index + 1.
int host_lu_solve(int n, std::vector<double> &A, std::vector<double> &b) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r + (size_t)n * c]) > std::fabs(A[piv + (size_t)n * c])) piv = r;
        if (A[piv + (size_t)n * c] == 0.0) return c + 1;
        if (piv != c) {
            for (int q = 0; q < n; ++q) std::swap(A[c + (size_t)n * q], A[piv + (size_t)n * q]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < n; ++r) {
            const double l = A[r + (size_t)n * c] / A[c + (size_t)n * c];
            A[r + (size_t)n * c] = l;
            for (int q = c + 1; q < n; ++q) A[r + (size_t)n * q] -= l * A[c + (size_t)n * q];
            b[r] -= l * b[c];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int q = r + 1; q < n; ++q) s -= A[r + (size_t)n * q] * b[q];
        b[r] = s / A[r + (size_t)n * r];
    }
    return 0;
}

unsigned stream_blocks(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

FieldTable table(const ttdft_mixer_s *M, bool grid) {
    FieldTable t;
    std::memset(&t, 0, sizeof(t));
    for (size_t j = 0; j < M->hist.size(); ++j) {
        t.in[j] = grid ? M->hist[j].grid_in : M->hist[j].nodal_in;
        t.out[j] = grid ? M->hist[j].grid_out : M->hist[j].nodal_out;
    }
    return t;
}

template <int MODE>
int combine(ttdft_mixer_s *M, bool grid, const MixCoef &cf, double *rho) {
    const int64_t n = grid ? M->n_grid : M->n_nodal;
    if (n == 0) return TTDFT_OK;
    hipLaunchKernelGGL(mix_combine_kernel<MODE>, dim3(stream_blocks(n)), dim3(256), 0,
                       M->ctx->stream, table(M, grid), (int)M->hist.size(), cf, M->alpha, n,
                       rho);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

// AndersonMixing.cc:150-262 on the grid history
int mixing_constants(ttdft_mixer_s *M) {
    ttdft_ctx ctx = M->ctx;
    const int h = (int)M->hist.size();
    const int m1 = h - 1;
    M->constants.assign(std::max(m1, 0), 0.0);
    if (m1 <= 0) return TTDFT_OK;
    const int mp = (m1 + 15) / 16 * 16, kp = (m1 + 1 + 15) / 16 * 16;
    const int64_t nchunks = (M->n_grid + kChunk - 1) / kChunk;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kGramBlocks, nchunks));
    const size_t lds = sizeof(double) * (size_t)(mp + kp) * kLdsStride;
    const size_t need = sizeof(double) * (size_t)nb * 4 * mp * kp;
    if (M->partial_bytes < need) {
        if (M->partial) TT_HIP(hipFree(M->partial));
        M->partial = nullptr;
        M->partial_bytes = 0;
        TT_TRY(dev_malloc_retry(ctx, (void **)&M->partial, need));
        M->partial_bytes = need;
    }
    PendingTiming tm;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_MIXING, &tm));
    const int rpw = (m1 + 3) / 4;  // history rows staged per wave
    auto kern = rpw <= 1 ? mix_gram_kernel<1>
                : rpw <= 2 ? mix_gram_kernel<2>
                : rpw <= 4 ? mix_gram_kernel<4>
                : rpw <= 8 ? mix_gram_kernel<8> : mix_gram_kernel<16>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), lds, ctx->stream, table(M, true), h, mp, kp,
                       M->jw, M->n_grid, M->partial);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(mix_reduce_kernel, dim3(mp * kp), dim3(256), 0, ctx->stream, M->partial,
                       4 * nb, mp * kp, M->gram);
    TT_LAUNCH_CHECK();
    // bytes: 2h grid fields and jw read once; flops: the Gram on the MFMA
    TT_TRY(timing_end(ctx, &tm, 8.0 * (2.0 * h + 1.0) * (double)M->n_grid,
                      2.0 * mp * kp * (double)M->n_grid));
    std::vector<double> g((size_t)mp * kp);
    TT_HIP(hipMemcpyAsync(g.data(), M->gram, sizeof(double) * g.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<double> A((size_t)m1 * m1), b(m1);
    for (int m = 0; m < m1; ++m) {
        for (int k = 0; k < m1; ++k) A[m + (size_t)k * m1] = g[(size_t)m * kp + k];
        b[m] = g[(size_t)m * kp + m1];
    }
    const int info = host_lu_solve(m1, A, b);
    TT_CHECK(info == 0, TTDFT_ENOCONV, "mixer: singular mixing system (pivot %d)", info);
    M->constants = b;
    return TTDFT_OK;
}

void free_entry(const ttdft_mixer_s::Entry &e) {
    for (double *p : {e.grid_in, e.grid_out, e.nodal_in, e.nodal_out})
        if (p) {
            hipError_t r = hipFree(p);
            (void)r;
        }
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" {

int ttdft_mixer_create(ttdft_ctx ctx, int64_t n_grid, int64_t n_nodal, const double *d_jacob,
                       const double *d_weight, double alpha, int32_t max_history,
                       ttdft_mixer *out) {
    clear_error();
    TT_CHECK(ctx && out, TTDFT_EBADARG, "mixer_create: NULL ctx or out");
    *out = nullptr;
    TT_CHECK(n_grid > 0 && n_nodal >= 0, TTDFT_EBADARG, "mixer_create: bad field sizes");
    TT_CHECK(d_jacob && d_weight, TTDFT_EBADARG, "mixer_create: NULL jacobian or weights");
    TT_CHECK(max_history >= 1 && max_history <= kMixMaxHistory, TTDFT_EBADARG,
             "mixer_create: max_history %d outside [1, %d]", max_history, kMixMaxHistory);
    TT_HIP(hipSetDevice(ctx->device));
    ttdft_mixer M = new ttdft_mixer_s();
    M->ctx = ctx;
    ctx_attach(ctx);
    M->n_grid = n_grid;
    M->n_nodal = n_nodal;
    M->alpha = alpha;
    M->max_history = max_history;
    hipError_t e = hipMalloc(&M->jw, sizeof(double) * n_grid);
    if (e == hipSuccess) e = hipMalloc(&M->gram, sizeof(double) * 64 * 80);
    if (e != hipSuccess) {
        ttdft_mixer_destroy(M);
        set_error("mixer_create: hipMalloc failed: %s", hipGetErrorString(e));
        return TTDFT_ENOMEM;
    }
    hipLaunchKernelGGL(mix_weights_kernel, dim3((unsigned)((n_grid + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_jacob, d_weight, n_grid, M->jw);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ttdft_mixer_destroy(M);
        set_error("mixer_create: %s", hipGetErrorString(e));
        return TTDFT_EHIP;
    }
    *out = M;
    return TTDFT_OK;
}

int ttdft_mixer_destroy(ttdft_mixer M) {
    if (!M) return TTDFT_OK;
    hipError_t e = hipStreamSynchronize(M->ctx->stream);
    (void)e;
    for (auto &x : M->hist) free_entry(x);
    for (auto &x : M->spare) free_entry(x);
    for (double *p : {M->jw, M->partial, M->gram})
        if (p) {
            e = hipFree(p);
            (void)e;
        }
    ttdft_ctx ctx = M->ctx;
    delete M;
    ctx_detach(ctx);
    return TTDFT_OK;
}

int ttdft_mixer_update(ttdft_mixer M, const double *d_rho_nodal_in, const double *d_rho_grid_in,
                       const double *d_rho_nodal_out, const double *d_rho_grid_out,
                       int32_t scf_iter) {
    clear_error();
    TT_CHECK(M, TTDFT_EBADARG, "mixer is NULL");
    TT_CHECK(d_rho_grid_in && d_rho_grid_out, TTDFT_EBADARG, "mixer_update: NULL grid density");
    TT_CHECK(M->n_nodal == 0 || (d_rho_nodal_in && d_rho_nodal_out), TTDFT_EBADARG,
             "mixer_update: NULL nodal density");
    (void)scf_iter;  // the reference names its history files after it; nothing is written here
    ttdft_ctx ctx = M->ctx;
    ttdft_mixer_s::Entry en;
    if (!M->spare.empty()) {
        en = M->spare.back();
        M->spare.pop_back();
    } else {
        const size_t g = sizeof(double) * M->n_grid, nd = sizeof(double) * M->n_nodal;
        int rc = dev_malloc_retry(ctx, (void **)&en.grid_in, g);
        if (rc == TTDFT_OK) rc = dev_malloc_retry(ctx, (void **)&en.grid_out, g);
        if (rc == TTDFT_OK && nd) rc = dev_malloc_retry(ctx, (void **)&en.nodal_in, nd);
        if (rc == TTDFT_OK && nd) rc = dev_malloc_retry(ctx, (void **)&en.nodal_out, nd);
        if (rc != TTDFT_OK) {
            const std::string msg = ttdft_last_error();
            free_entry(en);
            set_error("mixer_update: history entry: %s", msg.c_str());
            return TTDFT_ENOMEM;
        }
    }
    const size_t g = sizeof(double) * M->n_grid, nd = sizeof(double) * M->n_nodal;
    TT_HIP(hipMemcpyAsync(en.grid_in, d_rho_grid_in, g, hipMemcpyDeviceToDevice, ctx->stream));
    TT_HIP(hipMemcpyAsync(en.grid_out, d_rho_grid_out, g, hipMemcpyDeviceToDevice, ctx->stream));
    if (nd) {
        TT_HIP(hipMemcpyAsync(en.nodal_in, d_rho_nodal_in, nd, hipMemcpyDeviceToDevice,
                              ctx->stream));
        TT_HIP(hipMemcpyAsync(en.nodal_out, d_rho_nodal_out, nd, hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    M->hist.push_back(en);
    if ((int32_t)M->hist.size() > M->max_history) {  // :402-414
        M->spare.push_back(M->hist.front());
        M->hist.pop_front();
    }
    return TTDFT_OK;
}

int ttdft_mixer_compute_rho_in(ttdft_mixer M, int32_t scf_iter, double *d_rho_nodal_in,
                               double *d_rho_grid_in) {
    clear_error();
    TT_CHECK(M, TTDFT_EBADARG, "mixer is NULL");
    TT_CHECK(d_rho_grid_in, TTDFT_EBADARG, "mixer_compute_rho_in: NULL grid density");
    TT_CHECK(M->n_nodal == 0 || d_rho_nodal_in, TTDFT_EBADARG,
             "mixer_compute_rho_in: NULL nodal density");
    const int32_t off = scf_iter - M->history_offset;  // :331
    MixCoef cf;
    std::memset(&cf, 0, sizeof(cf));
    if (off == 0) {  // :333-339
        TT_TRY(combine<0>(M, true, cf, d_rho_grid_in));
        TT_TRY(combine<0>(M, false, cf, d_rho_nodal_in));
    } else if (off == 1) {  // :340-350
        TT_CHECK(!M->hist.empty(), TTDFT_EBADARG,
                 "mixer_compute_rho_in: no history for iteration %d", scf_iter);
        TT_TRY(combine<1>(M, true, cf, d_rho_grid_in));
        TT_TRY(combine<1>(M, false, cf, d_rho_nodal_in));
    } else {  // :351-372
        TT_CHECK(!M->hist.empty(), TTDFT_EBADARG,
                 "mixer_compute_rho_in: no history for iteration %d", scf_iter);
        TT_TRY(mixing_constants(M));
        for (size_t k = 0; k < M->constants.size(); ++k) cf.c[k] = M->constants[k];
        PendingTiming tm;
        TT_TRY(timing_begin(M->ctx, TTDFT_KSTAT_MIXING, &tm));
        TT_TRY(combine<2>(M, true, cf, d_rho_grid_in));
        TT_TRY(combine<2>(M, false, cf, d_rho_nodal_in));
        const double h = (double)M->hist.size();
        TT_TRY(timing_end(M->ctx, &tm, 8.0 * (2.0 * h + 1.0) * (double)(M->n_grid + M->n_nodal),
                          4.0 * h * (double)(M->n_grid + M->n_nodal)));
    }
    TT_HIP(hipStreamSynchronize(M->ctx->stream));
    return TTDFT_OK;
}

int ttdft_mixer_constants(ttdft_mixer M, double *constants, int32_t *count) {
    clear_error();
    TT_CHECK(M && count, TTDFT_EBADARG, "mixer_constants: NULL argument");
    *count = (int32_t)M->constants.size();
    if (constants) std::copy(M->constants.begin(), M->constants.end(), constants);
    return TTDFT_OK;
}

int ttdft_mixer_clear_history(ttdft_mixer M, int32_t scf_iter) {
    clear_error();
    TT_CHECK(M, TTDFT_EBADARG, "mixer is NULL");
    TT_HIP(hipStreamSynchronize(M->ctx->stream));
    for (auto &x : M->hist) M->spare.push_back(x);
    M->hist.clear();
    M->history_offset = scf_iter + 1;  // :445-461
    return TTDFT_OK;
}

int ttdft_mixer_history_size(ttdft_mixer M, int32_t *size) {
    TT_CHECK(M && size, TTDFT_EBADARG, "mixer_history_size: NULL argument");
    *size = (int32_t)M->hist.size();
    return TTDFT_OK;
}

}  // extern "C"
