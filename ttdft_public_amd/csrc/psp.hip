// Tabulated pseudopotentials on the device and the restart files of a run.
//
// The radial tables of an atom type (r, V_loc, V_l, R_l; atoms/NonLocalPSPData.cc)
// are interpolated by cubic splines whose first and last segments are
// parabolas: the C2 spline with S''' = 0 on both end segments, which is what
// alglib's spline1dbuildcubic builds when no end condition is given
// (:384-502).  The spline is built on the host (a tridiagonal solve over the
// second derivatives, a few hundred knots) and evaluated on the device, one
// grid point per lane with a binary search over the shared knots.
//
// Local part (computeLocalPart, :214-275): a superposition over the atoms
// of the spline, switched to -Z/d beyond d = 75.001.
// Projector data (the constructor, :33-212): dV_l = V_l - V_loc formed on the
// table, phi_lm = R_l Y_lm with real harmonics, and the integrals
// C_lm = int phi_lm^2 dV_l, summed exactly over the quadrature grid with a
// fixed block count and a fixed-order host sum (deterministic).
//
// Restart files (main.cc:541-592, 748-806, 1146-1170, 1212-1230): raw FP64 in
// big-endian byte order without a header (PetscBinaryWrite / PetscBinaryRead
// of PETSC_DOUBLE data) in the MATDENSE column-major order, converted from and
// to the engine's basis-major blocks, and cheby_bounds.txt with six values.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

constexpr int kPspBlocks = 256;
constexpr int kPspMaxL = 4;  // l = 0..3, 16 (l, m) fields

// C2 cubic spline through (x_i, y_i) with parabolic end segments, from its
// second derivatives M_i: interior rows
//   h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (s_i - s_{i-1}),
// end rows M_0 = M_1 and M_{n-1} = M_{n-2} (S'' constant on the end segments).
// Segment s: c0 + c1 t + c2 t^2 + c3 t^3, t = x - x_s.  Two knots give the
// straight line.
void spline_parabolic(int n, const double *x, const double *y, double *coef) {
    if (n == 2) {
        coef[0] = y[0];
        coef[1] = (y[1] - y[0]) / (x[1] - x[0]);
        coef[2] = coef[3] = 0.0;
        return;
    }
    std::vector<double> h(n - 1), lo(n), di(n), up(n), rhs(n), M(n);
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    di[0] = 1.0;
    up[0] = -1.0;
    rhs[0] = 0.0;
    lo[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        lo[i] = h[i - 1];
        di[i] = 2.0 * (h[i - 1] + h[i]);
        up[i] = h[i];
        rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
    }
    lo[n - 1] = -1.0;
    di[n - 1] = 1.0;
    up[n - 1] = 0.0;
    rhs[n - 1] = 0.0;
    // Thomas elimination.  The first and last rows are weakly dominant only;
    // for n >= 3 the eliminated pivots stay positive (row 1 gains h_0 from
    // row 0, the last row keeps 1 + up/di of its neighbour > 0).
    for (int i = 1; i < n; ++i) {
        const double w = lo[i] / di[i - 1];
        di[i] -= w * up[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    M[n - 1] = rhs[n - 1] / di[n - 1];
    for (int i = n - 2; i >= 0; --i) M[i] = (rhs[i] - up[i] * M[i + 1]) / di[i];
    for (int i = 0; i < n - 1; ++i) {
        coef[4 * i + 0] = y[i];
        coef[4 * i + 1] = (y[i + 1] - y[i]) / h[i] - h[i] * (2.0 * M[i] + M[i + 1]) / 6.0;
        coef[4 * i + 2] = 0.5 * M[i];
        coef[4 * i + 3] = (M[i + 1] - M[i]) / (6.0 * h[i]);
    }
}

// segment of rr among knots r[0..nr-1]; the end segments extend past the table
__device__ __forceinline__ int knot_segment(const double *__restrict__ r, int nr, double rr) {
    int lo = 0, hi = nr - 1;
    while (lo != hi - 1) {
        const int m = (lo + hi) >> 1;
        if (r[m] >= rr)
            hi = m;
        else
            lo = m;
    }
    return lo;
}

__device__ __forceinline__ double seg_eval(const double *__restrict__ coef, int s, double t) {
    const double *c = coef + 4 * s;
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
}

__global__ __launch_bounds__(256) void psp_local_kernel(
    int32_t natoms, const double *__restrict__ atoms, int32_t nr, const double *__restrict__ r,
    const double *__restrict__ coef, int32_t px, int32_t py, int32_t pz,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    double *__restrict__ v, int add) {
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
            const double *at = atoms + 7 * a;
            const double dx = x - at[1];
            const double dy = y - at[2];
            const double dz = z - at[3];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            const double d = dist + 1.0e-12;
            if (d > 75.001) {
                s += -at[0] / d;
            } else {
                const int seg = knot_segment(r, nr, dist);
                s += seg_eval(coef, seg, dist - r[seg]);
            }
        }
        v[idx] = add ? v[idx] + s : s;
    }
}

// Real spherical harmonics of degree < L at (cos theta = u, azimuth phi), in
// the order (l, m = -l..l): m < 0: sqrt2 N P_l^|m| sin(|m| phi), m > 0:
// sqrt2 N P_l^m cos(m phi), m = 0: N P_l^0, N^2 = (2l+1)/(4 pi) (l-m)!/(l+m)!,
// P_l^m with the Condon-Shortley phase (-1)^m.
template <int L>
__device__ __forceinline__ void real_harmonics(double u, double phi, double *Y) {
    const double s = sqrt(fmax(0.0, 1.0 - u * u));
    constexpr double kFourPi = 12.566370614359172954;
#pragma unroll
    for (int m = 0; m < L; ++m) {
        // P_m^m = (-1)^m (2m-1)!! s^m
        double pmm = 1.0;
        for (int i = 1; i <= m; ++i) pmm *= -(2.0 * i - 1.0) * s;
        double cm = 1.0, sm = 0.0;
        if (m > 0) sincos(m * phi, &sm, &cm);
        double p_lm2 = 0.0, p_lm1 = pmm;
        for (int l = m; l < L; ++l) {
            double p;
            if (l == m) {
                p = pmm;
            } else if (l == m + 1) {
                p = u * (2.0 * m + 1.0) * pmm;
            } else {
                p = ((2.0 * l - 1.0) * u * p_lm1 - (l + m - 1.0) * p_lm2) / (double)(l - m);
            }
            if (l > m) {
                p_lm2 = p_lm1;
                p_lm1 = p;
            }
            // (l-m)!/(l+m)!
            double ratio = 1.0;
            for (int i = l - m + 1; i <= l + m; ++i) ratio /= (double)i;
            const double nrm = sqrt((2.0 * l + 1.0) / kFourPi * ratio);
            const int base = l * l + l;  // index of m = 0
            if (m == 0) {
                Y[base] = nrm * p;
            } else {
                Y[base + m] = 1.4142135623730951 * nrm * p * cm;
                Y[base - m] = 1.4142135623730951 * nrm * p * sm;
            }
        }
    }
}

// One lane per grid point: dV_l and R_l from the splines (one knot search),
// phi_lm = R_l Y_lm; writes dV_l phi_lm and phi_lm where asked and adds
// w phi_lm^2 dV_l into the block's partial sums.
template <int L>
__global__ __launch_bounds__(256) void psp_projector_kernel(
    int32_t nr, const double *__restrict__ r, const double *__restrict__ cdv,
    const double *__restrict__ crl, int32_t px, int32_t py, int32_t pz,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    const double *__restrict__ wx, const double *__restrict__ wy, const double *__restrict__ wz,
    double *__restrict__ field, double *__restrict__ phi, double *__restrict__ partial) {
    constexpr int F = L * L;
    const int64_t total = (int64_t)px * py * pz;
    const int64_t seg_stride = 4 * (int64_t)(nr - 1);
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        const int qy = (int)(rest % py);
        const int qz = (int)(rest / py);
        const double x = gx[qx], y = gy[qy], z = gz[qz];
        const double rr = sqrt(x * x + y * y + z * z);
        const int seg = knot_segment(r, nr, rr);
        const double t = rr - r[seg];
        double Y[F];
        real_harmonics<L>(z / (rr + 1.0e-12), atan2(y, x), Y);
        const double w = partial ? wx[qx] * wy[qy] * wz[qz] : 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const double dv = seg_eval(cdv + l * seg_stride, seg, t);
            const double rl = seg_eval(crl + l * seg_stride, seg, t);
#pragma unroll
            for (int m = -l; m <= l; ++m) {
                const int f = l * l + l + m;
                const double p = rl * Y[f];
                if (field) field[(int64_t)f * total + idx] = dv * p;
                if (phi) phi[(int64_t)f * total + idx] = p;
                acc[f] += w * p * p * dv;
            }
        }
    }
    if (!partial) return;
    __shared__ double red[256];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        red[threadIdx.x] = acc[f];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * F + f] = red[0];
        __syncthreads();
    }
}

int check_knots(int32_t nr, const double *r, const char *who) {
    TT_CHECK(nr >= 2 && r, TTDFT_EBADARG, "%s: need at least two radii", who);
    for (int i = 0; i + 1 < nr; ++i)
        TT_CHECK(r[i + 1] > r[i], TTDFT_EBADARG, "%s: radii not increasing at row %d", who, i + 1);
    return TTDFT_OK;
}

void swap_bytes(double *v, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        uint64_t b;
        std::memcpy(&b, v + i, 8);
        b = __builtin_bswap64(b);
        std::memcpy(v + i, &b, 8);
    }
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" int ttdft_spline_parabolic(int32_t n, const double *x, const double *y, double *coef) {
    clear_error();
    TT_CHECK(x && y && coef, TTDFT_EBADARG, "spline_parabolic: NULL argument");
    TT_TRY(check_knots(n, x, "spline_parabolic"));
    spline_parabolic(n, x, y, coef);
    return TTDFT_OK;
}

extern "C" int ttdft_psp_local(ttdft_ctx ctx, int32_t natoms, const double *atoms, int32_t nr,
                               const double *r, const double *vloc, int32_t px, int32_t py,
                               int32_t pz, const double *d_gx, const double *d_gy,
                               const double *d_gz, double *d_v, int add) {
    clear_error();
    TT_CHECK(ctx && atoms && vloc && d_gx && d_gy && d_gz && d_v && natoms >= 1 && px >= 1 &&
                 py >= 1 && pz >= 1,
             TTDFT_EBADARG, "psp_local: bad arguments");
    TT_TRY(check_knots(nr, r, "psp_local"));
    std::vector<double> host(7 * (size_t)natoms + nr + 4 * (size_t)(nr - 1));
    std::memcpy(host.data(), atoms, sizeof(double) * 7 * natoms);
    std::memcpy(host.data() + 7 * natoms, r, sizeof(double) * nr);
    spline_parabolic(nr, r, vloc, host.data() + 7 * natoms + nr);
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * host.size()));
    double *d = (double *)ctx->scratch[6].ptr;
    TT_HIP(hipMemcpyAsync(d, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    const int64_t total = (int64_t)px * py * pz;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(psp_local_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, natoms,
                       d, nr, d + 7 * natoms, d + 7 * natoms + nr, px, py, pz, d_gx, d_gy, d_gz,
                       d_v, add);
    TT_LAUNCH_CHECK();
    TT_HIP(hipStreamSynchronize(ctx->stream));  // host staging goes out of scope
    return TTDFT_OK;
}

extern "C" int ttdft_psp_projectors(ttdft_ctx ctx, int32_t nr, const double *r,
                                    const double *vloc, const double *vl, const double *rl,
                                    int32_t lmax, int32_t px, int32_t py, int32_t pz,
                                    const double *d_gx, const double *d_gy, const double *d_gz,
                                    const double *d_wx, const double *d_wy, const double *d_wz,
                                    double *d_field, double *d_phi, double *clm) {
    clear_error();
    TT_CHECK(ctx && vloc && vl && rl && d_gx && d_gy && d_gz && px >= 1 && py >= 1 && pz >= 1,
             TTDFT_EBADARG, "psp_projectors: bad arguments");
    TT_CHECK(lmax >= 1 && lmax <= kPspMaxL, TTDFT_EBADARG,
             "psp_projectors: lmax %d outside 1..%d", lmax, kPspMaxL);
    TT_CHECK(!clm || (d_wx && d_wy && d_wz), TTDFT_EBADARG,
             "psp_projectors: C_lm needs the quadrature weights");
    TT_TRY(check_knots(nr, r, "psp_projectors"));
    const size_t seg = 4 * (size_t)(nr - 1);
    const int F = lmax * lmax;
    // device staging: knots, lmax dV_l splines, lmax R_l splines, partial sums
    std::vector<double> host(nr + 2 * lmax * seg);
    std::memcpy(host.data(), r, sizeof(double) * nr);
    std::vector<double> dv(nr);
    for (int l = 0; l < lmax; ++l) {
        for (int i = 0; i < nr; ++i) dv[i] = vl[(size_t)l * nr + i] - vloc[i];
        spline_parabolic(nr, r, dv.data(), host.data() + nr + l * seg);
        spline_parabolic(nr, r, rl + (size_t)l * nr, host.data() + nr + (lmax + l) * seg);
    }
    const size_t part = (size_t)kPspBlocks * F;
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * (host.size() + part)));
    double *d = (double *)ctx->scratch[6].ptr;
    double *d_part = clm ? d + host.size() : nullptr;
    TT_HIP(hipMemcpyAsync(d, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    const double *d_r = d, *d_cdv = d + nr, *d_crl = d + nr + lmax * seg;
#define TT_PSP_LAUNCH(L)                                                                      \
    hipLaunchKernelGGL(psp_projector_kernel<L>, dim3(kPspBlocks), dim3(256), 0, ctx->stream, \
                       nr, d_r, d_cdv, d_crl, px, py, pz, d_gx, d_gy, d_gz, d_wx, d_wy, d_wz,  \
                       d_field, d_phi, d_part)
    switch (lmax) {
        case 1: TT_PSP_LAUNCH(1); break;
        case 2: TT_PSP_LAUNCH(2); break;
        case 3: TT_PSP_LAUNCH(3); break;
        default: TT_PSP_LAUNCH(4); break;
    }
#undef TT_PSP_LAUNCH
    TT_LAUNCH_CHECK();
    if (clm) {
        std::vector<double> h(part);
        TT_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(double) * part, hipMemcpyDeviceToHost,
                              ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
        for (int f = 0; f < F; ++f) {
            double s = 0.0;
            for (int b = 0; b < kPspBlocks; ++b) s += h[(size_t)b * F + f];
            clm[f] = s;
        }
    }
    TT_HIP(hipStreamSynchronize(ctx->stream));  // host staging goes out of scope
    return TTDFT_OK;
}

extern "C" int ttdft_restart_write(ttdft_ctx ctx, const char *path, const void *dX, int64_t ldx,
                                   int dtype, int64_t n, int32_t k) {
    clear_error();
    TT_CHECK(ctx && path && dX && n >= 1 && k >= 1 && ldx >= k &&
                 (dtype == TTDFT_F64 || dtype == TTDFT_F32),
             TTDFT_EBADARG, "restart_write: bad arguments");
    // the reference's column-major FP64 layout (n values of state 0, then state 1, ...)
    std::vector<double> h((size_t)n * k);
    TT_TRY(ttdft_block_download_colmajor(ctx, dX, ldx, dtype, n, k, h.data(), n));
    swap_bytes(h.data(), h.size());
    FILE *f = std::fopen(path, "wb");
    TT_CHECK(f, TTDFT_EBADARG, "restart_write: cannot open %s: %s", path, std::strerror(errno));
    const size_t put = std::fwrite(h.data(), sizeof(double), h.size(), f);
    const int closed = std::fclose(f);
    TT_CHECK(put == h.size() && closed == 0, TTDFT_EBADARG, "restart_write: short write to %s",
             path);
    return TTDFT_OK;
}

extern "C" int ttdft_restart_read(ttdft_ctx ctx, const char *path, void *dX, int64_t ldx,
                                  int dtype, int64_t n, int32_t k) {
    clear_error();
    TT_CHECK(ctx && path && dX && n >= 1 && k >= 1 && ldx >= k &&
                 (dtype == TTDFT_F64 || dtype == TTDFT_F32),
             TTDFT_EBADARG, "restart_read: bad arguments");
    FILE *f = std::fopen(path, "rb");
    TT_CHECK(f, TTDFT_EBADARG, "restart_read: cannot open %s: %s", path, std::strerror(errno));
    std::vector<double> h((size_t)n * k);
    const size_t got = std::fread(h.data(), sizeof(double), h.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    TT_CHECK(got == h.size() && !extra, TTDFT_EBADARG,
             "restart_read: %s holds %s than the %lld values asked for", path,
             extra ? "more" : "fewer", (long long)h.size());
    swap_bytes(h.data(), h.size());
    return ttdft_block_upload_colmajor(ctx, h.data(), n, n, k, dX, ldx, dtype);
}

extern "C" int ttdft_restart_write_bounds(const char *path, const ttdft_chfsi_state *st,
                                          double energy) {
    clear_error();
    TT_CHECK(path && st, TTDFT_EBADARG, "restart_write_bounds: NULL argument");
    FILE *f = std::fopen(path, "w");
    TT_CHECK(f, TTDFT_EBADARG, "restart_write_bounds: cannot open %s: %s", path,
             std::strerror(errno));
    const int put = std::fprintf(f, "%.16e\t%.16e\t%.16e\t%.16e\t%.16e\t%.16e", st->upper,
                                 st->lower_unwanted, st->lower_wanted, st->err_upper,
                                 st->occ_energy, energy);
    const int closed = std::fclose(f);
    TT_CHECK(put > 0 && closed == 0, TTDFT_EBADARG, "restart_write_bounds: write to %s failed", path);
    return TTDFT_OK;
}

extern "C" int ttdft_restart_read_bounds(const char *path, ttdft_chfsi_state *st, double *energy) {
    clear_error();
    TT_CHECK(path && st, TTDFT_EBADARG, "restart_read_bounds: NULL argument");
    std::ifstream in(path);
    TT_CHECK(in.good(), TTDFT_EBADARG, "restart_read_bounds: cannot open %s", path);
    double v[6];
    for (int i = 0; i < 6; ++i)
        TT_CHECK(static_cast<bool>(in >> v[i]), TTDFT_EBADARG,
                 "restart_read_bounds: %s holds %d of the six values", path, i);
    st->upper = v[0];
    st->lower_unwanted = v[1];
    st->lower_wanted = v[2];
    st->err_upper = v[3];
    st->occ_energy = v[4];
    if (energy) *energy = v[5];
    return TTDFT_OK;
}
