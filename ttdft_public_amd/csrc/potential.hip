// Pieces of the effective potential and grid utilities of the SCF loop
// (dft/KSDFTPotential.cc:127-267, main.cc:882-960):
//   - LDA exchange (Slater) and correlation (Perdew-Zunger 1981 fit of
//     Ceperley-Alder) potentials and energy densities per particle;
//   - a local pseudopotential -Z erf(r / rc) / r summed over the atoms;
//   - the 3-D quadrature weight of a tensor grid, weighted dot products;
//   - trilinear resampling from a uniform grid and Gaussian splats of atom
//     positions on one (the learned initial density works on a uniform grid).
// Tensor grids are x fastest: f[qx + px (qy + py qz)].
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace ttdft {

int grid_reduce(ttdft_ctx ctx, const double *f, const double *g, int32_t px, int32_t py,
                int32_t pz, const double *wx, const double *wy, const double *wz, int squared,
                double *result);
int dev_axpby(ttdft_ctx ctx, double *y, double a, const double *x, double b, int64_t n);

namespace {

struct Dev {
    void *p = nullptr;
    ~Dev() {
        if (p) {
            hipError_t e = hipFree(p);
            (void)e;
        }
    }
    int alloc(size_t bytes) {
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return TTDFT_ENOMEM;
        }
        return TTDFT_OK;
    }
    double *d() const { return static_cast<double *>(p); }
};

unsigned nblk(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// Perdew-Zunger 1981 parameters of the Ceperley-Alder correlation energy
// (Phys. Rev. B 23, 5048, table XII, unpolarised)
constexpr double kGamma = -0.1423, kBeta1 = 1.0529, kBeta2 = 0.3334;
constexpr double kA = 0.0311, kB = -0.048, kC = 0.002, kD = -0.0116;

__global__ __launch_bounds__(256) void lda_xc_kernel(const double *__restrict__ rho, int64_t n,
                                                     double *__restrict__ vxc,
                                                     double *__restrict__ exc, int add,
                                                     double lin_coef) {
    const double cx = -0.75 * cbrt(3.0 / M_PI);  // e_x = cx rho^(1/3), v_x = 4/3 e_x
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double r = rho[i];
        double v = 0.0, e = 0.0;
        if (r > 1e-30) {
            const double c13 = cbrt(r);
            const double ex = cx * c13;
            const double rs = cbrt(3.0 / (4.0 * M_PI * r));
            double vc, ec;
            if (rs >= 1.0) {
                const double sq = sqrt(rs);
                const double den = 1.0 + kBeta1 * sq + kBeta2 * rs;
                ec = kGamma / den;
                vc = kGamma * (1.0 + (7.0 / 6.0) * kBeta1 * sq + (4.0 / 3.0) * kBeta2 * rs) /
                     (den * den);
            } else {
                const double lr = log(rs);
                ec = kA * lr + kB + kC * rs * lr + kD * rs;
                vc = kA * lr + (kB - kA / 3.0) + (2.0 / 3.0) * kC * rs * lr +
                     lin_coef * (2.0 * kD - kC) * rs;
            }
            v = (4.0 / 3.0) * ex + vc;
            e = ex + ec;
        }
        if (vxc) vxc[i] = add ? vxc[i] + v : v;
        if (exc) exc[i] = e;
    }
}

__global__ __launch_bounds__(256) void local_psp_kernel(int32_t natoms,
                                                        const double *__restrict__ atoms,
                                                        double rc, int32_t px, int32_t py,
                                                        int32_t pz, const double *__restrict__ gx,
                                                        const double *__restrict__ gy,
                                                        const double *__restrict__ gz,
                                                        double *__restrict__ V, int add) {
    const int64_t total = (int64_t)px * py * pz;
    const double inv_rc = 1.0 / rc, at0 = 2.0 / (sqrt(M_PI) * rc);
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        const double x = gx[qx], y = gy[rest % py], z = gz[rest / py];
        double s = 0.0;
        for (int a = 0; a < natoms; ++a) {
            const double *at = atoms + 4 * a;
            const double dx = x - at[0], dy = y - at[1], dz = z - at[2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            // erf(r/rc)/r -> 2/(sqrt(pi) rc) as r -> 0
            s -= at[3] * (r > 1e-10 ? erf(r * inv_rc) / r : at0);
        }
        V[idx] = add ? V[idx] + s : s;
    }
}

__global__ __launch_bounds__(256) void grid_weight_kernel(int32_t px, int32_t py, int32_t pz,
                                                          const double *__restrict__ wx,
                                                          const double *__restrict__ wy,
                                                          const double *__restrict__ wz,
                                                          double *__restrict__ w) {
    const int64_t total = (int64_t)px * py * pz;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        w[idx] = wx[qx] * wy[rest % py] * wz[rest / py];
    }
}

struct Uniform {
    int32_t n[3];
    double o[3], h[3];
};

// trilinear interpolation of a uniform-grid field (x fastest); points
// outside [o, o + (n - 1) h] on any axis get 0
__global__ __launch_bounds__(256) void resample_kernel(const double *__restrict__ src, Uniform u,
                                                       int32_t px, int32_t py, int32_t pz,
                                                       const double *__restrict__ gx,
                                                       const double *__restrict__ gy,
                                                       const double *__restrict__ gz,
                                                       double *__restrict__ dst, int add) {
    const int64_t total = (int64_t)px * py * pz;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int qx = (int)(idx % px);
        const int64_t rest = idx / px;
        const double p[3] = {gx[qx], gy[rest % py], gz[rest / py]};
        int i0[3];
        double t[3];
        bool inside = true;
        for (int d = 0; d < 3; ++d) {
            const double s = (p[d] - u.o[d]) / u.h[d];
            if (!(s >= 0.0 && s <= (double)(u.n[d] - 1))) inside = false;
            int i = (int)floor(s);
            i = min(max(i, 0), max(u.n[d] - 2, 0));
            i0[d] = i;
            t[d] = u.n[d] > 1 ? s - i : 0.0;
        }
        double v = 0.0;
        if (inside) {
            for (int c = 0; c < 8; ++c) {
                const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                const int ix = min(i0[0] + dx, u.n[0] - 1), iy = min(i0[1] + dy, u.n[1] - 1),
                          iz = min(i0[2] + dz, u.n[2] - 1);
                const double wgt = (dx ? t[0] : 1.0 - t[0]) * (dy ? t[1] : 1.0 - t[1]) *
                                   (dz ? t[2] : 1.0 - t[2]);
                v += wgt * src[ix + (int64_t)u.n[0] * (iy + (int64_t)u.n[1] * iz)];
            }
        }
        dst[idx] = add ? dst[idx] + v : v;
    }
}

__global__ __launch_bounds__(256) void splat_kernel(int32_t natoms,
                                                    const double *__restrict__ atoms,
                                                    double inv2s2, Uniform u,
                                                    double *__restrict__ out) {
    const int64_t total = (int64_t)u.n[0] * u.n[1] * u.n[2];
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int ix = (int)(idx % u.n[0]);
        const int64_t rest = idx / u.n[0];
        const double x = u.o[0] + u.h[0] * ix, y = u.o[1] + u.h[1] * (rest % u.n[1]),
                     z = u.o[2] + u.h[2] * (rest / u.n[1]);
        double s = 0.0;
        for (int a = 0; a < natoms; ++a) {
            const double *at = atoms + 4 * a;
            const double dx = x - at[0], dy = y - at[1], dz = z - at[2];
            s += at[3] * exp(-(dx * dx + dy * dy + dz * dz) * inv2s2);
        }
        out[idx] = s;
    }
}

__global__ __launch_bounds__(256) void clamp_kernel(double *__restrict__ f, int64_t n,
                                                    double lo) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        f[i] = fmax(f[i], lo);
}

int upload_atoms(ttdft_ctx ctx, int32_t natoms, const double *atoms, Dev &d) {
    TT_TRY(d.alloc(sizeof(double) * 4 * (size_t)std::max(natoms, 1)));
    if (natoms > 0)
        TT_HIP(hipMemcpyAsync(d.p, atoms, sizeof(double) * 4 * natoms, hipMemcpyHostToDevice,
                              ctx->stream));
    return TTDFT_OK;
}

int check_uniform(const int32_t n[3], const double o[3], const double h[3], Uniform &u) {
    TT_CHECK(n && o && h, TTDFT_EBADARG, "uniform grid: NULL shape");
    for (int d = 0; d < 3; ++d) {
        TT_CHECK(n[d] >= 1 && h[d] > 0.0, TTDFT_EBADARG, "uniform grid: bad axis %d", d);
        u.n[d] = n[d];
        u.o[d] = o[d];
        u.h[d] = h[d];
    }
    return TTDFT_OK;
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" {

int ttdft_lda_xc(ttdft_ctx ctx, const double *d_rho, int64_t n, double *d_vxc, double *d_exc,
                 int add, int variant) {
    clear_error();
    TT_CHECK(ctx && d_rho && n >= 0 && (d_vxc || d_exc), TTDFT_EBADARG, "lda_xc: bad arguments");
    TT_CHECK(variant == TTDFT_LDA_PZ81 || variant == TTDFT_LDA_PZ81_REFERENCE, TTDFT_EBADARG,
             "lda_xc: unknown variant %d", variant);
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    const double lin = variant == TTDFT_LDA_PZ81 ? 1.0 / 3.0 : 1.0 / 33.0;
    hipLaunchKernelGGL(lda_xc_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_rho, n, d_vxc,
                       d_exc, add, lin);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_local_pseudopotential(ttdft_ctx ctx, int32_t natoms, const double *atoms, double rc,
                                int32_t px, int32_t py, int32_t pz, const double *d_gx,
                                const double *d_gy, const double *d_gz, double *d_V, int add) {
    clear_error();
    TT_CHECK(ctx && natoms >= 0 && (natoms == 0 || atoms) && rc > 0.0 && px > 0 && py > 0 &&
                 pz > 0 && d_gx && d_gy && d_gz && d_V,
             TTDFT_EBADARG, "local_pseudopotential: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    Dev da;
    TT_TRY(upload_atoms(ctx, natoms, atoms, da));
    const int64_t total = (int64_t)px * py * pz;
    hipLaunchKernelGGL(local_psp_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream, natoms,
                       da.d(), rc, px, py, pz, d_gx, d_gy, d_gz, d_V, add);
    TT_LAUNCH_CHECK();
    TT_HIP(hipStreamSynchronize(ctx->stream));  // da is freed at scope end
    return TTDFT_OK;
}

int ttdft_grid_weights(ttdft_ctx ctx, int32_t px, int32_t py, int32_t pz, const double *d_wx,
                       const double *d_wy, const double *d_wz, double *d_w) {
    clear_error();
    TT_CHECK(ctx && px > 0 && py > 0 && pz > 0 && d_wx && d_wy && d_wz && d_w, TTDFT_EBADARG,
             "grid_weights: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    const int64_t total = (int64_t)px * py * pz;
    hipLaunchKernelGGL(grid_weight_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream, px, py,
                       pz, d_wx, d_wy, d_wz, d_w);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_grid_dot(ttdft_ctx ctx, const double *d_f, const double *d_g, int32_t px, int32_t py,
                   int32_t pz, const double *d_wx, const double *d_wy, const double *d_wz,
                   double *result) {
    clear_error();
    TT_CHECK(ctx && d_f && d_g && d_wx && d_wy && d_wz && result && px > 0 && py > 0 && pz > 0,
             TTDFT_EBADARG, "grid_dot: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    return grid_reduce(ctx, d_f, d_g, px, py, pz, d_wx, d_wy, d_wz, 2, result);
}

int ttdft_grid_resample(ttdft_ctx ctx, const double *d_src, const int32_t n[3],
                        const double origin[3], const double h[3], int32_t px, int32_t py,
                        int32_t pz, const double *d_gx, const double *d_gy, const double *d_gz,
                        double *d_dst, int add) {
    clear_error();
    Uniform u;
    TT_TRY(check_uniform(n, origin, h, u));
    TT_CHECK(ctx && d_src && d_dst && d_gx && d_gy && d_gz && px > 0 && py > 0 && pz > 0,
             TTDFT_EBADARG, "grid_resample: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    const int64_t total = (int64_t)px * py * pz;
    hipLaunchKernelGGL(resample_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream, d_src, u,
                       px, py, pz, d_gx, d_gy, d_gz, d_dst, add);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ttdft_gaussian_splat(ttdft_ctx ctx, int32_t natoms, const double *atoms, double sigma,
                         const int32_t n[3], const double origin[3], const double h[3],
                         double *d_out) {
    clear_error();
    Uniform u;
    TT_TRY(check_uniform(n, origin, h, u));
    TT_CHECK(ctx && natoms >= 0 && (natoms == 0 || atoms) && sigma > 0.0 && d_out,
             TTDFT_EBADARG, "gaussian_splat: bad arguments");
    TT_HIP(hipSetDevice(ctx->device));
    Dev da;
    TT_TRY(upload_atoms(ctx, natoms, atoms, da));
    const int64_t total = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(splat_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream, natoms,
                       da.d(), 0.5 / (sigma * sigma), u, d_out);
    TT_LAUNCH_CHECK();
    TT_HIP(hipStreamSynchronize(ctx->stream));  // da is freed at scope end
    return TTDFT_OK;
}

int ttdft_grid_axpby(ttdft_ctx ctx, double *d_y, double a, const double *d_x, double b,
                     int64_t n) {
    clear_error();
    TT_CHECK(ctx && d_y && n >= 0 && (d_x || b == 0.0), TTDFT_EBADARG, "grid_axpby: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    return dev_axpby(ctx, d_y, a, d_x, b, n);
}

int ttdft_grid_clamp(ttdft_ctx ctx, double *d_f, int64_t n, double lo) {
    clear_error();
    TT_CHECK(ctx && d_f && n >= 0, TTDFT_EBADARG, "grid_clamp: bad arguments");
    if (n == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(clamp_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, d_f, n, lo);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // extern "C"
