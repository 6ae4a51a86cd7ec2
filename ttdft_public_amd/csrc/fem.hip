// One-dimensional finite elements, the Tucker-basis 1-D matrices and the
// input files of a run (fem/FEM.cc, fem/QuadratureRule.cc,
// dft/ProjectHamiltonianSparse.cc:577-775, basis/BasisFunctionReader.h,
// atoms/AtomInformation.cc).
//
// Elements carry npe equispaced nodes on the reference interval [-1, 1]
// (FEM.cc:418-458) with Lagrange shape functions; neighbouring elements share
// their end nodes, so a mesh of nel elements has nel (npe - 1) + 1 nodes.  A
// quadrature rule (points xi_q and weights w_q on [-1, 1], in the caller's
// order) gives every element nq points: value sum_i N_i(xi_q) u_i, derivative
// sum_i N_i'(xi_q) u_i (with respect to xi, as FEM.cc:927-946), jacobian
// (x_last - x_first) / 2 and the integration weights jac w_q and w_q / jac
// (FEM.cc:541-568).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

#define TT_BLAS(call)                                                                     \
    do {                                                                                  \
        const rocblas_status st_ = (call);                                                \
        TT_CHECK(st_ == rocblas_status_success, TTDFT_ELIB, "%s: rocBLAS status %d", #call, \
                 (int)st_);                                                               \
    } while (0)

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
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65535));
}

// Lagrange shape function i of npe equispaced nodes on [-1, 1] and its
// derivative at t (product form).
void lagrange(int npe, int i, double t, double *val, double *der) {
    std::vector<double> xs(npe);
    for (int k = 0; k < npe; ++k) xs[k] = -1.0 + k * (2.0 / (npe - 1));
    double v = 1.0, dv = 0.0;
    for (int j = 0; j < npe; ++j) {
        if (j == i) continue;
        const double f = (t - xs[j]) / (xs[i] - xs[j]);
        dv = dv * f + v / (xs[i] - xs[j]);  // product rule, term by term
        v *= f;
    }
    *val = v;
    *der = dv;
}

// out[(f nel + e) nq + q] = sum_i S[q npe + i] u[f nn + e (npe - 1) + i]
__global__ __launch_bounds__(256) void elem_interp_kernel(const double *__restrict__ u, int nfun,
                                                          int nel, int npe, int nq,
                                                          const double *__restrict__ S,
                                                          double *__restrict__ out) {
    const int64_t nn = (int64_t)nel * (npe - 1) + 1;
    const int64_t total = (int64_t)nfun * nel * nq;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(t % nq);
        const int64_t fe = t / nq;
        const int e = (int)(fe % nel);
        const int64_t f = fe / nel;
        const double *ue = u + f * nn + (int64_t)e * (npe - 1);
        double s = 0.0;
        for (int i = 0; i < npe; ++i) s += S[q * npe + i] * ue[i];
        out[t] = s;
    }
}

// W[i + n f] = A[i + n f] * w[i]  (rows of a column-major n x m matrix scaled)
__global__ __launch_bounds__(256) void scale_rows_kernel(const double *__restrict__ A,
                                                         const double *__restrict__ w, int64_t n,
                                                         int64_t m, double *__restrict__ W) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * m;
         t += (int64_t)gridDim.x * blockDim.x)
        W[t] = A[t] * w[t % n];
}

// C[q + nq (I r + J)] = B[q + nq I] B[q + nq J]  (pairwise products)
__global__ __launch_bounds__(256) void pair_products_kernel(const double *__restrict__ B,
                                                            int64_t nq, int r,
                                                            double *__restrict__ C) {
    const int64_t total = nq * r * r;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t % nq, ij = t / nq;
        const int I = (int)(ij / r), J = (int)(ij % r);
        C[t] = B[q + nq * I] * B[q + nq * J];
    }
}

// M[b][I][J] = M[b][J][I] for I > J (the reference computes I <= J and mirrors)
__global__ __launch_bounds__(256) void mirror_kernel(double *__restrict__ M, int r, int nb) {
    const int64_t total = (int64_t)nb * r * r;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / ((int64_t)r * r);
        const int ij = (int)(t % ((int64_t)r * r));
        const int I = ij / r, J = ij % r;
        if (I > J) M[t] = M[b * r * r + (int64_t)J * r + I];
    }
}

// whitespace-separated numbers of each non-empty line
int read_rows(const char *path, std::vector<std::vector<double>> &rows) {
    TT_CHECK(path, TTDFT_EBADARG, "NULL path");
    std::ifstream fin(path);
    TT_CHECK(fin.good(), TTDFT_EBADARG, "cannot open %s", path);
    std::string line;
    while (std::getline(fin, line)) {
        std::istringstream ss(line);
        std::vector<double> v;
        double d;
        while (ss >> d) v.push_back(d);
        TT_CHECK(ss.eof(), TTDFT_EBADARG, "%s: not a number in line %zu", path, rows.size() + 1);
        if (!v.empty()) rows.push_back(v);
    }
    return TTDFT_OK;
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" {

int ttdft_fem_gauss_legendre(int32_t n, double *xi, double *w) {
    clear_error();
    TT_CHECK(n >= 1 && n <= 256 && xi && w, TTDFT_EBADARG, "gauss_legendre: bad arguments");
    // roots of P_n by Newton's method from cos(pi (i + 3/4) / (n + 1/2)), the
    // three-term recurrence for P_n and P_n' = n (x P_n - P_{n-1}) / (x^2 - 1);
    // w = 2 / ((1 - x^2) P_n'(x)^2).  Points ascending.
    for (int i = 0; i < n; ++i) {
        double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double pn = n == 1 ? x : p1, pm = n == 1 ? 1.0 : p0;
            dp = n * (x * pn - pm) / ((x - 1.0) * (x + 1.0));
            const double dx = pn / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double pn = n == 1 ? x : p1, pm = n == 1 ? 1.0 : p0;
        dp = n * (x * pn - pm) / ((x - 1.0) * (x + 1.0));
        xi[n - 1 - i] = x;
        w[n - 1 - i] = 2.0 / (((1.0 - x) * (1.0 + x)) * dp * dp);
    }
    return TTDFT_OK;
}

int ttdft_fem_quadrature_rule(int32_t n, double *xi, double *w) {
    clear_error();
    TT_CHECK(n >= 2 && n <= 20 && xi && w, TTDFT_EBADARG,
             "fem_quadrature_rule: the rules have 2 to 20 points (QuadratureRule.cc:26-560)");
    double x[20], v[20];
    TT_TRY(ttdft_fem_gauss_legendre(n, x, v));
    // Order of the points (indices into the ascending rule) as the rule tables
    // list them: ascending up to 8 points; the 9-point table's own order; from
    // 10 to 15 points outward from the centre, the negative point of each pair
    // first; from 16 points descending.
    static const int nine[9] = {4, 1, 7, 0, 8, 3, 5, 2, 6};
    int order[20];
    if (n <= 8) {
        for (int i = 0; i < n; ++i) order[i] = i;
    } else if (n == 9) {
        for (int i = 0; i < 9; ++i) order[i] = nine[i];
    } else if (n <= 15) {
        int o = 0;
        if (n & 1) order[o++] = n / 2;
        for (int k = 1; o < n; ++k) {
            const int lo = (n - 1) / 2 - k + ((n & 1) ? 0 : 1), hi = n - 1 - lo;
            order[o++] = lo;
            order[o++] = hi;
        }
    } else {
        for (int i = 0; i < n; ++i) order[i] = n - 1 - i;
    }
    for (int i = 0; i < n; ++i) {
        xi[i] = x[order[i]];
        w[i] = v[order[i]];
    }
    return TTDFT_OK;
}

namespace {
// evenly spaced points a, a + h, a + 2h, ... by repeated addition
// (FEM.cc:86-96), the last one dropped; written into out[0 .. npe - 2]
void element_nodes(double a, double b, int npe, double *out) {
    const double h = (b - a) / (npe - 1.0);
    double x = a;
    for (int j = 0; j < npe - 1; ++j) {
        out[j] = x;
        x += h;
    }
}
}  // namespace

int ttdft_fem_mesh(int32_t type, int32_t nel, int32_t npe, double start, double end,
                   double inner_size, double inner_h, double coarsening, double *nodes) {
    clear_error();
    TT_CHECK(nel >= 1 && npe >= 2 && nodes, TTDFT_EBADARG, "fem_mesh: bad arguments");
    const int64_t nn = (int64_t)nel * (npe - 1) + 1;  // FEM.cc:159-161
    if (type == TTDFT_MESH_UNIFORM) {                  // FEM.cc:168-175
        const double interval = std::fabs(end - start) / (nn - 1);
        for (int64_t i = 0; i < nn; ++i) nodes[i] = start + interval * i;
        return TTDFT_OK;
    }
    TT_CHECK(type == TTDFT_MESH_ADAPTIVE || type == TTDFT_MESH_MANUAL1 ||
                 type == TTDFT_MESH_HARTREE_OUTER,
             TTDFT_EBADARG, "fem_mesh: unknown mesh type %d", type);
    TT_CHECK(inner_h > 0.0 && inner_size > 0.0, TTDFT_EBADARG,
             "fem_mesh: inner domain size and mesh size must be positive");
    // element end points: an evenly spaced inner domain about the centre and
    // geometrically growing elements on both sides (FEM.cc:176-218, 255-317,
    // 319-368)
    const int inner = (int)std::floor(inner_size / inner_h);
    TT_CHECK(inner <= nel && ((nel - inner) % 2) == 0, TTDFT_EBADARG,
             "fem_mesh: %d elements leave %d outer elements, which must be even and not "
             "negative (an odd count leaves the last element end unset)",
             nel, nel - inner);
    const int outer = (nel - inner) / 2;
    const double centre = (start + end) / 2;
    const double inner_start = centre - inner_size / 2.0;
    std::vector<double> lin(nel + 1, 0.0);
    for (int c = 0; c <= inner; ++c) lin[outer + c] = inner_start + inner_h * c;
    double coef = coarsening, h = inner_h;
    if (type == TTDFT_MESH_MANUAL1) {
        coef = 1.5;
        h = 1.5;
    } else if (type == TTDFT_MESH_HARTREE_OUTER) {
        double ratio = 0.0;
        for (int i = 1; i <= outer; ++i) ratio += std::pow(coef, i);
        h = ((end - start) - inner_size) / 2.0 / ratio;
    }
    for (int i = 1; i <= outer; ++i) {
        h *= coef;
        lin[outer - i] = lin[outer - i + 1] - h;
        lin[outer + inner + i] = lin[outer + inner + i - 1] + h;
    }
    for (int e = 0; e < nel; ++e) element_nodes(lin[e], lin[e + 1], npe, nodes + (int64_t)e * (npe - 1));
    nodes[nn - 1] = type == TTDFT_MESH_ADAPTIVE ? end : lin[nel];
    return TTDFT_OK;
}

int ttdft_fem_interpolate(ttdft_ctx ctx, int32_t nel, int32_t npe, int32_t nq, const double *xi,
                          const double *w, const double *nodes, int32_t nfun,
                          const double *d_nodal, double *d_val, double *d_dval, double *jw,
                          double *inv_jw, double *xq) {
    clear_error();
    TT_CHECK(ctx && nel >= 1 && npe >= 2 && nq >= 1 && xi && w && nodes, TTDFT_EBADARG,
             "fem_interpolate: bad arguments");
    TT_CHECK(nfun == 0 || (d_nodal && d_val), TTDFT_EBADARG, "fem_interpolate: NULL field");
    std::vector<double> S((size_t)nq * npe), D((size_t)nq * npe);
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < npe; ++i) lagrange(npe, i, xi[q], &S[q * npe + i], &D[q * npe + i]);
    for (int e = 0; e < nel; ++e) {
        const double *xe = nodes + (int64_t)e * (npe - 1);
        const double jac = (xe[npe - 1] - xe[0]) / 2.0;
        for (int q = 0; q < nq; ++q) {
            const int64_t t = (int64_t)e * nq + q;
            double x = 0.0;
            for (int i = 0; i < npe; ++i) x += S[q * npe + i] * xe[i];
            if (xq) xq[t] = x;
            if (jw) jw[t] = jac * w[q];
            if (inv_jw) inv_jw[t] = w[q] / jac;
        }
    }
    if (nfun == 0) return TTDFT_OK;
    TT_HIP(hipSetDevice(ctx->device));
    Dev dS, dD;
    TT_TRY(dS.alloc(sizeof(double) * S.size()));
    TT_TRY(dD.alloc(sizeof(double) * D.size()));
    TT_HIP(hipMemcpyAsync(dS.p, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    TT_HIP(hipMemcpyAsync(dD.p, D.data(), sizeof(double) * D.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    const int64_t total = (int64_t)nfun * nel * nq;
    hipLaunchKernelGGL(elem_interp_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream, d_nodal,
                       nfun, nel, npe, nq, dS.d(), d_val);
    TT_LAUNCH_CHECK();
    if (d_dval) {
        hipLaunchKernelGGL(elem_interp_kernel, dim3(nblk(total)), dim3(256), 0, ctx->stream,
                           d_nodal, nfun, nel, npe, nq, dD.d(), d_dval);
        TT_LAUNCH_CHECK();
    }
    TT_HIP(hipStreamSynchronize(ctx->stream));  // S, D are freed at scope end
    return TTDFT_OK;
}

int ttdft_tucker_basis_matrices(ttdft_ctx ctx, int32_t nq, int32_t r, const double *d_B,
                                const double *d_DB, const double *d_jw, const double *d_inv_jw,
                                int32_t rv, const double *d_U, double *d_MatX, double *d_MatDX,
                                double *d_MatPot) {
    clear_error();
    TT_CHECK(ctx && nq >= 1 && r >= 1 && d_B && d_jw, TTDFT_EBADARG,
             "tucker_basis_matrices: bad arguments");
    TT_CHECK(!d_MatDX || (d_DB && d_inv_jw), TTDFT_EBADARG,
             "tucker_basis_matrices: derivatives or weights missing");
    TT_CHECK(!d_MatPot || (rv >= 1 && d_U), TTDFT_EBADARG,
             "tucker_basis_matrices: Veff factor missing");
    TT_CHECK((int64_t)nq * r * r < ((int64_t)1 << 31), TTDFT_EBADARG,
             "tucker_basis_matrices: nq r^2 too large");
    TT_HIP(hipSetDevice(ctx->device));
    const double one = 1.0, zero = 0.0;
    // B is row-major r x nq (one basis function per row) = column-major nq x r
    Dev W;
    TT_TRY(W.alloc(sizeof(double) * (size_t)nq * std::max(r, rv)));
    if (d_MatX) {  // MatX[I][J] = sum_q B_I B_J jw  (ProjectHamiltonianSparse.cc:595-612)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk((int64_t)nq * r)), dim3(256), 0,
                           ctx->stream, d_B, d_jw, (int64_t)nq, (int64_t)r, W.d());
        TT_LAUNCH_CHECK();
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, r,
                              r, nq, &one, W.d(), nq, d_B, nq, &zero, d_MatX, r));
        hipLaunchKernelGGL(mirror_kernel, dim3(nblk((int64_t)r * r)), dim3(256), 0, ctx->stream,
                           d_MatX, r, 1);
        TT_LAUNCH_CHECK();
    }
    if (d_MatDX) {  // MatDX[I][J] = sum_q B'_I B'_J w / jac  (:606-612)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk((int64_t)nq * r)), dim3(256), 0,
                           ctx->stream, d_DB, d_inv_jw, (int64_t)nq, (int64_t)r, W.d());
        TT_LAUNCH_CHECK();
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, r,
                              r, nq, &one, W.d(), nq, d_DB, nq, &zero, d_MatDX, r));
        hipLaunchKernelGGL(mirror_kernel, dim3(nblk((int64_t)r * r)), dim3(256), 0, ctx->stream,
                           d_MatDX, r, 1);
        TT_LAUNCH_CHECK();
    }
    if (d_MatPot) {  // MatPot[p][I][J] = sum_q U_p B_I B_J jw  (:678-775)
        Dev C;
        TT_TRY(C.alloc(sizeof(double) * (size_t)nq * r * r));
        hipLaunchKernelGGL(pair_products_kernel, dim3(nblk((int64_t)nq * r * r)), dim3(256), 0,
                           ctx->stream, d_B, (int64_t)nq, r, C.d());
        TT_LAUNCH_CHECK();
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk((int64_t)nq * rv)), dim3(256), 0,
                           ctx->stream, d_U, d_jw, (int64_t)nq, (int64_t)rv, W.d());
        TT_LAUNCH_CHECK();
        TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
                              r * r, rv, nq, &one, C.d(), nq, W.d(), nq, &zero, d_MatPot,
                              r * r));
        hipLaunchKernelGGL(mirror_kernel, dim3(nblk((int64_t)rv * r * r)), dim3(256), 0,
                           ctx->stream, d_MatPot, r, rv);
        TT_LAUNCH_CHECK();
        TT_HIP(hipStreamSynchronize(ctx->stream));  // C is freed at scope end
    }
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

int ttdft_read_basis_file(const char *path, int32_t *nfun, int32_t *nnodes, double *values) {
    clear_error();
    TT_CHECK(nfun && nnodes, TTDFT_EBADARG, "read_basis_file: NULL size outputs");
    std::vector<std::vector<double>> rows;
    TT_TRY(read_rows(path, rows));
    TT_CHECK(!rows.empty(), TTDFT_EBADARG, "%s: no basis functions", path);
    const size_t m = rows[0].size();
    for (auto &r : rows)
        TT_CHECK(r.size() == m, TTDFT_EBADARG, "%s: rows of different length", path);
    *nfun = (int32_t)rows.size();
    *nnodes = (int32_t)(m + 2);  // zero at both ends (BasisFunctionReader.h:44,48)
    if (!values) return TTDFT_OK;
    for (size_t f = 0; f < rows.size(); ++f) {
        double *v = values + f * (m + 2);
        v[0] = 0.0;
        std::copy(rows[f].begin(), rows[f].end(), v + 1);
        v[m + 1] = 0.0;
    }
    return TTDFT_OK;
}

int ttdft_read_table(const char *path, int32_t *nrows, int32_t *ncols, double *values) {
    clear_error();
    TT_CHECK(nrows && ncols, TTDFT_EBADARG, "read_table: NULL size outputs");
    std::vector<std::vector<double>> rows;
    TT_TRY(read_rows(path, rows));
    TT_CHECK(!rows.empty(), TTDFT_EBADARG, "%s: empty table", path);
    const size_t m = rows[0].size();
    for (auto &r : rows)
        TT_CHECK(r.size() == m, TTDFT_EBADARG, "%s: rows of different length", path);
    *nrows = (int32_t)rows.size();
    *ncols = (int32_t)m;
    if (values)
        for (size_t i = 0; i < rows.size(); ++i) std::copy(rows[i].begin(), rows[i].end(), values + i * m);
    return TTDFT_OK;
}

int ttdft_read_atom_file(const char *path, int32_t *ntype, int32_t *natoms, int32_t *nelec,
                         int32_t *lmax, double *atoms) {
    clear_error();
    TT_CHECK(path && ntype && natoms && nelec, TTDFT_EBADARG, "read_atom_file: bad arguments");
    std::ifstream fin(path);
    TT_CHECK(fin.good(), TTDFT_EBADARG, "cannot open %s", path);
    std::string line;
    TT_CHECK((bool)std::getline(fin, line), TTDFT_EBADARG, "%s: empty", path);
    {
        std::istringstream ss(line);  // "ntype natoms nelectrons" (AtomInformation.cc:33-36)
        TT_CHECK((bool)(ss >> *ntype >> *natoms >> *nelec), TTDFT_EBADARG,
                 "%s: bad first line", path);
    }
    TT_CHECK(*ntype >= 1 && *natoms >= 0, TTDFT_EBADARG, "%s: bad counts", path);
    TT_CHECK((bool)std::getline(fin, line), TTDFT_EBADARG, "%s: no lmax line", path);
    {
        std::istringstream ss(line);  // lmax per atom type (:37-43)
        for (int t = 0; t < *ntype; ++t) {
            int32_t l = 0;
            TT_CHECK((bool)(ss >> l), TTDFT_EBADARG, "%s: bad lmax line", path);
            if (lmax) lmax[t] = l;
        }
    }
    std::vector<std::array<double, 7>> rows;
    while (std::getline(fin, line)) {  // "Z, x, y, z, ., Zatomic, type" (:47-59)
        std::replace(line.begin(), line.end(), ',', ' ');
        std::istringstream ss(line);
        std::array<double, 7> a;
        int k = 0;
        while (k < 7 && (ss >> a[k])) ++k;
        if (k == 0) continue;
        TT_CHECK(k == 7, TTDFT_EBADARG, "%s: atom line %zu has %d fields", path, rows.size() + 1,
                 k);
        TT_CHECK(a[6] >= 0 && a[6] < *ntype, TTDFT_EBADARG, "%s: atom type %g", path, a[6]);
        rows.push_back(a);
    }
    TT_CHECK((int64_t)rows.size() == *natoms, TTDFT_EBADARG,
             "%s: %zu atom lines, header says %d (:60-63)", path, rows.size(), *natoms);
    if (atoms) {  // grouped by type, file order within a type (:79-87)
        int64_t o = 0;
        for (int t = 0; t < *ntype; ++t)
            for (auto &a : rows)
                if ((int)a[6] == t) std::copy(a.begin(), a.end(), atoms + 7 * o++);
    }
    return TTDFT_OK;
}

}  // extern "C"
