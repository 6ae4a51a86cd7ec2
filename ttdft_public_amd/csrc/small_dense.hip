// Dense k x k factorisations of the subspace iteration for small blocks
// (k <= 64), each in one workgroup with the matrix held in LDS: the
// Cholesky factor's inverse for Cholesky-QR (MultTm.cu:256-419, the
// SOLVE_WITH_INVERSE variant) and the symmetric eigensolve of Rayleigh-Ritz
// (EigenSolver.cc:30-184).  At these sizes the library routines spend their
// time in launch latency (hundreds of small kernels per call); one launch
// does the whole factorisation here.
//
// Matrices are column-major k x k, FP64.
#include <cmath>

#include "internal.h"

namespace ttdft {
namespace {

constexpr int kMax = 64;
constexpr int kLd = kMax + 1;

// QT = L^{-T} with S = L L^T (lower Cholesky of the symmetric S; only its
// lower triangle is read); QT's strictly lower part is zero.  info = j + 1
// when the leading minor of order j + 1 is not positive definite.
__global__ __launch_bounds__(256) void small_chol_inv_kernel(const double *__restrict__ S, int k,
                                                             double *__restrict__ QT,
                                                             int *__restrict__ info) {
    __shared__ double A[kMax * kLd];
    __shared__ double Li[kMax * kLd];
    __shared__ int bad;
    const int t = threadIdx.x;
    for (int e = t; e < k * k; e += 256) {
        const int i = e % k, j = e / k;
        A[i + kLd * j] = i >= j ? S[i + (int64_t)k * j] : 0.0;
    }
    if (t == 0) bad = 0;
    __syncthreads();
    // right-looking Cholesky, column j per step
    for (int j = 0; j < k; ++j) {
        const double d = A[j + kLd * j];
        if (!(d > 0.0)) {  // uniform across the workgroup
            if (t == 0) bad = j + 1;
            break;
        }
        const double ljj = sqrt(d);
        __syncthreads();
        if (t == 0) A[j + kLd * j] = ljj;
        for (int i = j + 1 + t; i < k; i += 256) A[i + kLd * j] /= ljj;
        __syncthreads();
        // trailing update of the lower triangle: A[i][l] -= L[i][j] L[l][j], j < l <= i
        const int m = k - j - 1;
        for (int e = t; e < m * m; e += 256) {
            const int i = j + 1 + e % m, l = j + 1 + e / m;
            if (l <= i) A[i + kLd * l] -= A[i + kLd * j] * A[l + kLd * j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) {
        if (t == 0) *info = bad;
        return;
    }
    // Linv by forward substitution, one column per thread: L Linv[:, c] = e_c
    if (t < k) {
        const int c = t;
        for (int i = 0; i < k; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int l = c; l < i; ++l) s -= A[i + kLd * l] * Li[l + kLd * c];
            Li[i + kLd * c] = i < c ? 0.0 : s / A[i + kLd * i];
        }
    }
    __syncthreads();
    // QT[i][j] = Linv[j][i]
    for (int e = t; e < k * k; e += 256) {
        const int i = e % k, j = e / k;
        QT[i + (int64_t)k * j] = Li[j + kLd * i];
    }
    if (t == 0) *info = 0;
}

// Eigenvalues (ascending) and eigenvectors (columns of Q) of the symmetric A
// (lower triangle read) by parallel cyclic Jacobi: each sweep runs k - 1
// rounds of k/2 disjoint rotations in the round-robin order; sweeps stop
// when the off-diagonal norm falls below 4 k eps |A| or after 50.
// info = 1 if that did not happen.
__global__ __launch_bounds__(256) void small_jacobi_kernel(const double *__restrict__ Ain, int k,
                                                           double *__restrict__ evals,
                                                           double *__restrict__ Q,
                                                           int *__restrict__ info) {
    __shared__ double A[kMax * kLd];
    __shared__ double V[kMax * kLd];
    __shared__ double cs[kMax / 2], sn[kMax / 2];
    __shared__ int pp[kMax / 2], qq[kMax / 2], order[kMax];
    __shared__ double red[256];
    __shared__ int done;
    const int t = threadIdx.x;
    const int m = (k + 1) & ~1;  // players; index k (odd k) is a bye
    for (int e = t; e < k * k; e += 256) {
        const int i = e % k, j = e / k;
        const double v = i >= j ? Ain[i + (int64_t)k * j] : Ain[j + (int64_t)k * i];
        A[i + kLd * j] = v;
        V[i + kLd * j] = i == j ? 1.0 : 0.0;
    }
    if (t < m) order[t] = t;
    __syncthreads();
    // |A|^2 for the stopping test
    double f = 0.0;
    for (int e = t; e < k * k; e += 256) {
        const double v = A[e % k + kLd * (e / k)];
        f += v * v;
    }
    red[t] = f;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    // converged when the off-diagonal norm is within (4 k eps) of |A|: rounding
    // keeps it near k eps |A| once the rotations have converged
    const double tol = (4.0 * k * 2.220446049250313e-16) * (4.0 * k * 2.220446049250313e-16) * red[0];
    __syncthreads();
    int sweep = 0;
    for (; sweep < 50; ++sweep) {
        // off-diagonal mass
        double off = 0.0;
        for (int e = t; e < k * k; e += 256) {
            const int i = e % k, j = e / k;
            if (i != j) off += A[i + kLd * j] * A[i + kLd * j];
        }
        red[t] = off;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) red[t] += red[t + h];
            __syncthreads();
        }
        if (t == 0) done = red[0] <= tol;
        __syncthreads();
        if (done) break;
        for (int round = 0; round < m - 1; ++round) {
            if (t < m / 2) {
                int p = order[t], q = order[m - 1 - t];
                if (p > q) {
                    const int x = p;
                    p = q;
                    q = x;
                }
                double c = 1.0, s = 0.0;
                if (q < k) {
                    const double apq = A[p + kLd * q];
                    if (apq != 0.0) {
                        const double tau = (A[q + kLd * q] - A[p + kLd * p]) / (2.0 * apq);
                        const double tt = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                                     : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + tt * tt);
                        s = tt * c;
                    }
                }
                pp[t] = p;
                qq[t] = q;
                cs[t] = c;
                sn[t] = s;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = t; e < (m / 2) * k; e += 256) {
                const int j = e / k, l = e % k;
                const int p = pp[j], q = qq[j];
                if (q >= k) continue;
                const double c = cs[j], s = sn[j];
                const double ap = A[p + kLd * l], aq = A[q + kLd * l];
                A[p + kLd * l] = c * ap - s * aq;
                A[q + kLd * l] = s * ap + c * aq;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J
            for (int e = t; e < (m / 2) * k; e += 256) {
                const int j = e / k, l = e % k;
                const int p = pp[j], q = qq[j];
                if (q >= k) continue;
                const double c = cs[j], s = sn[j];
                const double ap = A[l + kLd * p], aq = A[l + kLd * q];
                A[l + kLd * p] = c * ap - s * aq;
                A[l + kLd * q] = s * ap + c * aq;
                const double vp = V[l + kLd * p], vq = V[l + kLd * q];
                V[l + kLd * p] = c * vp - s * vq;
                V[l + kLd * q] = s * vp + c * vq;
            }
            __syncthreads();
            // next round: keep order[0], rotate the rest by one
            int nxt = 0;
            if (t < m) nxt = t == 0 ? order[0] : order[t == 1 ? m - 1 : t - 1];
            __syncthreads();
            if (t < m) order[t] = nxt;
            __syncthreads();
        }
    }
    // ascending order (ties by index)
    if (t < k) {
        const double w = A[t + kLd * t];
        int rank = 0;
        for (int j = 0; j < k; ++j) {
            const double u = A[j + kLd * j];
            rank += (u < w) || (u == w && j < t);
        }
        evals[rank] = w;
        for (int i = 0; i < k; ++i) Q[i + (int64_t)k * rank] = V[i + kLd * t];
    }
    if (t == 0) *info = sweep >= 50 ? 1 : 0;
}

// Orthonormal basis of the columns of Y (n x b column-major, in place) by
// classical Gram-Schmidt with one reorthogonalisation pass (CGS2), the block
// held in LDS; a column that vanishes against the earlier ones is left 0.
constexpr int kOrthMax = 18432;  // doubles of LDS for the block (n b)

__global__ __launch_bounds__(256) void small_orth_kernel(double *__restrict__ Y, int n, int b) {
    __shared__ double Ys[kOrthMax];
    __shared__ double dots[kMax];
    __shared__ double red[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int e = t; e < n * b; e += 256) Ys[e] = Y[e];
    __syncthreads();
    for (int j = 0; j < b; ++j) {
        double *y = Ys + (int64_t)n * j;
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = wave; i < j; i += 4) {  // dots[i] = q_i . y_j, one wave per i
                const double *q = Ys + (int64_t)n * i;
                double s = 0.0;
                for (int r = lane; r < n; r += 64) s += q[r] * y[r];
                for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
                if (lane == 0) dots[i] = s;
            }
            __syncthreads();
            for (int r = t; r < n; r += 256) {
                double v = y[r];
                for (int i = 0; i < j; ++i) v -= dots[i] * Ys[(int64_t)n * i + r];
                y[r] = v;
            }
            __syncthreads();
        }
        double s = 0.0;
        for (int r = t; r < n; r += 256) s += y[r] * y[r];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        const double nrm = sqrt((red[0] + red[1]) + (red[2] + red[3]));
        const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        for (int r = t; r < n; r += 256) y[r] *= inv;
        __syncthreads();
    }
    for (int e = t; e < n * b; e += 256) Y[e] = Ys[e];
}

}  // namespace

bool small_dense_enabled(int32_t k) {
    static const int off = [] {
        const char *e = getenv("TTDFT_SMALL_DENSE");
        return e && e[0] == '0';
    }();
    return !off && k >= 1 && k <= kMax;
}

int small_chol_inv(ttdft_ctx ctx, const double *S, int32_t k, double *QT) {
    hipLaunchKernelGGL(small_chol_inv_kernel, dim3(1), dim3(256), 0, ctx->stream, S, k, QT,
                       ctx->d_info);
    TT_LAUNCH_CHECK();
    int info = 0;
    TT_HIP(hipMemcpyAsync(&info, ctx->d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(info == 0, TTDFT_ENOTSPD, "overlap matrix not positive definite (minor %d)", info);
    return TTDFT_OK;
}

int small_syev(ttdft_ctx ctx, const double *A, int32_t k, double *evals, double *Q) {
    hipLaunchKernelGGL(small_jacobi_kernel, dim3(1), dim3(256), 0, ctx->stream, A, k, evals, Q,
                       ctx->d_info);
    TT_LAUNCH_CHECK();
    int info = 0;
    TT_HIP(hipMemcpyAsync(&info, ctx->d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(info == 0, TTDFT_ENOCONV, "Jacobi eigensolver did not converge in 50 sweeps");
    return TTDFT_OK;
}

}  // namespace ttdft

namespace ttdft {

bool small_orth_fits(int32_t n, int32_t b) {
    return n >= 1 && b >= 1 && b <= kMax && (int64_t)n * b <= kOrthMax;
}

int small_orth(ttdft_ctx ctx, double *Y, int32_t n, int32_t b) {
    TT_CHECK(small_orth_fits(n, b), TTDFT_EBADARG, "small_orth: %d x %d block too large", n, b);
    hipLaunchKernelGGL(small_orth_kernel, dim3(1), dim3(256), 0, ctx->stream, Y, n, b);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // namespace ttdft
