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

// Pair t of round r of the round-robin (circle) order on m players: player 0
// stays, players 1..m-1 shift by one place per round; position t meets
// position m - 1 - t.  Returned with p < q.
__device__ __forceinline__ void jacobi_pair(int m, int r, int t, int &p, int &q) {
    auto at = [&](int pos) { return pos == 0 ? 0 : ((pos - 1 - r) % (m - 1) + (m - 1)) % (m - 1) + 1; };
    p = at(t);
    q = at(m - 1 - t);
    if (p > q) {
        const int x = p;
        p = q;
        q = x;
    }
}

// Eigenvalues (ascending) and eigenvectors (columns of Q) of the symmetric A
// (lower triangle read) by parallel cyclic Jacobi: each sweep runs m - 1
// rounds (m = k rounded up to even; index k is a bye for odd k) of m/2
// disjoint rotations in the round-robin order; sweeps stop when the
// off-diagonal norm falls below 4 k eps |A| or after 50 (info = 1).
//
// A round is two phases: the rotation of each pair from the current A, then
// one pass that applies all of them, each thread taking whole 2x2 blocks
// (rows of pair j, columns of pair j') of A <- J^T A J (rows first, then
// columns, the same operations as two separate passes) and pairs of columns
// of V <- V J.
constexpr int kJacobiThreads = 1024;

__global__ __launch_bounds__(kJacobiThreads) void small_jacobi_kernel(
    const double *__restrict__ Ain, int k, double *__restrict__ evals, double *__restrict__ Q,
    int *__restrict__ info) {
    __shared__ double A[kMax * kLd];
    __shared__ double V[kMax * kLd];
    __shared__ double cs[kMax / 2], sn[kMax / 2];
    __shared__ int pp[kMax / 2], qq[kMax / 2];
    __shared__ double red[kJacobiThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int nw = kJacobiThreads / 64;
    const int m = (k + 1) & ~1;  // players; index k (odd k) is a bye
    const int np = m / 2;
    for (int e = t; e < k * k; e += kJacobiThreads) {
        const int i = e % k, j = e / k;
        const double v = i >= j ? Ain[i + (int64_t)k * j] : Ain[j + (int64_t)k * i];
        A[i + kLd * j] = v;
        V[i + kLd * j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // sum over the workgroup (every thread gets the total)
    auto block_sum = [&](double v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave] = v;
        __syncthreads();
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < nw; ++w) s += red[w];
        __syncthreads();
        return s;
    };
    double f = 0.0;
    for (int e = t; e < k * k; e += kJacobiThreads) {
        const double v = A[e % k + kLd * (e / k)];
        f += v * v;
    }
    // converged when the off-diagonal norm is within (4 k eps) of |A|: rounding
    // keeps it near k eps |A| once the rotations have converged
    const double eps4k = 4.0 * k * 2.220446049250313e-16;
    const double tol = eps4k * eps4k * block_sum(f);
    int sweep = 0;
    for (; sweep < 50; ++sweep) {
        double off = 0.0;
        for (int e = t; e < k * k; e += kJacobiThreads) {
            const int i = e % k, j = e / k;
            if (i != j) off += A[i + kLd * j] * A[i + kLd * j];
        }
        if (block_sum(off) <= tol) break;  // uniform across the workgroup
        for (int round = 0; round < m - 1; ++round) {
            if (t < np) {
                int p, q;
                jacobi_pair(m, round, t, p, q);
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
            for (int e = t; e < np * np + np * k; e += kJacobiThreads) {
                if (e < np * np) {
                    const int j = e / np, j2 = e % np;
                    const int p = pp[j], q = qq[j], p2 = pp[j2], q2 = qq[j2];
                    const bool hq = q < k, hq2 = q2 < k;
                    const double c = cs[j], s = sn[j], c2 = cs[j2], s2 = sn[j2];
                    // rows p, q of the block (columns p2, q2), row rotation first
                    const double b00 = A[p + kLd * p2];
                    const double b01 = hq2 ? A[p + kLd * q2] : 0.0;
                    const double b10 = hq ? A[q + kLd * p2] : 0.0;
                    const double b11 = hq && hq2 ? A[q + kLd * q2] : 0.0;
                    double r00 = b00, r01 = b01, r10 = b10, r11 = b11;
                    if (hq) {
                        r00 = c * b00 - s * b10;
                        r10 = s * b00 + c * b10;
                        r01 = c * b01 - s * b11;
                        r11 = s * b01 + c * b11;
                    }
                    if (hq2) {
                        A[p + kLd * p2] = c2 * r00 - s2 * r01;
                        A[p + kLd * q2] = s2 * r00 + c2 * r01;
                        if (hq) {
                            A[q + kLd * p2] = c2 * r10 - s2 * r11;
                            A[q + kLd * q2] = s2 * r10 + c2 * r11;
                        }
                    } else {
                        A[p + kLd * p2] = r00;
                        if (hq) A[q + kLd * p2] = r10;
                    }
                } else {
                    const int e2 = e - np * np;
                    const int j = e2 / k, l = e2 % k;
                    const int p = pp[j], q = qq[j];
                    if (q >= k) continue;
                    const double c = cs[j], s = sn[j];
                    const double vp = V[l + kLd * p], vq = V[l + kLd * q];
                    V[l + kLd * p] = c * vp - s * vq;
                    V[l + kLd * q] = s * vp + c * vq;
                }
            }
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

// Orthonormal basis of the columns of Y (n x b column-major, in place, b <= n)
// by classical Gram-Schmidt with one reorthogonalisation pass (CGS2).  A
// column whose residual vanishes against the earlier ones (|residual|^2 at
// most 1e-40 of the column's own |y|^2, exact zeros included; rounding noise
// is far above that and is orthonormalised like any other direction) is
// replaced by the next unit vector e_c and orthogonalised again, so the result
// always has b orthonormal columns.
//
// The block sits in LDS row-major with an odd leading dimension ld (row r's b
// entries contiguous; the lanes of a wave reading one entry of consecutive
// rows, or consecutive entries of one row, meet no bank conflicts).  Columns
// are kept unnormalised until the end, with 1 / |u_i|^2 folded into the
// projection coefficients.  Per column and pass: (1) each wave sums u_i . y
// over its rows, lane i for column i; (2) the eight partial sums are added;
// (3) each thread updates whole rows, y_r -= sum_i c_i u_{r,i} (the second
// pass also forms the partial |y|^2).
constexpr int kOrthMax = 18432;     // n b for the callers' check
constexpr int kOrthLds = 19456;     // doubles of LDS for the block (n ld)
constexpr int kOrthThreads = 512;

__global__ __launch_bounds__(kOrthThreads) void small_orth_kernel(double *__restrict__ Y, int n,
                                                                  int b) {
    __shared__ double Ys[kOrthLds];
    __shared__ double part[kOrthThreads / 64][kMax];
    __shared__ double coef[kMax];
    __shared__ double inv[kMax];
    __shared__ double orig[kMax];
    __shared__ double nrm[kOrthThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int nw = kOrthThreads / 64;
    const int ld = b | 1;
    for (int e = t; e < n * b; e += kOrthThreads) {
        const int r = e % n, c = e / n;
        Ys[r * ld + c] = Y[e];
    }
    __syncthreads();
    // the input columns' |y|^2
    if (lane < b) {
        double s = 0.0;
        for (int r = wave; r < n; r += nw) s += Ys[r * ld + lane] * Ys[r * ld + lane];
        part[wave][lane] = s;
    }
    __syncthreads();
    if (t < b) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < nw; ++w) s += part[w][t];
        orig[t] = s;
    }
    __syncthreads();
    int cand = 0;  // next unit vector for a vanished column (uniform)
    for (int j = 0; j < b; ++j) {
        double ref = orig[j];
        for (int attempt = 0;; ++attempt) {
            for (int pass = 0; pass < (j > 0 ? 2 : 1); ++pass) {
                if (j > 0) {
                    if (lane < j) {  // four independent sums keep 8 LDS reads in flight
                        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                        int r = wave;
                        for (; r + 3 * nw < n; r += 4 * nw) {
                            s0 += Ys[r * ld + lane] * Ys[r * ld + j];
                            s1 += Ys[(r + nw) * ld + lane] * Ys[(r + nw) * ld + j];
                            s2 += Ys[(r + 2 * nw) * ld + lane] * Ys[(r + 2 * nw) * ld + j];
                            s3 += Ys[(r + 3 * nw) * ld + lane] * Ys[(r + 3 * nw) * ld + j];
                        }
                        for (; r < n; r += nw) s0 += Ys[r * ld + lane] * Ys[r * ld + j];
                        part[wave][lane] = (s0 + s1) + (s2 + s3);
                    }
                    __syncthreads();
                    if (t < j) {
                        double s = 0.0;
#pragma unroll
                        for (int w = 0; w < nw; ++w) s += part[w][t];
                        coef[t] = s * inv[t] * inv[t];
                    }
                    __syncthreads();
                }
                double sq = 0.0;
                for (int r = t; r < n; r += kOrthThreads) {
                    const double *row = Ys + r * ld;
                    double v = row[j];
                    if (j > 0) {
                        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
                        int i = 0;
                        for (; i + 3 < j; i += 4) {
                            a0 += coef[i] * row[i];
                            a1 += coef[i + 1] * row[i + 1];
                            a2 += coef[i + 2] * row[i + 2];
                            a3 += coef[i + 3] * row[i + 3];
                        }
                        for (; i < j; ++i) a0 += coef[i] * row[i];
                        v -= (a0 + a1) + (a2 + a3);
                        Ys[r * ld + j] = v;
                    }
                    sq += v * v;
                }
                if (pass == (j > 0 ? 1 : 0)) {
                    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
                    if (lane == 0) nrm[wave] = sq;
                }
                __syncthreads();
            }
            const double nn = ((nrm[0] + nrm[1]) + (nrm[2] + nrm[3])) +
                              ((nrm[4] + nrm[5]) + (nrm[6] + nrm[7]));
            if (nn > 1e-40 * ref || attempt >= n) {
                if (t == 0) inv[j] = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
                break;
            }
            for (int r = t; r < n; r += kOrthThreads) Ys[r * ld + j] = r == cand ? 1.0 : 0.0;
            cand = cand + 1 < n ? cand + 1 : 0;
            ref = 1.0;
            __syncthreads();
        }
    }
    __syncthreads();
    for (int e = t; e < n * b; e += kOrthThreads) {
        const int r = e % n, c = e / n;
        Y[e] = Ys[r * ld + c] * inv[c];
    }
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
    hipLaunchKernelGGL(small_jacobi_kernel, dim3(1), dim3(kJacobiThreads), 0, ctx->stream, A, k,
                       evals, Q, ctx->d_info);
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
    return n >= 1 && b >= 1 && b <= kMax && (int64_t)n * b <= kOrthMax &&
           (int64_t)n * (b | 1) <= kOrthLds;
}

int small_orth(ttdft_ctx ctx, double *Y, int32_t n, int32_t b) {
    TT_CHECK(small_orth_fits(n, b), TTDFT_EBADARG, "small_orth: %d x %d block too large", n, b);
    hipLaunchKernelGGL(small_orth_kernel, dim3(1), dim3(kOrthThreads), 0, ctx->stream, Y, n,
                       b);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

}  // namespace ttdft
