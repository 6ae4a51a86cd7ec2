/*
 * CPU FP64 oracle for Anderson density mixing (test infrastructure only; see
 * chfsi_oracle.h for the rule that only tests, the smoke check and the
 * benchmark's CPU baseline may load it).
 *
 * Restates dft/AndersonMixing.cc: the mixing constants of
 * computeMixingConstants (:150-262) and the mixed density of computeRhoIn
 * (:264-323).  History entries are stored oldest first, h of them, each a
 * field of n values: in[j * n + i], out[j * n + i].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chfsi_oracle.h"

/* Solve A x = b (A n x n column-major, overwritten; b overwritten by x) by LU
 * with partial pivoting, the factorisation LAPACK's dgesv performs.  Returns
 * 0, or i + 1 when the pivot of column i is exactly zero. */
static int lu_solve(int n, double *A, double *b) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r + (size_t)n * c]) > fabs(A[piv + (size_t)n * c])) piv = r;
        if (A[piv + (size_t)n * c] == 0.0) return c + 1;
        if (piv != c) {
            for (int q = 0; q < n; ++q) {
                const double t = A[c + (size_t)n * q];
                A[c + (size_t)n * q] = A[piv + (size_t)n * q];
                A[piv + (size_t)n * q] = t;
            }
            const double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
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

/* AndersonMixing.cc:150-262.  F_j = out_j - in_j; with n the newest entry and
 * d_m = F_n - F_{h-2-m} (m = 0 .. h-2):
 *   coeff[m + k (h-1)] = sum_i d_k[i] d_m[i] jacob[i] weight[i]     (:188-210)
 *   rhs[m]             = sum_i d_m[i] F_n[i] jacob[i] weight[i]     (:211-218)
 * then coeff c = rhs (dgesv, :238-250).  Returns the dgesv-style info. */
int orc_anderson_constants(int h, int64_t n, const double *in, const double *out,
                           const double *jacob, const double *weight, double *constants) {
    const int m1 = h - 1;
    if (m1 <= 0) return 0;
    double *coeff = calloc((size_t)m1 * m1, sizeof(double));
    double *rhs = calloc((size_t)m1, sizeof(double));
    const double *inn = in + (size_t)(h - 1) * n, *outn = out + (size_t)(h - 1) * n;
    for (int m = 0; m < m1; ++m) {
        const double *inm = in + (size_t)(h - m - 2) * n, *outm = out + (size_t)(h - m - 2) * n;
        for (int k = 0; k < m1; ++k) {
            const double *ink = in + (size_t)(h - k - 2) * n;
            const double *outk = out + (size_t)(h - k - 2) * n;
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                const double fn = outn[i] - inn[i];
                const double dnm = fn - (outm[i] - inm[i]);
                const double dnk = fn - (outk[i] - ink[i]);
                s += dnk * dnm * jacob[i] * weight[i];
            }
            coeff[m + (size_t)k * m1] = s;
        }
        double r = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double fn = outn[i] - inn[i];
            const double dnm = fn - (outm[i] - inm[i]);
            r += dnm * fn * jacob[i] * weight[i];
        }
        rhs[m] = r;
    }
    const int info = lu_solve(m1, coeff, rhs);
    memcpy(constants, rhs, sizeof(double) * m1);
    free(rhs);
    free(coeff);
    return info;
}

/* AndersonMixing.cc:264-323: with n the newest entry,
 *   in_mix  = sum_k c_k (in_{h-2-k} - in_n) + in_n,
 *   out_mix = sum_k c_k (out_{h-2-k} - out_n) + out_n,
 *   rho     = |alpha out_mix + (1 - alpha) in_mix|. */
void orc_anderson_mix(int h, int64_t n, const double *in, const double *out,
                      const double *constants, double alpha, double *rho) {
    const double *inn = in + (size_t)(h - 1) * n, *outn = out + (size_t)(h - 1) * n;
    for (int64_t i = 0; i < n; ++i) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < h - 1; ++k) {
            a += constants[k] * (in[(size_t)(h - k - 2) * n + i] - inn[i]);
            b += constants[k] * (out[(size_t)(h - k - 2) * n + i] - outn[i]);
        }
        a += inn[i];
        b += outn[i];
        rho[i] = fabs(alpha * b + (1.0 - alpha) * a);
    }
}
