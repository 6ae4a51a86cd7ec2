/*
 * CPU FP64 oracle for the ChFSI hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load this
 * code, always as the checker (or as the timed CPU baseline), never as part of
 * the engine.  Each function restates a reference routine; the file:line it
 * follows is given at the function.  Build with -ffp-contract=off so that the
 * Fermi-level Newton iteration reproduces the reference's printed values
 * bit-for-bit.
 */
#include "chfsi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* random numbers: splitmix64 seeding + xoshiro256**                         */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

typedef struct { uint64_t s[4]; } xoshiro;

static void xo_seed(xoshiro *g, uint64_t seed) {
    uint64_t sm = seed;
    for (int i = 0; i < 4; ++i) g->s[i] = splitmix64(&sm);
}

static uint64_t xo_next(xoshiro *g) {
    uint64_t *s = g->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

static double xo_uniform(xoshiro *g) { return (double)(xo_next(g) >> 11) * 0x1.0p-53; }

void orc_rng_uniform(uint64_t seed, int64_t count, double *out) {
    xoshiro g;
    xo_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = xo_uniform(&g);
}

void orc_rng_normal(uint64_t seed, int64_t count, double *out) {
    /* Box-Muller on pairs of uniforms; u1 is shifted into (0,1]. */
    xoshiro g;
    xo_seed(&g, seed);
    for (int64_t i = 0; i < count; i += 2) {
        double u1 = 1.0 - xo_uniform(&g);
        double u2 = xo_uniform(&g);
        double r = sqrt(-2.0 * log(u1));
        out[i] = r * cos(2.0 * M_PI * u2);
        if (i + 1 < count) out[i + 1] = r * sin(2.0 * M_PI * u2);
    }
}

/* ------------------------------------------------------------------------ */
/* operator application                                                      */
/* ------------------------------------------------------------------------ */
void orc_csr_spmm(int64_t n, const int64_t *rowptr, const int32_t *col,
                  const double *val, const double *X, int64_t ldx, int k,
                  double *Y, int64_t ldy) {
    /* reference: AX::perform_ax's intended H_loc X, eigensolver/MultCuda.cu:491-820 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *y = Y + i * ldy;
        for (int j = 0; j < k; ++j) y[j] = 0.0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const double a = val[e];
            const double *x = X + (int64_t)col[e] * ldx;
            for (int j = 0; j < k; ++j) y[j] += a * x[j];
        }
    }
}

void orc_ham_apply(const orc_ham *H, const double *X, int64_t ldx, int k,
                   double *Y, int64_t ldy) {
    /* ChebyShevFilter.cc:102-126: Yp = C_nloc_trans (C_nloc X); Yp += H_loc X.
     * C_nloc carries the 1/C_lm row scale (ProjectHamiltonianSparse.cc:323-325),
     * C_nloc_trans is the unscaled transpose (:316-318). */
    const int64_t n = H->n;
    const int p = H->p;
    double *W = (double *)calloc((size_t)(p > 0 ? p : 1) * (size_t)k, sizeof(double));
    for (int r = 0; r < p; ++r) {
        double *w = W + (int64_t)r * k;
        for (int64_t e = H->c_rowptr[r]; e < H->c_rowptr[r + 1]; ++e) {
            const double a = H->c_val[e] * H->inv_c_lm[r];
            const double *x = X + (int64_t)H->c_col[e] * ldx;
            for (int j = 0; j < k; ++j) w[j] += a * x[j];
        }
    }
    /* nonlocal part: Y_i = sum_r C[r][i] W_r, gathered by scanning C rows */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *y = Y + i * ldy;
        for (int j = 0; j < k; ++j) y[j] = 0.0;
    }
    for (int r = 0; r < p; ++r) {
        const double *w = W + (int64_t)r * k;
        for (int64_t e = H->c_rowptr[r]; e < H->c_rowptr[r + 1]; ++e) {
            const double a = H->c_val[e];
            double *y = Y + (int64_t)H->c_col[e] * ldy;
            for (int j = 0; j < k; ++j) y[j] += a * w[j];
        }
    }
    free(W);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *y = Y + i * ldy;
        double acc[64];
        for (int j0 = 0; j0 < k; j0 += 64) {
            const int jn = (k - j0) < 64 ? (k - j0) : 64;
            for (int j = 0; j < jn; ++j) acc[j] = 0.0;
            for (int64_t e = H->rowptr[i]; e < H->rowptr[i + 1]; ++e) {
                const double a = H->val[e];
                const double *x = X + (int64_t)H->col[e] * ldx + j0;
                for (int j = 0; j < jn; ++j) acc[j] += a * x[j];
            }
            for (int j = 0; j < jn; ++j) y[j0 + j] = y[j0 + j] + acc[j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Chebyshev filter                                                          */
/* ------------------------------------------------------------------------ */
int orc_cheb_filter(const orc_ham *H, double *X, int64_t ldx, int k, int m,
                    double a, double b, double a0, int num_blocks,
                    int convention) {
    /* eigensolver/ChebyShevFilter.cc:34-188 (Zhou-Saad Algorithm 4.3).
     * The reference's Xp aliases the caller's X (:79-85); after its loop Xp
     * holds Y_{m-1} and Yp holds Y_m, and Xp is what is written back
     * (:163-176).  `target` is the index of the iterate returned. */
    if (m < 1 || num_blocks < 1 || num_blocks > k) return -1;
    const int64_t n = H->n;
    const int target = (convention == 0) ? m - 1 : m;
    int col0 = 0;
    for (int blk = 0; blk < num_blocks; ++blk) {
        /* chunk sizes as ChebyShevFilter.cc:57-60 */
        const int cols = k / num_blocks + (blk < k % num_blocks ? 1 : 0);
        double *Xc = X + col0;
        col0 += cols;
        if (target == 0) continue; /* Y_0 = X */
        const double e = (b - a) / 2;
        const double c = (b + a) / 2;
        double sigma = e / (a0 - c);
        const double sigma1 = sigma;
        const double gamma = 2.0 / sigma1;
        double alpha = sigma1 / e;
        double beta = alpha * (-c);
        double *Xp = (double *)malloc((size_t)n * cols * sizeof(double));
        double *Yp = (double *)malloc((size_t)n * cols * sizeof(double));
        double *Yn = (double *)malloc((size_t)n * cols * sizeof(double));
        for (int64_t i = 0; i < n; ++i)
            memcpy(Xp + i * cols, Xc + i * ldx, sizeof(double) * cols);
        /* Y_1 = alpha (H X) + beta X  (:102-126) */
        orc_ham_apply(H, Xp, cols, cols, Yp, cols);
        for (int64_t t = 0; t < n * cols; ++t) {
            Yp[t] = Yp[t] * alpha;
            Yp[t] = Yp[t] + beta * Xp[t];
        }
        /* Y_i = alpha (H Y_{i-1}) + beta Y_{i-1} + delta Y_{i-2}  (:129-170) */
        for (int it = 2; it <= target; ++it) {
            const double sigma2 = 1.0 / (gamma - sigma);
            alpha = 2.0 * sigma2 / e;
            beta = -c * alpha;
            const double delta = -sigma * sigma2;
            orc_ham_apply(H, Yp, cols, cols, Yn, cols);
            for (int64_t t = 0; t < n * cols; ++t) {
                Yn[t] = Yn[t] * alpha;
                Yn[t] = Yn[t] + beta * Yp[t];
                Yn[t] = Yn[t] + delta * Xp[t];
            }
            double *tmp = Xp;
            Xp = Yp;
            Yp = Yn;
            Yn = tmp;
            sigma = sigma2;
        }
        for (int64_t i = 0; i < n; ++i)
            memcpy(Xc + i * ldx, Yp + i * cols, sizeof(double) * cols);
        free(Xp);
        free(Yp);
        free(Yn);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Gram, Cholesky-QR, rotation                                              */
/* ------------------------------------------------------------------------ */
void orc_gram(int64_t n, int k, const double *X, int64_t ldx, const double *Y,
              int64_t ldy, double *S) {
    /* MultTM::mult: S = X^T Y (dgemm T,N), eigensolver/MultTm.cu:177-190.
     * The basis rows go in a fixed number of chunks, each summed into its own
     * k x k partial (64 rows at a time, 8 states of X against a whole row of
     * Y, so the partial rows stay in cache); the partials are added in chunk
     * order, so the result does not depend on the thread count. */
    enum { NCH = 64, RB = 64, AB = 8 };
    const size_t kk = (size_t)k * k;
    double *P = (double *)calloc((size_t)NCH * kk, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < NCH; ++c) {
        double *Pc = P + (size_t)c * kk; /* Pc[a k + b] */
        const int64_t c0 = n * c / NCH, c1 = n * (c + 1) / NCH;
        for (int64_t i0 = c0; i0 < c1; i0 += RB) {
            const int64_t i1 = i0 + RB < c1 ? i0 + RB : c1;
            for (int a0 = 0; a0 < k; a0 += AB) {
                const int na = a0 + AB <= k ? AB : k - a0;
                for (int64_t i = i0; i < i1; ++i) {
                    const double *x = X + i * ldx + a0, *y = Y + i * ldy;
                    for (int u = 0; u < na; ++u) {
                        const double xa = x[u];
                        double *row = Pc + (size_t)(a0 + u) * k;
                        for (int b = 0; b < k; ++b) row[b] += xa * y[b];
                    }
                }
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            double s = 0.0;
            for (int c = 0; c < NCH; ++c) s += P[(size_t)c * kk + (size_t)a * k + b];
            S[a + (int64_t)k * b] = s;
        }
    free(P);
}

int orc_cholesky_lower(int k, double *S) {
    /* potrf(LOWER), eigensolver/MultTm.cu:256-290 */
    for (int j = 0; j < k; ++j) {
        double d = S[j + (int64_t)k * j];
        for (int t = 0; t < j; ++t) d -= S[j + (int64_t)k * t] * S[j + (int64_t)k * t];
        if (!(d > 0.0)) return -(j + 1);
        d = sqrt(d);
        S[j + (int64_t)k * j] = d;
        for (int i = j + 1; i < k; ++i) {
            double s = S[i + (int64_t)k * j];
            for (int t = 0; t < j; ++t) s -= S[i + (int64_t)k * t] * S[j + (int64_t)k * t];
            S[i + (int64_t)k * j] = s / d;
        }
    }
    for (int j = 1; j < k; ++j)
        for (int i = 0; i < j; ++i) S[i + (int64_t)k * j] = 0.0;
    return 0;
}

int orc_orth(int64_t n, int k, double *X, int64_t ldx) {
    /* MultTM::orth: S = X^T X, L = chol(S), X <- X L^{-T} (trsm RIGHT, LOWER,
     * TRANS), eigensolver/MultTm.cu:226-467 */
    double *S = (double *)malloc(sizeof(double) * k * k);
    orc_gram(n, k, X, ldx, X, ldx, S);
    int st = orc_cholesky_lower(k, S);
    if (st != 0) {
        free(S);
        return st;
    }
    /* row-wise forward substitution: solve z L^T = x for each basis row
     * (L copied row-major so the inner sum runs along contiguous memory) */
    double *Lr = (double *)malloc(sizeof(double) * k * k);
    for (int j = 0; j < k; ++j)
        for (int t = 0; t <= j; ++t) Lr[(int64_t)j * k + t] = S[j + (int64_t)k * t];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *x = X + i * ldx;
        for (int j = 0; j < k; ++j) {
            const double *l = Lr + (int64_t)j * k;
            double s = x[j];
            for (int t = 0; t < j; ++t) s -= x[t] * l[t];
            x[j] = s / l[j];
        }
    }
    free(Lr);
    free(S);
    return 0;
}

void orc_rotate(int64_t n, int k, double *X, int64_t ldx, const double *Q) {
    /* main.cc:1185-1207: X <- X Q */
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * k);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double *x = X + i * ldx;
            for (int b = 0; b < k; ++b) {
                double s = 0.0;
                for (int a = 0; a < k; ++a) s += x[a] * Q[a + (int64_t)k * b];
                tmp[b] = s;
            }
            memcpy(x, tmp, sizeof(double) * k);
        }
        free(tmp);
    }
}

/* ------------------------------------------------------------------------ */
/* symmetric eigensolver (cyclic Jacobi)                                    */
/* ------------------------------------------------------------------------ */
int orc_syev(int k, const double *A, double *evals, double *Q) {
    double *M = (double *)malloc(sizeof(double) * k * k);
    /* symmetric matrix from the lower triangle (LAPACK uplo = L) */
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            const int r = i > j ? i : j, c = i > j ? j : i;
            M[i + (int64_t)k * j] = A[r + (int64_t)k * c];
        }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) Q[i + (int64_t)k * j] = (i == j) ? 1.0 : 0.0;
    double frob = 0.0;
    for (int64_t t = 0; t < (int64_t)k * k; ++t) frob += M[t] * M[t];
    const double tol2 = frob * DBL_EPSILON * DBL_EPSILON * 1e-4;
    int sweep = 0;
    for (; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 1; q < k; ++q)
            for (int p = 0; p < q; ++p) off += M[p + (int64_t)k * q] * M[p + (int64_t)k * q];
        if (off <= tol2) break;
        for (int p = 0; p < k - 1; ++p) {
            for (int q = p + 1; q < k; ++q) {
                const double apq = M[p + (int64_t)k * q];
                if (apq == 0.0) continue;
                const double app = M[p + (int64_t)k * p];
                const double aqq = M[q + (int64_t)k * q];
                const double theta = (aqq - app) / (2.0 * apq);
                double t;
                if (fabs(theta) > 1e150)
                    t = 0.5 / theta;
                else
                    t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = t * c;
                double *cp = M + (int64_t)k * p, *cq = M + (int64_t)k * q;
                for (int r = 0; r < k; ++r) {
                    const double xp = cp[r], xq = cq[r];
                    cp[r] = c * xp - s * xq;
                    cq[r] = s * xp + c * xq;
                }
                for (int r = 0; r < k; ++r) {
                    const double xp = M[p + (int64_t)k * r], xq = M[q + (int64_t)k * r];
                    M[p + (int64_t)k * r] = c * xp - s * xq;
                    M[q + (int64_t)k * r] = s * xp + c * xq;
                }
                M[p + (int64_t)k * q] = 0.0;
                M[q + (int64_t)k * p] = 0.0;
                double *vp = Q + (int64_t)k * p, *vq = Q + (int64_t)k * q;
                for (int r = 0; r < k; ++r) {
                    const double xp = vp[r], xq = vq[r];
                    vp[r] = c * xp - s * xq;
                    vq[r] = s * xp + c * xq;
                }
            }
        }
    }
    /* ascending order (EPS_SMALLEST_REAL, EigenSolver.cc:85-86) */
    int *perm = (int *)malloc(sizeof(int) * k);
    for (int i = 0; i < k; ++i) perm[i] = i;
    for (int i = 1; i < k; ++i) { /* insertion sort, stable */
        int v = perm[i];
        int j = i - 1;
        while (j >= 0 && M[perm[j] + (int64_t)k * perm[j]] > M[v + (int64_t)k * v]) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = v;
    }
    double *Qs = (double *)malloc(sizeof(double) * k * k);
    for (int i = 0; i < k; ++i) {
        evals[i] = M[perm[i] + (int64_t)k * perm[i]];
        memcpy(Qs + (int64_t)k * i, Q + (int64_t)k * perm[i], sizeof(double) * k);
    }
    memcpy(Q, Qs, sizeof(double) * k * k);
    free(Qs);
    free(perm);
    free(M);
    return sweep < 100 ? sweep : -1;
}

/* ------------------------------------------------------------------------ */
/* Lanczos bounds                                                            */
/* ------------------------------------------------------------------------ */
static double dotv(int64_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

int orc_lanczos_bounds(const orc_ham *H, int maxit, uint64_t seed,
                       double *upper, double *lower) {
    /* eigensolver/EigenSolver.cc:186-388 */
    const int64_t n = H->n;
    if (maxit < 1) return -1;
    double *v = (double *)malloc(sizeof(double) * n);
    double *v0 = (double *)malloc(sizeof(double) * n);
    double *f = (double *)malloc(sizeof(double) * n);
    double *T = (double *)calloc((size_t)maxit * maxit, sizeof(double));
    orc_rng_uniform(seed, n, v);
    const double nv = sqrt(dotv(n, v, v));
    for (int64_t i = 0; i < n; ++i) v[i] *= 1.0 / nv;
    orc_ham_apply(H, v, 1, 1, f, 1);
    double alpha = dotv(n, f, v);
    for (int64_t i = 0; i < n; ++i) f[i] += -alpha * v[i];
    T[0] = alpha;
    double beta = 0.0;
    for (int it = 1; it < maxit; ++it) {
        beta = sqrt(dotv(n, f, f));
        for (int64_t i = 0; i < n; ++i) f[i] *= 1.0 / beta;
        memcpy(v0, v, sizeof(double) * n);
        memcpy(v, f, sizeof(double) * n);
        orc_ham_apply(H, v, 1, 1, f, 1);
        for (int64_t i = 0; i < n; ++i) f[i] += -beta * v0[i];
        alpha = dotv(n, v, f);
        for (int64_t i = 0; i < n; ++i) f[i] += -alpha * v[i];
        T[it + (int64_t)maxit * (it - 1)] = beta;
        T[(it - 1) + (int64_t)maxit * it] = beta;
        T[it + (int64_t)maxit * it] = alpha;
    }
    beta = sqrt(dotv(n, f, f));
    double *ev = (double *)malloc(sizeof(double) * maxit);
    double *Q = (double *)malloc(sizeof(double) * maxit * maxit);
    int st = orc_syev(maxit, T, ev, Q);
    *upper = beta + fabs(ev[maxit - 1]);
    *lower = ev[0];
    free(ev);
    free(Q);
    free(T);
    free(f);
    free(v0);
    free(v);
    return st < 0 ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Fermi level, occupancies, band energy                                     */
/* ------------------------------------------------------------------------ */
double orc_fermi_energy(const double *evals, int k, int nelec, double kb,
                        double T, int maxit) {
    /* dft/KSDFTEnergyFunctional.cc:76-124, same expression order.  The
     * reference build contracts "s + a*b" into one fused multiply-add (gcc
     * with an FMA-capable -march); the oracle spells those FMAs out so that it
     * reproduces the printed Fermi levels digit for digit. */
    int occNum = nelec / 2;
    if (nelec % 2 == 1) occNum += 1;
    double ef = evals[occNum - 1];
    for (int it = 0; it < maxit; ++it) {
        double func = 0, funcd = 0, sumfi = 0;
        for (int idx = 0; idx < k; ++idx) {
            double arg = (evals[idx] - ef) / (kb * T);
            if (arg <= 0.0) {
                double term = 1.0 / (1.0 + exp(arg));
                sumfi = fma(2.0, term, sumfi);
                funcd = fma(2.0 * (exp(arg) / (kb * T)) * term, term, funcd);
            } else {
                double term = 1.0 / (1.0 + exp(-arg));
                sumfi = fma(2.0 * exp(-arg), term, sumfi);
                funcd = fma(2.0 * (exp(-arg) / (kb * T)) * term, term, funcd);
            }
        }
        func = sumfi - nelec;
        ef = ef - func / funcd;
        if (fabs(func) <= 1.0e-8) break;
    }
    return ef;
}

void orc_occupancy(const double *evals, int k, double ef, double kb, double T,
                   double *occ) {
    /* main.cc:1253-1261 */
    for (int i = 0; i < k; ++i) {
        double term = (evals[i] - ef) / (kb * T);
        if (term <= 0.0)
            occ[i] = 1.0 / (1.0 + exp(term));
        else
            occ[i] = exp(-term) / (1.0 + exp(-term));
    }
}

double orc_band_energy(const double *evals, const double *occ, int k) {
    /* main.cc:1315-1318 */
    double e = 0.0;
    for (int i = 0; i < k; ++i) e = fma(2 * occ[i], evals[i], e); /* contracted as above */
    return e;
}

/* ------------------------------------------------------------------------ */
/* density                                                                   */
/* ------------------------------------------------------------------------ */
void orc_density(const double *X, int64_t ldx, int k, const double *occ,
                 int rx, int ry, int rz, const double *Bx, const double *By,
                 const double *Bz, int px, int py, int pz, double *rho) {
    /* dft/ProjectHamiltonianSparse.cc:1021-1076: per state, ttm over z, x, y
     * then rho += 2 f psi^2 */
    const int64_t np = (int64_t)px * py * pz;
    memset(rho, 0, sizeof(double) * np);
    double *T1 = (double *)malloc(sizeof(double) * (size_t)rx * ry * pz);
    double *T2 = (double *)malloc(sizeof(double) * (size_t)px * ry * pz);
    double *psi = (double *)malloc(sizeof(double) * np);
    for (int s = 0; s < k; ++s) {
        /* T1[a,b,qz] = sum_c C[a,b,c] Bz[qz,c] */
#pragma omp parallel for schedule(static)
        for (int qz = 0; qz < pz; ++qz)
            for (int b = 0; b < ry; ++b)
                for (int a = 0; a < rx; ++a) {
                    double t = 0.0;
                    for (int c = 0; c < rz; ++c)
                        t += X[((int64_t)a + (int64_t)rx * b + (int64_t)rx * ry * c) * ldx + s] *
                             Bz[qz + (int64_t)pz * c];
                    T1[a + (int64_t)rx * b + (int64_t)rx * ry * qz] = t;
                }
        /* T2[qx,b,qz] = sum_a Bx[qx,a] T1[a,b,qz] */
#pragma omp parallel for schedule(static)
        for (int qz = 0; qz < pz; ++qz)
            for (int b = 0; b < ry; ++b)
                for (int qx = 0; qx < px; ++qx) {
                    double t = 0.0;
                    for (int a = 0; a < rx; ++a)
                        t += Bx[qx + (int64_t)px * a] * T1[a + (int64_t)rx * b + (int64_t)rx * ry * qz];
                    T2[qx + (int64_t)px * b + (int64_t)px * ry * qz] = t;
                }
        /* psi[qx,qy,qz] = sum_b By[qy,b] T2[qx,b,qz]; rho += 2 f psi^2 */
        const double w = 2.0 * occ[s];
#pragma omp parallel for schedule(static)
        for (int qz = 0; qz < pz; ++qz)
            for (int qy = 0; qy < py; ++qy)
                for (int qx = 0; qx < px; ++qx) {
                    double t = 0.0;
                    for (int b = 0; b < ry; ++b)
                        t += By[qy + (int64_t)py * b] * T2[qx + (int64_t)px * b + (int64_t)px * ry * qz];
                    rho[qx + (int64_t)px * qy + (int64_t)px * py * qz] += w * t * t;
                }
    }
    free(psi);
    free(T2);
    free(T1);
}

/* ------------------------------------------------------------------------ */
/* clamped cubic spline, initial density                                     */
/* ------------------------------------------------------------------------ */
int orc_spline_build(int n, const double *x, const double *y, double *coef) {
    /* second-derivative form of the clamped spline, end slopes 0 */
    if (n < 2) return -1;
    double *h = (double *)malloc(sizeof(double) * n);
    double *d = (double *)malloc(sizeof(double) * n);  /* diagonal */
    double *rhs = (double *)malloc(sizeof(double) * n);
    double *M = (double *)malloc(sizeof(double) * n);
    double *up = (double *)malloc(sizeof(double) * n); /* super/sub diagonal */
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    d[0] = 2.0 * h[0];
    up[0] = h[0];
    rhs[0] = 6.0 * ((y[1] - y[0]) / h[0] - 0.0);
    for (int i = 1; i < n - 1; ++i) {
        d[i] = 2.0 * (h[i - 1] + h[i]);
        up[i] = h[i];
        rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
    }
    d[n - 1] = 2.0 * h[n - 2];
    rhs[n - 1] = 6.0 * (0.0 - (y[n - 1] - y[n - 2]) / h[n - 2]);
    /* Thomas algorithm; sub-diagonal entry of row i is h[i-1] = up[i-1] */
    for (int i = 1; i < n; ++i) {
        double w = up[i - 1] / d[i - 1];
        d[i] -= w * up[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    M[n - 1] = rhs[n - 1] / d[n - 1];
    for (int i = n - 2; i >= 0; --i) M[i] = (rhs[i] - up[i] * M[i + 1]) / d[i];
    for (int i = 0; i < n - 1; ++i) {
        coef[4 * i + 0] = y[i];
        coef[4 * i + 1] = (y[i + 1] - y[i]) / h[i] - h[i] * (2.0 * M[i] + M[i + 1]) / 6.0;
        coef[4 * i + 2] = 0.5 * M[i];
        coef[4 * i + 3] = (M[i + 1] - M[i]) / (6.0 * h[i]);
    }
    free(up);
    free(M);
    free(rhs);
    free(d);
    free(h);
    return 0;
}

double orc_spline_eval(int n, const double *x, const double *coef, double r) {
    /* interval search over x[0..n-2] with the end segments extended */
    int l = 0, h = n - 1;
    while (l != h - 1) {
        int m = (l + h) / 2;
        if (x[m] >= r)
            h = m;
        else
            l = m;
    }
    const double t = r - x[l];
    const double *c = coef + 4 * l;
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
}

void orc_initial_density(int natoms, const double *xyz, int nr,
                         const double *r, const double *rho_r, int px, int py,
                         int pz, const double *gx, const double *gy,
                         const double *gz, double *rho) {
    /* dft/InitialGuess.cc:647-663 (and :665-680 for quadrature points) */
    double *coef = (double *)malloc(sizeof(double) * 4 * (nr - 1));
    orc_spline_build(nr, r, rho_r, coef);
#pragma omp parallel for schedule(static)
    for (int qz = 0; qz < pz; ++qz)
        for (int qy = 0; qy < py; ++qy)
            for (int qx = 0; qx < px; ++qx) {
                double s = 0.0;
                for (int a = 0; a < natoms; ++a) {
                    const double dx = gx[qx] - xyz[3 * a + 0];
                    const double dy = gy[qy] - xyz[3 * a + 1];
                    const double dz = gz[qz] - xyz[3 * a + 2];
                    const double rr = sqrt(dx * dx + dy * dy + dz * dz);
                    s += orc_spline_eval(nr, r, coef, rr);
                }
                rho[qx + (int64_t)px * qy + (int64_t)px * py * qz] = s;
            }
    free(coef);
}

double orc_grid_integral(const double *f, int px, int py, int pz,
                         const double *wx, const double *wy, const double *wz) {
    double total = 0.0;
    for (int qz = 0; qz < pz; ++qz)
        for (int qy = 0; qy < py; ++qy)
            for (int qx = 0; qx < px; ++qx)
                total += f[qx + (int64_t)px * qy + (int64_t)px * py * qz] * (wx[qx] * wy[qy] * wz[qz]);
    return total;
}

/* ------------------------------------------------------------------------ */
/* ChFSI driver block                                                        */
/* ------------------------------------------------------------------------ */
int orc_chfsi_step(const orc_ham *H, double *X, int64_t ldx, int k, int m,
                   int num_blocks, int restart_times, int occ_index,
                   int scf_iter, int start_iter, int maxit_lanczos,
                   uint64_t seed, int convention, double *evals,
                   orc_chfsi_state *st) {
    const int64_t n = H->n;
    st->passes = 0;
    st->lanczos_runs = 0;
    double temp_lower = 0.0;
    /* bound update, main.cc:987-1003 */
    if (st->err_upper > 0.01) {
        double cur;
        if (orc_lanczos_bounds(H, maxit_lanczos, seed, &cur, &temp_lower) != 0) return -4;
        st->lanczos_runs++;
        st->err_upper = (st->upper - cur) / st->upper;
        if (st->err_upper > 0.01)
            st->upper = cur;
        else
            st->upper = cur * 1.05;
    }
    /* main.cc:1007-1014 */
    if (scf_iter == 0) {
        st->lower_wanted = temp_lower;
        st->lower_unwanted = 0.0;
    } else if (scf_iter != start_iter) {
        st->lower_wanted = evals[0];
        st->lower_unwanted = evals[k - 1];
    }
    double *HX = (double *)malloc(sizeof(double) * (size_t)n * k);
    double *S = (double *)malloc(sizeof(double) * (size_t)k * k);
    double *Q = (double *)malloc(sizeof(double) * (size_t)k * k);
    int have_q = 0;
    int rc = 0;
    for (int r = 0; r < restart_times; ++r) {
        if (orc_cheb_filter(H, X, ldx, k, m, st->lower_unwanted, st->upper, st->lower_wanted,
                            num_blocks, convention) != 0) {
            rc = -1;
            break;
        }
        if (orc_orth(n, k, X, ldx) != 0) {
            rc = -3;
            break;
        }
        orc_ham_apply(H, X, ldx, k, HX, k);
        orc_gram(n, k, X, ldx, HX, k, S);
        if (orc_syev(k, S, evals, Q) < 0) {
            rc = -4;
            break;
        }
        have_q = 1;
        st->passes++;
        st->lower_wanted = evals[0];
        st->lower_unwanted = evals[k - 1];
        const double cur = evals[occ_index];
        const double err = fabs((cur - st->occ_energy) / st->occ_energy);
        st->occ_energy = cur;
        if (err < 1.0e-3) break;
    }
    if (rc == 0 && have_q) orc_rotate(n, k, X, ldx, Q);
    free(Q);
    free(S);
    free(HX);
    return rc;
}
