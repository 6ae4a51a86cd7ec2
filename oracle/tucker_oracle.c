/*
 * CPU FP64 oracle for the Tucker recompression of the density and the
 * Hartree potential by kernel expansion (test infrastructure only; see
 * chfsi_oracle.h for the rule that only tests, the smoke check and the
 * benchmark's CPU baseline may load it).
 *
 * Tensors are stored first index fastest: T[i0 + n0 (i1 + n1 i2)].  Factor
 * matrices are column-major: U[i + n r].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chfsi_oracle.h"

/* Leading r eigenvectors (descending eigenvalues) of the n x n Gram matrix
 * S, each with its largest-magnitude entry made positive (first such entry
 * on ties).  U is n x r column-major; evals gets the r eigenvalues. */
static int leading_eigvecs(int n, const double *S, int r, double *U, double *evals) {
    double *w = malloc(sizeof(double) * n);
    double *Q = malloc(sizeof(double) * (size_t)n * n);
    const int sweeps = orc_syev(n, S, w, Q);
    for (int j = 0; j < r; ++j) {
        const int src = n - 1 - j;  /* ascending order from orc_syev */
        const double *q = Q + (size_t)n * src;
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (fabs(q[i]) > fabs(q[imax])) imax = i;
        const double s = q[imax] < 0.0 ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) U[i + (size_t)n * j] = s * q[i];
        if (evals) evals[j] = w[src];
    }
    free(Q);
    free(w);
    return sweeps < 0 ? -1 : 0;
}

/* Sequentially truncated HOSVD with fixed ranks, modes 0, 1, 2 in order: the
 * algorithm of TuckerMPI::STHOSVD (the third-party TuckerMPI library that
 * tensor/TensorUtils.cc:245-336 calls; Vannieuwenhoven, Vandebril and
 * Meerbergen, SIAM J. Sci. Comput. 34 (2012)).  For each mode n: the Gram
 * matrix of the mode-n unfolding of the current tensor, its leading r_n
 * eigenvectors U_n, then the tensor is replaced by its product with U_n^T in
 * mode n.  The core is the final tensor.  Sign convention: the largest
 * |entry| of each column of U_n is positive.  evals (r0 + r1 + r2, may be
 * NULL) receives the kept Gram eigenvalues, mode by mode. */
int orc_sthosvd(int n0, int n1, int n2, int r0, int r1, int r2, const double *T, double *core,
                double *U0, double *U1, double *U2, double *evals) {
    const int n[3] = {n0, n1, n2}, r[3] = {r0, r1, r2};
    double *U[3] = {U0, U1, U2};
    int dims[3] = {n0, n1, n2};
    size_t cur_size = (size_t)n0 * n1 * n2;
    double *Y = malloc(sizeof(double) * cur_size);
    memcpy(Y, T, sizeof(double) * cur_size);
    int eoff = 0, status = 0;
    for (int mode = 0; mode < 3; ++mode) {
        const int nm = n[mode];
        /* strides of the current tensor */
        const size_t st[3] = {1, (size_t)dims[0], (size_t)dims[0] * dims[1]};
        double *S = calloc((size_t)nm * nm, sizeof(double));
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                for (int a = 0; a < dims[(mode + 1) % 3]; ++a)
                    for (int b = 0; b < dims[(mode + 2) % 3]; ++b) {
                        const size_t off = a * st[(mode + 1) % 3] + b * st[(mode + 2) % 3];
                        s += Y[off + i * st[mode]] * Y[off + j * st[mode]];
                    }
                S[i + (size_t)nm * j] = s;
                S[j + (size_t)nm * i] = s;
            }
        if (leading_eigvecs(nm, S, r[mode], U[mode], evals ? evals + eoff : NULL) < 0)
            status = -1;
        free(S);
        eoff += r[mode];
        /* Y <- Y x_mode U^T */
        int nd[3] = {dims[0], dims[1], dims[2]};
        nd[mode] = r[mode];
        const size_t nsize = (size_t)nd[0] * nd[1] * nd[2];
        double *Z = calloc(nsize, sizeof(double));
        const size_t sn[3] = {1, (size_t)nd[0], (size_t)nd[0] * nd[1]};
        for (int i0 = 0; i0 < nd[0]; ++i0)
            for (int i1 = 0; i1 < nd[1]; ++i1)
                for (int i2 = 0; i2 < nd[2]; ++i2) {
                    int idx[3] = {i0, i1, i2};
                    const int k = idx[mode];
                    double s = 0.0;
                    for (int q = 0; q < nm; ++q) {
                        idx[mode] = q;
                        s += U[mode][q + (size_t)nm * k] *
                             Y[idx[0] * st[0] + idx[1] * st[1] + idx[2] * st[2]];
                    }
                    Z[i0 * sn[0] + i1 * sn[1] + i2 * sn[2]] = s;
                }
        free(Y);
        Y = Z;
        dims[mode] = r[mode];
        cur_size = nsize;
    }
    memcpy(core, Y, sizeof(double) * cur_size);
    free(Y);
    return status;
}

/* dft/KSDFTPotential.cc:269-340, with computeExpansionIntegral (:440-599) and
 * computeHatreePotFromExpansionIntegralNew (:601-668):
 *   F_d,i[q][a] = sum_e exp(-alpha_i (xq_d[q] - xe_d[e])^2) U_d[e][a] jw_d[e]
 *                 (FEM::integrate_by_quad_values = ddot with jacobian x weight)
 *   V += sum_i omega_i  G x_0 F_0,i x_1 F_1,i x_2 F_2,i            (daxpy per term)
 * alpha / omega are used as given (the reference scales them by 1/A^2 and
 * 1/A after reading them, :281-284).  V is p0 x p1 x p2, first index fastest;
 * add = 0 overwrites it (INSERT_VALUES), 1 accumulates. */
void orc_hartree_expansion(int r0, int r1, int r2, const double *G, const double *const U[3],
                           const int ne[3], const double *const xe[3], const double *const jw[3],
                           const int p[3], const double *const xq[3], int nterms,
                           const double *alpha, const double *omega, double *V, int add) {
    const int r[3] = {r0, r1, r2};
    const size_t nv = (size_t)p[0] * p[1] * p[2];
    if (!add) memset(V, 0, sizeof(double) * nv);
    double *F[3];
    for (int d = 0; d < 3; ++d) F[d] = malloc(sizeof(double) * (size_t)p[d] * r[d]);
    double *t1 = malloc(sizeof(double) * (size_t)p[0] * r1 * r2);
    double *t2 = malloc(sizeof(double) * (size_t)p[0] * p[1] * r2);
    for (int it = 0; it < nterms; ++it) {
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < r[d]; ++a)
                for (int q = 0; q < p[d]; ++q) {
                    double s = 0.0;
                    for (int e = 0; e < ne[d]; ++e) {
                        const double dx = xq[d][q] - xe[d][e];
                        s += exp(-alpha[it] * dx * dx) * U[d][e + (size_t)ne[d] * a] * jw[d][e];
                    }
                    F[d][q + (size_t)p[d] * a] = s;
                }
        /* mode 0: t1[x + p0 (b + r1 c)] = sum_a F0[x][a] G[a + r0 (b + r1 c)] */
        for (int bc = 0; bc < r1 * r2; ++bc)
            for (int x = 0; x < p[0]; ++x) {
                double s = 0.0;
                for (int a = 0; a < r0; ++a) s += F[0][x + (size_t)p[0] * a] * G[a + (size_t)r0 * bc];
                t1[x + (size_t)p[0] * bc] = s;
            }
        /* mode 1: t2[x + p0 (y + p1 c)] = sum_b F1[y][b] t1[x + p0 (b + r1 c)] */
        for (int c = 0; c < r2; ++c)
            for (int y = 0; y < p[1]; ++y)
                for (int x = 0; x < p[0]; ++x) {
                    double s = 0.0;
                    for (int b = 0; b < r1; ++b)
                        s += F[1][y + (size_t)p[1] * b] * t1[x + (size_t)p[0] * (b + (size_t)r1 * c)];
                    t2[x + (size_t)p[0] * (y + (size_t)p[1] * c)] = s;
                }
        /* mode 2 and the daxpy: V[xy + p0 p1 z] += omega_i sum_c F2[z][c] t2[xy + p0 p1 c] */
        const size_t pxy = (size_t)p[0] * p[1];
        for (int z = 0; z < p[2]; ++z)
            for (size_t xy = 0; xy < pxy; ++xy) {
                double s = 0.0;
                for (int c = 0; c < r2; ++c) s += F[2][z + (size_t)p[2] * c] * t2[xy + pxy * c];
                V[xy + pxy * z] += omega[it] * s;
            }
    }
    free(t2);
    free(t1);
    for (int d = 0; d < 3; ++d) free(F[d]);
}
