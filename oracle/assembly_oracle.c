/*
 * CPU FP64 oracle for the Hamiltonian assembly (test infrastructure only; see
 * chfsi_oracle.h for the conventions and the rule that only tests, the smoke
 * check and the benchmark's CPU baseline may load it).
 *
 * Restates the reference's per-row construction of H_loc and of the projector
 * matrix C: a dense row ("dyad") over all n = rx ry rz basis functions,
 * x fastest, then the entries with |v| > truncation in column order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chfsi_oracle.h"

/* dft/ProjectHamiltonianSparse.cc:330-575 (Create_Hloc) and :1223-1592
 * (Compute_nnz_loc, the same rows counted):
 *   dyad[x,y,z] = 0.5 (DX[xI][x] Y[yI][y] Z[zI][z] + X[xI][x] DY[yI][y] Z[zI][z]
 *                      + X[xI][x] Y[yI][y] DZ[zI][z])                    (:477-490)
 *   dyad += G x_0 PXrow x_1 PYrow x_2 PZrow, PXrow[x][p] = PX[p][xI][x]   (:491-540)
 *   keep |dyad| > truncation                                            (:542-551) */
int64_t orc_assemble_hloc_rows(int rx, int ry, int rz, int rvx, int rvy, int rvz,
                               const double *X, const double *DX, const double *Y,
                               const double *DY, const double *Z, const double *DZ,
                               const double *PX, const double *PY, const double *PZ,
                               const double *G, double truncation, int64_t row0,
                               int64_t row1, int64_t *rowptr, int32_t *col, double *val) {
    const int64_t n = (int64_t)rx * ry * rz;
    double *dyad = malloc(sizeof(double) * n);
    double *t1 = malloc(sizeof(double) * (size_t)rx * rvy * rvz);  /* [r][q][x] */
    double *t2 = malloc(sizeof(double) * (size_t)rx * ry * rvz);   /* [r][y][x] */
    int64_t nnz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int64_t row = row0; row < row1; ++row) {
        const int xI = (int)(row % rx), yI = (int)((row / rx) % ry);
        const int zI = (int)(row / ((int64_t)rx * ry));
        const double *lx = X + (size_t)xI * rx, *ldx = DX + (size_t)xI * rx;
        const double *ly = Y + (size_t)yI * ry, *ldy = DY + (size_t)yI * ry;
        const double *lz = Z + (size_t)zI * rz, *ldz = DZ + (size_t)zI * rz;
        int64_t c = 0;
        for (int z = 0; z < rz; ++z)
            for (int y = 0; y < ry; ++y)
                for (int x = 0; x < rx; ++x) {
                    double d = ldx[x] * ly[y] * lz[z] + lx[x] * ldy[y] * lz[z] +
                               lx[x] * ly[y] * ldz[z];
                    dyad[c++] = 0.5 * d;
                }
        /* mode 0: t1[r][q][x] = sum_p PX[p][xI][x] G[p,q,r] */
        for (int r = 0; r < rvz; ++r)
            for (int q = 0; q < rvy; ++q)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int p = 0; p < rvx; ++p)
                        s += PX[((size_t)p * rx + xI) * rx + x] *
                             G[p + (size_t)rvx * (q + (size_t)rvy * r)];
                    t1[((size_t)r * rvy + q) * rx + x] = s;
                }
        /* mode 1: t2[r][y][x] = sum_q PY[q][yI][y] t1[r][q][x] */
        for (int r = 0; r < rvz; ++r)
            for (int y = 0; y < ry; ++y)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int q = 0; q < rvy; ++q)
                        s += PY[((size_t)q * ry + yI) * ry + y] *
                             t1[((size_t)r * rvy + q) * rx + x];
                    t2[((size_t)r * ry + y) * rx + x] = s;
                }
        /* mode 2: dyad[z][y][x] += sum_r PZ[r][zI][z] t2[r][y][x] */
        for (int z = 0; z < rz; ++z)
            for (int y = 0; y < ry; ++y)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int r = 0; r < rvz; ++r)
                        s += PZ[((size_t)r * rz + zI) * rz + z] *
                             t2[((size_t)r * ry + y) * rx + x];
                    dyad[x + (int64_t)rx * (y + (int64_t)ry * z)] += s;
                }
        for (int64_t j = 0; j < n; ++j)
            if (fabs(dyad[j]) > truncation) {
                if (col) {
                    col[nnz] = (int32_t)j;
                    val[nnz] = dyad[j];
                }
                ++nnz;
            }
        if (rowptr) rowptr[row - row0 + 1] = nnz;
    }
    free(t2);
    free(t1);
    free(dyad);
    return nnz;
}

int64_t orc_assemble_hloc(int rx, int ry, int rz, int rvx, int rvy, int rvz,
                          const double *X, const double *DX, const double *Y,
                          const double *DY, const double *Z, const double *DZ,
                          const double *PX, const double *PY, const double *PZ,
                          const double *G, double truncation, int64_t *rowptr,
                          int32_t *col, double *val) {
    return orc_assemble_hloc_rows(rx, ry, rz, rvx, rvy, rvz, X, DX, Y, DY, Z, DZ, PX, PY, PZ, G,
                                  truncation, 0, (int64_t)rx * ry * rz, rowptr, col, val);
}

/* dft/ProjectHamiltonianSparse.cc:196-278 (Create_Cnloc): projector row i is
 * the Tucker reconstruction S_i x_0 MX_i x_1 MY_i x_2 MZ_i, core S_i of size
 * r_i^3 (mode 0 fastest), factors column-major (MX_i[x + rx a], rx x r_i, as
 * Tucker::Matrix), entries |v| > truncation (:260-266).  Cores and factors of
 * consecutive rows are concatenated. */
int64_t orc_assemble_projectors(int p, const int32_t *ranks, int rx, int ry, int rz,
                                const double *cores, const double *mx, const double *my,
                                const double *mz, double truncation, int64_t *rowptr,
                                int32_t *col, double *val) {
    const int64_t n = (int64_t)rx * ry * rz;
    double *dense = malloc(sizeof(double) * n);
    int64_t nnz = 0;
    size_t oc = 0, ox = 0, oy = 0, oz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int i = 0; i < p; ++i) {
        const int r = ranks[i];
        const double *S = cores + oc, *A = mx + ox, *B = my + oy, *C = mz + oz;
        double *u = malloc(sizeof(double) * (size_t)rx * r * r);  /* [c][b][x] */
        double *w = malloc(sizeof(double) * (size_t)rx * ry * r); /* [c][y][x] */
        for (int c = 0; c < r; ++c)
            for (int b = 0; b < r; ++b)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int a = 0; a < r; ++a)
                        s += A[x + (size_t)rx * a] * S[a + (size_t)r * (b + (size_t)r * c)];
                    u[((size_t)c * r + b) * rx + x] = s;
                }
        for (int c = 0; c < r; ++c)
            for (int y = 0; y < ry; ++y)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int b = 0; b < r; ++b)
                        s += B[y + (size_t)ry * b] * u[((size_t)c * r + b) * rx + x];
                    w[((size_t)c * ry + y) * rx + x] = s;
                }
        for (int z = 0; z < rz; ++z)
            for (int y = 0; y < ry; ++y)
                for (int x = 0; x < rx; ++x) {
                    double s = 0.0;
                    for (int c = 0; c < r; ++c)
                        s += C[z + (size_t)rz * c] * w[((size_t)c * ry + y) * rx + x];
                    dense[x + (int64_t)rx * (y + (int64_t)ry * z)] = s;
                }
        for (int64_t j = 0; j < n; ++j)
            if (fabs(dense[j]) > truncation) {
                if (col) {
                    col[nnz] = (int32_t)j;
                    val[nnz] = dense[j];
                }
                ++nnz;
            }
        if (rowptr) rowptr[i + 1] = nnz;
        free(w);
        free(u);
        oc += (size_t)r * r * r;
        ox += (size_t)rx * r;
        oy += (size_t)ry * r;
        oz += (size_t)rz * r;
    }
    free(dense);
    return nnz;
}
