/*
 * CPU FP64 oracle for the one-dimensional finite elements and the Tucker-basis
 * 1-D matrices (test infrastructure only; see chfsi_oracle.h for the rule
 * that only tests, the smoke check and the benchmark's CPU baseline may load
 * it).
 *
 * Nodal fields are stored one field per row (nfun x nnodes), values at the
 * quadrature points one field per row (nfun x nel nq), points of element e at
 * e nq .. e nq + nq - 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chfsi_oracle.h"

/* fem/FEM.cc:86-96: start, start + h, start + 2h ... by repeated addition. */
static void linspace_acc(double a, double b, int n, double *v) {
    const double h = (b - a) / (n - 1.0);
    v[0] = a;
    for (int i = 1; i < n; ++i) v[i] = v[i - 1] + h;
}

/* fem/FEM.cc:159-375 (generateNumberNodes, generateNodes) for the uniform,
 * adaptive, manual1 and hartree_outer meshes (type 0, 1, 3, 4).  Returns the
 * number of nodes, or -1 for another type. */
int64_t orc_fem_mesh(int type, int nel, int npe, double start, double end, double inner_size,
                     double inner_h, double coarsening, double *nodes) {
    const int64_t nn = (int64_t)nel * (npe - 1) + 1;
    if (type == 0) {
        const double interval = fabs(end - start) / (nn - 1);
        for (int64_t i = 0; i < nn; ++i) nodes[i] = start + interval * i;
        return nn;
    }
    if (type != 1 && type != 3 && type != 4) return -1;
    const int n_in = (int)floor(inner_size / inner_h);
    const int n_out = (nel - n_in) / 2;
    const double centre = (start + end) / 2;
    const double in_start = centre - inner_size / 2.0;
    double *lin = calloc((size_t)nel + 1, sizeof(double));
    int cnt = 0;
    for (int i = n_out; i <= n_out + n_in; ++i) lin[i] = in_start + inner_h * cnt++;
    double grow = coarsening, size = inner_h;
    if (type == 3) { /* manual1: :255-291 */
        grow = 1.5;
        size = 1.5;
    } else if (type == 4) { /* hartree_outer: :292-339 */
        double g = 0.0;
        for (int i = 1; i <= n_out; ++i) g += pow(grow, i);
        size = ((end - start) - inner_size) / 2.0 / g;
    }
    for (int i = 1; i <= n_out; ++i) {
        size = size * grow;
        lin[n_out - i] = lin[n_out - i + 1] - size;
        lin[n_out + n_in + i] = lin[n_out + n_in + i - 1] + size;
    }
    double *tmp = malloc(sizeof(double) * npe);
    for (int i = 0; i < nel; ++i) {
        linspace_acc(lin[i], lin[i + 1], npe, tmp);
        for (int j = 0; j < npe - 1; ++j) nodes[(int64_t)i * (npe - 1) + j] = tmp[j];
    }
    nodes[nn - 1] = type == 1 ? end : lin[nel];
    free(tmp);
    free(lin);
    return nn;
}

/* Shape functions of npe equispaced nodes on [-1, 1] at the nq points xi:
 * S[i nq + q] = N_i(xi_q) as the product of (xi_q - x_k) / (x_i - x_k) in k
 * order (fem/FEM.cc:418-458); D[i nq + q] = N_i'(xi_q) from the monomial
 * coefficients of N_i, differentiated and summed as c_m x^m with pow
 * (fem/FEM.cc:484-528 with poly_multiply, poly_derivative, poly_eval,
 * :49-121). */
void orc_fem_shape(int npe, int nq, const double *xi, double *S, double *D) {
    double *xs = malloc(sizeof(double) * npe);
    double *c = malloc(sizeof(double) * (npe + 1));
    double *t = malloc(sizeof(double) * (npe + 1));
    for (int i = 0; i < npe; ++i) xs[i] = -1.0 + i * (2.0 / (npe - 1));
    for (int i = 0; i < npe; ++i) {
        for (int q = 0; q < nq; ++q) {
            double s = 1.0;
            for (int k = 0; k < npe; ++k)
                if (k != i) s = s * ((xi[q] - xs[k]) / (xs[i] - xs[k]));
            S[(size_t)i * nq + q] = s;
        }
        /* coefficients of prod_{j != i} (-x_j / (x_i - x_j) + x / (x_i - x_j)) */
        int deg = 0;
        c[0] = 1.0;
        for (int j = 0; j < npe; ++j) {
            if (j == i) continue;
            const double a0 = -xs[j] / (xs[i] - xs[j]), a1 = 1.0 / (xs[i] - xs[j]);
            for (int m = 0; m <= deg + 1; ++m) t[m] = 0.0;
            /* poly_multiply sums the partial products of the shorter factor's
             * terms in order: first c * a0, then (shifted) c * a1 */
            for (int m = 0; m <= deg; ++m) t[m] += c[m] * a0;
            for (int m = 0; m <= deg; ++m) t[m + 1] += c[m] * a1;
            ++deg;
            memcpy(c, t, sizeof(double) * (deg + 1));
        }
        for (int q = 0; q < nq; ++q) {
            double v = 0.0;
            for (int m = 1; m <= deg; ++m) v += (double)m * c[m] * pow(xi[q], (double)(m - 1));
            D[(size_t)i * nq + q] = v;
        }
    }
    free(t);
    free(c);
    free(xs);
}

/* Values and xi-derivatives at the quadrature points of nfun nodal fields by
 * the per-element products of fem/FEM.cc:906-946 (y = sum_i N_i u_i, i in
 * order), and the per-point data of FEM::generateQuadPointData (:552-571):
 * jac = (x_last - x_first) / 2, jw = jac w, inv_jw = w / jac, positions.
 * val, dval, jw, inv_jw, xq may be NULL. */
void orc_fem_interpolate(int nel, int npe, int nq, const double *xi, const double *w,
                         const double *nodes, int nfun, const double *nodal, double *val,
                         double *dval, double *jw, double *inv_jw, double *xq) {
    double *S = malloc(sizeof(double) * (size_t)npe * nq);
    double *D = malloc(sizeof(double) * (size_t)npe * nq);
    orc_fem_shape(npe, nq, xi, S, D);
    const int64_t nn = (int64_t)nel * (npe - 1) + 1;
    for (int e = 0; e < nel; ++e) {
        const double *xe = nodes + (int64_t)e * (npe - 1);
        const double jac = (xe[npe - 1] - xe[0]) / 2.0;
        for (int q = 0; q < nq; ++q) {
            const int64_t t = (int64_t)e * nq + q;
            double x = 0.0;
            for (int i = 0; i < npe; ++i) x += S[(size_t)i * nq + q] * xe[i];
            if (xq) xq[t] = x;
            if (jw) jw[t] = jac * w[q];
            if (inv_jw) inv_jw[t] = w[q] / jac;
        }
        for (int f = 0; f < nfun; ++f) {
            const double *u = nodal + (int64_t)f * nn + (int64_t)e * (npe - 1);
            for (int q = 0; q < nq; ++q) {
                double s = 0.0, d = 0.0;
                for (int i = 0; i < npe; ++i) {
                    s += S[(size_t)i * nq + q] * u[i];
                    d += D[(size_t)i * nq + q] * u[i];
                }
                const int64_t o = (int64_t)f * nel * nq + (int64_t)e * nq + q;
                if (val) val[o] = s;
                if (dval) dval[o] = d;
            }
        }
    }
    free(D);
    free(S);
}

/* dft/ProjectHamiltonianSparse.cc:577-722: for I <= J the pointwise product
 * of the two functions (times U_p for MatPot), integrated by the dot product
 * with jw (inv_jw for MatDX, fem/FEM.cc:978-1006); I > J mirrored.  B, DB: r
 * x nq row-major; U: nq x rv column-major; outputs row-major r x r (MatPot:
 * rv of them).  Any output may be NULL. */
void orc_tucker_basis_matrices(int nq, int r, const double *B, const double *DB,
                               const double *jw, const double *inv_jw, int rv, const double *U,
                               double *MatX, double *MatDX, double *MatPot) {
    for (int I = 0; I < r; ++I)
        for (int J = 0; J < r; ++J) {
            const size_t o = (size_t)I * r + J, m = (size_t)J * r + I;
            if (I > J) {
                if (MatX) MatX[o] = MatX[m];
                if (MatDX) MatDX[o] = MatDX[m];
                continue;
            }
            const double *bi = B + (size_t)I * nq, *bj = B + (size_t)J * nq;
            if (MatX) {
                double s = 0.0;
                for (int q = 0; q < nq; ++q) s += (bi[q] * bj[q]) * jw[q];
                MatX[o] = s;
            }
            if (MatDX) {
                const double *di = DB + (size_t)I * nq, *dj = DB + (size_t)J * nq;
                double s = 0.0;
                for (int q = 0; q < nq; ++q) s += (di[q] * dj[q]) * inv_jw[q];
                MatDX[o] = s;
            }
        }
    if (!MatPot) return;
    for (int p = 0; p < rv; ++p) {
        const double *u = U + (size_t)nq * p;
        double *M = MatPot + (size_t)p * r * r;
        for (int I = 0; I < r; ++I)
            for (int J = 0; J < r; ++J) {
                if (I > J) {
                    M[(size_t)I * r + J] = M[(size_t)J * r + I];
                    continue;
                }
                const double *bi = B + (size_t)I * nq, *bj = B + (size_t)J * nq;
                double s = 0.0;
                for (int q = 0; q < nq; ++q) s += (u[q] * (bi[q] * bj[q])) * jw[q];
                M[(size_t)I * r + J] = s;
            }
    }
}
