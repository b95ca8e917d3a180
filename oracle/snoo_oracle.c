/*
 * CPU FP64 oracle for the nonlinear reduced-vector-potential solve.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU
 * baseline timed by bench.py.  The product never links or calls this file.
 *
 * Each function restates the reference algorithm in the reference's loop
 * order (citations are to MelvinLie/snoo.py).  Pinning of every function is
 * described in the header comments of oracle/oracle.py.
 */
#include "snoo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.141592653589793238462643383279502984

/* ---------------------------------------------------------------------- */
/* Reluctivity: fem_c/ctools.c:130-204                                     */
/* ---------------------------------------------------------------------- */

double or_pchip_eval(int n_int, const double *c, const double *x, double B)
{
    /* n_int == -1 encodes a constant (ctools.c:134-136). */
    if (n_int == -1) return c[0];
    double f = 0.0;
    /* interval i holds x_i <= B < x_{i+1} (ctools.c:140-151) */
    for (int i = 0; i < n_int; ++i) {
        if (x[i] <= B && B < x[i + 1]) {
            for (int j = 0; j < 4; ++j) f += c[n_int * j + i] * pow(B - x[i], 3 - j);
        }
    }
    /* at or beyond the last knot the last cubic is extrapolated (ctools.c:153-160) */
    if (x[n_int] <= B) {
        for (int j = 0; j < 4; ++j)
            f += c[n_int * (j + 1) - 1] * pow(B - x[n_int - 1], 3 - j);
    }
    return f;
}

double or_pchip_deriv(int n_int, const double *c, const double *x, double B)
{
    if (n_int == -1) return 0.0;          /* ctools.c:172-174 */
    double f = 0.0;
    for (int i = 0; i < n_int; ++i) {      /* ctools.c:178-189 */
        if (x[i] <= B && B < x[i + 1]) {
            for (int j = 0; j < 3; ++j) f += (3 - j) * c[n_int * j + i] * pow(B - x[i], 2 - j);
        }
    }
    if (x[n_int] <= B) {                   /* ctools.c:191-198 */
        for (int j = 0; j < 3; ++j)
            f += (3 - j) * c[n_int * (j + 1) - 1] * pow(B - x[n_int - 1], 2 - j);
    }
    return f;
}

void or_nu_eval(int type_spec, int n_int, const double *c, const double *x,
                double B, double *nu, double *dnu)
{
    if (type_spec == OR_NU_BSQ) {
        /* nu(B) = P(B^2); dnu/dB = 2 B P'(B^2) */
        double s = B * B;
        *nu = or_pchip_eval(n_int, c, x, s);
        *dnu = 2.0 * B * or_pchip_deriv(n_int, c, x, s);
    } else {
        *nu = or_pchip_eval(n_int, c, x, B);
        *dnu = or_pchip_deriv(n_int, c, x, B);
    }
}

/* ---------------------------------------------------------------------- */
/* Geometry helpers: fem_c/ctools.c:373-478                                */
/* ---------------------------------------------------------------------- */

/* J[q*9+k*3+l] = sum_a X[n_a,k] dphi_a/du_l(xi_q), ctools.c:373-413 */
static void or_jacobian(double *J, const int *el, const double *nodes,
                        int Q, const double *d_phi)
{
    for (int i = 0; i < 9 * Q; ++i) J[i] = 0.0;
    for (int a = 0; a < 4; ++a)
        for (int q = 0; q < Q; ++q)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    J[q * 9 + k * 3 + l] += nodes[3 * el[a] + k] * d_phi[q * 12 + 3 * a + l];
}

/* signed determinant by cofactor expansion, ctools.c:440-445 */
static double or_det3(const double *m)
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) -
           m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* result = mat * vec (row-major 3x3), ctools.c:415-420 */
static void or_matvec(double *r, const double *m, const double *v)
{
    r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

/* B_q = Bs_q + sum_k (J_q c_k) x[g_k] / det_q, ctools.c:922-958 */
static void or_B_in_element(double *B, const double *Bs, int Q, const int *g,
                            const double *x, const double *cw, const double *J,
                            const double *det)
{
    double mv[3];
    for (int m = 0; m < Q; ++m) {
        B[3 * m] = Bs[3 * m];
        B[3 * m + 1] = Bs[3 * m + 1];
        B[3 * m + 2] = Bs[3 * m + 2];
        for (int k = 0; k < 6; ++k) {
            or_matvec(mv, &J[9 * m], &cw[3 * (m * 6 + k)]);
            B[3 * m] += mv[0] * x[g[k]] / det[m];
            B[3 * m + 1] += mv[1] * x[g[k]] / det[m];
            B[3 * m + 2] += mv[2] * x[g[k]] / det[m];
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Element kernels                                                          */
/* ---------------------------------------------------------------------- */

/* shared body of ctools.c:1195-1388 (with_source=1), :960-1193 (0) and the
 * diluted variant ctools.c:1565-1759 (dil != NULL: s at the quadrature
 * point's z from evaluate_finite_element, ctools.c:480-504, and
 * eval_dilution, ctools.c:124-128; nu = s nu(s|B|), nu' = s^2 nu'(s|B|)) */
static void or_K_dK_body(int with_source, int num_el, const double *nodes,
                         const int *cells, const int *glob_ids,
                         const int *orient, const double *curls,
                         const double *d_phi, const double *w, int Q,
                         int type_spec, int n_int, const double *c,
                         const double *xk, const double *x, const double *Bs,
                         const double *phi, const double *dil,
                         int *rows, int *cols, double *vK, double *vdK,
                         double *rhs)
{
    const double nu0 = 1.0 / 4.0 / OR_PI * 1e7;   /* ctools.c:1230 */
    double *J = (double *)calloc(9 * Q, sizeof(double));
    double *det = (double *)calloc(Q, sizeof(double));
    double *B = (double *)calloc(3 * Q, sizeof(double));
    double *Bmag = (double *)calloc(Q, sizeof(double));
    double *nu = (double *)calloc(Q, sizeof(double));
    double *dnu = (double *)calloc(Q, sizeof(double));
    double *zeros = (double *)calloc(3 * Q, sizeof(double));
    double cu[3], cv[3];
    long long t = 0;

    for (int i = 0; i < num_el; ++i) {
        const int *el = &cells[4 * i];
        const int *g = &glob_ids[6 * i];
        or_jacobian(J, el, nodes, Q, d_phi);
        for (int q = 0; q < Q; ++q) det[q] = or_det3(&J[9 * q]);
        /* orientation offset into the curl table, ctools.c:1280 */
        const double *cw = &curls[(long long)3 * Q * 6 * orient[i]];
        const double *bs = with_source ? &Bs[(long long)i * 3 * Q] : zeros;
        or_B_in_element(B, bs, Q, g, x, cw, J, det);
        for (int q = 0; q < Q; ++q) {             /* ctools.c:1287-1300 */
            Bmag[q] = sqrt(B[3 * q] * B[3 * q] + B[3 * q + 1] * B[3 * q + 1] +
                           B[3 * q + 2] * B[3 * q + 2]);
            if (dil) {                            /* ctools.c:1666-1682 */
                double z = 0.0;
                for (int a = 0; a < 4; ++a) z += phi[4 * q + a] * nodes[3 * el[a] + 2];
                const double s = (dil[0] + z * dil[1]) / (dil[2] + z * dil[3]);
                or_nu_eval(type_spec, n_int, c, xk, s * Bmag[q], &nu[q], &dnu[q]);
                nu[q] *= s;
                dnu[q] *= s * s;
            } else {
                or_nu_eval(type_spec, n_int, c, xk, Bmag[q], &nu[q], &dnu[q]);
            }
        }
        for (int j = 0; j < 6; ++j) {
            for (int k = j; k < 6; ++k) {         /* ctools.c:1303-1361 */
                double iK = 0.0, idK = 0.0;
                for (int q = 0; q < Q; ++q) {
                    or_matvec(cu, &J[9 * q], &cw[3 * (q * 6 + j)]);
                    or_matvec(cv, &J[9 * q], &cw[3 * (q * 6 + k)]);
                    iK += nu[q] * (cu[0] * cv[0] + cu[1] * cv[1] + cu[2] * cv[2]) * w[q] / det[q];
                    if (Bmag[q] >= 1e-14) {
                        double pu = cu[0] * B[3 * q] + cu[1] * B[3 * q + 1] + cu[2] * B[3 * q + 2];
                        double pv = cv[0] * B[3 * q] + cv[1] * B[3 * q + 1] + cv[2] * B[3 * q + 2];
                        idK += dnu[q] * pu * pv * w[q] / det[q] / Bmag[q];
                    }
                    if (with_source && j == k) {  /* ctools.c:1332-1335 */
                        const double *s = &Bs[3 * ((long long)i * Q + q)];
                        rhs[g[j]] += (nu0 - nu[q]) * (cu[0] * s[0] + cu[1] * s[1] + cu[2] * s[2]) * w[q];
                    }
                }
                rows[t] = g[j]; cols[t] = g[k]; vK[t] = iK; vdK[t] = idK; ++t;
                if (j != k) {
                    rows[t] = g[k]; cols[t] = g[j]; vK[t] = iK; vdK[t] = idK; ++t;
                }
            }
        }
    }
    free(J); free(det); free(B); free(Bmag); free(nu); free(dnu); free(zeros);
}

void or_K_dK_red(int num_el, const double *nodes, const int *cells,
                 const int *glob_ids, const int *orient,
                 const double *curls, const double *d_phi,
                 const double *w, int Q,
                 int type_spec, int n_int, const double *c, const double *xk,
                 const double *x, const double *Bs,
                 int *rows, int *cols, double *vK, double *vdK, double *rhs)
{
    or_K_dK_body(1, num_el, nodes, cells, glob_ids, orient, curls, d_phi, w, Q,
                 type_spec, n_int, c, xk, x, Bs, NULL, NULL, rows, cols, vK, vdK, rhs);
}

void or_K_dK_red_dil(int num_el, const double *nodes, const int *cells,
                     const int *glob_ids, const int *orient,
                     const double *curls, const double *phi, const double *d_phi,
                     const double *w, int Q,
                     int type_spec, int n_int, const double *c, const double *xk,
                     const double *x, const double *Bs, const double *dil,
                     int *rows, int *cols, double *vK, double *vdK, double *rhs)
{
    or_K_dK_body(1, num_el, nodes, cells, glob_ids, orient, curls, d_phi, w, Q,
                 type_spec, n_int, c, xk, x, Bs, phi, dil, rows, cols, vK, vdK, rhs);
}

void or_K_dK_air(int num_el, const double *nodes, const int *cells,
                 const int *glob_ids, const int *orient,
                 const double *curls, const double *d_phi,
                 const double *w, int Q,
                 int type_spec, int n_int, const double *c, const double *xk,
                 const double *x,
                 int *rows, int *cols, double *vK, double *vdK)
{
    or_K_dK_body(0, num_el, nodes, cells, glob_ids, orient, curls, d_phi, w, Q,
                 type_spec, n_int, c, xk, x, NULL, NULL, NULL, rows, cols, vK, vdK, NULL);
}

/* ctools.c:1500-1563 (with evaluate_finite_element, ctools.c:480-504) */
void or_B_Hcurl(int num_el, const double *nodes, const int *cells,
                const int *glob_ids, const int *orient,
                const double *curls, const double *phi, const double *d_phi,
                int Q, const double *x, double *points, double *B)
{
    double *J = (double *)calloc(9 * Q, sizeof(double));
    double *det = (double *)calloc(Q, sizeof(double));
    double *zeros = (double *)calloc(3 * Q, sizeof(double));
    for (int i = 0; i < num_el; ++i) {
        const int *el = &cells[4 * i];
        or_jacobian(J, el, nodes, Q, d_phi);
        for (int q = 0; q < Q; ++q) det[q] = or_det3(&J[9 * q]);
        const double *cw = &curls[(long long)3 * Q * 6 * orient[i]];
        double *pts = &points[(long long)i * 3 * Q];
        for (int q = 0; q < Q; ++q) {
            pts[3 * q] = pts[3 * q + 1] = pts[3 * q + 2] = 0.0;
            for (int a = 0; a < 4; ++a) {
                pts[3 * q] += phi[4 * q + a] * nodes[3 * el[a]];
                pts[3 * q + 1] += phi[4 * q + a] * nodes[3 * el[a] + 1];
                pts[3 * q + 2] += phi[4 * q + a] * nodes[3 * el[a] + 2];
            }
        }
        or_B_in_element(&B[(long long)i * 3 * Q], zeros, Q, &glob_ids[6 * i], x, cw, J, det);
    }
    free(J); free(det); free(zeros);
}

/* ---------------------------------------------------------------------- */
/* Biot-Savart, fem_c/ctools.c:1761-1923                                   */
/* ---------------------------------------------------------------------- */

void or_biot_savart(const double *src, int num_src, const double *tar,
                    int num_tar, double current, double rad, double *B,
                    int num_threads)
{
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < num_tar; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        const double *t = &tar[3 * i];
        for (int j = 0; j < num_src; ++j) {
            const double *s = &src[6 * j];
            double n[3] = {s[3] - s[0], s[4] - s[1], s[5] - s[2]};
            double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            double d1[3] = {s[0] - t[0], s[1] - t[1], s[2] - t[2]};
            double d2[3] = {s[3] - t[0], s[4] - t[1], s[5] - t[2]};
            double l1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
            double l2 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
            /* foot point parameter and perpendicular offset, ctools.c:1850-1866 */
            double tp = -1.0 * (d1[0] * n[0] + d1[1] * n[1] + d1[2] * n[2]) / nn;
            double p[3] = {t[0] - (s[0] + tp * n[0]), t[1] - (s[1] + tp * n[1]),
                           t[2] - (s[2] + tp * n[2])};
            double dd = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            if ((dd < rad * rad && tp <= 1.0 && tp >= 0.0) || l1 < rad || l2 < rad) {
                /* inside the core: mu0 I r / (2 pi R^2) along n x p, ctools.c:1869-1884 */
                double bd[3] = {n[1] * p[2] - n[2] * p[1], n[2] * p[0] - n[0] * p[2],
                                n[0] * p[1] - n[1] * p[0]};
                double nb = sqrt(bd[0] * bd[0] + bd[1] * bd[1] + bd[2] * bd[2]);
                double f = 2.0 * sqrt(dd) / rad / rad / nb;
                acc[0] += bd[0] * f; acc[1] += bd[1] * f; acc[2] += bd[2] * f;
            } else {
                /* finite segment, Russenschuck (5.51), ctools.c:1886-1902 */
                double d12 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
                double f = (l1 + l2) / (l1 * l2 + d12) / l1 / l2;
                acc[0] += (d1[1] * d2[2] - d1[2] * d2[1]) * f;
                acc[1] += (d1[2] * d2[0] - d1[0] * d2[2]) * f;
                acc[2] += (d1[0] * d2[1] - d1[1] * d2[0]) * f;
            }
        }
        /* scale once per target, ctools.c:1908-1912 */
        B[3 * i] = 0.0 + acc[0] * current * 1e-7;
        B[3 * i + 1] = 0.0 + acc[1] * current * 1e-7;
        B[3 * i + 2] = 0.0 + acc[2] * current * 1e-7;
    }
}

/* ---------------------------------------------------------------------- */
/* Sparse algebra                                                           */
/* ---------------------------------------------------------------------- */

typedef struct { int col; long long ord; double v; } or_ent;

static int or_ent_cmp(const void *a, const void *b)
{
    const or_ent *x = (const or_ent *)a, *y = (const or_ent *)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

long long or_coo_to_csr(int n, long long nt, const int *ri, const int *ci,
                        const double *v, int *indptr, int *indices,
                        double *data)
{
    /* bucket by row in input order (scipy coo_tocsr) */
    long long *cnt = (long long *)calloc(n + 1, sizeof(long long));
    for (long long t = 0; t < nt; ++t) cnt[ri[t] + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    or_ent *e = (or_ent *)malloc((nt > 0 ? nt : 1) * sizeof(or_ent));
    long long *pos = (long long *)malloc((n + 1) * sizeof(long long));
    memcpy(pos, cnt, (n + 1) * sizeof(long long));
    for (long long t = 0; t < nt; ++t) {
        long long p = pos[ri[t]]++;
        e[p].col = ci[t]; e[p].ord = t; e[p].v = v[t];
    }
    /* sort each row by column, then sum duplicates (csr_sum_duplicates) */
    long long nnz = 0;
    indptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        long long a = cnt[i], b = cnt[i + 1];
        qsort(e + a, (size_t)(b - a), sizeof(or_ent), or_ent_cmp);
        long long k = a;
        while (k < b) {
            int col = e[k].col;
            double s = e[k].v;
            ++k;
            while (k < b && e[k].col == col) { s += e[k].v; ++k; }
            indices[nnz] = col; data[nnz] = s; ++nnz;
        }
        indptr[i + 1] = (int)nnz;
    }
    free(cnt); free(e); free(pos);
    return nnz;
}

void or_spmv(int n, const int *indptr, const int *indices, const double *data,
             const double *x, double *y)
{
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) s += data[k] * x[indices[k]];
        y[i] = s;
    }
}

/* reduced-system matvec: rows and columns with mask==0 are dropped */
static void or_spmv_masked(int n, const int *indptr, const int *indices,
                           const double *data, const unsigned char *mask,
                           const double *x, double *y)
{
    for (int i = 0; i < n; ++i) {
        if (!mask[i]) { y[i] = 0.0; continue; }
        double s = 0.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (mask[indices[k]]) s += data[k] * x[indices[k]];
        y[i] = s;
    }
}

static double or_dot_masked(int n, const unsigned char *m, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) if (m[i]) s += a[i] * b[i];
    return s;
}

/* scipy/sparse/linalg/_isolve/iterative.py cg (scipy 1.15.3), driven as at
 * snoopy/red_mvp_solver.py:195-205 */
int or_pcg(int n, const int *indptr, const int *indices, const double *data,
           const unsigned char *mask, const double *b, double *x,
           double rtol, double atol, int maxiter, int *info)
{
    double *dinv = (double *)calloc(n, sizeof(double));
    double *r = (double *)calloc(n, sizeof(double));
    double *z = (double *)calloc(n, sizeof(double));
    double *p = (double *)calloc(n, sizeof(double));
    double *q = (double *)calloc(n, sizeof(double));
    for (int i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        double d = 0.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (indices[k] == i) d += data[k];
        dinv[i] = 1.0 / d;                        /* M = diags(1/J_c.diagonal()) */
    }
    double bnrm = sqrt(or_dot_masked(n, mask, b, b));
    double tol = atol > rtol * bnrm ? atol : rtol * bnrm;  /* _get_atol_rtol */
    int it = 0;
    *info = 0;
    if (bnrm == 0.0) {                            /* returns b itself */
        for (int i = 0; i < n; ++i) if (mask[i]) x[i] = b[i];
        goto done;
    }
    or_spmv_masked(n, indptr, indices, data, mask, x, q);
    for (int i = 0; i < n; ++i) r[i] = mask[i] ? b[i] - q[i] : 0.0;
    double rho_prev = 0.0;
    for (it = 0; it < maxiter; ++it) {
        if (sqrt(or_dot_masked(n, mask, r, r)) < tol) goto done;
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        double rho = or_dot_masked(n, mask, r, z);
        if (it > 0) {
            double beta = rho / rho_prev;
            for (int i = 0; i < n; ++i) { p[i] *= beta; p[i] += z[i]; }
        } else {
            for (int i = 0; i < n; ++i) p[i] = z[i];
        }
        or_spmv_masked(n, indptr, indices, data, mask, p, q);
        double alpha = rho / or_dot_masked(n, mask, p, q);
        for (int i = 0; i < n; ++i) if (mask[i]) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        rho_prev = rho;
    }
    *info = maxiter;
done:
    free(dinv); free(r); free(z); free(p); free(q);
    return it;
}
