/*
 * aerocfd CPU oracle (TEST INFRASTRUCTURE ONLY): flux Jacobian, LU-SGS
 * first-order implicit update, BDF2 line-sweep update.
 *
 * Pinned by: the compiled reference (oracle/_ref) on random states, plus
 * closed forms in test_oracle_closed_forms.py (Euler homogeneity
 * A(q) q = F(q); a uniform state with rhs = 0 gives dq = 0; a brute-force
 * dense Gauss-Seidel solve of the same lower/upper systems on a 4x4x4 block).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acfd_oracle.h"

#define IX(i, j, k) ora_ix(e, (i), (j), (k))
#define A(r, c) a[((c) - 1) * 5 + ((r) - 1)]

/* sub_timeadvancement_jac.f90:7-63 */
void ora_calc_jm(double *a, double q1, double q2, double q3, double q4, double q5,
                 double dxidx, double dxidy, double dxidz, double gamma)
{
    double a2 = gamma - 1.0, a3 = gamma - 2.0;
    double rho = q1, mu0 = q2, mv0 = q3, mw0 = q4, e0 = q5;
    double uu = (dxidx * mu0 + dxidy * mv0 + dxidz * mw0) / rho;
    double phi = 0.5 * a2 * (mu0 * mu0 + mv0 * mv0 + mw0 * mw0) / (rho * rho);
    double a1 = gamma * e0 / rho - phi;
    A(1, 1) = 0.0; A(1, 2) = dxidx; A(1, 3) = dxidy; A(1, 4) = dxidz; A(1, 5) = 0.0;
    A(2, 1) = phi * dxidx - uu * mu0 / rho;
    A(2, 2) = uu - a3 * dxidx * mu0 / rho;
    A(2, 3) = dxidy * mu0 / rho - a2 * dxidx * mv0 / rho;
    A(2, 4) = dxidz * mu0 / rho - a2 * dxidx * mw0 / rho;
    A(2, 5) = a2 * dxidx;
    A(3, 1) = phi * dxidy - uu * mv0 / rho;
    A(3, 2) = dxidx * mv0 / rho - a2 * dxidy * mu0 / rho;
    A(3, 3) = uu - a3 * dxidy * mv0 / rho;
    A(3, 4) = dxidz * mv0 / rho - a2 * dxidy * mw0 / rho;
    A(3, 5) = a2 * dxidy;
    A(4, 1) = phi * dxidz - uu * mw0 / rho;
    A(4, 2) = dxidx * mw0 / rho - a2 * dxidz * mu0 / rho;
    A(4, 3) = dxidy * mw0 / rho - a2 * dxidz * mv0 / rho;
    A(4, 4) = uu - a3 * dxidz * mw0 / rho;
    A(4, 5) = a2 * dxidz;
    A(5, 1) = (phi - a1) * uu;
    A(5, 2) = a1 * dxidx - a2 * uu * mu0 / rho;
    A(5, 3) = a1 * dxidy - a2 * uu * mv0 / rho;
    A(5, 4) = a1 * dxidz - a2 * uu * mw0 / rho;
    A(5, 5) = gamma * uu;
}

/* Split Jacobian at a neighbour node: 0.5*(A +- r I)*inv_j
 * (sub_timeadvancement_1stimplicit.f90:283-301 and :459-477). */
static void split_jac(double *a, const ora_ext *e, const double *q, const double *dxidx,
                      const double *inv_j, const double *r, double sgn, int c0, long n,
                      double gamma)
{
    const double *qq = q + 5 * n;
    const double *m = dxidx + 9 * n + c0;
    ora_calc_jm(a, qq[0], qq[1], qq[2], qq[3], qq[4], m[0], m[1], m[2], gamma);
    for (int d = 1; d <= 5; d++) {
        if (sgn > 0) A(d, d) = A(d, d) + r[n];
        else A(d, d) = A(d, d) - r[n];
    }
    for (int x = 0; x < 25; x++) a[x] = 0.5 * a[x] * inv_j[n];
    (void)e;
}

static inline double matrow(const double *a, int row, const double *v)
{
    return A(row, 1) * v[0] + A(row, 2) * v[1] + A(row, 3) * v[2] + A(row, 4) * v[3] + A(row, 5) * v[4];
}

/* sub_timeadvancement_1stimplicit.f90:7-585 (3D) */
void ora_implicit1st(const ora_ext *e, double *q, const double *dt, const double *pri_v,
                     double *rhs, const double *rhsi, const double *rhsv, const double *dxidx,
                     const double *inv_j, double gamma, double cfl)
{
    long N = ora_next(e);
    double *diag = calloc(N, sizeof(double));
    double *ra = calloc(N, sizeof(double)), *rb = calloc(N, sizeof(double)), *rc = calloc(N, sizeof(double));
    double *dqs = calloc(5 * N, sizeof(double)), *dq = calloc(5 * N, sizeof(double));
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) rhs[5 * n + v] = rhsi[5 * n + v] + rhsv[5 * n + v];
            }
    double phi = cfl >= 1.0e3 ? 0.0 : 1.0;
    for (int k = e->ks1 - 1; k <= e->ke1 + 1; k++)
        for (int j = e->js1 - 1; j <= e->je1 + 1; j++)
            for (int i = e->is1 - 1; i <= e->ie1 + 1; i++) {
                long n = IX(i, j, k);
                const double *pv = pri_v + 7 * n;
                double u = pv[1], v = pv[2], w = pv[3], c = pv[6];
                double *rr[3] = {ra, rb, rc};
                for (int d = 0; d < 3; d++) {
                    const double *m = dxidx + 9 * n + 3 * d;
                    double temp = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                    temp = sqrt(temp);
                    double cu = m[0] * u + m[1] * v + m[2] * w;
                    rr[d][n] = fabs(cu) + fabs(c * temp);
                }
                diag[n] = phi * 1.0 / dt[n] + ra[n] + rb[n] + rc[n];
            }
    double ja[25], jb[25], jc[25];
    /* forward sweep (:279-368) */
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                long ni = IX(i - 1, j, k), nj = IX(i, j - 1, k), nk = IX(i, j, k - 1);
                split_jac(ja, e, q, dxidx, inv_j, ra, 1.0, 0, ni, gamma);
                split_jac(jb, e, q, dxidx, inv_j, rb, 1.0, 3, nj, gamma);
                split_jac(jc, e, q, dxidx, inv_j, rc, 1.0, 6, nk, gamma);
                for (int r = 1; r <= 5; r++) {
                    double ta = matrow(ja, r, dqs + 5 * ni);
                    double tb = matrow(jb, r, dqs + 5 * nj);
                    double tc = matrow(jc, r, dqs + 5 * nk);
                    dqs[5 * n + r - 1] = (rhs[5 * n + r - 1] + (ta + tb + tc)) / diag[n] / inv_j[n];
                }
            }
    /* backward sweep (:455-546) */
    for (int k = e->ke1; k >= e->ks1; k--)
        for (int j = e->je1; j >= e->js1; j--)
            for (int i = e->ie1; i >= e->is1; i--) {
                long n = IX(i, j, k);
                long ni = IX(i + 1, j, k), nj = IX(i, j + 1, k), nk = IX(i, j, k + 1);
                split_jac(ja, e, q, dxidx, inv_j, ra, -1.0, 0, ni, gamma);
                split_jac(jb, e, q, dxidx, inv_j, rb, -1.0, 3, nj, gamma);
                split_jac(jc, e, q, dxidx, inv_j, rc, -1.0, 6, nk, gamma);
                for (int r = 1; r <= 5; r++) {
                    double ta = matrow(ja, r, dq + 5 * ni);
                    double tb = matrow(jb, r, dq + 5 * nj);
                    double tc = matrow(jc, r, dq + 5 * nk);
                    dq[5 * n + r - 1] = dqs[5 * n + r - 1] - (ta + tb + tc) / inv_j[n] / diag[n];
                }
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) q[5 * n + v] = q[5 * n + v] + dq[5 * n + v];
            }
    free(diag); free(ra); free(rb); free(rc); free(dqs); free(dq);
}

/* One step of the scalar-diagonal line recurrence used three times by
 * crank2nd (sub_timeadvancement_2ndcrank.f90:128-170 pattern). The
 * neighbour node nb supplies pri_v, metric row c0, rou and dq; sgn = +1 in
 * the forward and -1 in the backward sweep. Returns a1..a5 in a[]. */
static void crank_terms(double *a, const double *pri_v, const double *dxidx, const double *rou,
                        const double *dq, long nb, int c0, double hdv, double sgn, double gamma)
{
    const double *pv = pri_v + 7 * nb;
    double u = pv[1], v = pv[2], w = pv[3], c = pv[6];
    const double *m = dxidx + 9 * nb + c0;
    double q2 = 0.5 * (u * u + v * v + w * w);
    double r = rou[nb];
    double h = pow(c, 2.0) / (gamma - 1.0) + q2;
    double cu = m[0] * u + m[1] * v + m[2] * w;
    const double *s = dq + 5 * nb;
    double b1 = hdv * (-cu * s[0] + m[0] * s[1] + m[1] * s[2] + m[2] * s[3]);
    double b2 = hdv * (q2 * s[0] - u * s[1] - v * s[2] - w * s[3] + s[4]) * (gamma - 1.0);
    double b3 = sgn > 0 ? hdv * (cu + r) : hdv * (cu - r);
    a[0] = b1 + b3 * s[0];
    a[1] = b1 * u + b2 * m[0] + b3 * s[1];
    a[2] = b1 * v + b2 * m[1] + b3 * s[2];
    a[3] = b1 * w + b2 * m[2] + b3 * s[3];
    a[4] = b1 * h + b2 * cu + b3 * s[4];
}

static double rou_at(const double *pri_v, const double *dxidx, long n, int c0)
{
    const double *pv = pri_v + 7 * n;
    double u = pv[1], v = pv[2], w = pv[3], c = pv[6];
    const double *m = dxidx + 9 * n + c0;
    double temp = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    temp = sqrt(temp);
    double cu = u * m[0] + v * m[1] + w * m[2];
    return fabs(cu) + fabs(c * temp);
}

/* sub_timeadvancement_2ndcrank.f90:7-467 (3D) */
void ora_crank2nd(const ora_ext *e, double *q, const double *q0, const double *dt,
                  const double *pri_v, double *rhs, const double *rhsi, const double *rhsv,
                  const double *dxidx, const double *inv_j, double gamma)
{
    long N = ora_next(e);
    double *rou = calloc(N, sizeof(double)), *dtsub = calloc(N, sizeof(double));
    double *dq = calloc(5 * N, sizeof(double));
    const double alph = 0.5, alph_i = 1.0 / (1.0 + alph);
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) rhs[5 * n + v] = rhsi[5 * n + v] + rhsv[5 * n + v];
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) {
                    double qm1 = q0[10 * n + 2 * v + 0], qz = q0[10 * n + 2 * v + 1];
                    rhs[5 * n + v] = alph_i * (rhs[5 * n + v] * dt[n]
                                               - ((1.0 + alph) * q[5 * n + v] - (1.0 + 2.0 * alph) * qz
                                                  + (alph) * qm1) * inv_j[n]);
                }
            }
    for (int k = e->ks1 - 1; k <= e->ke1 + 1; k++)
        for (int j = e->js1 - 1; j <= e->je1 + 1; j++)
            for (int i = e->is1 - 1; i <= e->ie1 + 1; i++) {
                long n = IX(i, j, k);
                dtsub[n] = dt[n] / (1.0 + alph);
            }
    double a[5];
    /* xi lines (:107-216) */
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++) {
            for (int i = e->is1 - 1; i <= e->ie1 + 1; i++) rou[IX(i, j, k)] = rou_at(pri_v, dxidx, IX(i, j, k), 0);
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k), nb = IX(i - 1, j, k);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 0, 0.5 * dtsub[n], 1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = div * (rhs[5 * n + v] + a[v]);
            }
            for (int i = e->ie1; i >= e->is1; i--) {
                long n = IX(i, j, k), nb = IX(i + 1, j, k);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 0, 0.5 * dtsub[n], -1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = dq[5 * n + v] - div * a[v];
            }
        }
    /* eta lines (:222-333) */
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int i = e->is1; i <= e->ie1; i++) {
            for (int j = e->js1 - 1; j <= e->je1 + 1; j++) rou[IX(i, j, k)] = rou_at(pri_v, dxidx, IX(i, j, k), 3);
            for (int j = e->js1; j <= e->je1; j++) {
                long n = IX(i, j, k), nb = IX(i, j - 1, k);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 3, 0.5 * dtsub[n], 1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = div * (dq[5 * n + v] + a[v]);
            }
            for (int j = e->je1; j >= e->js1; j--) {
                long n = IX(i, j, k), nb = IX(i, j + 1, k);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 3, 0.5 * dtsub[n], -1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = dq[5 * n + v] - div * a[v];
            }
        }
    /* zeta lines, forward from ks1+1 and backward from ke1-1 (:339-449) */
    for (int j = e->js1; j <= e->je1; j++)
        for (int i = e->is1; i <= e->ie1; i++) {
            for (int k = e->ks1 - 1; k <= e->ke1 + 1; k++) rou[IX(i, j, k)] = rou_at(pri_v, dxidx, IX(i, j, k), 6);
            for (int k = e->ks1 + 1; k <= e->ke1; k++) {
                long n = IX(i, j, k), nb = IX(i, j, k - 1);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 6, 0.5 * dtsub[n], 1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = div * (dq[5 * n + v] + a[v]);
            }
            for (int k = e->ke1 - 1; k >= e->ks1; k--) {
                long n = IX(i, j, k), nb = IX(i, j, k + 1);
                crank_terms(a, pri_v, dxidx, rou, dq, nb, 6, 0.5 * dtsub[n], -1.0, gamma);
                double div = 1.0 / (1.0 + dtsub[n] * rou[n]);
                for (int v = 0; v < 5; v++) dq[5 * n + v] = dq[5 * n + v] - div * a[v];
            }
        }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) q[5 * n + v] = q[5 * n + v] + dq[5 * n + v] / inv_j[n];
            }
    free(rou); free(dtsub); free(dq);
}

/* main.f90:692-711; q0(-1:0,5,...) stored as q0[10*n + 2*v + (level+1)] */
void ora_update_q0(const ora_ext *e, double *q0, const double *q)
{
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                long n = IX(i, j, k);
                for (int v = 0; v < 5; v++) {
                    q0[10 * n + 2 * v + 0] = q0[10 * n + 2 * v + 1];
                    q0[10 * n + 2 * v + 1] = q[5 * n + v];
                }
            }
}
