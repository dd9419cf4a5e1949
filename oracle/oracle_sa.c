/*
 * aerocfd CPU oracle (TEST INFRASTRUCTURE ONLY): Spalart-Allmaras model
 * (length scale, convection, diffusion, source with its Jacobian, scalar
 * LU-SGS update).
 *
 * Pinned by: the compiled reference (oracle/_ref) on random states
 * (test_oracle_vs_ref.py), and closed forms in test_oracle_closed_forms.py
 * (uniform nut with zero velocity gives zero convection and diffusion).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acfd_oracle.h"

#define IX(i, j, k) ora_ix(e, (i), (j), (k))

/* module_tursa.f90:4-11 */
static const double dta = 2.0 / 3.0, kar = 0.41;
static const double cb1 = 0.1355, cb2 = 0.622;
#define CW1 (cb1 / kar / kar + (1.0 + cb2) / dta)
static const double cw2 = 0.3, cw3 = 2.0;
static const double ct3 = 1.2, ct4 = 0.5;
static const double cv1 = 7.1;

/* d/dxi etc. with one-sided ends at the physical range, central inside; the
 * operand is read as zero outside the computed range, as the reference's
 * zero-initialised temporaries are (sub_tur_sa.f90:276-291). */
static double dcomp(const ora_ext *e, const double *a, const unsigned char *valid, int i, int j,
                    int k, int axis)
{
    int a0 = axis == 0 ? e->is0 : axis == 1 ? e->js0 : e->ks0;
    int a1 = axis == 0 ? e->ie0 : axis == 1 ? e->je0 : e->ke0;
    int c = axis == 0 ? i : axis == 1 ? j : k;
    int di = axis == 0, dj = axis == 1, dk = axis == 2;
    long n = IX(i, j, k);
#define VAL(ii, jj, kk) (valid == NULL || valid[IX(ii, jj, kk)] ? a[IX(ii, jj, kk)] : 0.0)
    if (c == a0) return VAL(i + di, j + dj, k + dk) - a[n];
    if (c == a1) return a[n] - VAL(i - di, j - dj, k - dk);
    return (VAL(i + di, j + dj, k + dk) - VAL(i - di, j - dj, k - dk)) / 2.0;
#undef VAL
}

/* sub_tur_sa.f90:130-212 */
static void sa_rhs1(const ora_ext *e, double *sa_rhs, const double *nut, const double *pri_v,
                    const double *dxidx)
{
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                double u = pri_v[7 * n + 1], v = pri_v[7 * n + 2], w = pri_v[7 * n + 3];
                double r[3];
                for (int d = 0; d < 3; d++) {
                    const double *m = dxidx + 9 * n + 3 * d;
                    double uu = m[0] * u + m[1] * v + m[2] * w;
                    int di = d == 0, dj = d == 1, dk = d == 2;
                    double nutp = nut[IX(i + di, j + dj, k + dk)];
                    double nutc = nut[n];
                    double nutm = nut[IX(i - di, j - dj, k - dk)];
                    r[d] = -(0.5 * (uu + fabs(uu)) * (nutc - nutm) + 0.5 * (uu - fabs(uu)) * (nutp - nutc));
                }
                sa_rhs[n] = r[0] + r[1] + r[2];
            }
}

/* sub_tur_sa.f90:220-523 */
static void sa_rhs2(const ora_ext *e, double *sa_rhs, const double *nut, const double *amu,
                    const double *pri_v, const double *dxidx, const double *inv_j,
                    const double *alpha, const double *beta, const double *gam, double re)
{
    long N = ora_next(e);
    unsigned char *valid = calloc(N, 1);
    double *dn = calloc(3 * N, sizeof(double));      /* dnutdxi, dnutdeta, dnutdzeta */
    double *tabg = calloc(3 * N, sizeof(double));    /* tempalpha, tempbeta, tempgamma */
    double *temp = calloc(N, sizeof(double));
    double *tall = calloc(3 * N, sizeof(double));    /* work: one component at a time */
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) valid[IX(i, j, k)] = 1;
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int d = 0; d < 3; d++) dn[3 * n + d] = dcomp(e, nut, NULL, i, j, k, d);
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                const double *g = dn + 3 * n;
                tabg[n] = alpha[3 * n] * g[0] + alpha[3 * n + 1] * g[1] + alpha[3 * n + 2] * g[2];
                tabg[N + n] = beta[3 * n] * g[0] + beta[3 * n + 1] * g[1] + beta[3 * n + 2] * g[2];
                tabg[2 * N + n] = gam[3 * n] * g[0] + gam[3 * n + 1] * g[1] + gam[3 * n + 2] * g[2];
            }
    double *tempa = tall, *tempb = tall + N, *tempc = tall + 2 * N;
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                tempa[n] = dcomp(e, tabg, valid, i, j, k, 0);
                tempb[n] = dcomp(e, tabg + N, valid, i, j, k, 1);
                tempc[n] = dcomp(e, tabg + 2 * N, valid, i, j, k, 2);
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                double nut0 = nut[n];
                double nu0 = amu[3 * n] / pri_v[7 * n];
                temp[n] = nu0 + (1.0 + cb2) * nut0;
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                const double *m = dxidx + 9 * n;
                double txi = dcomp(e, temp, valid, i, j, k, 0);
                double teta = dcomp(e, temp, valid, i, j, k, 1);
                double tzeta = dcomp(e, temp, valid, i, j, k, 2);
                double tx = txi * m[0] + teta * m[3] + tzeta * m[6];
                double ty = txi * m[1] + teta * m[4] + tzeta * m[7];
                double tz = txi * m[2] + teta * m[5] + tzeta * m[8];
                double nut0 = nut[n];
                double inv_j0 = inv_j[n];
                double deltanut = (tempa[n] + tempb[n] + tempc[n]) / inv_j0;
                const double *g = dn + 3 * n;
                double dnutdx = g[0] * m[0] + g[1] * m[3] + g[2] * m[6];
                double dnutdy = g[0] * m[1] + g[1] * m[4] + g[2] * m[7];
                double dnutdz = g[0] * m[2] + g[1] * m[5] + g[2] * m[8];
                double diff = -(cb2 / (re * dta)) * nut0 * deltanut
                            + (1.0 / (re * dta)) * (temp[n] * deltanut + tx * dnutdx + ty * dnutdy + tz * dnutdz);
                sa_rhs[n] = sa_rhs[n] + diff;
            }
    free(valid); free(dn); free(tabg); free(temp); free(tall);
}

/* sub_tur_sa.f90:532-629. r = nut0*div (multiplied, :579) as in the reference. */
static void sa_rhs3(const ora_ext *e, double *sa_rhs, const double *nut, const double *pri_v,
                    const double *amu, const double *vor, const double *length, double *diasd,
                    double re)
{
    const double cw1 = CW1;
    const double kar2 = kar * kar;
    const double cv13 = cv1 * cv1 * cv1;
    const double cw36 = cw3 * cw3 * cw3 * cw3 * cw3 * cw3;
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                double d = pri_v[7 * n];
                double nut0 = nut[n];
                double nu0 = amu[3 * n] / d;
                double omega = fabs(vor[n]);
                double dist0 = length[n];
                double leix = fmax(nut0 / nu0, 0.001);
                double leix2 = leix * leix, leix3 = leix * leix * leix;
                double fv1 = leix3 / (leix3 + cv13);
                double fv2 = 1.0 - leix / (1.0 + leix * fv1);
                double dist2 = dist0 * dist0;
                double s = omega + nut0 * fv2 / (re * kar2 * dist2);
                s = fmax(s, 0.3 * omega);
                double div = s * re * kar2 * dist2;
                div = fmax(div, 1.0e-6);
                double r = nut0 * div;
                r = fmin(r, 10.0);
                double r2 = r * r, r4 = r * r * r * r;
                double r6 = r * r * r * r * r * r;
                double g = r + cw2 * (r6 - r);
                double g6 = g * g * g * g * g * g;
                double fw = g * pow((1.0 + cw36) / (g6 + cw36), 1.0 / 6.0);
                double ft2 = ct3 * exp(-ct4 * leix2);
                double aaa = cb1 * ((1.0 - ft2) * fv2 + ft2) / kar2;
                double bbb = cw1 * fw;
                double ccc = nut0 / dist2;
                double sd1 = (aaa - bbb) * ccc / re;
                double sd = sd1 * nut0;
                double sp = cb1 * (1.0 - ft2) * omega * nut0;
                sa_rhs[n] = sa_rhs[n] + sd + sp;
                double leixn = 1.0 / nu0;
                double den = leix3 + cv13;
                double fv1n = (3.0 * leix2 * leixn * den - leix3 * (3.0 * leix2 * leixn)) / (den * den);
                double den2 = 1.0 + leix * fv1;
                double fv2n = (leixn * den2 - leix * (leixn * fv1 + leix * fv1n)) / (den2 * den2);
                double ft2n = -2.0 * ct4 * leix * leixn * ft2;
                double sn = 1.0 / (re * kar2 * dist2) * (fv2 + fv2n * nut0);
                div = (s * s) * (re * kar2 * dist2);
                div = fmax(div, 1.0e-6);
                double rn = (s - sn * nut0) / div;
                double r5 = r * r * r * r * r;
                double gn = (1.0 + cw2 * 6.0 * r5 - cw2) * rn;
                double cw3m6 = pow(cw3, -6.0);
                double fwn = -(1.0 / 6.0) * pow((pow(g, -6.0) + cw3m6) / (1.0 + cw3m6), -7.0 / 6.0);
                fwn = fwn / (1.0 + cw3m6) * (-6.0 * pow(g, -7.0)) * gn;
                double spn = cb1 * (1.0 - ft2) * omega + cb1 * (-ft2n) * omega * nut0;
                double sdn = (1.0 / re) * (cb1 * (fv2 - ft2 * fv2 + ft2) / kar2 - cw1 * fw) * 2.0 * nut0 / dist2;
                sdn = (1.0 / re) * (cb1 * (fv2n - ft2n * fv2 - ft2 * fv2n + ft2n) / kar2 - cw1 * fwn)
                      * ((nut0 * nut0) / dist2) + sdn;
                diasd[n] = fmin(spn + sdn, 0.0);
            }
}

/* sub_tur_sa.f90:638-825 (3D) */
static void sa_lusgs(const ora_ext *e, double *nut, const double *dt, const double *sa_rhs,
                     const double *pri_v, const double *dxidx, const double *diasd, double re,
                     double cfl, int iflag_time)
{
    long N = ora_next(e);
    double *dnut = calloc(N, sizeof(double)), *dnut1 = calloc(N, sizeof(double));
    double *diag = calloc(N, sizeof(double));
    double *cu = calloc(3 * N, sizeof(double));
    double phi = (cfl >= 1.0e3 && iflag_time == 0) ? 0.0 : 1.0;
    for (int k = e->ks1 - 1; k <= e->ke1 + 1; k++)
        for (int j = e->js1 - 1; j <= e->je1 + 1; j++)
            for (int i = e->is1 - 1; i <= e->ie1 + 1; i++) {
                long n = IX(i, j, k);
                double u = pri_v[7 * n + 1], v = pri_v[7 * n + 2], w = pri_v[7 * n + 3];
                const double *m = dxidx + 9 * n;
                double x2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                cu[3 * n] = m[0] * u + m[1] * v + m[2] * w;
                x2 = x2 + m[3] * m[3] + m[4] * m[4] + m[5] * m[5];
                cu[3 * n + 1] = m[3] * u + m[4] * v + m[5] * w;
                x2 = x2 + m[6] * m[6] + m[7] * m[7] + m[8] * m[8];
                cu[3 * n + 2] = m[6] * u + m[7] * v + m[8] * w;
                double tm = 4.0 * x2 * nut[n] / re;
                /* diasd is only defined on the computed range; the halo ring of
                 * diag is never read by the sweeps. */
                int inner = i >= e->is1 && i <= e->ie1 && j >= e->js1 && j <= e->je1 && k >= e->ks1 && k <= e->ke1;
                double ds = inner ? diasd[n] : 0.0;
                diag[n] = 1.0 / dt[n] * phi + fabs(cu[3 * n]) + fabs(cu[3 * n + 1]) + fabs(cu[3 * n + 2]) + tm - ds;
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                double ta = 0.5 * (cu[3 * n] + fabs(cu[3 * n])) * dnut1[IX(i - 1, j, k)];
                double tb = 0.5 * (cu[3 * n + 1] + fabs(cu[3 * n + 1])) * dnut1[IX(i, j - 1, k)];
                double tc = 0.5 * (cu[3 * n + 2] + fabs(cu[3 * n + 2])) * dnut1[IX(i, j, k - 1)];
                dnut1[n] = (sa_rhs[n] + (ta + tb + tc)) / diag[n];
            }
    for (int k = e->ke1; k >= e->ks1; k--)
        for (int j = e->je1; j >= e->js1; j--)
            for (int i = e->ie1; i >= e->is1; i--) {
                long n = IX(i, j, k);
                double ta = 0.5 * (cu[3 * n] - fabs(cu[3 * n])) * dnut[IX(i + 1, j, k)];
                double tb = 0.5 * (cu[3 * n + 1] - fabs(cu[3 * n + 1])) * dnut[IX(i, j + 1, k)];
                double tc = 0.5 * (cu[3 * n + 2] - fabs(cu[3 * n + 2])) * dnut[IX(i, j, k + 1)];
                dnut[n] = dnut1[n] - (ta + tb + tc) / diag[n];
            }
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                nut[n] = nut[n] + dnut[n];
            }
    free(dnut); free(dnut1); free(diag); free(cu);
}

/* sub_tur_sa.f90:7-122 */
void ora_sa_model(const ora_ext *e, double *nut, double *sa_rhs, const double *dt,
                  const double *pri_v, const double *amu, const double *dudx, const double *vor,
                  const double *dxidx, const double *inv_j, const double *alpha,
                  const double *beta, const double *gam, double *length, const double *dist,
                  const double *spac, double re, double cfl, int iflag_des, int iflag_time)
{
    long N = ora_next(e);
    const double cdes = 0.65;
    double *diasd = calloc(N, sizeof(double));
    memset(sa_rhs, 0, N * sizeof(double));
    for (long n = 0; n < N; n++) length[n] = dist[n];
    if (iflag_des == 1) {
        for (long n = 0; n < N; n++) length[n] = fmin(dist[n], cdes * spac[n]);
    } else if (iflag_des == 2) {
        /* dudx is declared dudx(3,is:ie,js:je,ks:ke,3) in sa_model
         * (sub_tur_sa.f90:29) while the caller holds dudx(9,...): element
         * (a,n,b) of the dummy is flat element (a-1) + 3n + 3N(b-1) of the
         * actual array. */
#define DU(a, n, b) dudx[(a) - 1 + 3 * (n) + 3 * N * ((b) - 1)]
        for (long n = 0; n < N; n++) {
            double uy = DU(1, n, 2), uz = DU(1, n, 3);
            double vx = DU(2, n, 1), vz = DU(2, n, 3);
            double wx = DU(3, n, 1), wy = DU(3, n, 2);
            double d0 = pri_v[7 * n], nut0 = nut[n];
            double nu0 = amu[3 * n] / d0;
            double dist0 = dist[n], spac0 = spac[n];
            double temp = uy * uy + uz * uz + vx * vx + vz * vz + wx * wx + wy * wy;
            temp = sqrt(temp);
            double rd = (nu0 + nut0) / (temp * re * (kar * kar) * (dist0 * dist0));
            double x = 8.0 * rd;
            double fd = 1.0 - tanh(x * x * x);
            length[n] = dist0 - fd * fmax(0.0, dist0 - cdes * spac0);
        }
#undef DU
    }
    sa_rhs1(e, sa_rhs, nut, pri_v, dxidx);
    sa_rhs2(e, sa_rhs, nut, amu, pri_v, dxidx, inv_j, alpha, beta, gam, re);
    sa_rhs3(e, sa_rhs, nut, pri_v, amu, vor, length, diasd, re);
    sa_lusgs(e, nut, dt, sa_rhs, pri_v, dxidx, diasd, re, cfl, iflag_time);
    free(diasd);
}
