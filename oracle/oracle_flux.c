/*
 * aerocfd CPU oracle (TEST INFRASTRUCTURE ONLY): primitive variables,
 * WENO3 reconstruction, Roe flux, inviscid and viscous right-hand sides,
 * local time step.
 *
 * Pinned by: the compiled reference (oracle/_ref, same inputs, tests
 * test_oracle_vs_ref.py), plus closed forms in test_oracle_closed_forms.py
 * (freestream gives rhsi = 0, WENO3 weights on linear data, Roe consistency
 * F(U,U) = physical flux, uniform flow gives rhsv = 0, Sutherland law).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acfd_oracle.h"

#define NI(e) ((e)->ie - (e)->is + 1)
#define NJ(e) ((e)->je - (e)->js + 1)
#define NK(e) ((e)->ke - (e)->ks + 1)

long ora_ix(const ora_ext *e, int i, int j, int k)
{
    return (long)(i - e->is) + (long)NI(e) * ((long)(j - e->js) + (long)NJ(e) * (long)(k - e->ks));
}

long ora_next(const ora_ext *e) { return (long)NI(e) * NJ(e) * NK(e); }

#define IX(i, j, k) ora_ix(e, (i), (j), (k))

/* sub_getpri_v.f90:33-59 */
long ora_getpri(const ora_ext *e, const double *q, double *pri_v, double gamma, double cv)
{
    long nneg = 0;
    for (int k = e->ks; k <= e->ke; k++)
        for (int j = e->js; j <= e->je; j++)
            for (int i = e->is; i <= e->ie; i++) {
                long n = IX(i, j, k);
                const double *qq = q + 5 * n;
                double d = qq[0];
                double u = qq[1] / d, v = qq[2] / d, w = qq[3] / d;
                double q2 = 0.5 * (u * u + v * v + w * w);
                double en = qq[4] / d;
                double t = (en - q2) / cv;
                double p = (qq[4] - q2 * d) * (gamma - 1.0);
                double c = sqrt(gamma * p / d);
                double *pv = pri_v + 7 * n;
                pv[0] = d; pv[1] = u; pv[2] = v; pv[3] = w; pv[4] = p; pv[5] = t; pv[6] = c;
                if (d < 0.0) nneg++;
            }
    return nneg;
}

/* WENO3 candidate pair, sub_inviscid_upwind.f90:32-41 (right) and :57-66
 * (left). For the right state u1,u2,u3 = f(i),f(i+1),f(i+2); for the left
 * state u1,u2,u3 = f(i-1),f(i),f(i+1). */
double ora_weno3_right(double u1, double u2, double u3)
{
    const double eps = 1.0e-6;
    double s1 = (u2 - u3) * (u2 - u3);
    double s2 = (u1 - u2) * (u1 - u2);
    double t1 = eps + s1, t2 = eps + s2;
    double a1 = 1.0 / (3.0 * (t1 * t1));
    double a2 = 2.0 / (3.0 * (t2 * t2));
    double w1 = a1 / (a1 + a2);
    double w2 = a2 / (a1 + a2);
    return w1 * (-u3 / 2.0 + 3.0 * u2 / 2.0) + w2 * (u2 / 2.0 + u1 / 2.0);
}

double ora_weno3_left(double u1, double u2, double u3)
{
    const double eps = 1.0e-6;
    double s1 = (u2 - u1) * (u2 - u1);
    double s2 = (u3 - u2) * (u3 - u2);
    double t1 = eps + s1, t2 = eps + s2;
    double a1 = 1.0 / (3.0 * (t1 * t1));
    double a2 = 2.0 / (3.0 * (t2 * t2));
    double w1 = a1 / (a1 + a2);
    double w2 = a2 / (a1 + a2);
    return w1 * (-u1 / 2.0 + 3.0 * u2 / 2.0) + w2 * (u2 / 2.0 + u3 / 2.0);
}

/* sub_inviscid_upwind.f90:1-180. The face i+1/2 is stored at node i, for i
 * from lo-1 to hi along coord, with (lo,hi) the computed range is1..ie1
 * passed as is0..ie0. First order at the extended edge (:25-27, :50-52). */
void ora_weno3(const ora_ext *e, double *df, const double *f, int coord, int direction)
{
    int di = coord == 1, dj = coord == 2, dk = coord == 3;
    int lo[3] = {e->is1, e->js1, e->ks1}, hi[3] = {e->ie1, e->je1, e->ke1};
    int ext_lo = coord == 1 ? e->is : coord == 2 ? e->js : e->ks;
    int ext_hi = coord == 1 ? e->ie : coord == 2 ? e->je : e->ke;
    lo[coord - 1] -= 1;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                int a = coord == 1 ? i : coord == 2 ? j : k;
                double *o = df + 5 * IX(i, j, k);
                if (direction == 1) {
                    if (a == ext_hi - 1) {
                        const double *p = f + 5 * IX(i + di, j + dj, k + dk);
                        for (int v = 0; v < 5; v++) o[v] = p[v];
                    } else {
                        const double *p1 = f + 5 * IX(i, j, k);
                        const double *p2 = f + 5 * IX(i + di, j + dj, k + dk);
                        const double *p3 = f + 5 * IX(i + 2 * di, j + 2 * dj, k + 2 * dk);
                        for (int v = 0; v < 5; v++) o[v] = ora_weno3_right(p1[v], p2[v], p3[v]);
                    }
                } else {
                    if (a == ext_lo) {
                        const double *p = f + 5 * IX(i, j, k);
                        for (int v = 0; v < 5; v++) o[v] = p[v];
                    } else {
                        const double *p1 = f + 5 * IX(i - di, j - dj, k - dk);
                        const double *p2 = f + 5 * IX(i, j, k);
                        const double *p3 = f + 5 * IX(i + di, j + dj, k + dk);
                        for (int v = 0; v < 5; v++) o[v] = ora_weno3_left(p1[v], p2[v], p3[v]);
                    }
                }
            }
}

/* Roe flux at one face with face metric (mx,my,mz); sub_inviscid_splitting.f90:65-162 */
void ora_roe_face(const double *qlf, const double *qrf, double mx, double my, double mz,
                     double gamma, double *out)
{
    const double delt = 0.06;
    double temp = sqrt(mx * mx + my * my + mz * mz);
    double nx = mx / temp, ny = my / temp, nz = mz / temp;
    double dl = qlf[0], ul = qlf[1], vl = qlf[2], wl = qlf[3], pl = qlf[4];
    double dr = qrf[0], ur = qrf[1], vr = qrf[2], wr = qrf[3], pr = qrf[4];
    double hl = gamma * pl / ((gamma - 1.0) * dl) + 0.5 * (ul * ul + vl * vl + wl * wl);
    double hr = gamma * pr / ((gamma - 1.0) * dr) + 0.5 * (ur * ur + vr * vr + wr * wr);
    double uul = ul * mx + vl * my + wl * mz;
    double uur = ur * mx + vr * my + wr * mz;
    double ubarl = uul / temp, ubarr = uur / temp;
    double fl[5], fr[5];
    fl[0] = dl * uul;
    fl[1] = dl * ul * uul + mx * pl;
    fl[2] = dl * vl * uul + my * pl;
    fl[3] = dl * wl * uul + mz * pl;
    fl[4] = dl * hl * uul;
    fr[0] = dr * uur;
    fr[1] = dr * ur * uur + mx * pr;
    fr[2] = dr * vr * uur + my * pr;
    fr[3] = dr * wr * uur + mz * pr;
    fr[4] = dr * hr * uur;
    double tmpp = sqrt(dr / dl);
    double avd = sqrt(dl * dr);
    double avu = (ul + ur * tmpp) / (1.0 + tmpp);
    double avv = (vl + vr * tmpp) / (1.0 + tmpp);
    double avw = (wl + wr * tmpp) / (1.0 + tmpp);
    double avh = (hl + hr * tmpp) / (1.0 + tmpp);
    double ava = sqrt((gamma - 1.0) * (avh - 0.5 * (avu * avu + avv * avv + avw * avw)));
    double avuu = avu * nx + avv * ny + avw * nz;
    double lam[5];
    lam[0] = fabs(avuu - ava);
    lam[1] = fabs(avuu);
    lam[2] = fabs(avuu);
    lam[3] = fabs(avuu);
    lam[4] = fabs(avuu + ava);
    for (int n = 0; n < 5; n++) {
        /* sign(1, delt - lamda) >= 0 selects the Harten value (:122-130) */
        if (delt - lam[n] >= 0.0) lam[n] = (lam[n] * lam[n] + delt * delt) / (2.0 * delt);
    }
    double deltad = dr - dl, deltap = pr - pl;
    double deltau = ur - ul, deltav = vr - vl, deltaw = wr - wl;
    double deltauu = ubarr - ubarl;
    double ava2 = ava * ava;
    double alpha1 = lam[1] * (deltad - deltap / ava2);
    double alpha2 = lam[4] * (deltap + avd * ava * deltauu) / (2.0 * ava2);
    double alpha3 = lam[0] * (deltap - avd * ava * deltauu) / (2.0 * ava2);
    double alpha4 = alpha1 + alpha2 + alpha3;
    double alpha5 = ava * (alpha2 - alpha3);
    double alpha6 = lam[1] * (avd * deltau - avd * deltauu * nx);
    double alpha7 = lam[1] * (avd * deltav - avd * deltauu * ny);
    double alpha8 = lam[1] * (avd * deltaw - avd * deltauu * nz);
    double f[5];
    f[0] = alpha4;
    f[1] = avu * alpha4 + alpha5 * nx + alpha6;
    f[2] = avv * alpha4 + alpha5 * ny + alpha7;
    f[3] = avw * alpha4 + alpha5 * nz + alpha8;
    f[4] = avh * alpha4 + avuu * alpha5 + avu * alpha6 + avv * alpha7 + avw * alpha8
         - ava2 * alpha1 / (gamma - 1.0);
    for (int v = 0; v < 5; v++) out[v] = 0.5 * (fl[v] + fr[v] - temp * f[v]);
}

/* sub_inviscid_splitting.f90:1-172 */
void ora_roe(const ora_ext *e, const double *ql, const double *qr, double *flux,
             const double *dxidx, const double *inv_j, double gamma, int coord)
{
    int di = coord == 1, dj = coord == 2, dk = coord == 3;
    int c0 = 3 * coord - 3;
    for (int k = e->ks1 - dk; k <= e->ke1; k++)
        for (int j = e->js1 - dj; j <= e->je1; j++)
            for (int i = e->is1 - di; i <= e->ie1; i++) {
                long n0 = IX(i, j, k), n1 = IX(i + di, j + dj, k + dk);
                double mx = 0.5 * (dxidx[9 * n0 + c0] * inv_j[n0] + dxidx[9 * n1 + c0] * inv_j[n1]);
                double my = 0.5 * (dxidx[9 * n0 + c0 + 1] * inv_j[n0] + dxidx[9 * n1 + c0 + 1] * inv_j[n1]);
                double mz = 0.5 * (dxidx[9 * n0 + c0 + 2] * inv_j[n0] + dxidx[9 * n1 + c0 + 2] * inv_j[n1]);
                ora_roe_face(ql + 5 * n0, qr + 5 * n0, mx, my, mz, gamma, flux + 5 * n0);
            }
}

/* conservative -> primitive (d,u,v,w,p) of a reconstructed state,
 * sub_inviscid1.f90:58-74 */
static inline void cons2prim(const double *u, double *o, double gamma)
{
    o[0] = u[0];
    o[1] = u[1] / u[0];
    o[2] = u[2] / u[0];
    o[3] = u[3] / u[0];
    o[4] = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]) * (gamma - 1.0);
}

/* sub_inviscid1.f90:8-189 (3D) */
void ora_inviscid(const ora_ext *e, double *rhsi, const double *q, const double *dxidx,
                  const double *inv_j, double gamma)
{
    long N = ora_next(e);
    double *ql = calloc(5 * N, sizeof(double)), *qr = calloc(5 * N, sizeof(double));
    double *ul = calloc(5 * N, sizeof(double)), *ur = calloc(5 * N, sizeof(double));
    double *flux = calloc(5 * N, sizeof(double));
    memset(rhsi, 0, 5 * N * sizeof(double));
    for (int coord = 1; coord <= 3; coord++) {
        int di = coord == 1, dj = coord == 2, dk = coord == 3;
        memset(ql, 0, 5 * N * sizeof(double));
        memset(qr, 0, 5 * N * sizeof(double));
        memset(ul, 0, 5 * N * sizeof(double));
        memset(ur, 0, 5 * N * sizeof(double));
        memset(flux, 0, 5 * N * sizeof(double));
        ora_weno3(e, ur, q, coord, 1);
        ora_weno3(e, ul, q, coord, -1);
        for (int k = e->ks1 - dk; k <= e->ke1; k++)
            for (int j = e->js1 - dj; j <= e->je1; j++)
                for (int i = e->is1 - di; i <= e->ie1; i++) {
                    long n = IX(i, j, k);
                    cons2prim(ul + 5 * n, ql + 5 * n, gamma);
                    cons2prim(ur + 5 * n, qr + 5 * n, gamma);
                }
        ora_roe(e, ql, qr, flux, dxidx, inv_j, gamma, coord);
        for (int k = e->ks1; k <= e->ke1; k++)
            for (int j = e->js1; j <= e->je1; j++)
                for (int i = e->is1; i <= e->ie1; i++) {
                    long n = IX(i, j, k), m = IX(i - di, j - dj, k - dk);
                    for (int v = 0; v < 5; v++) {
                        if (coord == 1)
                            rhsi[5 * n + v] = -(flux[5 * n + v] - flux[5 * m + v]);
                        else
                            rhsi[5 * n + v] = rhsi[5 * n + v] - (flux[5 * n + v] - flux[5 * m + v]);
                    }
                }
    }
    free(ql); free(qr); free(ul); free(ur); free(flux);
}

/* Node difference used by sub_viscous.f90:107-178 and :296-340: one-sided at
 * the physical range ends, central inside. */
static inline double ndiff(double fm, double f0, double fp, int a, int a0, int a1)
{
    if (a == a0) return fp - f0;
    if (a == a1) return f0 - fm;
    return (fp - fm) / 2.0;
}

/* sub_viscous.f90:9-378 (3D) */
void ora_viscous(const ora_ext *e, double *rhsv, const double *pri_v, double *amu, double *vor,
                 double *dudx, double *tao, const double *dxidx, const double *inv_j,
                 double tinf, double prl, double prt, double gamma, double cv, double re)
{
    long N = ora_next(e);
    double *u = calloc(N, sizeof(double)), *v = calloc(N, sizeof(double));
    double *w = calloc(N, sizeof(double)), *t = calloc(N, sizeof(double));
    double *g = calloc(12 * N, sizeof(double)); /* txi..wzeta: [t,u,v,w] x [xi,eta,zeta] */
    double *fv = calloc(15 * N, sizeof(double)); /* fv, gv, hv */
    for (int k = e->ks1 - 1; k <= e->ke1 + 1; k++)
        for (int j = e->js1 - 1; j <= e->je1 + 1; j++)
            for (int i = e->is1 - 1; i <= e->ie1 + 1; i++) {
                long n = IX(i, j, k);
                u[n] = pri_v[7 * n + 1];
                v[n] = pri_v[7 * n + 2];
                w[n] = pri_v[7 * n + 3];
                t[n] = pri_v[7 * n + 5];
            }
    const double *fld[4] = {t, u, v, w};
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                long n = IX(i, j, k);
                for (int f = 0; f < 4; f++) {
                    const double *a = fld[f];
                    double gx = 0, gy = 0, gz = 0;
                    if (i == e->is0) gx = a[IX(i + 1, j, k)] - a[n];
                    else if (i == e->ie0) gx = a[n] - a[IX(i - 1, j, k)];
                    else gx = (a[IX(i + 1, j, k)] - a[IX(i - 1, j, k)]) / 2.0;
                    if (j == e->js0) gy = a[IX(i, j + 1, k)] - a[n];
                    else if (j == e->je0) gy = a[n] - a[IX(i, j - 1, k)];
                    else gy = (a[IX(i, j + 1, k)] - a[IX(i, j - 1, k)]) / 2.0;
                    if (k == e->ks0) gz = a[IX(i, j, k + 1)] - a[n];
                    else if (k == e->ke0) gz = a[n] - a[IX(i, j, k - 1)];
                    else gz = (a[IX(i, j, k + 1)] - a[IX(i, j, k - 1)]) / 2.0;
                    g[12 * n + 3 * f + 0] = gx;
                    g[12 * n + 3 * f + 1] = gy;
                    g[12 * n + 3 * f + 2] = gz;
                }
            }
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                long n = IX(i, j, k);
                const double *m = dxidx + 9 * n;
                const double *gg = g + 12 * n;
                double u0 = u[n], v0 = v[n], w0 = w[n], t0 = t[n];
                double *am = amu + 3 * n;
                am[0] = (1.0 + 124.0 / tinf) / (t0 + 124.0 / tinf) * pow(t0, 1.5);
                am[2] = am[0] + am[1];
                double ki = gamma * cv * (am[0] / prl + am[1] / prt);
                double tx = gg[0] * m[0] + gg[1] * m[3] + gg[2] * m[6];
                double ty = gg[0] * m[1] + gg[1] * m[4] + gg[2] * m[7];
                double tz = gg[0] * m[2] + gg[1] * m[5] + gg[2] * m[8];
                double *du = dudx + 9 * n;
                for (int f = 0; f < 3; f++) {
                    const double *h = gg + 3 * (f + 1);
                    du[3 * f + 0] = h[0] * m[0] + h[1] * m[3] + h[2] * m[6];
                    du[3 * f + 1] = h[0] * m[1] + h[1] * m[4] + h[2] * m[7];
                    du[3 * f + 2] = h[0] * m[2] + h[1] * m[5] + h[2] * m[8];
                }
                double *ta = tao + 6 * n;
                double mu = am[2];
                ta[0] = mu * (4.0 / 3.0 * du[0] - 2.0 / 3.0 * (du[4] + du[8]));
                ta[1] = mu * (du[1] + du[3]);
                ta[2] = mu * (du[6] + du[2]);
                ta[3] = mu * (4.0 / 3.0 * du[4] - 2.0 / 3.0 * (du[0] + du[8]));
                ta[4] = mu * (du[7] + du[5]);
                ta[5] = mu * (4.0 / 3.0 * du[8] - 2.0 / 3.0 * (du[0] + du[4]));
                double bx = u0 * ta[0] + v0 * ta[1] + w0 * ta[2] + ki * tx;
                double by = u0 * ta[1] + v0 * ta[3] + w0 * ta[4] + ki * ty;
                double bz = u0 * ta[2] + v0 * ta[4] + w0 * ta[5] + ki * tz;
                vor[n] = sqrt((du[5] - du[7]) * (du[5] - du[7]) + (du[2] - du[6]) * (du[2] - du[6])
                              + (du[1] - du[3]) * (du[1] - du[3]));
                double ij = inv_j[n];
                double *F = fv + 15 * n;
                /* d = 0 (xi), 1 (eta), 2 (zeta): metric row 3d..3d+2 */
                for (int d = 0; d < 3; d++) {
                    const double *r = m + 3 * d;
                    F[5 * d + 0] = 0.0;
                    F[5 * d + 1] = (r[0] * ta[0] + r[1] * ta[1] + r[2] * ta[2]) * ij;
                    F[5 * d + 2] = (r[0] * ta[1] + r[1] * ta[3] + r[2] * ta[4]) * ij;
                    F[5 * d + 3] = (r[0] * ta[2] + r[1] * ta[4] + r[2] * ta[5]) * ij;
                    F[5 * d + 4] = (r[0] * bx + r[1] * by + r[2] * bz) * ij;
                }
            }
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                long n = IX(i, j, k);
                long im = IX(i - 1, j, k), ip = IX(i + 1, j, k);
                long jm = IX(i, j - 1, k), jp = IX(i, j + 1, k);
                long km = IX(i, j, k - 1), kp = IX(i, j, k + 1);
                for (int c = 0; c < 5; c++) {
                    double a = ndiff(i == e->is0 ? 0 : fv[15 * im + c], fv[15 * n + c],
                                     i == e->ie0 ? 0 : fv[15 * ip + c], i, e->is0, e->ie0);
                    double b = ndiff(j == e->js0 ? 0 : fv[15 * jm + 5 + c], fv[15 * n + 5 + c],
                                     j == e->je0 ? 0 : fv[15 * jp + 5 + c], j, e->js0, e->je0);
                    double h = ndiff(k == e->ks0 ? 0 : fv[15 * km + 10 + c], fv[15 * n + 10 + c],
                                     k == e->ke0 ? 0 : fv[15 * kp + 10 + c], k, e->ks0, e->ke0);
                    rhsv[5 * n + c] = (a + b + h) / re;
                }
            }
    free(u); free(v); free(w); free(t); free(g); free(fv);
}

/* sub_timeadvancement_finddt.f90:24-130. Statistics follow the reference's
 * if / else-if update (:81-85, :118-122). */
void ora_find_dt(const ora_ext *e, double *dt, double *stats, const double *pri_v,
                 const double *amu, const double *dxidx, double re, double *cfl, int iflag_time)
{
    double dtmin = 1.0e10, dtmax = 0.0, cflmin = 1.0e10, cflmax = 0.0;
    for (int k = e->ks; k <= e->ke; k++)
        for (int j = e->js; j <= e->je; j++)
            for (int i = e->is; i <= e->ie; i++) {
                long n = IX(i, j, k);
                const double *m = dxidx + 9 * n;
                double txi = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                double teta = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
                double tzeta = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
                const double *pv = pri_v + 7 * n;
                double d0 = pv[0], u0 = pv[1], v0 = pv[2], w0 = pv[3], c0 = pv[6];
                double uu = (m[0] * u0 + m[1] * v0 + m[2] * w0) / txi;
                double vv = (m[3] * u0 + m[4] * v0 + m[5] * w0) / teta;
                double ww = (m[6] * u0 + m[7] * v0 + m[8] * w0) / tzeta;
                double a3 = amu[3 * n + 2];
                double t1 = fabs(uu) + c0 + 2.0 * txi * a3 * 4.0 / 3.0 / re / d0;
                double t2 = fabs(vv) + c0 + 2.0 * teta * a3 * 4.0 / 3.0 / re / d0;
                double t3 = fabs(ww) + c0 + 2.0 * tzeta * a3 * 4.0 / 3.0 / re / d0;
                if (iflag_time == 0) {
                    dt[n] = *cfl / (txi * t1 + teta * t2 + tzeta * t3);
                    if (dt[n] >= dtmax) dtmax = dt[n];
                    else if (dt[n] <= dtmin) dtmin = dt[n];
                } else {
                    *cfl = dt[n] * (txi * t1 + teta * t2 + tzeta * t3);
                    if (*cfl >= cflmax) cflmax = *cfl;
                    else if (*cfl <= cflmin) cflmin = *cfl;
                }
            }
    if (stats) {
        stats[0] = dtmax; stats[1] = dtmin; stats[2] = cflmax; stats[3] = cflmin;
    }
}
