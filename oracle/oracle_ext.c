/* CPU oracle of the path's extensions that have no counterpart in the
 * reference (TEST INFRASTRUCTURE ONLY): Steger-Warming flux-vector
 * splitting with WENO5 flux reconstruction, a weighted compact nonlinear
 * scheme (WENO5 midpoint interpolation + Roe flux), and the three-stage
 * Shu-Osher Runge-Kutta time step.  These follow the published methods:
 *   Steger & Warming, J. Comput. Phys. 40 (1981) 263-293;
 *   Jiang & Shu, J. Comput. Phys. 126 (1996) 202-228 (WENO5 weights,
 *     smoothness indicators, epsilon = 1e-6);
 *   Deng & Zhang, J. Comput. Phys. 165 (2000) 22-44 (WCNS midpoint
 *     interpolation, linear weights 1/16, 10/16, 5/16);
 *   Shu & Osher, J. Comput. Phys. 77 (1988) 439-471 (TVD RK3).
 * Everything else (metrics, extended ranges, the Roe flux, boundary
 * conditions, halo exchange) is the reference's, reused from the other
 * oracle files.  Compiled -O2 -ffp-contract=off like the rest. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "acfd_oracle.h"

#define IX(i, j, k) ora_ix(e, (i), (j), (k))

/* Steger-Warming split fluxes of conservative state q through metric M
 * (grad xi_d / J):  F+- = rho/(2 gamma) [2(gamma-1) l1 (1,u,v,w,|u|^2/2)
 * + l4 (1, u+c n, H+c U~) + l5 (1, u-c n, H-c U~)], l = (l +- |l|)/2 of the
 * eigenvalues l1 = M.u, l4 = M.u + c|M|, l5 = M.u - c|M|; n = M/|M|. */
void ora_sw_split(const double *q, const double *M, double gamma, double *fp, double *fm)
{
    double rho = q[0], u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
    double v2 = u * u + v * v + w * w;
    double p = (gamma - 1.0) * (q[4] - 0.5 * rho * v2);
    double c = sqrt(gamma * p / rho);
    double mm = sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
    double nx = M[0] / mm, ny = M[1] / mm, nz = M[2] / mm;
    double ut = u * nx + v * ny + w * nz;
    double h = c * c / (gamma - 1.0) + 0.5 * v2;
    double l[3] = {mm * ut, mm * (ut + c), mm * (ut - c)};
    double s = rho / (2.0 * gamma);
    for (int side = 0; side < 2; side++) {
        double lp[3];
        for (int a = 0; a < 3; a++) lp[a] = side == 0 ? 0.5 * (l[a] + fabs(l[a])) : 0.5 * (l[a] - fabs(l[a]));
        double a1 = 2.0 * (gamma - 1.0) * lp[0];
        double *f = side == 0 ? fp : fm;
        f[0] = s * (a1 + lp[1] + lp[2]);
        f[1] = s * (a1 * u + lp[1] * (u + c * nx) + lp[2] * (u - c * nx));
        f[2] = s * (a1 * v + lp[1] * (v + c * ny) + lp[2] * (v - c * ny));
        f[3] = s * (a1 * w + lp[1] * (w + c * nz) + lp[2] * (w - c * nz));
        f[4] = s * (a1 * 0.5 * v2 + lp[1] * (h + c * ut) + lp[2] * (h - c * ut));
    }
}

/* WENO5 flux reconstruction at i+1/2 from f(i-2..i+2) (left-biased),
 * Jiang & Shu eqs. (2.10)-(2.17). */
double ora_weno5_rec(double a, double b, double c, double d, double e5)
{
    const double eps = 1.0e-6;
    double q0 = a / 3.0 - 7.0 * b / 6.0 + 11.0 * c / 6.0;
    double q1 = -b / 6.0 + 5.0 * c / 6.0 + d / 3.0;
    double q2 = c / 3.0 + 5.0 * d / 6.0 - e5 / 6.0;
    double b0 = 13.0 / 12.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 0.25 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
    double b1 = 13.0 / 12.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 0.25 * (b - d) * (b - d);
    double b2 = 13.0 / 12.0 * (c - 2.0 * d + e5) * (c - 2.0 * d + e5) + 0.25 * (3.0 * c - 4.0 * d + e5) * (3.0 * c - 4.0 * d + e5);
    double a0 = 0.1 / ((eps + b0) * (eps + b0));
    double a1 = 0.6 / ((eps + b1) * (eps + b1));
    double a2 = 0.3 / ((eps + b2) * (eps + b2));
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

/* WENO3 flux reconstruction at i+1/2 from f(i-1..i+1): candidates
 * -f(i-1)/2 + 3f(i)/2 and (f(i) + f(i+1))/2, linear weights 1/3, 2/3. */
double ora_weno3_rec(double a, double b, double c)
{
    const double eps = 1.0e-6;
    double q0 = -0.5 * a + 1.5 * b, q1 = 0.5 * (b + c);
    double b0 = (b - a) * (b - a), b1 = (c - b) * (c - b);
    double a0 = (1.0 / 3.0) / ((eps + b0) * (eps + b0));
    double a1 = (2.0 / 3.0) / ((eps + b1) * (eps + b1));
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}

/* WENO5 interpolation of node values to the midpoint i+1/2 from
 * u(i-2..i+2): candidates (3,-10,15)/8, (-1,6,3)/8, (3,6,-1)/8, linear
 * weights 1/16, 10/16, 5/16, Jiang-Shu smoothness indicators. */
double ora_weno5_interp(double a, double b, double c, double d, double e5)
{
    const double eps = 1.0e-6;
    double q0 = (3.0 * a - 10.0 * b + 15.0 * c) / 8.0;
    double q1 = (-b + 6.0 * c + 3.0 * d) / 8.0;
    double q2 = (3.0 * c + 6.0 * d - e5) / 8.0;
    double b0 = 13.0 / 12.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 0.25 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
    double b1 = 13.0 / 12.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 0.25 * (b - d) * (b - d);
    double b2 = 13.0 / 12.0 * (c - 2.0 * d + e5) * (c - 2.0 * d + e5) + 0.25 * (3.0 * c - 4.0 * d + e5) * (3.0 * c - 4.0 * d + e5);
    double a0 = (1.0 / 16.0) / ((eps + b0) * (eps + b0));
    double a1 = (10.0 / 16.0) / ((eps + b1) * (eps + b1));
    double a2 = (5.0 / 16.0) / ((eps + b2) * (eps + b2));
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

/* Face flux between positions a and a+1 of one line (axis positions of
 * the extended range lo..hi).  v(o) returns component x of the quantity at
 * offset o from a.  Stencil order drops where the extended range ends:
 * 6 points (a-2..a+3) -> 5th order, 4 points (a-1..a+2) -> 3rd order,
 * else first order. */
static int face_order(int a, int lo, int hi)
{
    if (a - 2 >= lo && a + 3 <= hi) return 5;
    if (a - 1 >= lo && a + 2 <= hi) return 3;
    return 1;
}

/* node metric M = grad xi_d / J */
static void node_metric(const ora_ext *e, const double *dxidx, const double *inv_j, long n, int d, double *M)
{
    for (int c = 0; c < 3; c++) M[c] = dxidx[9 * n + 3 * d + c] * inv_j[n];
}

static void conservative_to_primitive(const double *u, double *o, double gamma)
{
    o[0] = u[0];
    o[1] = u[1] / u[0];
    o[2] = u[2] / u[0];
    o[3] = u[3] / u[0];
    o[4] = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]) * (gamma - 1.0);
}

void ora_roe_face(const double *ql, const double *qr, double mx, double my, double mz, double gamma,
                  double *out);
double ora_weno3_left(double u1, double u2, double u3);
double ora_weno3_right(double u1, double u2, double u3);

/* Flux through face a+1/2 of one line; base points at node a of the line
 * (stride st), lo/hi = extended range of the line's axis, d = axis. */
static void face_flux_ext(const ora_ext *e, const double *q, const double *dxidx, const double *inv_j,
                          double gamma, int scheme, long base, long st, int a, int lo, int hi, int d,
                          double *F)
{
    int ord = face_order(a, lo, hi);
    if (scheme == 1) {
        double fp[6][5], fm[6][5];
        for (int o = -2; o <= 3; o++) {
            if (a + o < lo || a + o > hi) continue;
            double M[3];
            long no = base + o * st;
            node_metric(e, dxidx, inv_j, no, d, M);
            ora_sw_split(q + 5 * no, M, gamma, fp[o + 2], fm[o + 2]);
        }
        for (int x = 0; x < 5; x++) {
            double p, m;
            if (ord == 5) {
                p = ora_weno5_rec(fp[0][x], fp[1][x], fp[2][x], fp[3][x], fp[4][x]);
                m = ora_weno5_rec(fm[5][x], fm[4][x], fm[3][x], fm[2][x], fm[1][x]);
            } else if (ord == 3) {
                p = ora_weno3_rec(fp[1][x], fp[2][x], fp[3][x]);
                m = ora_weno3_rec(fm[4][x], fm[3][x], fm[2][x]);
            } else {
                p = fp[2][x];
                m = fm[3][x];
            }
            F[x] = p + m;
        }
        return;
    }
    double ul[5], ur[5], pl[5], pr[5];
    for (int x = 0; x < 5; x++) {
        double v[6];
        for (int o = -2; o <= 3; o++) v[o + 2] = (a + o >= lo && a + o <= hi) ? q[5 * (base + o * st) + x] : 0.0;
        if (ord == 5) {
            ul[x] = ora_weno5_interp(v[0], v[1], v[2], v[3], v[4]);
            ur[x] = ora_weno5_interp(v[5], v[4], v[3], v[2], v[1]);
        } else if (ord == 3) {
            /* the reference's WENO3 states near the range ends */
            ul[x] = ora_weno3_left(v[1], v[2], v[3]);
            ur[x] = ora_weno3_right(v[2], v[3], v[4]);
        } else {
            ul[x] = v[2];
            ur[x] = v[3];
        }
    }
    conservative_to_primitive(ul, pl, gamma);
    conservative_to_primitive(ur, pr, gamma);
    long n1 = base + st;
    double mx = 0.5 * (dxidx[9 * base + 3 * d] * inv_j[base] + dxidx[9 * n1 + 3 * d] * inv_j[n1]);
    double my = 0.5 * (dxidx[9 * base + 3 * d + 1] * inv_j[base] + dxidx[9 * n1 + 3 * d + 1] * inv_j[n1]);
    double mz = 0.5 * (dxidx[9 * base + 3 * d + 2] * inv_j[base] + dxidx[9 * n1 + 3 * d + 2] * inv_j[n1]);
    ora_roe_face(pl, pr, mx, my, mz, gamma, F);
}

/* order of the midpoint-to-node difference at node a (WCNS): 6th order
 * with faces a-5/2..a+5/2 all at fifth order, 4th order with faces
 * a-3/2..a+3/2 at third order or better, else 2nd order. */
int ora_wcns_diff_order(int a, int lo, int hi)
{
    if (a - 3 >= lo && a + 2 <= hi - 1 && face_order(a - 3, lo, hi) == 5 && face_order(a + 2, lo, hi) == 5) return 6;
    if (a - 2 >= lo && a + 1 <= hi - 1 && face_order(a - 2, lo, hi) >= 3 && face_order(a + 1, lo, hi) >= 3) return 4;
    return 2;
}

/* inviscid residual with scheme 1 (Steger-Warming + WENO5: conservative
 * finite difference of the reconstructed fluxes) or 2 (WCNS: WENO5
 * midpoint interpolation of the conservative variables, Roe flux at the
 * midpoint with the midpoint metric, explicit midpoint-to-node difference
 * of order 6 / 4 / 2 by distance to the extended range ends:
 * 75/64 d1 - 25/384 d3 + 3/640 d5, 9/8 d1 - 1/24 d3, d1, with
 * dn = F(a+n/2) - F(a-n/2)).  rhsi is zero outside the computed range. */
void ora_inviscid_ext(const ora_ext *e, double *rhsi, const double *q, const double *dxidx,
                      const double *inv_j, double gamma, int scheme)
{
    long N = ora_next(e);
    memset(rhsi, 0, 5 * N * sizeof(double));
    int lo3[3] = {e->is, e->js, e->ks}, hi3[3] = {e->ie, e->je, e->ke};
    for (int d = 0; d < 3; d++) {
        long st = d == 0 ? 1 : d == 1 ? (long)(e->ie - e->is + 1) : (long)(e->ie - e->is + 1) * (e->je - e->js + 1);
        int lo = lo3[d], hi = hi3[d];
        /* every node's fluxes are its own: the same results on any number
         * of threads (only the run time of the checker changes) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = e->ks1; k <= e->ke1; k++)
            for (int j = e->js1; j <= e->je1; j++)
                for (int i = e->is1; i <= e->ie1; i++) {
                    long n = IX(i, j, k);
                    int a = d == 0 ? i : d == 1 ? j : k;
                    double F[6][5]; /* faces a-3 .. a+2 (index f = face - a + 3) */
                    int dord = scheme == 2 ? ora_wcns_diff_order(a, lo, hi) : 2;
                    int f0 = dord == 6 ? 0 : dord == 4 ? 1 : 2, f1 = dord == 6 ? 5 : dord == 4 ? 4 : 3;
                    for (int f = f0; f <= f1; f++)
                        face_flux_ext(e, q, dxidx, inv_j, gamma, scheme, n + (long)(f - 3) * st, st, a + f - 3, lo, hi,
                                      d, F[f]);
                    for (int x = 0; x < 5; x++) {
                        double d1 = F[3][x] - F[2][x], dd;
                        if (dord == 6)
                            dd = 75.0 / 64.0 * d1 - 25.0 / 384.0 * (F[4][x] - F[1][x]) + 3.0 / 640.0 * (F[5][x] - F[0][x]);
                        else if (dord == 4)
                            dd = 9.0 / 8.0 * d1 - 1.0 / 24.0 * (F[4][x] - F[1][x]);
                        else
                            dd = d1;
                        rhsi[5 * n + x] -= dd;
                    }
                }
    }
}

/* Shu-Osher stage s (0,1,2): q = a_s qn + b_s (q + dt rhs / inv_j) on the
 * computed range, (a,b) = (0,1), (3/4,1/4), (1/3,2/3). */
void ora_rk3_stage(const ora_ext *e, double *q, const double *qn, const double *rhs, const double *dt,
                   const double *inv_j, int s)
{
    const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int k = e->ks1; k <= e->ke1; k++)
        for (int j = e->js1; j <= e->je1; j++)
            for (int i = e->is1; i <= e->ie1; i++) {
                long n = IX(i, j, k);
                for (int x = 0; x < 5; x++) {
                    double upd = q[5 * n + x] + dt[n] * rhs[5 * n + x] / inv_j[n];
                    q[5 * n + x] = A[s] * qn[5 * n + x] + B[s] * upd;
                }
            }
}

static void refresh(int nblk, ora_block *blk, const ora_config *c)
{
    ora_physical_bc(nblk, blk, c);
    ora_update_buffer(nblk, blk, c);
    ora_average(nblk, blk, c);
}

/* One explicit RK3 step with the given inviscid scheme (0 = the
 * reference's WENO3 + Roe, 1, 2 as ora_inviscid_ext).  dt is the block's
 * dt array: fixed (unsteady) or local (steady, find_dt with cfl) taken at
 * the first stage.  rhs keeps the first stage's residual (rhsi + rhsv). */
int ora_step_rk3(int nblk, ora_block *blk, ora_config *c, int scheme)
{
    double **qn = malloc(sizeof(double *) * nblk);
    double **tmp = malloc(sizeof(double *) * nblk);
    for (int m = 0; m < nblk; m++) {
        long N = ora_next(&blk[m].e);
        qn[m] = malloc(5 * N * sizeof(double));
        tmp[m] = malloc(5 * N * sizeof(double));
        memcpy(qn[m], blk[m].q, 5 * N * sizeof(double));
    }
    for (int s = 0; s < 3; s++) {
        if (s > 0) refresh(nblk, blk, c);
        for (int m = 0; m < nblk; m++) {
            ora_block *b = blk + m;
            const ora_ext *e = &b->e;
            long N = ora_next(e);
            ora_getpri(e, b->q, b->pri_v, c->gamma, c->cv);
            if (scheme == 0) ora_inviscid(e, b->rhsi, b->q, b->dxidx, b->inv_j, c->gamma);
            else ora_inviscid_ext(e, b->rhsi, b->q, b->dxidx, b->inv_j, c->gamma, scheme);
            if (c->iflag_solver == 0)
                ora_viscous(e, b->rhsv, b->pri_v, b->amu, b->vor, b->dudx, b->tao, b->dxidx, b->inv_j,
                            c->tinf, c->prl, c->prt, c->gamma, c->cv, c->re);
            else
                memset(b->rhsv, 0, 5 * N * sizeof(double));
            if (s == 0) {
                double st[4];
                ora_find_dt(e, b->dt, st, b->pri_v, b->amu, b->dxidx, c->re, &c->cfl, c->iflag_time);
                c->dtmax = st[0]; c->dtmin = st[1]; c->cflmax = st[2]; c->cflmin = st[3];
            }
            double *r = s == 0 ? b->rhs : tmp[m];
            memset(r, 0, 5 * N * sizeof(double));
            for (int k = e->ks1; k <= e->ke1; k++)
                for (int j = e->js1; j <= e->je1; j++)
                    for (int i = e->is1; i <= e->ie1; i++) {
                        long n = IX(i, j, k);
                        for (int x = 0; x < 5; x++) r[5 * n + x] = b->rhsi[5 * n + x] + b->rhsv[5 * n + x];
                    }
            ora_rk3_stage(e, b->q, qn[m], r, b->dt, b->inv_j, s);
        }
    }
    refresh(nblk, blk, c);
    for (int m = 0; m < nblk; m++) {
        ora_getpri(&blk[m].e, blk[m].q, blk[m].pri_v, c->gamma, c->cv);
        free(qn[m]);
        free(tmp[m]);
    }
    free(qn);
    free(tmp);
    return 0;
}
