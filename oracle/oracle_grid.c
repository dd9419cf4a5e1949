/*
 * Grid pre-processing and wall forces -- TEST INFRASTRUCTURE ONLY (see
 * acfd_oracle.h).
 *
 *   ora_gridtransformation  sub_init_gridtransformation.f90:27-874 (3-D branch)
 *   ora_spacing             sub_init_spacing.f90:16-110
 *   ora_walldist            sub_init_walldist.f90:9-85 (the distance part)
 *   ora_forces_block        sub_output_forceintegration.f90:182-509 (3-D)
 *
 * Pinning: tests/test_oracle_vs_ref.py requires bitwise equality with the
 * reference's own routines (compiled from its sources into oracle/_ref) on
 * curvilinear multi-block grids with halos and wall faces, and the lift and
 * drag coefficients the reference's clcd_output writes;
 * tests/test_golden.py holds the oracle to the committed fixture of the
 * same cases; closed forms (tests/test_oracle_grid.py): exact metrics of
 * affine grids, edge lengths of uniform grids, the wall distance of a plane
 * wall and by brute force.  ora_forces_block has no closed form to meet:
 * the reference's face-normal average mixes metric components (see the
 * function), so it is pinned by the reference's own coefficients only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "acfd_oracle.h"

#define NI(e) ((e)->ie - (e)->is + 1)
#define NJ(e) ((e)->je - (e)->js + 1)
#define NK(e) ((e)->ke - (e)->ks + 1)
#define IX(i, j, k) ora_ix(e, (i), (j), (k))

static int lo_of(const ora_ext *e, int a) { return a == 0 ? e->is : a == 1 ? e->js : e->ks; }
static int hi_of(const ora_ext *e, int a) { return a == 0 ? e->ie : a == 1 ? e->je : e->ke; }
static int lo0_of(const ora_ext *e, int a) { return a == 0 ? e->is0 : a == 1 ? e->js0 : e->ks0; }
static int hi0_of(const ora_ext *e, int a) { return a == 0 ? e->ie0 : a == 1 ? e->je0 : e->ke0; }

/* Number of coordinates of (i,j,k) outside the physical range. */
static int halo_count(const ora_ext *e, int i, int j, int k)
{
    return (i < e->is0 || i > e->ie0) + (j < e->js0 || j > e->je0) + (k < e->ks0 || k > e->ke0);
}

/* Derivative along axis `a` of the nc-component AoS field v at the nodes
 * where gridtransformation defines it (:103-352 for the coordinates,
 * :564-807 for the temporaries): along a line whose two other coordinates
 * are physical the whole extended line, one-sided at is/ie; on the face
 * halo slabs (exactly one other coordinate outside the physical range) the
 * physical range, one-sided at is0/ie0.  Other nodes are left untouched. */
static void deriv(const ora_ext *e, const double *v, int nc, int a, double *out)
{
    for (int k = e->ks; k <= e->ke; k++)
        for (int j = e->js; j <= e->je; j++)
            for (int i = e->is; i <= e->ie; i++) {
                const int p[3] = {i, j, k};
                const int others = halo_count(e, a == 0 ? e->is0 : i, a == 1 ? e->js0 : j, a == 2 ? e->ks0 : k);
                int lo, hi;
                if (others == 0) {
                    lo = lo_of(e, a);
                    hi = hi_of(e, a);
                } else if (others == 1) {
                    lo = lo0_of(e, a);
                    hi = hi0_of(e, a);
                    if (p[a] < lo || p[a] > hi) continue;
                } else {
                    continue;
                }
                const int di = a == 0, dj = a == 1, dk = a == 2;
                const long n = IX(i, j, k), np = IX(i + di, j + dj, k + dk), nm = IX(i - di, j - dj, k - dk);
                for (int c = 0; c < nc; c++) {
                    if (p[a] == lo) out[nc * n + c] = v[nc * np + c] - v[nc * n + c];
                    else if (p[a] == hi) out[nc * n + c] = v[nc * n + c] - v[nc * nm + c];
                    else out[nc * n + c] = (v[nc * np + c] - v[nc * nm + c]) / 2.0;
                }
            }
}

/* sub_init_gridtransformation.f90:27-874, 3-D.  x, y, z scalar fields;
 * dxidx(9,...), inv_j, alpha/beta/gamma(3,...) AoS.  Every output is
 * written at the nodes with at most one coordinate outside the physical
 * range (the box and its six face halo slabs); at the other (edge and
 * corner halo) nodes the reference works on uninitialised temporaries, and
 * here those outputs are left untouched. */
void ora_gridtransformation(const ora_ext *e, const double *x, const double *y, const double *z, double *dxidx,
                            double *inv_j, double *alpha, double *beta, double *gamma)
{
    const long N = ora_next(e);
    double *xyz = calloc(3 * N, sizeof(double));
    double *dxi = calloc(3 * N, sizeof(double)), *deta = calloc(3 * N, sizeof(double));
    double *dzeta = calloc(3 * N, sizeof(double));
    double *txi = calloc(6 * N, sizeof(double)), *teta = calloc(6 * N, sizeof(double));
    double *tzeta = calloc(6 * N, sizeof(double));
    double *txx = calloc(6 * N, sizeof(double)), *tee = calloc(6 * N, sizeof(double));
    double *tzz = calloc(6 * N, sizeof(double));
    for (long n = 0; n < N; n++) {
        xyz[3 * n] = x[n];
        xyz[3 * n + 1] = y[n];
        xyz[3 * n + 2] = z[n];
    }
    /* :103-352 */
    deriv(e, xyz, 3, 0, dxi);
    deriv(e, xyz, 3, 1, deta);
    deriv(e, xyz, 3, 2, dzeta);
    /* 1/J (:356-529): the box and the face halo slabs */
    for (int k = e->ks; k <= e->ke; k++)
        for (int j = e->js; j <= e->je; j++)
            for (int i = e->is; i <= e->ie; i++) {
                if (halo_count(e, i, j, k) > 1) continue;
                const long n = IX(i, j, k);
                const double dxdxi0 = dxi[3 * n], dxdeta0 = deta[3 * n], dxdzeta0 = dzeta[3 * n];
                const double dydxi0 = dxi[3 * n + 1], dydeta0 = deta[3 * n + 1], dydzeta0 = dzeta[3 * n + 1];
                const double dzdxi0 = dxi[3 * n + 2], dzdeta0 = deta[3 * n + 2], dzdzeta0 = dzeta[3 * n + 2];
                inv_j[n] = dxdxi0 * dydeta0 * dzdzeta0 + dxdeta0 * dydzeta0 * dzdxi0 + dxdzeta0 * dydxi0 * dzdeta0
                         - dxdzeta0 * dydeta0 * dzdxi0 - dxdxi0 * dydzeta0 * dzdeta0 - dxdeta0 * dydxi0 * dzdzeta0;
            }
    /* temporaries (:560-585) */
    for (long n = 0; n < N; n++) {
        const double *X = xyz + 3 * n;
        txi[6 * n + 0] = dzeta[3 * n + 1] * X[2];
        txi[6 * n + 1] = dzeta[3 * n + 2] * X[0];
        txi[6 * n + 2] = dzeta[3 * n + 0] * X[1];
        txi[6 * n + 3] = deta[3 * n + 1] * X[2];
        txi[6 * n + 4] = deta[3 * n + 2] * X[0];
        txi[6 * n + 5] = deta[3 * n + 0] * X[1];
        teta[6 * n + 0] = dzeta[3 * n + 1] * X[2];
        teta[6 * n + 1] = dzeta[3 * n + 2] * X[0];
        teta[6 * n + 2] = dzeta[3 * n + 0] * X[1];
        teta[6 * n + 3] = dxi[3 * n + 1] * X[2];
        teta[6 * n + 4] = dxi[3 * n + 2] * X[0];
        teta[6 * n + 5] = dxi[3 * n + 0] * X[1];
        tzeta[6 * n + 0] = deta[3 * n + 1] * X[2];
        tzeta[6 * n + 1] = deta[3 * n + 2] * X[0];
        tzeta[6 * n + 2] = deta[3 * n + 0] * X[1];
        tzeta[6 * n + 3] = dxi[3 * n + 1] * X[2];
        tzeta[6 * n + 4] = dxi[3 * n + 2] * X[0];
        tzeta[6 * n + 5] = dxi[3 * n + 0] * X[1];
    }
    /* :587-807 */
    deriv(e, txi, 6, 0, txx);
    deriv(e, teta, 6, 1, tee);
    deriv(e, tzeta, 6, 2, tzz);
    /* :809-835 and alpha/beta/gamma :840-866 */
    for (int k = e->ks; k <= e->ke; k++)
        for (int j = e->js; j <= e->je; j++)
            for (int i = e->is; i <= e->ie; i++) {
                if (halo_count(e, i, j, k) > 1) continue;
                const long n = IX(i, j, k);
                const double ij = inv_j[n];
                double *m = dxidx + 9 * n;
                m[0] = (tzz[6 * n + 0] - tee[6 * n + 0]) / ij;
                m[1] = (tzz[6 * n + 1] - tee[6 * n + 1]) / ij;
                m[2] = (tzz[6 * n + 2] - tee[6 * n + 2]) / ij;
                m[3] = (txx[6 * n + 0] - tzz[6 * n + 3]) / ij;
                m[4] = (txx[6 * n + 1] - tzz[6 * n + 4]) / ij;
                m[5] = (txx[6 * n + 2] - tzz[6 * n + 5]) / ij;
                m[6] = (tee[6 * n + 3] - txx[6 * n + 3]) / ij;
                m[7] = (tee[6 * n + 4] - txx[6 * n + 4]) / ij;
                m[8] = (tee[6 * n + 5] - txx[6 * n + 5]) / ij;
                if (alpha) {
                    const double a0 = m[0], a1 = m[1], a2 = m[2], b0 = m[3], b1 = m[4], b2 = m[5];
                    const double c0 = m[6], c1 = m[7], c2 = m[8];
                    alpha[3 * n + 0] = (a0 * a0 + a1 * a1 + a2 * a2) * ij;
                    alpha[3 * n + 1] = (a0 * b0 + a1 * b1 + a2 * b2) * ij;
                    alpha[3 * n + 2] = (a0 * c0 + a1 * c1 + a2 * c2) * ij;
                    beta[3 * n + 0] = (b0 * a0 + b1 * a1 + b2 * a2) * ij;
                    beta[3 * n + 1] = (b0 * b0 + b1 * b1 + b2 * b2) * ij;
                    beta[3 * n + 2] = (b0 * c0 + b1 * c1 + b2 * c2) * ij;
                    gamma[3 * n + 0] = (c0 * a0 + c1 * a1 + c2 * a2) * ij;
                    gamma[3 * n + 1] = (c0 * b0 + c1 * b1 + c2 * b2) * ij;
                    gamma[3 * n + 2] = (c0 * c0 + c1 * c1 + c2 * c2) * ij;
                }
            }
    free(xyz); free(dxi); free(deta); free(dzeta);
    free(txi); free(teta); free(tzeta); free(txx); free(tee); free(tzz);
}

/* sub_init_spacing.f90:37-108: edge lengths along xi, eta, zeta on the
 * physical range (forward differences, the last node copies its
 * neighbour), their maximum and the cell volume.  Outputs outside the
 * physical range are left untouched. */
void ora_spacing(const ora_ext *e, const double *x, const double *y, const double *z, double *spci, double *spcj,
                 double *spck, double *spacing, double *vol)
{
    const int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double *spc[3] = {spci, spcj, spck};
    for (int a = 0; a < 3; a++) {
        if (a == 2 && !(e->ke0 > e->ks0)) continue;
        const int hi = hi0_of(e, a);
        for (int k = e->ks0; k <= e->ke0; k++)
            for (int j = e->js0; j <= e->je0; j++)
                for (int i = e->is0; i <= e->ie0; i++) {
                    const int p = a == 0 ? i : a == 1 ? j : k;
                    const long n = IX(i, j, k);
                    if (p < hi) {
                        const long m = IX(i + d[a][0], j + d[a][1], k + d[a][2]);
                        const double dx = x[m] - x[n], dy = y[m] - y[n], dz = z[m] - z[n];
                        spc[a][n] = sqrt(dx * dx + dy * dy + dz * dz);
                    }
                }
        for (int k = e->ks0; k <= e->ke0; k++)
            for (int j = e->js0; j <= e->je0; j++)
                for (int i = e->is0; i <= e->ie0; i++) {
                    const int p = a == 0 ? i : a == 1 ? j : k;
                    if (p == hi) spc[a][IX(i, j, k)] = spc[a][IX(i - d[a][0], j - d[a][1], k - d[a][2])];
                }
    }
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                const long n = IX(i, j, k);
                double s = spci[n];
                if (spcj[n] > s) s = spcj[n];
                if (spck[n] > s) s = spck[n];
                spacing[n] = s;
                if (vol) vol[n] = e->ke0 == e->ks0 ? spci[n] * spcj[n] : spci[n] * spcj[n] * spck[n];
            }
}

/* sub_init_walldist.f90:30-85: the distance of every physical node of a
 * block to the nearest of nw wall points (xw, yw, zw: the face nodes of all
 * wall subfaces of all blocks); 1e9 when there are none.  Halo nodes are
 * left untouched. */
void ora_walldist(const ora_ext *e, const double *x, const double *y, const double *z, long nw, const double *xw,
                  const double *yw, const double *zw, double *dist)
{
    for (int k = e->ks0; k <= e->ke0; k++)
        for (int j = e->js0; j <= e->je0; j++)
            for (int i = e->is0; i <= e->ie0; i++) {
                const long n = IX(i, j, k);
                const double xx = x[n], yy = y[n], zz = z[n];
                double w = 1.0e9;
                for (long p = 0; p < nw; p++) {
                    const double dx = xx - xw[p], dy = yy - yw[p], dz = zz - zw[p];
                    const double t = sqrt(dx * dx + dy * dy + dz * dz);
                    if (t <= w) w = t;
                }
                dist[n] = w;
            }
}

/* force_compute (sub_output_forceintegration.f90:481-509) */
static void force_compute(double *fp, double *fv, double uc, double vc, double wc, double pc, double nx, double ny,
                          double nz, double s, double dn, double pinf, double miuc, double re)
{
    fp[0] = (pc - pinf) * s * nx;
    fp[1] = (pc - pinf) * s * ny;
    fp[2] = (pc - pinf) * s * nz;
    const double un = uc * nx + vc * ny + wc * nz;
    fv[0] = miuc * (uc - un * nx) / dn * s / re;
    fv[1] = miuc * (vc - un * ny) / dn * s / re;
    fv[2] = miuc * (wc - un * nz) / dn * s / re;
}

/* The wall forces of one block (clcd_output, sub_output_forceintegration.f90:
 * 182-453, 3-D): every cell of every wall subface (blk_t = -20), the
 * velocity from the first interior layer and the pressure and viscosity on
 * the wall, averaged over the cell's four nodes, against the face-normal
 * vector averaged from the unit metric of the four interior nodes -- with
 * the reference's own mix of components in that average (the second pair
 * of nodes takes the next component), reproduced as is.  Accumulates into
 * fp[3], fv[3] in the reference's order (subfaces, then the outer and inner
 * cell loops), and the magnitudes of the terms into afp[3], afv[3]. */
void ora_forces_block(const ora_ext *e, int nsub, const ora_subface *sub, const double *pri_v, const double *dxidx,
                      const double *amu, const double *spci, const double *spcj, const double *spck, double pinf,
                      double re, double *fp, double *fv, double *afp, double *afv)
{
    for (int s = 0; s < nsub; s++) {
        const ora_subface *b = sub + s;
        if (b->blk_t != -20) continue;
        const int face = b->face;
        const int d = (face - 1) / 2;          /* normal axis */
        const int c0 = 3 * d;                  /* its metric row */
        const double sg = (face & 1) ? -1.0 : 1.0;
        /* tangential axes t1 < t2 and their ranges */
        const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        const int r[3][2] = {{b->is, b->ie}, {b->js, b->je}, {b->ks, b->ke}};
        const int wall = (face & 1) ? r[d][0] : r[d][1];
        const int in = (face & 1) ? wall + 1 : wall - 1;
        for (int v = r[t2][0]; v <= r[t2][1] - 1; v++)
            for (int u = r[t1][0]; u <= r[t1][1] - 1; u++) {
                long W[4], I[4];  /* corners 00, 10, 01, 11 on the wall and interior layers */
                for (int c = 0; c < 4; c++) {
                    int p[3];
                    p[d] = wall;
                    p[t1] = u + (c & 1);
                    p[t2] = v + (c >> 1);
                    W[c] = ora_ix(e, p[0], p[1], p[2]);
                    p[d] = in;
                    I[c] = ora_ix(e, p[0], p[1], p[2]);
                }
                const double uc = (pri_v[7 * I[0] + 1] + pri_v[7 * I[1] + 1] + pri_v[7 * I[2] + 1] + pri_v[7 * I[3] + 1]) * 0.25;
                const double vc = (pri_v[7 * I[0] + 2] + pri_v[7 * I[1] + 2] + pri_v[7 * I[2] + 2] + pri_v[7 * I[3] + 2]) * 0.25;
                const double wc = (pri_v[7 * I[0] + 3] + pri_v[7 * I[1] + 3] + pri_v[7 * I[2] + 3] + pri_v[7 * I[3] + 3]) * 0.25;
                const double pc = (pri_v[7 * W[0] + 4] + pri_v[7 * W[1] + 4] + pri_v[7 * W[2] + 4] + pri_v[7 * W[3] + 4]) * 0.25;
                double t[4];
                for (int c = 0; c < 4; c++) {
                    const double *m = dxidx + 9 * I[c] + c0;
                    t[c] = 1.0 / sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                }
#define M(c, k) dxidx[9 * I[c] + c0 + (k)]
                const double nx = sg * (M(0, 0) * t[0] + M(1, 0) * t[1] + M(2, 1) * t[2] + M(3, 1) * t[3]) * 0.25;
                const double ny = sg * (M(0, 2) * t[0] + M(1, 2) * t[1] + M(2, 0) * t[2] + M(3, 0) * t[3]) * 0.25;
                const double nz = sg * (M(0, 1) * t[0] + M(1, 1) * t[1] + M(2, 2) * t[2] + M(3, 2) * t[3]) * 0.25;
#undef M
                const double miuc = (amu[3 * W[0] + 2] + amu[3 * W[1] + 2] + amu[3 * W[2] + 2] + amu[3 * W[3] + 2]) * 0.25;
                const long w0 = W[0];
                const double area = d == 0 ? spcj[w0] * spck[w0] : d == 1 ? spci[w0] * spck[w0] : spci[w0] * spcj[w0];
                const double dn = d == 0 ? spci[w0] : d == 1 ? spcj[w0] : spck[w0];
                double f1[3], f2[3];
                force_compute(f1, f2, uc, vc, wc, pc, nx, ny, nz, area, dn, pinf, miuc, re);
                for (int x = 0; x < 3; x++) {
                    fp[x] = fp[x] + f1[x];
                    fv[x] = fv[x] + f2[x];
                    afp[x] += fabs(f1[x]);
                    afv[x] += fabs(f2[x]);
                }
            }
    }
}
