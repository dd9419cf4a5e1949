/*
 * aerocfd CPU oracle (TEST INFRASTRUCTURE ONLY): physical boundary
 * conditions, inter-block halo fill, overlap averaging, residual norm and the
 * time-step driver.
 *
 * Pinned by: the compiled reference (oracle/_ref + oracle/ref_harness.f90
 * drive physical_bc, UpdateBuffer, average1/average2 and residual on the same
 * multi-block states; test_oracle_vs_ref.py), freestream invariance of the
 * farfield and inflow conditions, and the geometric pin: exchanging node
 * coordinates through the same mapping reproduces the neighbour's true
 * coordinates for every face x ori pairing (test_oracle_closed_forms.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acfd_oracle.h"

#define IXE(e, i, j, k) ora_ix((e), (i), (j), (k))

static const double nutinf = 1.341946; /* module_tursa.f90:10 */
static const double cv1 = 7.1;

/* Unit outward normal of a face from the metric row of its axis:
 * sub_bc_-10_farfield.f90:43-52 (face 1, negated) and :113-122 (face 2). */
static void face_normal(const double *dxidx, long n, int face, double *nx, double *ny, double *nz)
{
    int row = 3 * ((face - 1) / 2);
    const double *m = dxidx + 9 * n + row;
    double temp = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    temp = sqrt(temp);
    if (face % 2 == 1) {
        *nx = -m[0] / temp; *ny = -m[1] / temp; *nz = -m[2] / temp;
    } else {
        *nx = m[0] / temp; *ny = m[1] / temp; *nz = m[2] / temp;
    }
}

/* Iterate the face nodes of a subface. For face f the boundary index along
 * its axis is the subface's low (odd f) or high (even f) index; the interior
 * neighbour at depth d is one step further into the block. */
typedef struct { int i, j, k, di, dj, dk; } fnode;

static int face_nodes(int face, int is, int ie, int js, int je, int ks, int ke, fnode **out)
{
    int axis = (face - 1) / 2, minus = face % 2 == 1;
    int lo[3] = {is, js, ks}, hi[3] = {ie, je, ke};
    int fix = minus ? lo[axis] : hi[axis];
    lo[axis] = hi[axis] = fix;
    long cnt = (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    if (cnt <= 0) { *out = NULL; return 0; }
    fnode *f = malloc(cnt * sizeof(fnode));
    int s = minus ? 1 : -1;
    long c = 0;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                f[c].i = i; f[c].j = j; f[c].k = k;
                f[c].di = axis == 0 ? s : 0; f[c].dj = axis == 1 ? s : 0; f[c].dk = axis == 2 ? s : 0;
                c++;
            }
    *out = f;
    return (int)c;
}

/* sub_bc_-10_farfield.f90:7-459 */
void ora_bc_farfield(const ora_ext *e, double *q, double *tur, const double *dxidx,
                     const ora_config *c, int face, int is, int ie, int js, int je, int ks, int ke)
{
    double gamma = c->gamma;
    double d_inf = c->d0, u_inf = c->u0 * cos(c->aoa), v_inf = c->u0 * sin(c->aoa), w_inf = 0.0;
    double s_inf = c->p0 / pow(c->d0, gamma), c_inf = c->c0;
    fnode *fn; int nn = face_nodes(face, is, ie, js, je, ks, ke, &fn);
    (void)d_inf;
    for (int t = 0; t < nn; t++) {
        long nb = IXE(e, fn[t].i, fn[t].j, fn[t].k);
        long np = IXE(e, fn[t].i + fn[t].di, fn[t].j + fn[t].dj, fn[t].k + fn[t].dk);
        double nx, ny, nz;
        face_normal(dxidx, nb, face, &nx, &ny, &nz);
        const double *qp = q + 5 * np;
        double d_p = qp[0];
        double u_p = qp[1] / d_p, v_p = qp[2] / d_p, w_p = qp[3] / d_p;
        double p_p = (qp[4] - 0.5 * d_p * (u_p * u_p + v_p * v_p + w_p * w_p)) * (gamma - 1.0);
        double c_p = sqrt(gamma * p_p / d_p);
        double s_p = p_p / pow(d_p, gamma);
        double vn_inf = u_inf * nx + v_inf * ny + w_inf * nz;
        double vn_p = u_p * nx + v_p * ny + w_p * nz;
        double rm = vn_inf - 2.0 * c_inf / (gamma - 1.0);
        double rp = vn_p + 2.0 * c_p / (gamma - 1.0);
        double vn_b = 0.5 * (rp + rm);
        double s_b = 0, u_ref = 0, v_ref = 0, w_ref = 0, turval = 0;
        if (vn_b < 0.0) {
            s_b = s_inf; u_ref = u_inf; v_ref = v_inf; w_ref = w_inf; turval = nutinf;
        }
        if (vn_b >= 0.0) {
            s_b = s_p; u_ref = u_p; v_ref = v_p; w_ref = w_p; turval = tur ? tur[np] : 0.0;
        }
        double u_ref_t = u_ref * nx + v_ref * ny + w_ref * nz;
        double u_b = u_ref + (vn_b - u_ref_t) * nx;
        double v_b = v_ref + (vn_b - u_ref_t) * ny;
        double w_b = w_ref + (vn_b - u_ref_t) * nz;
        double d_b = pow(c_p * c_p / (gamma * s_b), 1.0 / (gamma - 1.0));
        double p_b = s_b * pow(d_b, gamma);
        double *qb = q + 5 * nb;
        qb[0] = d_b;
        qb[1] = d_b * u_b;
        qb[2] = d_b * v_b;
        qb[3] = d_b * w_b;
        qb[4] = p_b / (gamma - 1.0) + 0.5 * d_b * (u_b * u_b + v_b * v_b + w_b * w_b);
        if (c->iflag_turbulence >= 1 && tur) tur[nb] = turval;
    }
    free(fn);
}

/* sub_bc_-20_solidwall.f90:10-211 and sub_bc_-30_symm.f90:8-210 (same
 * arithmetic): copy the interior neighbour and remove its normal momentum. */
void ora_bc_slipwall(const ora_ext *e, double *q, double *tur, const double *dxidx, int turb,
                     int face, int is, int ie, int js, int je, int ks, int ke)
{
    fnode *fn; int nn = face_nodes(face, is, ie, js, je, ks, ke, &fn);
    for (int t = 0; t < nn; t++) {
        long nb = IXE(e, fn[t].i, fn[t].j, fn[t].k);
        long np = IXE(e, fn[t].i + fn[t].di, fn[t].j + fn[t].dj, fn[t].k + fn[t].dk);
        double nx, ny, nz;
        face_normal(dxidx, nb, face, &nx, &ny, &nz);
        const double *qp = q + 5 * np;
        double cu = qp[1] * nx + qp[2] * ny + qp[3] * nz;
        double *qb = q + 5 * nb;
        qb[0] = qp[0];
        qb[1] = qp[1] - cu * nx;
        qb[2] = qp[2] - cu * ny;
        qb[3] = qp[3] - cu * nz;
        qb[4] = qp[4];
        if (turb >= 1 && tur) tur[nb] = tur[np];
    }
    free(fn);
}

/* sub_bc_-20_solidwall.f90:213-326 */
void ora_bc_nonslipwall(const ora_ext *e, double *q, double *tur, int turb,
                        int face, int is, int ie, int js, int je, int ks, int ke)
{
    fnode *fn; int nn = face_nodes(face, is, ie, js, je, ks, ke, &fn);
    for (int t = 0; t < nn; t++) {
        const fnode *f = fn + t;
        long nb = IXE(e, f->i, f->j, f->k);
        long n1 = IXE(e, f->i + f->di, f->j + f->dj, f->k + f->dk);
        long n2 = IXE(e, f->i + 2 * f->di, f->j + 2 * f->dj, f->k + 2 * f->dk);
        long n3 = IXE(e, f->i + 3 * f->di, f->j + 3 * f->dj, f->k + 3 * f->dk);
        double *qb = q + 5 * nb;
        qb[0] = (11.0 * q[5 * n1] - 7.0 * q[5 * n2] + 2.0 * q[5 * n3]) / 6.0;
        qb[1] = 0.0;
        qb[2] = 0.0;
        qb[3] = 0.0;
        qb[4] = (11.0 * q[5 * n1 + 4] - 7.0 * q[5 * n2 + 4] + 2.0 * q[5 * n3 + 4]) / 6.0;
        if (turb >= 1 && tur) tur[nb] = 0.0;
    }
    free(fn);
}

/* sub_bc_-40_outflow.f90:9-129 */
void ora_bc_outflow(const ora_ext *e, double *q, double *tur, int turb,
                    int face, int is, int ie, int js, int je, int ks, int ke)
{
    fnode *fn; int nn = face_nodes(face, is, ie, js, je, ks, ke, &fn);
    for (int t = 0; t < nn; t++) {
        long nb = IXE(e, fn[t].i, fn[t].j, fn[t].k);
        long np = IXE(e, fn[t].i + fn[t].di, fn[t].j + fn[t].dj, fn[t].k + fn[t].dk);
        for (int v = 0; v < 5; v++) q[5 * nb + v] = q[5 * np + v];
        if (turb >= 1 && tur) tur[nb] = tur[np];
    }
    free(fn);
}

/* sub_bc_-50_inflow.f90:7-139 */
void ora_bc_inflow(const ora_ext *e, double *q, double *tur, const ora_config *c,
                   int face, int is, int ie, int js, int je, int ks, int ke)
{
    fnode *fn; int nn = face_nodes(face, is, ie, js, je, ks, ke, &fn);
    for (int t = 0; t < nn; t++) {
        long nb = IXE(e, fn[t].i, fn[t].j, fn[t].k);
        double *qb = q + 5 * nb;
        qb[0] = c->d0;
        qb[1] = c->d0 * c->u0 * cos(c->aoa);
        qb[2] = c->d0 * c->u0 * sin(c->aoa);
        qb[3] = 0.0;
        qb[4] = c->p0 / (c->gamma - 1.0) + 0.5 * (c->u0 * c->u0);
        if (c->iflag_turbulence >= 1 && tur) tur[nb] = nutinf;
    }
    free(fn);
}

/* sub_bc.f90:9-90 */
void ora_physical_bc(int nblk, ora_block *blk, const ora_config *c)
{
    for (int m = 0; m < nblk; m++) {
        ora_block *b = blk + m;
        const ora_ext *e = &b->e;
        for (int s = 0; s < b->nsub; s++) {
            const ora_subface *f = b->sub + s;
            switch (f->blk_t) {
            case -50: ora_bc_inflow(e, b->q, b->nut, c, f->face, f->is, f->ie, f->js, f->je, f->ks, f->ke); break;
            case -40: ora_bc_outflow(e, b->q, b->nut, c->iflag_turbulence, f->face, f->is, f->ie, f->js, f->je, f->ks, f->ke); break;
            case -30: ora_bc_slipwall(e, b->q, b->nut, b->dxidx, c->iflag_turbulence, f->face, f->is, f->ie, f->js, f->je, f->ks, f->ke); break;
            case -20:
                if (c->iflag_solver == 0)
                    ora_bc_nonslipwall(e, b->q, b->nut, c->iflag_turbulence, f->face, f->is, f->ie, f->js, f->je, f->ks, f->ke);
                else
                    ora_bc_slipwall(e, b->q, b->nut, b->dxidx, c->iflag_turbulence, f->face, f->is, f->ie, f->js, f->je, f->ks, f->ke);
                break;
            /* the farfield call passes the block's js for the subface's js
             * (sub_bc.f90:57-59) */
            case -10: ora_bc_farfield(e, b->q, b->nut, b->dxidx, c, f->face, f->is, f->ie, e->js, f->je, f->ks, f->ke); break;
            default: break;
            }
        }
    }
    if (c->iflag_turbulence == 2 && c->iflag_solver == 0) {
        for (int m = 0; m < nblk; m++) {
            ora_block *b = blk + m;
            long N = ora_next(&b->e);
            for (long n = 0; n < N; n++) {
                double d = b->pri_v[7 * n];
                double nut = b->nut[n];
                double leix = nut / b->amu[3 * n] * d;
                leix = fmax(leix, 0.001);
                double l3 = leix * leix * leix;
                double fv1 = l3 / (l3 + cv1 * cv1 * cv1);
                /* max(nut,1e-4) is overwritten by min(nut,1e6) of the old
                 * value (sub_bc.f90:80-81) */
                b->nut[n] = fmin(nut, 1.0e6);
                b->amu[3 * n + 1] = b->nut[n] * fv1 * d;
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * Inter-block pairing. For the receiving subface (m, s) the reference loads
 * the sender's buffer into ua(l,m) by the sender's face (face_t), permutes it
 * into ua2 by ori and stores ua2 into its own face
 * (sub_bc_updatebuffer.f90:261-539; identical in
 * sub_bc_updatebufferaverage.f90:244-431). This routine reproduces those
 * three steps on node coordinates: it returns, for every receiving face
 * node, the sender's face node. */
typedef struct { int ri, rj, rk; int si, sj, sk; } pairnode;

static int pair_nodes(const ora_block *blk, int m, int s, pairnode **out)
{
    const ora_subface *f = blk[m].sub + s;
    const ora_subface *t = blk[f->blk_t - 1].sub + (f->subface_t - 1);
    int n11 = t->ie - t->is + 1, n12 = t->je - t->js + 1, n13 = t->ke - t->ks + 1;
    int l1 = 0, m1 = 0;
    switch (t->face) {
    case 1: l1 = n13; m1 = n12; break;
    case 2: l1 = n12; m1 = n13; break;
    case 3: l1 = n11; m1 = n13; break;
    case 4: l1 = n13; m1 = n11; break;
    case 5: l1 = n12; m1 = n11; break;
    case 6: l1 = n11; m1 = n12; break;
    }
    /* ua(ll,mm) holds sender coordinates */
    int *ua = malloc(sizeof(int) * 3 * l1 * m1);
#define UA(ll, mm) (ua + 3 * (((mm) - 1) * l1 + ((ll) - 1)))
    for (int mm = 1; mm <= m1; mm++)
        for (int ll = 1; ll <= l1; ll++) {
            int *p = UA(ll, mm);
            switch (t->face) {
            case 1: p[0] = t->is; p[1] = t->js + mm - 1; p[2] = t->ks + ll - 1; break;
            case 2: p[0] = t->ie; p[1] = t->js + ll - 1; p[2] = t->ks + mm - 1; break;
            case 3: p[0] = t->is + ll - 1; p[1] = t->js; p[2] = t->ks + mm - 1; break;
            case 4: p[0] = t->is + mm - 1; p[1] = t->je; p[2] = t->ks + ll - 1; break;
            case 5: p[0] = t->is + mm - 1; p[1] = t->js + ll - 1; p[2] = t->ks; break;
            case 6: p[0] = t->is + ll - 1; p[1] = t->js + mm - 1; p[2] = t->ke; break;
            }
        }
    int l2 = (f->ori == 2 || f->ori == 4) ? m1 : l1;
    int m2 = (f->ori == 2 || f->ori == 4) ? l1 : m1;
    int *ua2 = malloc(sizeof(int) * 3 * l2 * m2);
#define UA2(ll, mm) (ua2 + 3 * (((mm) - 1) * l2 + ((ll) - 1)))
    for (int mm = 1; mm <= m2; mm++)
        for (int ll = 1; ll <= l2; ll++) {
            const int *src = NULL;
            switch (f->ori) {
            case 1: src = UA(ll, m1 - mm + 1); break;
            case 2: src = UA(mm, ll); break;
            case 3: src = UA(l1 - ll + 1, mm); break;
            case 4: src = UA(l1 - mm + 1, m1 - ll + 1); break;
            }
            memcpy(UA2(ll, mm), src, 3 * sizeof(int));
        }
    pairnode *pn = malloc(sizeof(pairnode) * l2 * m2);
    int c = 0;
    /* store loops in the reference's order (outer index first) */
    for (int b = 1; b <= m2; b++)
        for (int a = 1; a <= l2; a++) {
            pairnode *p = pn + c++;
            const int *src = UA2(a, b);
            p->si = src[0]; p->sj = src[1]; p->sk = src[2];
            switch (f->face) {
            case 1: p->ri = f->is; p->rj = f->js + b - 1; p->rk = f->ks + a - 1; break;
            case 2: p->ri = f->ie; p->rj = f->js + a - 1; p->rk = f->ks + b - 1; break;
            case 3: p->ri = f->is + a - 1; p->rj = f->js; p->rk = f->ks + b - 1; break;
            case 4: p->ri = f->is + b - 1; p->rj = f->je; p->rk = f->ks + a - 1; break;
            case 5: p->ri = f->is + b - 1; p->rj = f->js + a - 1; p->rk = f->ks; break;
            case 6: p->ri = f->is + a - 1; p->rj = f->js + b - 1; p->rk = f->ke; break;
            }
        }
#undef UA
#undef UA2
    free(ua); free(ua2);
    *out = pn;
    return c;
}

/* inward unit step of a face (direction from the face into its block) */
static void inward(int face, int *di, int *dj, int *dk)
{
    int axis = (face - 1) / 2, s = face % 2 == 1 ? 1 : -1;
    *di = axis == 0 ? s : 0; *dj = axis == 1 ? s : 0; *dk = axis == 2 ? s : 0;
}

/* sub_bc_updatebuffer.f90:10-548 on one rank: every pack happens before any
 * unpack, so the fill is a pure gather from the pre-exchange state. */
void ora_update_buffer(int nblk, ora_block *blk, const ora_config *c)
{
    int sa = c->iflag_turbulence == 2;
    /* gather phase into per-subface staging */
    typedef struct { int m, s, cnt; double *val; long *dst; } stage;
    int tot = 0;
    for (int m = 0; m < nblk; m++) tot += blk[m].nsub;
    stage *st = calloc(tot, sizeof(stage));
    int ns = 0;
    for (int m = 0; m < nblk; m++)
        for (int s = 0; s < blk[m].nsub; s++) {
            const ora_subface *f = blk[m].sub + s;
            if (f->blk_t <= 0) continue;
            pairnode *pn; int cnt = pair_nodes(blk, m, s, &pn);
            const ora_block *sb = blk + (f->blk_t - 1);
            const ora_subface *t = sb->sub + (f->subface_t - 1);
            int sdi, sdj, sdk, rdi, rdj, rdk;
            inward(t->face, &sdi, &sdj, &sdk);
            inward(f->face, &rdi, &rdj, &rdk);
            stage *g = st + ns++;
            g->m = m; g->s = s; g->cnt = cnt;
            g->val = malloc(sizeof(double) * cnt * 2 * 6);
            g->dst = malloc(sizeof(long) * cnt * 2);
            for (int x = 0; x < cnt; x++)
                for (int nb = 1; nb <= 2; nb++) {
                    long src = IXE(&sb->e, pn[x].si + nb * sdi, pn[x].sj + nb * sdj, pn[x].sk + nb * sdk);
                    long dst = IXE(&blk[m].e, pn[x].ri - nb * rdi, pn[x].rj - nb * rdj, pn[x].rk - nb * rdk);
                    double *v = g->val + 12 * x + 6 * (nb - 1);
                    for (int c5 = 0; c5 < 5; c5++) v[c5] = sb->q[5 * src + c5];
                    v[5] = sa ? sb->nut[src] : 0.0;
                    g->dst[2 * x + nb - 1] = dst;
                }
            free(pn);
        }
    for (int x = 0; x < ns; x++) {
        stage *g = st + x;
        ora_block *b = blk + g->m;
        for (int y = 0; y < g->cnt; y++)
            for (int nb = 0; nb < 2; nb++) {
                long dst = g->dst[2 * y + nb];
                const double *v = g->val + 12 * y + 6 * nb;
                for (int c5 = 0; c5 < 5; c5++) b->q[5 * dst + c5] = v[c5];
                if (sa) b->nut[dst] = v[5];
            }
        free(g->val); free(g->dst);
    }
    free(st);
}

/* sub_bc_updatebufferaverage.f90:10-562: average1 gathers the neighbour's
 * face node of every pairing, then average2 applies q = (q + buffer)*0.5
 * subface by subface in order, so that a node shared by several subfaces is
 * averaged repeatedly. */
void ora_average(int nblk, ora_block *blk, const ora_config *c)
{
    int sa = c->iflag_turbulence == 2;
    typedef struct { int m, cnt; double *val; long *dst; } stage;
    int tot = 0;
    for (int m = 0; m < nblk; m++) tot += blk[m].nsub;
    stage *st = calloc(tot, sizeof(stage));
    int ns = 0;
    for (int m = 0; m < nblk; m++)
        for (int s = 0; s < blk[m].nsub; s++) {
            const ora_subface *f = blk[m].sub + s;
            if (f->blk_t <= 0) continue;
            pairnode *pn; int cnt = pair_nodes(blk, m, s, &pn);
            const ora_block *sb = blk + (f->blk_t - 1);
            stage *g = st + ns++;
            g->m = m; g->cnt = cnt;
            g->val = malloc(sizeof(double) * cnt * 6);
            g->dst = malloc(sizeof(long) * cnt);
            for (int x = 0; x < cnt; x++) {
                long src = IXE(&sb->e, pn[x].si, pn[x].sj, pn[x].sk);
                g->dst[x] = IXE(&blk[m].e, pn[x].ri, pn[x].rj, pn[x].rk);
                for (int c5 = 0; c5 < 5; c5++) g->val[6 * x + c5] = sb->q[5 * src + c5];
                g->val[6 * x + 5] = sa ? sb->nut[src] : 0.0;
            }
            free(pn);
        }
    for (int x = 0; x < ns; x++) {
        stage *g = st + x;
        ora_block *b = blk + g->m;
        for (int y = 0; y < g->cnt; y++) {
            long d = g->dst[y];
            for (int c5 = 0; c5 < 5; c5++) b->q[5 * d + c5] = (b->q[5 * d + c5] + g->val[6 * y + c5]) * 0.5;
            if (sa) b->nut[d] = (b->nut[d] + g->val[6 * y + 5]) * 0.5;
        }
        free(g->val); free(g->dst);
    }
    free(st);
}

/* sub_output_residual.f90:24-56 */
void ora_residual(int nblk, const ora_block *blk, long totalnodes, double *out)
{
    double r[5] = {0, 0, 0, 0, 0}, rt = 0.0;
    for (int m = 0; m < nblk; m++) {
        const ora_block *b = blk + m;
        const ora_ext *e = &b->e;
        for (int k = e->ks1; k <= e->ke1; k++)
            for (int j = e->js1; j <= e->je1; j++)
                for (int i = e->is1; i <= e->ie1; i++) {
                    long n = IXE(e, i, j, k);
                    for (int v = 0; v < 5; v++) {
                        double x = b->rhs[5 * n + v] * b->dt[n] / b->inv_j[n];
                        r[v] = r[v] + x * x;
                    }
                    double y = (b->sa_rhs ? b->sa_rhs[n] : 0.0) * b->dt[n];
                    rt = rt + y * y;
                }
    }
    for (int v = 0; v < 5; v++) out[v] = sqrt(r[v] / (double)totalnodes);
    out[5] = sqrt(rt / (double)totalnodes);
    out[6] = 0.0;
}

static void halo_refresh(int nblk, ora_block *blk, const ora_config *c)
{
    ora_physical_bc(nblk, blk, c);
    ora_update_buffer(nblk, blk, c);
    ora_average(nblk, blk, c);
}

void ora_prepare(int nblk, ora_block *blk, ora_config *c)
{
    halo_refresh(nblk, blk, c);
}

static void residual_stage(ora_block *b, ora_config *c)
{
    const ora_ext *e = &b->e;
    ora_getpri(e, b->q, b->pri_v, c->gamma, c->cv);
    ora_inviscid(e, b->rhsi, b->q, b->dxidx, b->inv_j, c->gamma);
    if (c->iflag_solver == 0)
        ora_viscous(e, b->rhsv, b->pri_v, b->amu, b->vor, b->dudx, b->tao, b->dxidx, b->inv_j,
                    c->tinf, c->prl, c->prt, c->gamma, c->cv, c->re);
    double st[4];
    ora_find_dt(e, b->dt, st, b->pri_v, b->amu, b->dxidx, c->re, &c->cfl, c->iflag_time);
    c->dtmax = st[0]; c->dtmin = st[1]; c->cflmax = st[2]; c->cflmin = st[3];
}

/* main.f90:343-573 */
int ora_step(int nblk, ora_block *blk, ora_config *c)
{
    int sa = c->iflag_turbulence == 2 && c->iflag_solver == 0;
    if (c->iflag_timeadvance == 2) {
        for (int m = 0; m < nblk; m++) {
            ora_block *b = blk + m;
            residual_stage(b, c);
            ora_implicit1st(&b->e, b->q, b->dt, b->pri_v, b->rhs, b->rhsi, b->rhsv, b->dxidx,
                            b->inv_j, c->gamma, c->cfl);
            if (sa)
                ora_sa_model(&b->e, b->nut, b->sa_rhs, b->dt, b->pri_v, b->amu, b->dudx, b->vor,
                             b->dxidx, b->inv_j, b->aalpha, b->bbeta, b->ggamma, b->length,
                             b->dist, b->spacing, c->re, c->cfl, c->iflag_des, c->iflag_time);
        }
        halo_refresh(nblk, blk, c);
        for (int m = 0; m < nblk; m++) ora_getpri(&blk[m].e, blk[m].q, blk[m].pri_v, c->gamma, c->cv);
        return 0;
    }
    if (c->iflag_timeadvance == 3) {
        for (int sub = 1; sub <= 3; sub++) {
            halo_refresh(nblk, blk, c);
            for (int m = 0; m < nblk; m++) {
                ora_block *b = blk + m;
                residual_stage(b, c);
                ora_crank2nd(&b->e, b->q, b->q0, b->dt, b->pri_v, b->rhs, b->rhsi, b->rhsv,
                             b->dxidx, b->inv_j, c->gamma);
            }
        }
        for (int m = 0; m < nblk; m++) ora_update_q0(&blk[m].e, blk[m].q0, blk[m].q);
        if (sa)
            for (int m = 0; m < nblk; m++) {
                ora_block *b = blk + m;
                ora_sa_model(&b->e, b->nut, b->sa_rhs, b->dt, b->pri_v, b->amu, b->dudx, b->vor,
                             b->dxidx, b->inv_j, b->aalpha, b->bbeta, b->ggamma, b->length,
                             b->dist, b->spacing, c->re, c->cfl, c->iflag_des, c->iflag_time);
            }
        halo_refresh(nblk, blk, c);
        for (int m = 0; m < nblk; m++) ora_getpri(&blk[m].e, blk[m].q, blk[m].pri_v, c->gamma, c->cv);
        return 0;
    }
    return -1;
}
