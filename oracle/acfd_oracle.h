/*
 * aerocfd CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's per-block residual-and-update
 * path (jiamin0329/aerocfd, Fortran 90).  It is compiled with
 * -O2 -ffp-contract=off so that every expression is evaluated in the same
 * order, without fused multiply-adds, as the reference built for baseline
 * x86-64.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker: the product (aerocfd_amd) never
 * links or calls it.
 *
 * Layout: identical to the reference -- Fortran column-major, component
 * first (AoS): q(5,is:ie,js:je,ks:ke) is q[5*IX(i,j,k) + (v-1)].
 *
 * Pinning (what each function is checked against) is listed in the header
 * comment of each source file.
 */
#ifndef ACFD_ORACLE_H
#define ACFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Index ranges of one block (module_block.f90:23-25): extended (is..ke),
 * physical (is0..ke0) and computed (is1..ke1). */
typedef struct {
    int is, ie, js, je, ks, ke;
    int is0, ie0, js0, je0, ks0, ke0;
    int is1, ie1, js1, je1, ks1, ke1;
} ora_ext;

/* One subface record of bcinfo (sub_init_readbcinfo.f90:96-98). Block and
 * subface ids are 1-based as in the reference. */
typedef struct {
    int f_no, face;
    int is, ie, js, je, ks, ke;
    int blk_t, subface_t, ori;
} ora_subface;

/* Per-block state arrays (all AoS, extended extents). Any pointer may be
 * NULL when the configuration does not use it. */
typedef struct {
    ora_ext e;
    int nsub;
    const ora_subface *sub;
    double *q;      /* 5 */
    double *q0;     /* 2*5, q0(-1:0,5,...) */
    double *pri_v;  /* 7 */
    double *rhsi, *rhsv, *rhs; /* 5 each */
    double *dudx;   /* 9 */
    double *tao;    /* 6 */
    double *vor;    /* 1 */
    double *amu;    /* 3 */
    double *dt;     /* 1 */
    double *dxidx;  /* 9 */
    double *inv_j;  /* 1 */
    double *aalpha, *bbeta, *ggamma; /* 3 each */
    double *dist, *spacing, *length; /* 1 each */
    double *nut, *sa_rhs;            /* 1 each */
} ora_block;

/* Flow constants and flags (module_flags.f90, module_ns_consts.f90,
 * sub_init_flowfield.f90:30-39). */
typedef struct {
    double gamma, cv, re, tinf, prl, prt;
    double d0, u0, p0, c0, aoa;   /* non-dimensional freestream, aoa in rad */
    double cfl;                   /* mutated by find_dt in unsteady mode */
    int iflag_solver;             /* 0 = NS, 1 = Euler */
    int iflag_time;               /* 0 = steady, 1 = unsteady */
    int iflag_turbulence;         /* 0 = laminar, 2 = SA */
    int iflag_des;                /* 0, 1 = DES97, 2 = DDES */
    int iflag_timeadvance;        /* 2 = 1st implicit (LU-SGS), 3 = BDF2 */
    long totalnodes;
    /* statistics written by find_dt (main.f90:589-592) */
    double dtmax, dtmin, cflmax, cflmin;
} ora_config;

long ora_ix(const ora_ext *e, int i, int j, int k);
long ora_next(const ora_ext *e);

/* sub_getpri_v.f90:19-63; returns the number of nodes with d < 0. */
long ora_getpri(const ora_ext *e, const double *q, double *pri_v, double gamma, double cv);
/* sub_inviscid_upwind.f90:1-180 */
void ora_weno3(const ora_ext *e, double *df, const double *f, int coord, int direction);
/* sub_inviscid_splitting.f90:1-172 */
void ora_roe(const ora_ext *e, const double *ql, const double *qr, double *flux,
             const double *dxidx, const double *inv_j, double gamma, int coord);
/* sub_inviscid1.f90:8-189 */
void ora_inviscid(const ora_ext *e, double *rhsi, const double *q, const double *dxidx,
                  const double *inv_j, double gamma);
/* sub_viscous.f90:9-378 */
void ora_viscous(const ora_ext *e, double *rhsv, const double *pri_v, double *amu, double *vor,
                 double *dudx, double *tao, const double *dxidx, const double *inv_j,
                 double tinf, double prl, double prt, double gamma, double cv, double re);
/* sub_timeadvancement_finddt.f90:24-130; stats[4] = dtmax,dtmin,cflmax,cflmin */
void ora_find_dt(const ora_ext *e, double *dt, double *stats, const double *pri_v,
                 const double *amu, const double *dxidx, double re, double *cfl, int iflag_time);
/* sub_timeadvancement_jac.f90:7-63; a is a(5,5) column-major */
void ora_calc_jm(double *a, double q1, double q2, double q3, double q4, double q5,
                 double mx, double my, double mz, double gamma);
/* sub_timeadvancement_1stimplicit.f90:7-585 */
void ora_implicit1st(const ora_ext *e, double *q, const double *dt, const double *pri_v,
                     double *rhs, const double *rhsi, const double *rhsv, const double *dxidx,
                     const double *inv_j, double gamma, double cfl);
/* sub_timeadvancement_2ndcrank.f90:7-467 */
void ora_crank2nd(const ora_ext *e, double *q, const double *q0, const double *dt,
                  const double *pri_v, double *rhs, const double *rhsi, const double *rhsv,
                  const double *dxidx, const double *inv_j, double gamma);
/* main.f90:692-711 */
void ora_update_q0(const ora_ext *e, double *q0, const double *q);
/* sub_tur_sa.f90:7-825 (dudx read through the (3,...,3) dummy shape, as in
 * the reference) */
void ora_sa_model(const ora_ext *e, double *nut, double *sa_rhs, const double *dt,
                  const double *pri_v, const double *amu, const double *dudx, const double *vor,
                  const double *dxidx, const double *inv_j, const double *alpha,
                  const double *beta, const double *gam, double *length, const double *dist,
                  const double *spac, double re, double cfl, int iflag_des, int iflag_time);

/* Physical boundary conditions of one subface (sub_bc_-NN_*.f90). The range
 * arguments are the subface range exactly as passed by physical_bc. */
void ora_bc_farfield(const ora_ext *e, double *q, double *tur, const double *dxidx,
                     const ora_config *c, int face, int is, int ie, int js, int je, int ks, int ke);
void ora_bc_slipwall(const ora_ext *e, double *q, double *tur, const double *dxidx, int turb,
                     int face, int is, int ie, int js, int je, int ks, int ke);
void ora_bc_nonslipwall(const ora_ext *e, double *q, double *tur, int turb,
                        int face, int is, int ie, int js, int je, int ks, int ke);
void ora_bc_outflow(const ora_ext *e, double *q, double *tur, int turb,
                    int face, int is, int ie, int js, int je, int ks, int ke);
void ora_bc_inflow(const ora_ext *e, double *q, double *tur, const ora_config *c,
                   int face, int is, int ie, int js, int je, int ks, int ke);

/* sub_bc.f90:9-90 over all blocks */
void ora_physical_bc(int nblk, ora_block *blk, const ora_config *c);
/* sub_bc_updatebuffer.f90:10-548 (single rank) */
void ora_update_buffer(int nblk, ora_block *blk, const ora_config *c);
/* sub_bc_updatebufferaverage.f90:10-562 (average1 + average2) */
void ora_average(int nblk, ora_block *blk, const ora_config *c);
/* sub_output_residual.f90:8-69; out[0..4] main, out[5] SA, out[6] = 0 */
void ora_residual(int nblk, const ora_block *blk, long totalnodes, double *out);

/* One time step of main.f90:343-573 for iflag_timeadvance 2 or 3. */
int ora_step(int nblk, ora_block *blk, ora_config *c);
/* Initial refresh before the time loop (main.f90:309-318) */
void ora_prepare(int nblk, ora_block *blk, ora_config *c);

/* Extensions with no counterpart in the reference (oracle_ext.c): published
 * methods, pinned by closed forms. */
void ora_sw_split(const double *q, const double *M, double gamma, double *fp, double *fm);
double ora_weno5_rec(double a, double b, double c, double d, double e5);
double ora_weno3_rec(double a, double b, double c);
double ora_weno5_interp(double a, double b, double c, double d, double e5);
int ora_wcns_diff_order(int a, int lo, int hi);
/* scheme 1 = Steger-Warming + WENO5, 2 = WCNS (WENO5 midpoint interpolation + Roe) */
void ora_inviscid_ext(const ora_ext *e, double *rhsi, const double *q, const double *dxidx,
                      const double *inv_j, double gamma, int scheme);
void ora_rk3_stage(const ora_ext *e, double *q, const double *qn, const double *rhs, const double *dt,
                   const double *inv_j, int s);
/* one Shu-Osher RK3 step; scheme 0 = the reference's WENO3 + Roe */
int ora_step_rk3(int nblk, ora_block *blk, ora_config *c, int scheme);

/* Grid pre-processing (oracle_grid.c).  x, y, z are scalar fields over
 * the extended extents. */
/* sub_init_gridtransformation.f90:27-874 (3-D): written on the nodes with at
 * most one coordinate outside the physical range; alpha may be NULL */
void ora_gridtransformation(const ora_ext *e, const double *x, const double *y, const double *z, double *dxidx,
                            double *inv_j, double *alpha, double *beta, double *gamma);
/* sub_init_spacing.f90:16-110 (physical range); vol may be NULL */
void ora_spacing(const ora_ext *e, const double *x, const double *y, const double *z, double *spci, double *spcj,
                 double *spck, double *spacing, double *vol);
/* wall forces of one block (sub_output_forceintegration.f90:182-453, 3-D),
 * accumulated into fp[3], fv[3]; afp/afv accumulate the magnitudes */
void ora_forces_block(const ora_ext *e, int nsub, const ora_subface *sub, const double *pri_v, const double *dxidx,
                      const double *amu, const double *spci, const double *spcj, const double *spck, double pinf,
                      double re, double *fp, double *fv, double *afp, double *afv);
/* sub_init_walldist.f90:9-85 against nw wall points (physical range) */
void ora_walldist(const ora_ext *e, const double *x, const double *y, const double *z, long nw, const double *xw,
                  const double *yw, const double *zw, double *dist);

#ifdef __cplusplus
}
#endif
#endif
