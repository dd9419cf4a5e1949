/*
 * aerocfd_amd -- MI355X (gfx950) implementation of aerocfd's per-block
 * residual-and-update path behind a plain C ABI.
 *
 * The reference (jiamin0329/aerocfd, Fortran 90 + MPI) exposes this path as
 * explicit-shape external subroutines called from its time loop
 * (main.f90:342-613) and as argument-less routines that walk the block table
 * (physical_bc, UpdateBuffer, average1/2, residual).  This header replaces
 * both with:
 *   - a resident context that owns every block on one GPU (flags and
 *     constants explicit, no globals), and
 *   - stateless per-routine entry points for unit parity.
 * All arrays crossing the boundary are FP64 in the reference's own layout:
 * Fortran column-major, component first, e.g. q(5,is:ie,js:je,ks:ke).
 * Indices are 1-based as in the reference.  Every function returns 0 on
 * success and a negative code on error; acfd_last_error() returns the message.
 * Nothing here aborts the host.  A context is used by one host thread.
 */
#ifndef AEROCFD_AMD_H
#define AEROCFD_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define ACFD_OK            0
#define ACFD_EINVAL       -1   /* bad argument or configuration */
#define ACFD_EDEVICE      -2   /* HIP runtime error */
#define ACFD_ESTATE       -3   /* call out of order */
#define ACFD_ENUMERIC     -4   /* negative density found (sub_getpri_v.f90:54-56) */

/* Flags and constants.  Replaces the module state of module_flags.f90:4-27,
 * module_ns_consts.f90:4-34 and the control file read by
 * sub_init_readcontrolfile.f90:27-47.  Codes are the reference's. */
typedef struct acfd_config {
    double ma, re, tinf, aoa_deg;   /* freestream: Mach, Reynolds, T_inf [K], angle of attack [deg] */
    double cfl;                     /* CFL number (steady local dt) */
    double dt;                      /* time step (unsteady) */
    double prl, prt;                /* laminar / turbulent Prandtl numbers (reference: 0.72 / 0.90) */
    int iflag_solver;               /* 0 = Navier-Stokes, 1 = Euler */
    int iflag_time;                 /* 0 = steady (local dt), 1 = unsteady */
    int iflag_turbulence;           /* 0 = laminar, 2 = Spalart-Allmaras */
    int iflag_des;                  /* 0 = RANS, 1 = DES97, 2 = DDES */
    int iflag_timeadvance;          /* 2 = first-order implicit (LU-SGS), 3 = BDF2 sub-iterations,
                                       4 = explicit three-stage Shu-Osher RK3 (no counterpart in the
                                       reference; laminar or Euler) */
    int precision;                  /* 0 = FP64 everywhere; 1 = FP32 reconstruction/splitting/viscous, FP64 update */
    int lusgs_kernel;               /* 0 = hyperplane launches, 1 = pipelined tile sweep per block,
                                       2 = tile sweeps of all blocks in one launch (all the same results) */
    int inviscid_scheme;            /* 0 = WENO3 + Roe (the reference's inviscid_flux_1); no counterpart
                                       in the reference: 1 = Steger-Warming splitting + WENO5,
                                       2 = WCNS (WENO5 midpoint interpolation + Roe, 6th-order difference) */
    int use_graph;                  /* 1 = acfd_step replays one captured HIP graph per step (same results) */
    int batch_blocks;               /* 1 = each stage is one launch over all blocks (same results) */
    int reserved[2];
} acfd_config;

/* One subface record, as read from bcinfo (sub_init_readbcinfo.f90:96-98):
 * face 1..6 = i-,i+,j-,j+,k-,k+; blk_t > 0 names the neighbour block
 * (1-based) and subface_t its subface; blk_t < 0 is a physical condition
 * (-10 farfield, -20 wall, -30 symmetry, -40 outflow, -50 inflow); ori 1..4
 * is the face orientation code of sub_bc_updatebuffer.f90:358-395. */
typedef struct acfd_subface {
    int f_no, face;
    int is, ie, js, je, ks, ke;
    int blk_t, subface_t, ori;
} acfd_subface;

/* Field identifiers for upload/download (module_block.f90:27-88). */
enum acfd_field {
    ACFD_Q = 1, ACFD_PRI_V = 2, ACFD_RHSI = 3, ACFD_RHSV = 4, ACFD_RHS = 5, ACFD_DUDX = 6,
    ACFD_TAO = 7, ACFD_VOR = 8, ACFD_AMU = 9, ACFD_DT = 10, ACFD_DXIDX = 11, ACFD_INV_J = 12,
    ACFD_AALPHA = 13, ACFD_BBETA = 14, ACFD_GGAMMA = 15, ACFD_DIST = 16, ACFD_SPACING = 17,
    ACFD_NUT = 18, ACFD_SA_RHS = 19, ACFD_Q0 = 20, ACFD_LENGTH = 21,
    /* after acfd_grid_metrics: coordinates and edge lengths (sub_init_spacing.f90:29) */
    ACFD_X = 22, ACFD_Y = 23, ACFD_Z = 24, ACFD_SPCI = 25, ACFD_SPCJ = 26, ACFD_SPCK = 27
};

/* Stages of one time step, for acfd_run_stage (reference routine in brackets). */
enum acfd_stage {
    ACFD_STAGE_GETPRI = 1,        /* get_primitivevariables, sub_getpri_v.f90:19 */
    ACFD_STAGE_INVISCID = 2,      /* inviscid_flux_1, sub_inviscid1.f90:8 */
    ACFD_STAGE_VISCOUS = 3,       /* viscous_flux, sub_viscous.f90:9 */
    ACFD_STAGE_FIND_DT = 4,       /* find_dt, sub_timeadvancement_finddt.f90:24 */
    ACFD_STAGE_IMPLICIT1ST = 5,   /* implicit1st, sub_timeadvancement_1stimplicit.f90:7 */
    ACFD_STAGE_SA = 6,            /* sa_model, sub_tur_sa.f90:7 */
    ACFD_STAGE_PHYSICAL_BC = 7,   /* physical_bc, sub_bc.f90:9 */
    ACFD_STAGE_UPDATE_BUFFER = 8, /* UpdateBuffer, sub_bc_updatebuffer.f90:10 */
    ACFD_STAGE_AVERAGE = 9,       /* average1 + average2, sub_bc_updatebufferaverage.f90:10,442 */
    ACFD_STAGE_CRANK2ND = 10,     /* crank2nd, sub_timeadvancement_2ndcrank.f90:7 */
    ACFD_STAGE_UPDATE_Q0 = 11     /* update_q0, main.f90:692 */
};

typedef struct acfd_ctx acfd_ctx;

const char *acfd_last_error(void);

/* ---------------------------------------------------------------- context */
/* Create a context on the current HIP device with one stream. */
int acfd_create(const acfd_config *cfg, acfd_ctx **out);
int acfd_destroy(acfd_ctx *ctx);

/* Block table and interfaces.  dims[3*m..3*m+2] = ni,nj,nk of block m+1;
 * nsub[m] subfaces of block m+1 follow one another in subs.  Index ranges
 * are derived as in sub_init_allocatememory.f90:73-144 (2 halo layers on
 * interface faces) and the halo copy lists are built once.  Replaces
 * init_readbcinfo / init_allocatememory / init_readbcinfo2. */
int acfd_set_topology(acfd_ctx *ctx, int nblk, const int *dims, const int *nsub,
                      const acfd_subface *subs);
/* ranges[18] = is,ie,js,je,ks,ke, is0..ke0, is1..ke1 of block m (1-based) */
int acfd_block_ranges(acfd_ctx *ctx, int m, int *ranges);

/* Copy one field of block m between host (reference AoS layout, extended
 * extents) and the device-resident state.  The batched steady update forms
 * rhs = rhsi + rhsv without storing rhsv; rhsv is then formed again from
 * the viscous node fluxes by the first call that needs it (its download or
 * upload, acfd_run_stage, acfd_grid_metrics), so every call sees the value
 * of the last viscous evaluation. */
int acfd_upload(acfd_ctx *ctx, int m, int field, const double *host);
int acfd_download(acfd_ctx *ctx, int m, int field, double *host);

/* Run one stage over all blocks, in the reference's block order. */
int acfd_run_stage(acfd_ctx *ctx, int stage);
/* The boundary refresh before the time loop (main.f90:309-318). */
int acfd_prepare(acfd_ctx *ctx);
/* nsteps time steps of main.f90:343-573 (iflag_timeadvance 2 or 3). */
int acfd_step(acfd_ctx *ctx, int nsteps);
/* RMS residual of sub_output_residual.f90:24-56: out[0..4] main equations,
 * out[5] SA, out[6] = 0.  Every step forms it, as the reference's loop does
 * (output interval 1, main.f90:341,575-603), and this call reads it back;
 * after a call that writes fields it is formed from the current rhs, dt and
 * sa_rhs. */
int acfd_residual(acfd_ctx *ctx, double *out);
/* out = dtmax, dtmin, cflmax, cflmin of the last find_dt call and the
 * current cfl (main.f90:589-592). */
int acfd_stats(acfd_ctx *ctx, double *out);
/* Wait for all work of the context. */
int acfd_sync(acfd_ctx *ctx);
/* Per-stage GPU time accumulated with HIP events on the context's stream.
 * Returns the totals recorded so far (ms/calls may be NULL); then, for
 * enable >= 0, clears them and turns recording on (enable > 0) or off
 * (enable == 0); enable < 0 only reads.  ms and calls have 32 entries: [stage] for
 * stages 1..11, and sub-stages of ACFD_STAGE_IMPLICIT1ST: [12] forward
 * sweep, [13] backward sweep, [14] gather into the sweep layout (tile
 * sweeps), [15] rhs / spectral radii / diagonal, [16] q += dq.  With
 * batch_blocks / lusgs_kernel 2 each kernel of a batched stage is also
 * timed on its own: [17] SA forward sweep, [18] SA backward sweep, [19] SA
 * length scale / right-hand side / diagonal (with lusgs_kernel 2: the
 * first of its two per-node passes, [28] the second), [20] nut += dnut,
 * [21] rhsi cleared, [22..24] inviscid i, j, k direction, [25] viscous node
 * fluxes, [26] their difference, [27] the SA operands moved into their
 * sweep layout, [29] the viscous node fluxes with the steady local time
 * step (NS, steady, batched: find_dt folded into that pass).  Indices 0, 30
 * and 31 are unused.  The
 * reference's only instrument is a per-step wall time (main.f90:575-586). */
int acfd_timing(acfd_ctx *ctx, int enable, double *ms, long *calls);

/* ---------------------------------------------------- stateless routines */
/* Each runs one reference routine on one block given its 18 ranges
 * (is..ke, is0..ke0, is1..ke1), copying inputs in and outputs out. */

/* get_primitivevariables (sub_getpri_v.f90:19-63); returns ACFD_ENUMERIC
 * after filling pri_v when a density is negative. */
int acfd_getpri(const int *ranges, double gamma, double cv, const double *q, double *pri_v);
/* inviscid_flux_1 (sub_inviscid1.f90:8-189) */
int acfd_inviscid_rhs(const acfd_config *cfg, const int *ranges, const double *q,
                      const double *dxidx, const double *inv_j, double *rhsi);
/* viscous_flux (sub_viscous.f90:9-378); amu is read (component 2) and written
 * (components 1 and 3). */
int acfd_viscous_rhs(const acfd_config *cfg, const int *ranges, const double *pri_v,
                     double *amu, const double *dxidx, const double *inv_j, double *rhsv,
                     double *vor, double *dudx, double *tao);
/* find_dt (sub_timeadvancement_finddt.f90:24-130); *cfl is read (steady) or
 * overwritten (unsteady) as the reference does. */
int acfd_find_dt(const acfd_config *cfg, const int *ranges, const double *pri_v,
                 const double *amu, const double *dxidx, double *dt, double *cfl);
/* implicit1st (sub_timeadvancement_1stimplicit.f90:7-585): q updated in place. */
int acfd_lusgs(const acfd_config *cfg, const int *ranges, double *q, const double *dt,
               const double *pri_v, const double *rhsi, const double *rhsv, const double *dxidx,
               const double *inv_j, double cfl, double *rhs);
/* crank2nd (sub_timeadvancement_2ndcrank.f90:7-467): q updated in place. */
int acfd_bdf2(const acfd_config *cfg, const int *ranges, double *q, const double *q0,
              const double *dt, const double *pri_v, const double *rhsi, const double *rhsv,
              const double *dxidx, const double *inv_j, double *rhs);
/* sa_model (sub_tur_sa.f90:7-122): nut updated in place. */
int acfd_sa_step(const acfd_config *cfg, const int *ranges, double *nut, double *sa_rhs,
                 const double *dt, const double *pri_v, const double *amu, const double *dudx,
                 const double *vor, const double *dxidx, const double *inv_j,
                 const double *aalpha, const double *bbeta, const double *ggamma,
                 double *length, const double *dist, const double *spacing, double cfl);

/* No counterpart in the reference: viscous_flux (the same computation and
 * results as acfd_viscous_rhs) for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  It allocates nothing: every array is a DEVICE
 * pointer in the library's own layout (structure of arrays over the
 * extended extents: component c of node (i,j,k) at
 * a[c*n + (i-is) + ni*((j-js) + nj*(k-ks))], n = ni*nj*nk), and the node
 * fluxes are staged in a static device array of the library sized for at
 * most 524288 (2^19) extended nodes per call; larger blocks return
 * ACFD_EINVAL.  rhsv is written on the physical range is0..ie0 (elsewhere
 * untouched); amu components 1 and 3 are written there; vor (and dudx, tao,
 * which need vor) may be NULL.  Work is queued on `stream` (a hipStream_t,
 * NULL = the null stream) and not waited for.  Every call on one GPU
 * shares that GPU's staging array: calls that overlap on one GPU (other
 * threads, other streams without ordering) race. */
int acfd_viscous_rhs_noalloc(const acfd_config *cfg, const int *ranges, const double *pri_v,
                             double *amu, const double *dxidx, const double *inv_j, double *rhsv,
                             double *vor, double *dudx, double *tao, void *stream);

/* ------------------------------------------------------------- grid setup */
/* Halo coordinate exchange over the context's topology, as
 * UpdateBufferCoordinate (sub_bc_updatebuffercoordinate.f90:10-473), with an
 * optional per-subface translation (period) added to the received
 * coordinates: shift[3*s..] for the s-th subface in the order given to
 * acfd_set_topology (NULL = none).  xyz[m] points to 3 consecutive extended
 * scalar fields x,y,z of block m+1, updated in place on the host. */
int acfd_exchange_coordinates(acfd_ctx *ctx, double **xyz, const double *shift);

/* Grid pre-processing on the device, replacing the start-up calls of
 * main.f90:163-209: init_spacing (sub_init_spacing.f90:16-110), init_walldist
 * (sub_init_walldist.f90:9-85, when the SA model is on; wall points are the
 * face nodes of every subface with blk_t = -20) and gridtransformation
 * (sub_init_gridtransformation.f90:27-874, 3-D).  xyz[m] points to the x, y,
 * z fields of block m+1 over its extended extents with the halos filled
 * (acfd_exchange_coordinates).  Writes dxidx and inv_j -- and with SA
 * aalpha/bbeta/ggamma, spacing and dist -- of the resident blocks, and keeps
 * the coordinates and edge lengths on the device (fields ACFD_X..ACFD_SPCK).
 * Results equal the reference's bit for bit where it defines them; at halo
 * nodes with two or three coordinates outside the physical range (never
 * read by the path) the metrics are zero. */
int acfd_grid_metrics(acfd_ctx *ctx, double *const *xyz);

/* Wall forces and lift/drag of the resident state, replacing clcd_output
 * (sub_output_forceintegration.f90:8-509, 3-D) without downloading it:
 * every cell of every subface with blk_t = -20, evaluated as the reference
 * does -- including its mix of metric components in the averaged face
 * normal -- and summed in the reference's order.  Needs the edge lengths
 * (acfd_grid_metrics, or ACFD_SPCI..ACFD_SPCK uploaded).  sref: the
 * reference area of Control.in.  out[8] = pressure force x, y, z, viscous
 * force x, y, z, cl, cd. */
int acfd_forces(acfd_ctx *ctx, double sref, double *out);

/* --------------------------------------------------- learned WCNS sensor */
/* No counterpart in the reference.  A network 12 -> 32 -> 32 -> 1 reads the
 * smoothness indicators and jumps of density and pressure at the six nodes
 * of a WCNS interface and predicts whether the WENO5 nonlinear weights are
 * needed (label: they differ from the linear weights by more than 1e-3);
 * where it says no, the reconstruction uses the linear weights.  Inference
 * runs as BF16 MFMA with FP32 accumulation inside the WCNS kernels;
 * training runs on the GPU (FP32 forward/backward, Adam).  Parameters: 1505
 * floats W1[12][32], b1[32], W2[32][32], b2[32], w3[32], b3. */

/* get (set = 0) or set (set = 1) the FP32 parameters */
int acfd_sensor_weights(acfd_ctx *ctx, float *w, int set);
/* training samples from the current state: every face with a full
 * six-node stencil on every stride-th grid line, all blocks and
 * directions; *count = number of samples */
int acfd_sensor_collect(acfd_ctx *ctx, int stride, long *count);
/* iters Adam steps on minibatches of the samples (binary cross-entropy);
 * losses[it] receives each minibatch loss when not NULL */
int acfd_sensor_train(acfd_ctx *ctx, int iters, int batch, float lr, unsigned seed, float *losses);
/* out[4] on the collected samples at the given threshold: accuracy,
 * fraction given linear weights, fraction of nonlinear-labelled faces given
 * linear weights, fraction labelled nonlinear */
int acfd_sensor_evaluate(acfd_ctx *ctx, float threshold, double *out);
/* use the sensor in the WCNS residual (inviscid_scheme 2): linear weights
 * where its probability is below threshold; every audit_interval
 * evaluations of a block the plain scheme runs too, and a relative
 * deviation of the residual above audit_bound makes that block fall back
 * to plain WCNS for good (audit_interval 0: no audits) */
int acfd_sensor_enable(acfd_ctx *ctx, int on, float threshold, int audit_interval, double audit_bound);
/* reuse the network's decisions: it runs on every interval-th WCNS
 * evaluation of a block, which records its per-face decisions in a byte
 * mask of the block (3 bytes per node); the evaluations between read the
 * mask instead (interval 1, the default: the network runs every time).
 * The next evaluation of every block runs the network. */
int acfd_sensor_refresh(acfd_ctx *ctx, int interval);
/* out[6]: faces evaluated, faces given linear weights, audits, largest
 * audit deviation, blocks fallen back, 0 */
int acfd_sensor_stats(acfd_ctx *ctx, double *out);
/* stateless, for checking the training kernels: mean loss and its
 * gradient over n samples X[n][12], Y[n] */
int acfd_mlp_grad(const float *w, const float *X, const float *Y, long n, float *loss, float *grad);
/* stateless batch inference through the BF16 MFMA routine of the WCNS kernels */
int acfd_sensor_infer(const float *w, const float *X, long n, float *prob);

#ifdef __cplusplus
}
#endif
#endif
