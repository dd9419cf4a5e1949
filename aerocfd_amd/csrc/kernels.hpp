// Host-side launchers of the device kernels (one translation unit each).
#pragma once
#include "batch.hpp"
#include "common.hpp"

namespace acfd {

// k_flux.hip
int launch_getpri(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, unsigned int *nneg);
int launch_fill(hipStream_t s, double *p, long n, double v);
int launch_inviscid(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision);
int launch_viscous(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision,
                   int store_aux);
// unsteady: the local cfl of the block's last node goes to *cfl and, when
// not null, *cfl_blk
int launch_find_dt(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, double *cfl, double *cfl_blk,
                   unsigned long long *stats);
int launch_reset_stats(hipStream_t s, unsigned long long *stats);

// k_flux5.hip: schemes and time integration with no reference counterpart
// (scheme 1 Steger-Warming + WENO5, 2 WCNS; Shu-Osher RK3)
int launch_inviscid_ext(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision, int scheme);
int launch_rk3_stage(hipStream_t s, const Ext &e, const Fields &f, const double *qn, int stage, bool viscous);
int launch_copy(hipStream_t s, const double *a, double *b, long n);

// k_sensor.hip: learned WCNS sensor (training, data, inference)
struct SensorNet;
// the sensor's face counts are spread over slots on separate cache lines
// (one add per wavefront, to the slot of its workgroup and row), so that the
// adds of many wavefronts do not queue on one address
constexpr int kSensorStatSlots = 64, kSensorStatStride = 16;  // stride in counters (128 bytes)
struct SensorArgs {            // passed to the WCNS kernels when the sensor is on
    const SensorNet *net;
    float threshold;           // use the linear weights where p < threshold
    unsigned long long *stats; // per slot: [0] faces evaluated, [1] faces given linear weights
    // the block's decisions, one byte per face (1: linear weights) at
    // mask[d * n + node] for the face between node and its successor along
    // direction d; recorded when the network runs, read back when it does not
    unsigned char *mask;
    int replay;                // 1: take the decisions from mask instead of the network
};
int launch_mlp_grad(hipStream_t s, const float *w, const float *X, const float *Y, long n, long first, long count,
                    long stride, long offset, float *part, float *lpart, float *grad, float *loss);
long mlp_partials_needed(long count);
int launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, float lr, int t);
int launch_sensor_pack(hipStream_t s, const float *w, SensorNet *net);
int launch_sensor_infer(hipStream_t s, const SensorNet *net, const float *X, long n, float *prob);
long sensor_samples(const Ext &e, int axis, int stride);
int launch_sensor_collect(hipStream_t s, const Ext &e, const Fields &f, double gamma, int axis, int stride,
                          float *X, float *Y);
size_t sensor_net_bytes();
long diff_partials(const Ext &e);
int launch_diff_partial(hipStream_t s, const Ext &e, const double *a, const double *b, double *part);
int launch_inviscid_wcns_sensor(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision,
                                const SensorArgs &sa);

// batched launches over every block (block table, per-kind grids)
dim3 grid_1d(long n);                    // 256-thread workgroups over n items
dim3 inviscid_grid(const Ext &e, int axis);
dim3 viscous_grid(const Ext &e);
dim3 lu_prep_grid(const Ext &e);
int launch_getpri_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                        unsigned int *nneg);
// part -1: every kernel of the stage (inviscid -2: all but the clearing of
// rhsi); otherwise only one of them (inviscid: 0 clear rhsi, 1..3 the i, j,
// k directions; viscous: 1 node fluxes, 2 their difference), for per-kernel
// timing
int launch_inviscid_batch(hipStream_t s, const BlockRec *blks, const BatchGrid *bg, double *arena,
                          const Consts &c, int precision, int part = -1);
int launch_viscous_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                         int precision, int store_aux, int part = -1);
// the viscous stage with the steady find_dt folded into its node pass
// (statistics into stats from the last block; reset them first); with
// prep, also what launch_lu_prep_batch forms for the tile sweeps: the time
// term phi/dt (node pass) and rhs = rhsi + rhsv (difference pass, which
// then does not store rhsv: launch_viscous_batch part 2 forms it again from
// the node fluxes)
int launch_viscous_dt_batch(hipStream_t s, const BlockRec *blks, int nblk, const BatchGrid &bg, double *arena,
                            const Consts &c, int precision, int store_aux, const double *cfl,
                            unsigned long long *stats, int part = -1, int prep = 0);
int launch_find_dt_batch(hipStream_t s, const BlockRec *blks, int nblk, const BatchGrid &bg, double *arena,
                         const Consts &c, double *cfl, double *cfl_blk, unsigned long long *stats);
// rhs and the time term phi/dt of the batched tile sweeps
int launch_lu_prep_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const double *cfl,
                         int cfl_stride);
int launch_residual_batch(hipStream_t s, const BlockRec *blks, int nblk, double *arena, double *partial,
                          int with_sa);

// k_grid.hip
int launch_grid_metrics(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z, double *scratch9,
                        double *inv_j, double *dxidx, double *aal, double *bbe, double *gga);
int launch_grid_spacing(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z, double *spc,
                        double *spacing);
// wall points in chunks of kWallChunk; cbox: per chunk its bounding box
// (lo x, y, z, hi x, y, z)
constexpr int kWallChunk = 256;
int launch_walldist(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z,
                    const double *wall, long nw, const double *cbox, int nchunk, double *dist);
// one cell of a wall subface: block index, face, wall-plane index and the
// lower corner along the face's two tangential axes (in axis order)
struct WallCell {
    int m, face, wall, u, v, pad;
};
int launch_wall_forces(hipStream_t s, const BlockRec *blks, const WallCell *cells, int ncell, double *arena,
                       double *geo, const long *geo_off, double pinf, double re, double *work, double *out6);

// k_lusgs.hip
// workspace of the tile sweeps: sweep-layout operands (normal memory),
// exchange slots (uncached memory, sentinel-filled) and an error word
struct LuWs {
    double *sl;
    double *ex;
    int *err;
};
constexpr int kLusgsPhases = 5;
int launch_lusgs(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                 int variant, const LuWs *ws, int shape);
int launch_lusgs_phase(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                       int variant, const LuWs *ws, int phase, int shape, int geo_cached = 0);
int lusgs_tile_shapes();
long lusgs_tiles(const Ext &e, int shape);
// doubles of the sweep layout and of the exchange slots (max over shapes)
void lusgs_ws_sizes(const Ext &e, long *sl_n, long *ex_n);
// all blocks' tile sweeps in one launch per phase (one tile shape for all)
struct LuBatch {
    void *ents = nullptr;     // device table, one entry per block
    int *tile_ent = nullptr;  // device, entry of each flat tile number
    double *arena = nullptr, *sl = nullptr, *ex = nullptr;  // bases of the table's offsets
    int nent = 0, total = 0, shape = 0;
    int resident[2] = {0, 0};  // forward / backward sweep: every workgroup fits on the device at once
    int fuse_add = 0;          // the backward sweep stores q + dq (no add pass)
    int refill = 0;            // the exchange region is refilled with the sentinel after each sweep
    long ex_n = 0;             // doubles of the exchange region the batch uses
};
void lusgs_batch_ws(int nblk, const Ext *e, int shape, long *sl_n, long *ex_n);
// f[m]'s fields must lie in the allocation starting at arena
int lusgs_batch_make(int nblk, const Ext *e, const Fields *f, double *arena, int shape, double *sl, double *ex,
                     LuBatch *out);
void lusgs_batch_free(LuBatch *b);
int launch_lusgs_batch(hipStream_t s, const LuBatch &b, int phase, double gamma, int *err, int *ctr,
                       int geo_cached = 0);

int launch_fill_sentinel(hipStream_t s, double *p, long n);
// tile sweeps of the SA model's scalar LU-SGS (cu = a, b, c, diag)
int launch_sa_sweeps(hipStream_t s, const Ext &e, int shape, const double *cu, const double *rhs, double *dns,
                     double *dn, double *ex, int *err);
// sl: the SA sweep layout (sa_batch_sl doubles for this shape) or null
void sa_batch_sl(int nblk, const Ext *e, int shape, long *sl_n);
int sa_batch_make(int nblk, const Ext *e, const Fields *f, double *arena, int shape, double *ex, double *sl,
                  LuBatch *out);
// the batched SA sweeps over the SA sweep layout: phase 0 operands into
// the layout, 1 forward sweep, -1 backward sweep, 2 nut += dnut
int launch_sa_sl_batch(hipStream_t s, const LuBatch &b, int *err, int *ctr, int phase);
// dir 1: the forward sweep only, -1: the backward sweep only, 0: both
int launch_sa_batch(hipStream_t s, const LuBatch &b, int *err, int *ctr, int dir = 0);

// k_bdf2.hip
int launch_crank(hipStream_t s, const Ext &e, const Fields &f, const Consts &c);
int launch_update_q0(hipStream_t s, const Ext &e, const Fields &f);

// k_sa.hip
struct SaSweep {
    int shape;   // tile shape index
    double *ex;  // sentinel-filled exchange slots
    int *err;
};
int launch_sa(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
              const SaSweep *sweep = nullptr);
int launch_sa_pre(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl);
int launch_sa_post(hipStream_t s, const Ext &e, const Fields &f);
// length scale, right-hand side and the sweep operands of every block in two
// launches (linear-tile grid); block m reads its cfl at cfl[m * cfl_stride];
// part 1 / 2 launches only the first / second pass (per-kernel timing);
// derive_abg: alpha, beta, gamma are the products of dxidx and 1/J that the
// grid set-up stores (sa_metric_products), formed again instead of read;
// length_on: the first pass forms the length scale (off while it holds the
// flow-independent length of SA or DES97 already)
int launch_sa_pre_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                        const double *cfl, int cfl_stride, int part = 0, bool derive_abg = false,
                        bool length_on = true);

// k_bc.hip
struct BcTask {       // one physical subface
    int block;        // 0-based
    int face, kind;   // kind = blk_t (-10 .. -50)
    int is, ie, js, je, ks, ke;  // loop range actually used (farfield js quirk applied)
};
int launch_bc(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const BcTask &t);
int launch_sa_clamp(hipStream_t s, const Ext &e, const Fields &f);
long bc_nodes(const BcTask &t);  // face nodes of a subface
// tasks[m]: one subface of block tasks[m].block, m over the grid's entries
int launch_bc_batch(hipStream_t s, const BcTask *tasks, const BatchGrid &bg, const BlockRec *blks, double *arena,
                    const Consts &c);
int launch_sa_clamp_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena);

// halo copy lists (device arrays of node references)
struct CopyList {
    int count;
    const int *src_blk, *dst_blk;     // 0-based block ids
    const long *src, *dst;            // node indices in the blocks
};
int launch_halo_copy(hipStream_t s, const CopyList &cl, double *const *q_tab, double *const *nut_tab,
                     const long *n_tab, int with_nut);
// average: gather into buf (6 per entry), then apply in order per node
struct AvgList {
    int count;                        // entries
    const int *src_blk;
    const long *src;
    int ngroups;                      // distinct destination nodes
    const int *grp_blk;               // block of each group
    const long *grp_dst;              // node index of each group
    const int *grp_start;             // first entry of each group (ngroups+1)
};
int launch_average(hipStream_t s, const AvgList &al, double *const *q_tab, double *const *nut_tab,
                   const long *n_tab, double *buf, int with_nut);

int launch_residual(hipStream_t s, const Ext &e, const Fields &f, double *partial, int with_sa);
int residual_partials(const Ext &e);
int launch_residual_final(hipStream_t s, const double *partial, int count, double *out);

// layout transposition between reference AoS and device SoA
int launch_aos_to_soa(hipStream_t s, const double *aos, double *soa, long n, int ncomp);
int launch_soa_to_aos(hipStream_t s, const double *soa, double *aos, long n, int ncomp);

}  // namespace acfd
