// Host side of the C ABI: the resident context (block table, device
// arrays, halo copy lists, step orchestration) and the stateless per-routine
// entry points.  The orchestration follows the reference time loop
// (main.f90:343-573); the topology rules follow sub_init_allocatememory.f90
// and sub_init_readbcinfo.f90.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aerocfd_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "sensor.hpp"

using namespace acfd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e__ = (expr);                                                       \
        if (e__ != hipSuccess)                                                         \
            return fail(ACFD_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e__)); \
    } while (0)

#define KRN_TRY(expr)                                                                          \
    do {                                                                                       \
        int r__ = (expr);                                                                      \
        if (r__ != 0)                                                                          \
            return fail(ACFD_EDEVICE, std::string(#expr ": ") + hipGetErrorString((hipError_t)r__)); \
    } while (0)

// Flow constants of a configuration: the non-dimensional freestream of
// sub_init_flowfield.f90:30-39 and the flags.
Consts consts_of(const acfd_config &cfg)
{
    const double gamma = 1.4, pi = 3.1415926535897932;
    Consts c;
    c.gamma = gamma;
    c.d0 = 1.0;
    c.u0 = 1.0;
    c.p0 = 1.0 * 1.0 / gamma / cfg.ma / cfg.ma;
    c.c0 = std::sqrt(1.0) / cfg.ma;
    c.cv = 1.0 / gamma / (gamma - 1.0) / cfg.ma / cfg.ma;
    c.re = cfg.re;
    c.tinf = cfg.tinf;
    c.prl = cfg.prl > 0 ? cfg.prl : 0.72;
    c.prt = cfg.prt > 0 ? cfg.prt : 0.90;
    c.aoa = cfg.aoa_deg * pi / 180.0;
    c.solver = cfg.iflag_solver;
    c.itime = cfg.iflag_time;
    c.turb = cfg.iflag_turbulence;
    c.des = cfg.iflag_des;
    return c;
}

const int kNcomp[28] = {0, 5, 7, 5, 5, 5, 9, 6, 1, 3, 1, 9, 1, 3, 3, 3, 1, 1, 1, 1, 10, 1, 1, 1, 1, 1, 1, 1};
constexpr int kLastField = 27;

constexpr int kTimeSlots = 32;  // acfd_timing entries
constexpr int kSensorParams = 1505, kSensorFeatures = 12;  // the sensor network's size
constexpr size_t kSensorStatBytes = (size_t)kSensorStatSlots * kSensorStatStride * sizeof(unsigned long long);

struct Block {
    Ext e;
    Fields f;
    long arena_off = 0;     // doubles from the context's arena base
    double *geo = nullptr;  // acfd_grid_metrics: x, y, z, spci, spcj, spck (n each)
    int tile_shape = -1;    // chosen by the autotuner on first use
    std::vector<acfd_subface> sub;
};

// Uncached device memory for the tile sweeps' exchange arrays and counters.
// Memory of this type that a kernel has read is never returned to the
// runtime before the process ends; chunks are recycled between contexts.
struct UcChunk {
    void *p;
    size_t n;
    bool used;
};
std::vector<UcChunk> g_uc;

void *uc_acquire(size_t n)
{
    for (auto &c : g_uc)
        if (!c.used && c.n >= n) {
            c.used = true;
            return c.p;
        }
    void *p = nullptr;
    if (hipExtMallocWithFlags(&p, n, hipDeviceMallocUncached) != hipSuccess) return nullptr;
    g_uc.push_back({p, n, true});
    return p;
}

void uc_release(void *p)
{
    for (auto &c : g_uc)
        if (c.p == p) c.used = false;
}

}  // namespace

struct acfd_ctx {
    acfd_config cfg;
    Consts c;
    hipStream_t stream = nullptr;
    std::vector<Block> blk;
    double *arena = nullptr;  // every block's fields
    long arena_n = 0;
    double *geo = nullptr;    // acfd_grid_metrics: coordinates and edge lengths of every block
    // acfd_forces: wall cells (reference order), per-block offsets into geo, work
    WallCell *d_wcells = nullptr;
    int n_wcells = -1;
    long *d_geo_off = nullptr;
    double *d_fwork = nullptr;
    long totalnodes = 0;
    bool topo = false;
    // device scalars
    double *d_cfl = nullptr;
    unsigned long long *d_stats = nullptr;
    unsigned int *d_nneg = nullptr;
    double *d_res = nullptr;       // 6 sums
    double *d_partial = nullptr;   // residual partials
    LuWs lu{nullptr, nullptr, nullptr};  // tile-sweep workspace (lusgs_kernel 1, 2)
    long lu_ex_n = 0;
    LuBatch lub[4];          // lusgs_kernel 2: every block in one launch, per tile shape
    LuBatch sab[4];          // the same for the SA model's scalar sweeps
    int lub_shape = -1;      // chosen by the autotuner on first use
    // what the sweep layout holds of the grid operands: the block of the last
    // gather (-1: the batch, -2: nothing valid) and its tiling
    int sl_block = -2, sl_shape = -1;
    bool capturing = false;
    // the captured step gathers without the grid operands: it replays only
    // onto the contents it was captured on (sl_block, sl_shape at its start)
    // and leaves those of its end
    bool graph_geo_cached = false;
    int graph_sl_in[2] = {-2, -1}, graph_sl_out[2] = {-2, -1};
    // pri_v holds getpri of the current q of every block: set by the closing
    // getpri of a step, cleared by every call that may write q or pri_v.
    // The next step's residual then starts without its getpri; a captured
    // step that did so replays only while it holds.
    bool pri_current = false;
    bool graph_needs_pri = false;
    // rhsi is zero outside every block's computed range (the arena starts
    // zeroed and the inviscid kernels store only inside it): the batched
    // inviscid stage then skips clearing it.  An upload of rhsi clears the
    // flag (and drops a captured step that relied on it).
    bool rhsi_clean = true;
    // the last residual stage also formed implicit1st's rhs, spectral radii
    // and diagonal (batched, steady NS): the next batched update skips its
    // preparation pass
    bool lu_prep_ready = false;
    // the last residual stage formed rhs = rhsi + rhsv without storing rhsv
    // (the same fold): rhsv_refresh forms it from the node fluxes, which
    // stay in w_flux until the next inviscid stage or grid set-up; a
    // captured step that does so sets it on every replay (graph_rhsv)
    bool rhsv_stale = false, graph_rhsv = false;
    // alpha, beta, gamma are the products of dxidx and 1/J that
    // acfd_grid_metrics stored (an upload of any of the five clears it): the
    // batched SA pass forms them again instead of reading them
    bool abg_derived = false;
    // the SA length scale holds min(d, C_des spacing) (DES97) or d (plain
    // SA), which no step changes: the batched SA pass skips forming it
    // again (an upload of dist, spacing or length, or acfd_grid_metrics,
    // clears it)
    bool sa_length_ready = false;
    // d_res holds the norms of the current rhs, dt and sa_rhs (set by
    // acfd_step, cleared by every call that writes fields)
    bool res_current = false;
    int *lu_ctr = nullptr;   // ticket counters of the batched sweeps
    // SA's update runs beside the flow's LU-SGS update on a second stream
    // (batched sweeps, SA): its own exchange slots and ticket counters
    bool sa_concurrent = false;
    // the residual's viscous node pass runs on the second stream beside the
    // inviscid kernels (batched NS stages, reference scheme)
    bool res_concurrent = false;
    hipEvent_t ev_rfork = nullptr, ev_rjoin = nullptr;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double *sa_ex = nullptr;
    int *sa_ctr = nullptr;
    // batched SA update (lusgs_kernel 2): the SA sweep layout and the
    // linear-tile grid of its two per-node passes
    double *sa_sl = nullptr;
    BatchGrid bg_sa;
    double *d_cfl_blk = nullptr;  // cfl after each block's find_dt
    // batch_blocks: block table and the per-kind grids of the batched launches
    BlockRec *d_blocks = nullptr;
    BatchGrid bg_getpri, bg_inv[4], bg_visc, bg_dt, bg_prep;
    // physical boundary conditions: round r holds the r-th subface of every block
    std::vector<BcTask *> d_bc_round;
    std::vector<BatchGrid> bg_bc;
    bool batch_ready = false;
    double *d_stage = nullptr;     // upload/download staging
    long stage_n = 0;
    // pointer tables for the multi-block copy kernels
    double **d_qtab = nullptr, **d_nuttab = nullptr;
    long *d_ntab = nullptr;
    // halo copy list
    int halo_n = 0;
    int *d_h_sb = nullptr, *d_h_db = nullptr;
    long *d_h_sn = nullptr, *d_h_dn = nullptr;
    std::vector<int> h_sb, h_db;
    std::vector<long> h_sn, h_dn;
    std::vector<int> h_sub;   // receiving subface (global order) of every entry
    // average list
    AvgList avg{};
    int *d_a_sb = nullptr, *d_a_gb = nullptr, *d_a_gs = nullptr;
    long *d_a_sn = nullptr, *d_a_gd = nullptr;
    double *d_a_buf = nullptr;
    std::vector<BcTask> bc;
    // timing: event pairs recorded on the stream without waiting; read
    // back (and the pool recycled) when full or when the totals are asked for
    bool timing = false;
    double stage_ms[kTimeSlots] = {0};
    long stage_calls[kTimeSlots] = {0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_stage;   // stage of each recorded pair
    size_t ev_used = 0;          // pairs in use
    // learned WCNS sensor: FP32 parameters and Adam moments, the BF16
    // inference copy, training samples and statistics
    float *s_w = nullptr, *s_m = nullptr, *s_v = nullptr, *s_grad = nullptr, *s_loss = nullptr;
    float *s_part = nullptr, *s_lpart = nullptr;
    long s_part_cap = 0;
    int s_adam_t = 0;
    SensorNet *s_net = nullptr;
    float *s_X = nullptr, *s_Y = nullptr;
    long s_n = 0, s_cap = 0;
    unsigned long long *s_stats = nullptr;   // faces evaluated, faces linear
    bool s_on = false;
    float s_thr = 0.5f;
    int s_audit = 0;             // audit every s_audit inviscid evaluations of a block (0: never)
    double s_bound = 1e-6;       // largest relative deviation an audit accepts
    std::vector<long> s_calls;   // inviscid evaluations per block
    std::vector<char> s_off;     // blocks that fell back to plain WCNS
    // the network decides on every s_refresh-th evaluation of a block; the
    // evaluations between take its decisions from the block's byte mask
    // (3 n bytes, one per face and direction)
    int s_refresh = 1;
    std::vector<unsigned char *> s_mask;
    int s_audits = 0;
    double s_audit_max = 0.0;
    hipGraphExec_t graph_exec = nullptr;  // one captured step (use_graph)
    double *s_dpart = nullptr;
    long s_dpart_n = 0;
};

namespace {


template <typename T>
int dev_upload(T **dst, const std::vector<T> &src)
{
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    if (src.empty()) return 0;
    HIP_TRY(hipMalloc((void **)dst, src.size() * sizeof(T)));
    HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// Ranges of sub_init_allocatememory.f90:73-144 (3D, bufferLength = 2).
void block_ranges(int ni, int nj, int nk, const acfd_subface *s, int ns, Ext &e)
{
    const int bl = 2;
    e.is0 = 1; e.ie0 = ni; e.js0 = 1; e.je0 = nj; e.ks0 = 1; e.ke0 = nk;
    e.is = e.is0; e.ie = e.ie0; e.js = e.js0; e.je = e.je0; e.ks = e.ks0; e.ke = e.ke0;
    for (int x = 0; x < ns; x++) {
        if (s[x].blk_t <= 0) continue;
        switch (s[x].face) {
        case 1: e.is = e.is0 - bl; break;
        case 2: e.ie = e.ie0 + bl; break;
        case 3: e.js = e.js0 - bl; break;
        case 4: e.je = e.je0 + bl; break;
        case 5: e.ks = e.ks0 - bl; break;
        case 6: e.ke = e.ke0 + bl; break;
        }
    }
    e.is1 = e.is0 == e.is ? e.is0 + 1 : e.is0;
    e.ie1 = e.ie0 == e.ie ? e.ie0 - 1 : e.ie0;
    e.js1 = e.js0 == e.js ? e.js0 + 1 : e.js0;
    e.je1 = e.je0 == e.je ? e.je0 - 1 : e.je0;
    e.ks1 = e.ks0 == e.ks ? e.ks0 + 1 : e.ks0;
    e.ke1 = e.ke0 == e.ke ? e.ke0 - 1 : e.ke0;
    e.ni = e.ie - e.is + 1;
    e.nj = e.je - e.js + 1;
    e.nk = e.ke - e.ks + 1;
    e.n = (long)e.ni * e.nj * e.nk;
}

struct PairNode {
    int ri, rj, rk, si, sj, sk;
};

// For the receiving subface f of block m, the sender's face node of every
// receiving face node: the face_t load (sub_bc_updatebuffer.f90:265-354),
// the ori permutation (:358-395) and the face store (:398-539), done on
// coordinates.  Returns false when the faces do not match in size.
bool pair_nodes(const std::vector<Block> &blk, const acfd_subface &f, std::vector<PairNode> &out)
{
    const acfd_subface &t = blk[f.blk_t - 1].sub[f.subface_t - 1];
    int n11 = t.ie - t.is + 1, n12 = t.je - t.js + 1, n13 = t.ke - t.ks + 1;
    int l1 = 0, m1 = 0;
    switch (t.face) {
    case 1: l1 = n13; m1 = n12; break;
    case 2: l1 = n12; m1 = n13; break;
    case 3: l1 = n11; m1 = n13; break;
    case 4: l1 = n13; m1 = n11; break;
    case 5: l1 = n12; m1 = n11; break;
    case 6: l1 = n11; m1 = n12; break;
    default: return false;
    }
    auto ua = [&](int ll, int mm, int *p) {
        switch (t.face) {
        case 1: p[0] = t.is; p[1] = t.js + mm - 1; p[2] = t.ks + ll - 1; break;
        case 2: p[0] = t.ie; p[1] = t.js + ll - 1; p[2] = t.ks + mm - 1; break;
        case 3: p[0] = t.is + ll - 1; p[1] = t.js; p[2] = t.ks + mm - 1; break;
        case 4: p[0] = t.is + mm - 1; p[1] = t.je; p[2] = t.ks + ll - 1; break;
        case 5: p[0] = t.is + mm - 1; p[1] = t.js + ll - 1; p[2] = t.ks; break;
        case 6: p[0] = t.is + ll - 1; p[1] = t.js + mm - 1; p[2] = t.ke; break;
        }
    };
    bool tr = f.ori == 2 || f.ori == 4;
    int l2 = tr ? m1 : l1, m2 = tr ? l1 : m1;
    int n1 = f.ie - f.is + 1, n2 = f.je - f.js + 1, n3 = f.ke - f.ks + 1;
    // receiving face extents that the store loops walk (a over l2, b over m2)
    int ra = 0, rb = 0;
    switch (f.face) {
    case 1: ra = n3; rb = n2; break;
    case 2: ra = n2; rb = n3; break;
    case 3: ra = n1; rb = n3; break;
    case 4: ra = n3; rb = n1; break;
    case 5: ra = n2; rb = n1; break;
    case 6: ra = n1; rb = n2; break;
    default: return false;
    }
    if (ra != l2 || rb != m2 || f.ori < 1 || f.ori > 4) return false;
    out.clear();
    for (int b = 1; b <= m2; b++)
        for (int a = 1; a <= l2; a++) {
            int ll = 0, mm = 0;
            switch (f.ori) {
            case 1: ll = a; mm = m1 - b + 1; break;
            case 2: ll = b; mm = a; break;
            case 3: ll = l1 - a + 1; mm = b; break;
            case 4: ll = l1 - b + 1; mm = m1 - a + 1; break;
            }
            PairNode p;
            int sp[3];
            ua(ll, mm, sp);
            p.si = sp[0]; p.sj = sp[1]; p.sk = sp[2];
            switch (f.face) {
            case 1: p.ri = f.is; p.rj = f.js + b - 1; p.rk = f.ks + a - 1; break;
            case 2: p.ri = f.ie; p.rj = f.js + a - 1; p.rk = f.ks + b - 1; break;
            case 3: p.ri = f.is + a - 1; p.rj = f.js; p.rk = f.ks + b - 1; break;
            case 4: p.ri = f.is + b - 1; p.rj = f.je; p.rk = f.ks + a - 1; break;
            case 5: p.ri = f.is + b - 1; p.rj = f.js + a - 1; p.rk = f.ks; break;
            case 6: p.ri = f.is + a - 1; p.rj = f.js + b - 1; p.rk = f.ke; break;
            }
            out.push_back(p);
        }
    return true;
}

void inward(int face, int &di, int &dj, int &dk)
{
    int axis = (face - 1) / 2, s = (face & 1) ? 1 : -1;
    di = axis == 0 ? s : 0;
    dj = axis == 1 ? s : 0;
    dk = axis == 2 ? s : 0;
}

// Field pointers of block b from `base` (null: only count); returns the
// block's doubles, rounded up to 512-byte multiples.
long place_block(acfd_ctx *ctx, Block &b, double *base)
{
    const Ext &e = b.e;
    const bool sa = ctx->c.turb == 2 && ctx->c.solver == 0;
    const bool bdf = ctx->cfg.iflag_timeadvance == 3;
    // the tile sweeps exchange dq* and dq through uncached memory
    // field table: pointer slot and component count
    struct Slot {
        double **p;
        int nc;
    } slots[] = {
        {&b.f.q, 5}, {&b.f.pri, 7}, {&b.f.rhsi, 5}, {&b.f.rhsv, 5}, {&b.f.rhs, 5}, {&b.f.dxidx, 9},
        {&b.f.inv_j, 1}, {&b.f.amu, 3}, {&b.f.dt, 1}, {&b.f.dudx, 9}, {&b.f.tao, 6}, {&b.f.vor, 1},
        {&b.f.aal, sa ? 3 : 0}, {&b.f.bbe, sa ? 3 : 0}, {&b.f.gga, sa ? 3 : 0}, {&b.f.dist, sa ? 1 : 0},
        {&b.f.spacing, sa ? 1 : 0}, {&b.f.length, sa ? 1 : 0}, {&b.f.nut, 1}, {&b.f.sa_rhs, 1},
        {&b.f.q0, bdf ? 10 : 0}, {&b.f.w_flux, 12}, {&b.f.w_dqs, 5}, {&b.f.w_dq, 5},
        {&b.f.w_diag, 1}, {&b.f.w_r, 3}, {&b.f.w_sa, sa ? 9 : 0},
    };
    long off = 0;
    for (auto &s : slots) {
        if (base) *s.p = s.nc ? base + off : nullptr;
        off += ((long)s.nc * e.n + 63) / 64 * 64;
    }
    return off;
}

// Every block's fields in one allocation (the arena), so device tables can
// address any block's field as an offset from one base pointer.
int alloc_blocks(acfd_ctx *ctx)
{
    long tot = 0;
    for (auto &b : ctx->blk) {
        b.arena_off = tot;
        tot += place_block(ctx, b, nullptr);
    }
    HIP_TRY(hipMalloc((void **)&ctx->arena, tot * sizeof(double)));
    HIP_TRY(hipMemsetAsync(ctx->arena, 0, tot * sizeof(double), ctx->stream));
    ctx->arena_n = tot;
    for (auto &b : ctx->blk) place_block(ctx, b, ctx->arena + b.arena_off);
    return 0;
}

double *field_ptr(Block &b, int field)
{
    switch (field) {
    case ACFD_Q: return b.f.q;
    case ACFD_PRI_V: return b.f.pri;
    case ACFD_RHSI: return b.f.rhsi;
    case ACFD_RHSV: return b.f.rhsv;
    case ACFD_RHS: return b.f.rhs;
    case ACFD_DUDX: return b.f.dudx;
    case ACFD_TAO: return b.f.tao;
    case ACFD_VOR: return b.f.vor;
    case ACFD_AMU: return b.f.amu;
    case ACFD_DT: return b.f.dt;
    case ACFD_DXIDX: return b.f.dxidx;
    case ACFD_INV_J: return b.f.inv_j;
    case ACFD_AALPHA: return b.f.aal;
    case ACFD_BBETA: return b.f.bbe;
    case ACFD_GGAMMA: return b.f.gga;
    case ACFD_DIST: return b.f.dist;
    case ACFD_SPACING: return b.f.spacing;
    case ACFD_NUT: return b.f.nut;
    case ACFD_SA_RHS: return b.f.sa_rhs;
    case ACFD_Q0: return b.f.q0;
    case ACFD_LENGTH: return b.f.length;
    case ACFD_X: case ACFD_Y: case ACFD_Z: case ACFD_SPCI: case ACFD_SPCJ: case ACFD_SPCK:
        return b.geo ? b.geo + (long)(field - ACFD_X) * b.e.n : nullptr;
    }
    return nullptr;
}

int ensure_stage(acfd_ctx *ctx, long n)
{
    if (ctx->stage_n >= n) return 0;
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    ctx->d_stage = nullptr;
    HIP_TRY(hipMalloc((void **)&ctx->d_stage, n * sizeof(double)));
    ctx->stage_n = n;
    return 0;
}

// One batched grid from the blocks' own grids (device arrays).
int make_grid(BatchGrid &bg, const std::vector<dim3> &g)
{
    std::vector<KGrid> kg;
    std::vector<int> map;
    for (size_t m = 0; m < g.size(); m++) {
        KGrid k{(int)map.size(), (int)g[m].x, (int)g[m].y, (int)g[m].z};
        const long cnt = (long)g[m].x * g[m].y * g[m].z;
        if (map.size() + cnt > (size_t)INT32_MAX) return fail(ACFD_EINVAL, "batched grid too large");
        for (long x = 0; x < cnt; x++) map.push_back((int)m);
        kg.push_back(k);
    }
    bg.total = (int)map.size();
    if (bg.total == 0) return 0;
    int rc = dev_upload(&bg.kg, kg);
    if (!rc) rc = dev_upload(&bg.map, map);
    return rc;
}

void free_grid(BatchGrid &bg)
{
    if (bg.kg) (void)hipFree(bg.kg);
    if (bg.map) (void)hipFree(bg.map);
    bg = BatchGrid();
}

long field_off(const acfd_ctx *ctx, const double *p) { return p ? (long)(p - ctx->arena) : -1; }

std::vector<BlockRec> block_records(const acfd_ctx *ctx)
{
    std::vector<BlockRec> rec;
    for (auto &b : ctx->blk) {
        const Fields &f = b.f;
        BlockRec r;
        r.e = b.e;
        FieldOff &o = r.o;
        o.q = field_off(ctx, f.q); o.pri = field_off(ctx, f.pri); o.rhsi = field_off(ctx, f.rhsi);
        o.rhsv = field_off(ctx, f.rhsv); o.rhs = field_off(ctx, f.rhs); o.dxidx = field_off(ctx, f.dxidx);
        o.inv_j = field_off(ctx, f.inv_j); o.amu = field_off(ctx, f.amu); o.dt = field_off(ctx, f.dt);
        o.dudx = field_off(ctx, f.dudx); o.tao = field_off(ctx, f.tao); o.vor = field_off(ctx, f.vor);
        o.aal = field_off(ctx, f.aal); o.bbe = field_off(ctx, f.bbe); o.gga = field_off(ctx, f.gga);
        o.dist = field_off(ctx, f.dist); o.spacing = field_off(ctx, f.spacing); o.length = field_off(ctx, f.length);
        o.nut = field_off(ctx, f.nut); o.sa_rhs = field_off(ctx, f.sa_rhs); o.q0 = field_off(ctx, f.q0);
        o.w_flux = field_off(ctx, f.w_flux); o.w_dqs = field_off(ctx, f.w_dqs); o.w_dq = field_off(ctx, f.w_dq);
        o.w_diag = field_off(ctx, f.w_diag); o.w_r = field_off(ctx, f.w_r); o.w_sa = field_off(ctx, f.w_sa);
        rec.push_back(r);
    }
    return rec;
}

int ensure_block_table(acfd_ctx *ctx)
{
    if (ctx->d_blocks) return 0;
    return dev_upload(&ctx->d_blocks, block_records(ctx));
}

// Block table and grids of the batched stage launches (batch_blocks).
int build_batch(acfd_ctx *ctx)
{
    if (!ctx->cfg.batch_blocks) return 0;
    std::vector<dim3> g_pri, g_zero, g_inv[3], g_visc, g_prep;
    for (auto &b : ctx->blk) {
        g_pri.push_back(grid_1d(b.e.n));
        g_zero.push_back(grid_1d(5 * b.e.n));
        for (int a = 0; a < 3; a++) g_inv[a].push_back(inviscid_grid(b.e, a));
        g_visc.push_back(viscous_grid(b.e));
        g_prep.push_back(lu_prep_grid(b.e));
    }
    int rc = ensure_block_table(ctx);
    if (!rc) rc = make_grid(ctx->bg_getpri, g_pri);
    if (!rc) rc = make_grid(ctx->bg_inv[0], g_zero);
    for (int a = 0; a < 3 && !rc; a++) rc = make_grid(ctx->bg_inv[a + 1], g_inv[a]);
    if (!rc) rc = make_grid(ctx->bg_visc, g_visc);
    if (!rc) rc = make_grid(ctx->bg_dt, g_pri);
    if (!rc) rc = make_grid(ctx->bg_prep, g_prep);
    // physical boundary rounds; a block's subfaces keep the order of ctx->bc
    for (size_t r = 0; r < ctx->bg_bc.size(); r++) {
        if (ctx->d_bc_round[r]) (void)hipFree(ctx->d_bc_round[r]);
        free_grid(ctx->bg_bc[r]);
    }
    ctx->d_bc_round.clear();
    ctx->bg_bc.clear();
    std::vector<std::vector<BcTask>> per(ctx->blk.size());
    for (const BcTask &t : ctx->bc) per[t.block].push_back(t);
    size_t rounds = 0;
    for (auto &v : per) rounds = std::max(rounds, v.size());
    for (size_t r = 0; r < rounds && !rc; r++) {
        std::vector<BcTask> tasks;
        std::vector<dim3> g;
        for (auto &v : per)
            if (r < v.size()) {
                tasks.push_back(v[r]);
                g.push_back(dim3((unsigned)((std::max(bc_nodes(v[r]), 0L) + 255) / 256)));
            }
        ctx->d_bc_round.push_back(nullptr);
        ctx->bg_bc.push_back(BatchGrid());
        rc = dev_upload(&ctx->d_bc_round.back(), tasks);
        if (!rc) rc = make_grid(ctx->bg_bc.back(), g);
    }
    ctx->batch_ready = rc == 0;
    return rc;
}

bool batched(const acfd_ctx *ctx) { return ctx->cfg.batch_blocks && ctx->batch_ready; }

// Coordinates and edge lengths of every block (x, y, z, spci, spcj, spck).
int ensure_geo(acfd_ctx *ctx)
{
    if (ctx->geo) return 0;
    long tot = 0;
    for (auto &b : ctx->blk) tot += 6 * b.e.n;
    HIP_TRY(hipMalloc((void **)&ctx->geo, tot * sizeof(double)));
    // on the context's stream: a null-stream memset is not ordered with the
    // (non-blocking) context stream that then uploads the coordinates
    HIP_TRY(hipMemsetAsync(ctx->geo, 0, tot * sizeof(double), ctx->stream));
    long off = 0;
    for (auto &b : ctx->blk) {
        b.geo = ctx->geo + off;
        off += 6 * b.e.n;
    }
    return 0;
}

// Workspace of the tile sweeps, sized for the largest block and every tile
// shape: the sweep-layout operand array (ordinary device memory), the
// exchange slots (uncached memory, filled with the sentinel; every sweep
// leaves them that way) and an error word.  Nothing for hyperplane sweeps.
int alloc_lu_ws(acfd_ctx *ctx)
{
    const int nb = (int)ctx->blk.size();
    HIP_TRY(hipMalloc((void **)&ctx->d_cfl_blk, nb * sizeof(double)));
    for (int m = 0; m < nb; m++) HIP_TRY(hipMemcpy(ctx->d_cfl_blk + m, &ctx->cfg.cfl, sizeof(double), hipMemcpyHostToDevice));
    if (ctx->cfg.lusgs_kernel != 1 && ctx->cfg.lusgs_kernel != 2) return 0;
    const bool batch = ctx->cfg.lusgs_kernel == 2;
    std::vector<Ext> es;
    std::vector<Fields> fs;
    for (auto &b : ctx->blk) {
        es.push_back(b.e);
        fs.push_back(b.f);
    }
    long sl_n = 0, ex_n = 0;
    if (batch) {
        // every block has its own part of the workspace; sized for the
        // largest total over the tile shapes
        for (int sh = 0; sh < lusgs_tile_shapes(); sh++) {
            long a = 0, x = 0;
            lusgs_batch_ws(nb, es.data(), sh, &a, &x);
            sl_n = std::max(sl_n, a);
            ex_n = std::max(ex_n, x);
        }
    } else {
        for (auto &b : ctx->blk) {
            long a = 0, x = 0;
            lusgs_ws_sizes(b.e, &a, &x);
            sl_n = std::max(sl_n, a);
            ex_n = std::max(ex_n, x);
        }
    }
    if (sl_n <= 0) return 0;
    HIP_TRY(hipMalloc((void **)&ctx->lu.sl, (size_t)sl_n * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&ctx->lu.err, sizeof(int)));
    HIP_TRY(hipMemsetAsync(ctx->lu.err, 0, sizeof(int), ctx->stream));
    ctx->lu.ex = (double *)uc_acquire((size_t)ex_n * sizeof(double));
    if (!ctx->lu.ex) return fail(ACFD_EDEVICE, "uncached device memory for the tile exchange is unavailable");
    ctx->lu_ex_n = ex_n;
    KRN_TRY(launch_fill_sentinel(ctx->stream, ctx->lu.ex, ex_n));
    if (batch) {
        HIP_TRY(hipMalloc((void **)&ctx->lu_ctr, 2 * sizeof(int)));
        HIP_TRY(hipMemsetAsync(ctx->lu_ctr, 0, 2 * sizeof(int), ctx->stream));
        const bool sa = ctx->c.turb == 2 && ctx->c.solver == 0;
        const char *cc = std::getenv("ACFD_SA_CONCURRENT");
        ctx->sa_concurrent = sa && !(cc && std::atoi(cc) == 0);
        if (ctx->sa_concurrent) {
            // the SA sweeps need fewer slots than the flow sweeps (one value
            // per slot against five)
            ctx->sa_ex = (double *)uc_acquire((size_t)ex_n * sizeof(double));
            if (!ctx->sa_ex) return fail(ACFD_EDEVICE, "uncached device memory for the tile exchange is unavailable");
            KRN_TRY(launch_fill_sentinel(ctx->stream, ctx->sa_ex, ex_n));
            HIP_TRY(hipMalloc((void **)&ctx->sa_ctr, 2 * sizeof(int)));
            HIP_TRY(hipMemsetAsync(ctx->sa_ctr, 0, 2 * sizeof(int), ctx->stream));
        }
        const char *rc_env = std::getenv("ACFD_RES_CONCURRENT");
        ctx->res_concurrent = ctx->c.solver == 0 && !(rc_env && std::atoi(rc_env) == 0);
        if (ctx->sa_concurrent || ctx->res_concurrent) {
            HIP_TRY(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ctx->ev_rfork, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ctx->ev_rjoin, hipEventDisableTiming));
        }
        if (sa) {
            // the SA sweep layout, sized for the largest tile shape, and
            // the grid of the SA model's per-node passes
            long sa_sl_n = 0;
            for (int sh = 0; sh < lusgs_tile_shapes(); sh++) {
                long a = 0;
                sa_batch_sl(nb, es.data(), sh, &a);
                sa_sl_n = std::max(sa_sl_n, a);
            }
            if (sa_sl_n > 0) HIP_TRY(hipMalloc((void **)&ctx->sa_sl, (size_t)sa_sl_n * sizeof(double)));
            std::vector<dim3> g;
            for (auto &b : ctx->blk) g.push_back(lin_grid(b.e));
            int rc = ensure_block_table(ctx);
            if (!rc) rc = make_grid(ctx->bg_sa, g);
            if (rc) return rc;
        }
        for (int sh = 0; sh < lusgs_tile_shapes(); sh++) {
            if (lusgs_batch_make(nb, es.data(), fs.data(), ctx->arena, sh, ctx->lu.sl, ctx->lu.ex, &ctx->lub[sh]))
                return fail(ACFD_EDEVICE, "batched LU-SGS tables: device allocation failed");
            if (sa && sa_batch_make(nb, es.data(), fs.data(), ctx->arena, sh, ctx->sa_concurrent ? ctx->sa_ex : ctx->lu.ex,
                                    ctx->sa_sl, &ctx->sab[sh]))
                return fail(ACFD_EDEVICE, "batched SA sweep tables: device allocation failed");
            if (ctx->sa_concurrent) {
                // two persistent launches share the device: tiles by ticket
                // (each waits only on tiles that are already running)
                ctx->lub[sh].resident[0] = ctx->lub[sh].resident[1] = 0;
                ctx->sab[sh].resident[0] = ctx->sab[sh].resident[1] = 0;
            }
        }
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

// A tile sweep that gave up waiting leaves its results invalid and some
// exchange slots written: report it and restore the slots.
int tile_error(acfd_ctx *ctx)
{
    if (!ctx->lu.err) return 0;
    int err = 0;
    HIP_TRY(hipMemcpy(&err, ctx->lu.err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        (void)launch_fill_sentinel(ctx->stream, ctx->lu.ex, ctx->lu_ex_n);
        if (ctx->sa_ex) (void)launch_fill_sentinel(ctx->stream, ctx->sa_ex, ctx->lu_ex_n);
        (void)hipMemsetAsync(ctx->lu.err, 0, sizeof(int), ctx->stream);
        if (ctx->lu_ctr) (void)hipMemsetAsync(ctx->lu_ctr, 0, 2 * sizeof(int), ctx->stream);
        if (ctx->sa_ctr) (void)hipMemsetAsync(ctx->sa_ctr, 0, 2 * sizeof(int), ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
        return fail(ACFD_EDEVICE, "pipelined LU-SGS sweep timed out waiting for a neighbouring tile");
    }
    return 0;
}

constexpr size_t kEventPairs = 2048;

void timing_flush(acfd_ctx *ctx)
{
    if (!ctx->ev_used) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (size_t x = 0; x < ctx->ev_used; x++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev_pool[2 * x], ctx->ev_pool[2 * x + 1]) == hipSuccess) {
            ctx->stage_ms[ctx->ev_stage[x]] += ms;
            ctx->stage_calls[ctx->ev_stage[x]] += 1;
        }
    }
    ctx->ev_used = 0;
}

// Records an event pair around a stage on the context's stream (no wait).
struct StageTimer {
    acfd_ctx *ctx;
    size_t slot = 0;
    bool on;
    StageTimer(acfd_ctx *c, int s) : ctx(c), on(c->timing)
    {
        if (!on) return;
        if (ctx->ev_pool.empty()) {
            ctx->ev_pool.resize(2 * kEventPairs);
            ctx->ev_stage.resize(kEventPairs);
            for (auto &ev : ctx->ev_pool) (void)hipEventCreate(&ev);
        }
        if (ctx->ev_used == kEventPairs) timing_flush(ctx);
        slot = ctx->ev_used++;
        ctx->ev_stage[slot] = s;
        (void)hipEventRecord(ctx->ev_pool[2 * slot], ctx->stream);
    }
    ~StageTimer()
    {
        if (on) (void)hipEventRecord(ctx->ev_pool[2 * slot + 1], ctx->stream);
    }
};

int check_nneg(acfd_ctx *ctx)
{
    unsigned int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, ctx->d_nneg, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (h) {
        HIP_TRY(hipMemsetAsync(ctx->d_nneg, 0, sizeof(unsigned int), ctx->stream));
        return fail(ACFD_ENUMERIC, "negative density found at " + std::to_string(h) + " nodes");
    }
    return 0;
}

// ---------------------------------------------------------------- stages
int st_getpri(acfd_ctx *ctx, int m)
{
    StageTimer t(ctx, ACFD_STAGE_GETPRI);
    KRN_TRY(launch_getpri(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, ctx->d_nneg));
    return 0;
}

// WCNS with the learned sensor on block m; every s_audit evaluations the
// plain scheme runs too and a relative deviation above s_bound makes the
// block fall back to plain WCNS for good.
int inviscid_with_sensor(acfd_ctx *ctx, int m)
{
    Block &b = ctx->blk[m];
    const long call = ctx->s_calls[m]++;
    SensorArgs sa{ctx->s_net, ctx->s_thr, ctx->s_stats, nullptr, 0};
    if (ctx->s_refresh > 1) {
        sa.mask = ctx->s_mask[m];
        sa.replay = call % ctx->s_refresh != 0;
    }
    if (ctx->s_audit > 0 && call % ctx->s_audit == 0) {
        KRN_TRY(launch_inviscid_ext(ctx->stream, b.e, b.f, ctx->c, ctx->cfg.precision, 2));
        KRN_TRY(launch_copy(ctx->stream, b.f.rhsi, b.f.w_flux, 5 * b.e.n));
        KRN_TRY(launch_inviscid_wcns_sensor(ctx->stream, b.e, b.f, ctx->c, ctx->cfg.precision, sa));
        const long np = diff_partials(b.e);
        if (ctx->s_dpart_n < 2 * np) {
            if (ctx->s_dpart) (void)hipFree(ctx->s_dpart);
            ctx->s_dpart = nullptr;
            HIP_TRY(hipMalloc((void **)&ctx->s_dpart, 2 * np * sizeof(double)));
            ctx->s_dpart_n = 2 * np;
        }
        KRN_TRY(launch_diff_partial(ctx->stream, b.e, b.f.rhsi, b.f.w_flux, ctx->s_dpart));
        std::vector<double> h(2 * np);
        HIP_TRY(hipMemcpyAsync(h.data(), ctx->s_dpart, 2 * np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double d2 = 0.0, r2 = 0.0;
        for (long x = 0; x < np; x++) {
            d2 += h[2 * x];
            r2 += h[2 * x + 1];
        }
        const double dev = r2 > 0.0 ? std::sqrt(d2 / r2) : std::sqrt(d2);
        ctx->s_audits++;
        ctx->s_audit_max = std::max(ctx->s_audit_max, dev);
        if (!(dev <= ctx->s_bound)) {
            ctx->s_off[m] = 1;
            KRN_TRY(launch_copy(ctx->stream, b.f.w_flux, b.f.rhsi, 5 * b.e.n));
        }
        return 0;
    }
    KRN_TRY(launch_inviscid_wcns_sensor(ctx->stream, b.e, b.f, ctx->c, ctx->cfg.precision, sa));
    return 0;
}

int st_inviscid(acfd_ctx *ctx, int m)
{
    StageTimer t(ctx, ACFD_STAGE_INVISCID);
    if (ctx->s_on && ctx->cfg.inviscid_scheme == 2 && !ctx->s_off[m]) return inviscid_with_sensor(ctx, m);
    if (ctx->cfg.inviscid_scheme == 0)
        KRN_TRY(launch_inviscid(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, ctx->cfg.precision));
    else
        KRN_TRY(launch_inviscid_ext(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, ctx->cfg.precision,
                                    ctx->cfg.inviscid_scheme));
    return 0;
}

int st_viscous(acfd_ctx *ctx, int m, int store_aux)
{
    StageTimer t(ctx, ACFD_STAGE_VISCOUS);
    KRN_TRY(launch_viscous(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, ctx->cfg.precision, store_aux));
    return 0;
}

int st_find_dt(acfd_ctx *ctx, int m)
{
    StageTimer t(ctx, ACFD_STAGE_FIND_DT);
    if (ctx->c.itime == 0) KRN_TRY(launch_reset_stats(ctx->stream, ctx->d_stats));
    KRN_TRY(launch_find_dt(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, ctx->d_cfl, ctx->d_cfl_blk + m,
                           ctx->d_stats));
    return 0;
}

// sub-stage timing ids of the LU-SGS phases
constexpr int kTimeLuForward = 12, kTimeLuBackward = 13, kTimeLuGather = 14, kTimeLuOther = 15,
              kTimeLuAdd = 16;
// single-kernel timers of the batched stages (recorded only with timing on,
// when each stage's kernels are launched one by one)
constexpr int kTimeSaForward = 17, kTimeSaBackward = 18, kTimeSaPre = 19, kTimeSaAdd = 20, kTimeInvZero = 21,
              kTimeInvI = 22, kTimeInvJ = 23, kTimeInvK = 24, kTimeViscNode = 25, kTimeViscDiff = 26,
              kTimeSaGather = 27, kTimeSaPreB = 28, kTimeViscNodeDt = 29;

// With batched stages, NS and the steady local time step, find_dt runs in
// the viscous node pass (ACFD_VISCOUS_DT_FUSED=0: its own pass; the same
// results either way).
bool viscous_dt_fused_on()
{
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("ACFD_VISCOUS_DT_FUSED");
        v = e && std::atoi(e) == 0 ? 0 : 1;
    }
    return v == 1;
}

// The batched SA update moves its sweep operands into a sweep layout
// (ACFD_SA_LAYOUT=0: the sweeps read them in place, as the per-block sweeps
// do; the same results either way).
bool sa_layout_on()
{
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("ACFD_SA_LAYOUT");
        v = e && std::atoi(e) == 0 ? 0 : 1;
    }
    return v == 1;
}

// Autotuner of the tile sweep: on a block's first update, time the gather
// and forward sweep once per tile shape (they recompute dq* in the sweep
// layout from unchanged inputs, so repeating them changes nothing) and keep
// the fastest shape.
bool geo_cache_on()
{
    static const bool on = [] {
        const char *v = std::getenv("ACFD_SL_GEO_CACHE");
        return !(v && std::atoi(v) == 0);
    }();
    return on;
}

// Whether the sweep layout already holds the grid operands (dxidx, inv_j)
// of `block` (-1: the batch) in tiling `shape`, so that the gather about to
// be launched may leave them in place; records that it holds them after.
// The sweeps write only the result fields, and the grid operands change
// only through sl_invalidate's callers.  ACFD_SL_GEO_CACHE=0 gathers every
// field every time.
int sl_geo_cached(acfd_ctx *ctx, int block, int shape)
{
    const bool same = ctx->sl_block == block && ctx->sl_shape == shape && geo_cache_on();
    ctx->sl_block = block;
    ctx->sl_shape = shape;
    if (same && ctx->capturing) ctx->graph_geo_cached = true;
    return same ? 1 : 0;
}

// the metrics of some block were written: the next gathers move every field
void sl_invalidate(acfd_ctx *ctx)
{
    ctx->sl_block = -2;
    ctx->sl_shape = -1;
}

int tune_tiles(acfd_ctx *ctx, Block &b)
{
    if (const char *fix = std::getenv("ACFD_TILE_SHAPE")) {
        int s = std::atoi(fix);
        if (s >= 0 && s < lusgs_tile_shapes()) {
            b.tile_shape = s;
            return 0;
        }
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int best_s = 0;
    for (int s = 0; s < lusgs_tile_shapes(); s++) {
        if (lusgs_tiles(b.e, s) <= 0) continue;
        (void)hipEventRecord(e0, ctx->stream);
        int rc = launch_lusgs_phase(ctx->stream, b.e, b.f, ctx->c, ctx->d_cfl, 1, &ctx->lu, 1, s,
                                    sl_geo_cached(ctx, (int)(&b - ctx->blk.data()), s));
        if (!rc) rc = launch_lusgs_phase(ctx->stream, b.e, b.f, ctx->c, ctx->d_cfl, 1, &ctx->lu, 2, s);
        (void)hipEventRecord(e1, ctx->stream);
        if (rc) return fail(ACFD_EDEVICE, "tile sweep launch failed during tuning");
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (std::getenv("ACFD_VERBOSE"))
            std::fprintf(stderr, "acfd: tile sweep shape %d (%ld tiles): %.3f ms\n", s, lusgs_tiles(b.e, s), ms);
        if (ms < best) {
            best = ms;
            best_s = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b.tile_shape = best_s;
    return 0;
}

int st_implicit(acfd_ctx *ctx, int m, const double *cfl)
{
    StageTimer t(ctx, ACFD_STAGE_IMPLICIT1ST);
    Block &b = ctx->blk[m];
    const int variant = ctx->cfg.lusgs_kernel;
    for (int ph = 0; ph < kLusgsPhases; ph++) {
        if (ph == 1 && variant == 1 && b.tile_shape < 0) {
            int rc = tune_tiles(ctx, b);
            if (rc) return rc;
        }
        StageTimer ts(ctx, ph == 2 ? kTimeLuForward : ph == 3 ? kTimeLuBackward
                           : ph == 1 ? kTimeLuGather : ph == 4 ? kTimeLuAdd : kTimeLuOther);
        const int shape = b.tile_shape < 0 ? 0 : b.tile_shape;
        KRN_TRY(launch_lusgs_phase(ctx->stream, b.e, b.f, ctx->c, cfl, variant, &ctx->lu, ph, shape,
                                   ph == 1 && variant == 1 ? sl_geo_cached(ctx, m, shape) : 0));
    }
    return 0;
}

// Autotuner of the batched sweep: the same timing as tune_tiles over the
// whole batch, one shape for all blocks.
int tune_batch(acfd_ctx *ctx)
{
    if (const char *fix = std::getenv("ACFD_TILE_SHAPE")) {
        int s = std::atoi(fix);
        if (s >= 0 && s < lusgs_tile_shapes()) {
            ctx->lub_shape = s;
            return 0;
        }
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int best_s = 0;
    for (int s = 0; s < lusgs_tile_shapes(); s++) {
        const LuBatch &lb = ctx->lub[s];
        if (lb.total <= 0) continue;
        (void)hipEventRecord(e0, ctx->stream);
        int rc = launch_lusgs_batch(ctx->stream, lb, 1, ctx->c.gamma, ctx->lu.err, ctx->lu_ctr,
                                    sl_geo_cached(ctx, -1, s));
        if (!rc) rc = launch_lusgs_batch(ctx->stream, lb, 2, ctx->c.gamma, ctx->lu.err, ctx->lu_ctr);
        (void)hipEventRecord(e1, ctx->stream);
        if (rc) return fail(ACFD_EDEVICE, "batched tile sweep launch failed during tuning");
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (std::getenv("ACFD_VERBOSE"))
            std::fprintf(stderr, "acfd: batched sweep shape %d (%d tiles, %d blocks): %.3f ms\n", s, lb.total,
                         lb.nent, ms);
        if (ms < best) {
            best = ms;
            best_s = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->lub_shape = best_s;
    return 0;
}

// implicit1st of every block, the sweeps batched over the blocks.  Blocks
// only meet at the halo refresh, so updating all of them after all their
// residuals equals updating each after its own; block m's update reads the
// cfl its find_dt left (cfl + m), or the global one when cfl_per_block is
// null.
//
// The batched update never writes w_dqs / w_dq (its preparation does not
// clear them, and its sweeps keep dq* and dq in the sweep layout).  The SA
// update uses those two work fields for its cu and dnut, and runs beside
// this update on a second stream only because of that.
constexpr bool kBatchedUpdateWritesDq = false;
int st_implicit_all(acfd_ctx *ctx, const double *cfl_per_block)
{
    const int nb = (int)ctx->blk.size();
    if (ctx->cfg.lusgs_kernel != 2) {
        for (int m = 0; m < nb; m++) {
            int rc = st_implicit(ctx, m, cfl_per_block ? cfl_per_block + m : ctx->d_cfl);
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_IMPLICIT1ST);
    const bool prep_ready = ctx->lu_prep_ready;
    ctx->lu_prep_ready = false;
    if (prep_ready) {
        // formed by the residual's viscous passes
    } else if (batched(ctx)) {
        StageTimer ts(ctx, kTimeLuOther);
        KRN_TRY(launch_lu_prep_batch(ctx->stream, ctx->d_blocks, ctx->bg_prep, ctx->arena,
                                     cfl_per_block ? cfl_per_block : ctx->d_cfl, cfl_per_block ? 1 : 0));
    } else {
        StageTimer ts(ctx, kTimeLuOther);
        for (int m = 0; m < nb; m++) {
            const Block &b = ctx->blk[m];
            KRN_TRY(launch_lusgs_phase(ctx->stream, b.e, b.f, ctx->c, cfl_per_block ? cfl_per_block + m : ctx->d_cfl,
                                       1, &ctx->lu, 0, 0));
        }
    }
    if (ctx->lub_shape < 0) {
        int rc = tune_batch(ctx);
        if (rc) return rc;
    }
    const LuBatch &lb = ctx->lub[ctx->lub_shape];
    for (int ph = 1; ph < kLusgsPhases; ph++) {
        StageTimer ts(ctx, ph == 2 ? kTimeLuForward : ph == 3 ? kTimeLuBackward : ph == 1 ? kTimeLuGather : kTimeLuAdd);
        KRN_TRY(launch_lusgs_batch(ctx->stream, lb, ph, ctx->c.gamma, ctx->lu.err, ctx->lu_ctr,
                                   ph == 1 ? sl_geo_cached(ctx, -1, ctx->lub_shape) : 0));
    }
    return 0;
}

int st_sa(acfd_ctx *ctx, int m, const double *cfl)
{
    StageTimer t(ctx, ACFD_STAGE_SA);
    const Block &b = ctx->blk[m];
    if ((ctx->cfg.lusgs_kernel == 1 || ctx->cfg.lusgs_kernel == 2) && ctx->lu.ex) {
        // tile sweeps in the flow sweep's exchange slots (idle, sentinel-filled)
        const int shape = b.tile_shape >= 0 ? b.tile_shape : ctx->lub_shape >= 0 ? ctx->lub_shape : 0;
        const SaSweep sw{shape, ctx->sa_ex ? ctx->sa_ex : ctx->lu.ex, ctx->lu.err};
        KRN_TRY(launch_sa(ctx->stream, b.e, b.f, ctx->c, cfl, &sw));
    } else {
        KRN_TRY(launch_sa(ctx->stream, b.e, b.f, ctx->c, cfl));
    }
    return 0;
}

// sa_model of every block, the sweeps batched over the blocks with
// lusgs_kernel 2 (block m reads cfl_of(m)).
template <typename F>
int st_sa_all(acfd_ctx *ctx, F cfl_of)
{
    const int nb = (int)ctx->blk.size();
    // the scalar SA sweeps run 8x8 tiles whatever the flow sweeps chose:
    // measured the fastest of the four shapes on the channel (SA stage
    // 32.2 ms against 34.0 at 16x8 and 36.4 at 16x16) and the flat plate
    int shape = 2;
    if (const char *v = std::getenv("ACFD_SA_TILE_SHAPE")) {
        const int s = std::atoi(v);
        if (s >= 0 && s < 4) shape = s;
    }
    if (ctx->cfg.lusgs_kernel != 2 || ctx->sab[shape].total <= 0) {
        for (int m = 0; m < nb; m++) {
            int rc = st_sa(ctx, m, cfl_of(m));
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_SA);
    int *ctr = ctx->sa_ctr ? ctx->sa_ctr : ctx->lu_ctr;
    if (ctx->sa_sl && ctx->bg_sa.total > 0 && sa_layout_on()) {
        // two per-node passes over all blocks, the operands into the SA
        // sweep layout, the two sweeps over it, nut += dnut
        const LuBatch &sb = ctx->sab[shape];
        const double *c0 = cfl_of(0);
        const int stride = nb > 1 && cfl_of(1) != c0 ? 1 : 0;
        for (int part = ctx->timing ? 1 : 0; part <= (ctx->timing ? 2 : 0); part++) {
            StageTimer ts(ctx, part == 2 ? kTimeSaPreB : kTimeSaPre);
            KRN_TRY(launch_sa_pre_batch(ctx->stream, ctx->d_blocks, ctx->bg_sa, ctx->arena, ctx->c, c0, stride, part,
                                        ctx->abg_derived, !ctx->sa_length_ready));
        }
        if (ctx->c.des != 2) ctx->sa_length_ready = true;  // DDES's depends on the flow
        {
            StageTimer ts(ctx, kTimeSaGather);
            KRN_TRY(launch_sa_sl_batch(ctx->stream, sb, ctx->lu.err, ctr, 0));
        }
        {
            StageTimer ts(ctx, kTimeSaForward);
            KRN_TRY(launch_sa_sl_batch(ctx->stream, sb, ctx->lu.err, ctr, 1));
        }
        {
            StageTimer ts(ctx, kTimeSaBackward);
            KRN_TRY(launch_sa_sl_batch(ctx->stream, sb, ctx->lu.err, ctr, -1));
        }
        StageTimer ts(ctx, kTimeSaAdd);
        KRN_TRY(launch_sa_sl_batch(ctx->stream, sb, ctx->lu.err, ctr, 2));
        return 0;
    }
    {
        StageTimer ts(ctx, kTimeSaPre);
        for (int m = 0; m < nb; m++)
            KRN_TRY(launch_sa_pre(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c, cfl_of(m)));
    }
    if (ctx->timing) {
        {
            StageTimer ts(ctx, kTimeSaForward);
            KRN_TRY(launch_sa_batch(ctx->stream, ctx->sab[shape], ctx->lu.err, ctr, 1));
        }
        StageTimer ts(ctx, kTimeSaBackward);
        KRN_TRY(launch_sa_batch(ctx->stream, ctx->sab[shape], ctx->lu.err, ctr, -1));
    } else {
        KRN_TRY(launch_sa_batch(ctx->stream, ctx->sab[shape], ctx->lu.err, ctr, 0));
    }
    StageTimer ts(ctx, kTimeSaAdd);
    for (int m = 0; m < nb; m++) KRN_TRY(launch_sa_post(ctx->stream, ctx->blk[m].e, ctx->blk[m].f));
    return 0;
}


int st_crank(acfd_ctx *ctx, int m)
{
    StageTimer t(ctx, ACFD_STAGE_CRANK2ND);
    KRN_TRY(launch_crank(ctx->stream, ctx->blk[m].e, ctx->blk[m].f, ctx->c));
    return 0;
}

int st_update_q0(acfd_ctx *ctx, int m)
{
    StageTimer t(ctx, ACFD_STAGE_UPDATE_Q0);
    KRN_TRY(launch_update_q0(ctx->stream, ctx->blk[m].e, ctx->blk[m].f));
    return 0;
}

int st_physical_bc(acfd_ctx *ctx)
{
    StageTimer t(ctx, ACFD_STAGE_PHYSICAL_BC);
    if (batched(ctx)) {
        for (size_t r = 0; r < ctx->bg_bc.size(); r++)
            KRN_TRY(launch_bc_batch(ctx->stream, ctx->d_bc_round[r], ctx->bg_bc[r], ctx->d_blocks, ctx->arena,
                                    ctx->c));
        if (ctx->c.turb == 2 && ctx->c.solver == 0)
            KRN_TRY(launch_sa_clamp_batch(ctx->stream, ctx->d_blocks, ctx->bg_getpri, ctx->arena));
        return 0;
    }
    for (const BcTask &task : ctx->bc) {
        Block &b = ctx->blk[task.block];
        KRN_TRY(launch_bc(ctx->stream, b.e, b.f, ctx->c, task));
    }
    if (ctx->c.turb == 2 && ctx->c.solver == 0)
        for (auto &b : ctx->blk) KRN_TRY(launch_sa_clamp(ctx->stream, b.e, b.f));
    return 0;
}

int st_update_buffer(acfd_ctx *ctx)
{
    StageTimer t(ctx, ACFD_STAGE_UPDATE_BUFFER);
    CopyList cl{ctx->halo_n, ctx->d_h_sb, ctx->d_h_db, ctx->d_h_sn, ctx->d_h_dn};
    KRN_TRY(launch_halo_copy(ctx->stream, cl, ctx->d_qtab, ctx->d_nuttab, ctx->d_ntab, ctx->c.turb == 2));
    return 0;
}

int st_average(acfd_ctx *ctx)
{
    StageTimer t(ctx, ACFD_STAGE_AVERAGE);
    KRN_TRY(launch_average(ctx->stream, ctx->avg, ctx->d_qtab, ctx->d_nuttab, ctx->d_ntab, ctx->d_a_buf,
                           ctx->c.turb == 2));
    return 0;
}

int refresh(acfd_ctx *ctx)
{
    int rc = st_physical_bc(ctx);
    if (!rc) rc = st_update_buffer(ctx);
    if (!rc) rc = st_average(ctx);
    return rc;
}

int residual_stage(acfd_ctx *ctx, int m, int store_aux, bool pri_ok = false)
{
    int rc = pri_ok ? 0 : st_getpri(ctx, m);
    if (!rc) rc = st_inviscid(ctx, m);
    if (!rc && ctx->c.solver == 0) rc = st_viscous(ctx, m, store_aux);
    if (!rc) rc = st_find_dt(ctx, m);
    return rc;
}

// A stage over every block: one launch per kernel with batch_blocks, else
// block after block.  Blocks only meet at the halo refresh, so running a
// stage for all blocks before the next stage equals running each block's
// stages back to back.
int st_getpri_all(acfd_ctx *ctx)
{
    if (!batched(ctx)) {
        for (int m = 0; m < (int)ctx->blk.size(); m++) {
            int rc = st_getpri(ctx, m);
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_GETPRI);
    KRN_TRY(launch_getpri_batch(ctx->stream, ctx->d_blocks, ctx->bg_getpri, ctx->arena, ctx->c, ctx->d_nneg));
    return 0;
}

int st_inviscid_all(acfd_ctx *ctx)
{
    // the batched launch covers the reference scheme without the sensor
    if (!batched(ctx) || ctx->cfg.inviscid_scheme != 0) {
        for (int m = 0; m < (int)ctx->blk.size(); m++) {
            int rc = st_inviscid(ctx, m);
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_INVISCID);
    // rhsi outside the computed ranges needs clearing only after an upload
    const int first = ctx->rhsi_clean ? 1 : 0;
    ctx->rhsi_clean = true;
    if (ctx->timing) {
        for (int part = first; part < 4; part++) {
            StageTimer ts(ctx, kTimeInvZero + part);
            KRN_TRY(launch_inviscid_batch(ctx->stream, ctx->d_blocks, ctx->bg_inv, ctx->arena, ctx->c,
                                          ctx->cfg.precision, part));
        }
        return 0;
    }
    KRN_TRY(launch_inviscid_batch(ctx->stream, ctx->d_blocks, ctx->bg_inv, ctx->arena, ctx->c, ctx->cfg.precision,
                                  first ? -2 : -1));
    return 0;
}

int st_viscous_all(acfd_ctx *ctx, int store_aux)
{
    if (!batched(ctx)) {
        for (int m = 0; m < (int)ctx->blk.size(); m++) {
            int rc = st_viscous(ctx, m, store_aux);
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_VISCOUS);
    if (ctx->timing) {
        for (int part = 1; part <= 2; part++) {
            StageTimer ts(ctx, part == 1 ? kTimeViscNode : kTimeViscDiff);
            KRN_TRY(launch_viscous_batch(ctx->stream, ctx->d_blocks, ctx->bg_visc, ctx->arena, ctx->c,
                                         ctx->cfg.precision, store_aux, part));
        }
        return 0;
    }
    KRN_TRY(launch_viscous_batch(ctx->stream, ctx->d_blocks, ctx->bg_visc, ctx->arena, ctx->c, ctx->cfg.precision,
                                 store_aux));
    return 0;
}

int st_find_dt_all(acfd_ctx *ctx)
{
    if (!batched(ctx)) {
        for (int m = 0; m < (int)ctx->blk.size(); m++) {
            int rc = st_find_dt(ctx, m);
            if (rc) return rc;
        }
        return 0;
    }
    StageTimer t(ctx, ACFD_STAGE_FIND_DT);
    if (ctx->c.itime == 0) KRN_TRY(launch_reset_stats(ctx->stream, ctx->d_stats));
    KRN_TRY(launch_find_dt_batch(ctx->stream, ctx->d_blocks, (int)ctx->blk.size(), ctx->bg_dt, ctx->arena, ctx->c,
                                 ctx->d_cfl, ctx->d_cfl_blk, ctx->d_stats));
    return 0;
}

// viscous_flux and the steady find_dt of every block, the find_dt folded
// into the viscous node pass (the unsteady find_dt evaluates one node per
// block, nothing to fold)
// parts -1: both passes; 1: the node pass; 2: the difference pass
int st_viscous_dt_all(acfd_ctx *ctx, int store_aux, int parts = -1)
{
    StageTimer t(ctx, ACFD_STAGE_VISCOUS);
    const int nb = (int)ctx->blk.size();
    // the batched LU-SGS update's time term and rhs come from these two
    // passes too (implicit1st then skips its own preparation)
    const int prep = ctx->cfg.lusgs_kernel == 2 && ctx->cfg.iflag_timeadvance == 2 ? 1 : 0;
    if (parts >= 0) {
        if (parts == 1) KRN_TRY(launch_reset_stats(ctx->stream, ctx->d_stats));
        KRN_TRY(launch_viscous_dt_batch(ctx->stream, ctx->d_blocks, nb, ctx->bg_visc, ctx->arena, ctx->c,
                                        ctx->cfg.precision, store_aux, ctx->d_cfl, ctx->d_stats, parts, prep));
        if (parts == 2) {
            ctx->lu_prep_ready = prep == 1;
            if (prep) ctx->rhsv_stale = true;
        }
        return 0;
    }
    KRN_TRY(launch_reset_stats(ctx->stream, ctx->d_stats));
    if (ctx->timing) {
        for (int part = 1; part <= 2; part++) {
            StageTimer ts(ctx, part == 1 ? kTimeViscNodeDt : kTimeViscDiff);
            KRN_TRY(launch_viscous_dt_batch(ctx->stream, ctx->d_blocks, nb, ctx->bg_visc, ctx->arena, ctx->c,
                                            ctx->cfg.precision, store_aux, ctx->d_cfl, ctx->d_stats, part, prep));
        }
    } else {
        KRN_TRY(launch_viscous_dt_batch(ctx->stream, ctx->d_blocks, nb, ctx->bg_visc, ctx->arena, ctx->c,
                                        ctx->cfg.precision, store_aux, ctx->d_cfl, ctx->d_stats, -1, prep));
    }
    ctx->lu_prep_ready = prep == 1;
    if (prep) ctx->rhsv_stale = true;
    return 0;
}

// rhsv of the last residual stage where that stage did not store it
int rhsv_refresh(acfd_ctx *ctx)
{
    if (!ctx->rhsv_stale) return 0;
    ctx->rhsv_stale = false;
    KRN_TRY(launch_viscous_batch(ctx->stream, ctx->d_blocks, ctx->bg_visc, ctx->arena, ctx->c, ctx->cfg.precision, 0,
                                 2));
    return 0;
}

int residual_stage_all(acfd_ctx *ctx, int store_aux, bool pri_ok = false)
{
    int rc = pri_ok ? 0 : st_getpri_all(ctx);
    const bool fused = ctx->c.solver == 0 && ctx->c.itime == 0 && batched(ctx) && viscous_dt_fused_on();
    if (!rc && fused && ctx->res_concurrent && ctx->cfg.inviscid_scheme == 0 && !ctx->timing) {
        // the inviscid kernels read q and the metrics and write rhsi; the
        // viscous node pass reads pri_v, amu and the metrics and writes
        // neither: it runs on the second stream meanwhile, and the
        // difference pass (rhs = rhsi + rhsv) after both
        HIP_TRY(hipEventRecord(ctx->ev_rfork, ctx->stream));
        HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_rfork, 0));
        std::swap(ctx->stream, ctx->stream2);
        rc = st_viscous_dt_all(ctx, store_aux, 1);
        std::swap(ctx->stream, ctx->stream2);
        if (!rc) rc = st_inviscid_all(ctx);
        HIP_TRY(hipEventRecord(ctx->ev_rjoin, ctx->stream2));
        HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_rjoin, 0));
        return rc ? rc : st_viscous_dt_all(ctx, store_aux, 2);
    }
    if (!rc) rc = st_inviscid_all(ctx);
    if (!rc && fused) return st_viscous_dt_all(ctx, store_aux);
    if (!rc && ctx->c.solver == 0) rc = st_viscous_all(ctx, store_aux);
    if (!rc) rc = st_find_dt_all(ctx);
    return rc;
}

int launch_residual_all(acfd_ctx *ctx)
{
    if (batched(ctx)) {
        const int nb = (int)ctx->blk.size();
        KRN_TRY(launch_residual_batch(ctx->stream, ctx->d_blocks, nb, ctx->arena, ctx->d_partial,
                                      ctx->c.turb == 2 && ctx->c.solver == 0));
        KRN_TRY(launch_residual_final(ctx->stream, ctx->d_partial, nb * residual_partials(ctx->blk[0].e), ctx->d_res));
        return 0;
    }
    int parts = 0;
    for (auto &b : ctx->blk) {
        int p = residual_partials(b.e);
        KRN_TRY(launch_residual(ctx->stream, b.e, b.f, ctx->d_partial + 6 * parts,
                                ctx->c.turb == 2 && ctx->c.solver == 0));
        parts += p;
    }
    KRN_TRY(launch_residual_final(ctx->stream, ctx->d_partial, parts, ctx->d_res));
    return 0;
}

int one_step(acfd_ctx *ctx)
{
    const bool sa = ctx->c.turb == 2 && ctx->c.solver == 0;
    const int aux = sa ? (ctx->c.des == 2 ? 2 : 1) : 0;
    const int nb = (int)ctx->blk.size();
    int rc = 0;
    // block m's update reads the cfl its own find_dt left (the reference
    // runs each block's find_dt, implicit1st and sa_model back to back)
    const double *cfl = ctx->c.itime == 0 ? nullptr : ctx->d_cfl_blk;
    auto cfl_of = [&](int m) { return cfl ? cfl + m : ctx->d_cfl; };
    const bool stagewise = ctx->cfg.batch_blocks || ctx->cfg.lusgs_kernel == 2;
    // pri_v of every block is getpri of its q when the last step's closing
    // getpri is the latest write (a block's q changes only in its own update,
    // after its residual)
    const bool pri_ok = ctx->pri_current;
    ctx->pri_current = false;
    if (pri_ok && ctx->capturing && (ctx->cfg.iflag_timeadvance == 2 || ctx->cfg.iflag_timeadvance == 4))
        ctx->graph_needs_pri = true;
    if (ctx->cfg.iflag_timeadvance == 2 && stagewise) {
        rc = residual_stage_all(ctx, aux, pri_ok);
        static_assert(!kBatchedUpdateWritesDq, "the concurrent SA update owns w_dqs / w_dq");
        if (!rc && sa && ctx->sa_concurrent && ctx->lub_shape >= 0 && !ctx->timing) {
            // sa_model reads nothing implicit1st writes (pri_v, amu, vor, dt
            // and the cfl are from the residual) and implicit1st reads no
            // SA field: the SA update runs on the second stream meanwhile
            // (not under the per-stage timers, whose stage times would then
            // include each other's)
            HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
            HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
            std::swap(ctx->stream, ctx->stream2);
            rc = st_sa_all(ctx, cfl_of);
            std::swap(ctx->stream, ctx->stream2);
            if (!rc) rc = st_implicit_all(ctx, cfl ? cfl : nullptr);
            HIP_TRY(hipEventRecord(ctx->ev_join, ctx->stream2));
            HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
        } else {
            if (!rc) rc = st_implicit_all(ctx, cfl ? cfl : nullptr);
            if (!rc && sa) rc = st_sa_all(ctx, cfl_of);
        }
        if (!rc) rc = refresh(ctx);
        if (!rc) rc = st_getpri_all(ctx);
    } else if (ctx->cfg.iflag_timeadvance == 3 && stagewise) {
        for (int sub = 0; sub < 3 && !rc; sub++) {
            rc = refresh(ctx);
            if (!rc) rc = residual_stage_all(ctx, aux);
            for (int m = 0; m < nb && !rc; m++) rc = st_crank(ctx, m);
        }
        for (int m = 0; m < nb && !rc; m++) rc = st_update_q0(ctx, m);
        // after the sub-iterations the reference's cfl is the last block's
        for (int m = 0; m < nb && !rc && sa; m++) rc = st_sa(ctx, m, ctx->d_cfl);
        if (!rc) rc = refresh(ctx);
        if (!rc) rc = st_getpri_all(ctx);
    } else if (ctx->cfg.iflag_timeadvance == 2) {
        for (int m = 0; m < nb && !rc; m++) {
            rc = residual_stage(ctx, m, aux, pri_ok);
            if (!rc) rc = st_implicit(ctx, m, cfl_of(m));
            if (!rc && sa) rc = st_sa(ctx, m, cfl_of(m));
        }
        if (!rc) rc = refresh(ctx);
        for (int m = 0; m < nb && !rc; m++) rc = st_getpri(ctx, m);
    } else if (ctx->cfg.iflag_timeadvance == 3) {
        for (int sub = 0; sub < 3 && !rc; sub++) {
            rc = refresh(ctx);
            for (int m = 0; m < nb && !rc; m++) {
                rc = residual_stage(ctx, m, aux);
                if (!rc) rc = st_crank(ctx, m);
            }
        }
        for (int m = 0; m < nb && !rc; m++) rc = st_update_q0(ctx, m);
        // after the sub-iterations the reference's cfl is the last block's
        for (int m = 0; m < nb && !rc && sa; m++) rc = st_sa(ctx, m, ctx->d_cfl);
        if (!rc) rc = refresh(ctx);
        for (int m = 0; m < nb && !rc; m++) rc = st_getpri(ctx, m);
    } else if (ctx->cfg.iflag_timeadvance == 4) {
        // explicit Shu-Osher RK3: q^n is kept in w_dqs (free in this mode)
        for (int m = 0; m < nb; m++) {
            const Block &b = ctx->blk[m];
            KRN_TRY(launch_copy(ctx->stream, b.f.q, b.f.w_dqs, 5 * b.e.n));
        }
        for (int stage = 0; stage < 3 && !rc; stage++) {
            if (stage > 0) rc = refresh(ctx);
            if (stagewise) {
                if (!rc && !(stage == 0 && pri_ok)) rc = st_getpri_all(ctx);
                if (!rc) rc = st_inviscid_all(ctx);
                if (!rc && ctx->c.solver == 0) rc = st_viscous_all(ctx, 0);
                if (!rc && stage == 0) rc = st_find_dt_all(ctx);
                for (int m = 0; m < nb && !rc; m++) {
                    const Block &b = ctx->blk[m];
                    rc = launch_rk3_stage(ctx->stream, b.e, b.f, b.f.w_dqs, stage, ctx->c.solver == 0)
                             ? fail(ACFD_EDEVICE, "rk3 stage launch failed") : 0;
                }
                continue;
            }
            for (int m = 0; m < nb && !rc; m++) {
                rc = stage == 0 && pri_ok ? 0 : st_getpri(ctx, m);
                if (!rc) rc = st_inviscid(ctx, m);
                if (!rc && ctx->c.solver == 0) rc = st_viscous(ctx, m, 0);
                if (!rc && stage == 0) rc = st_find_dt(ctx, m);
                if (!rc) {
                    const Block &b = ctx->blk[m];
                    rc = launch_rk3_stage(ctx->stream, b.e, b.f, b.f.w_dqs, stage, ctx->c.solver == 0)
                             ? fail(ACFD_EDEVICE, "rk3 stage launch failed") : 0;
                }
            }
        }
        if (!rc) rc = refresh(ctx);
        if (!rc) rc = st_getpri_all(ctx);
    } else {
        return fail(ACFD_EINVAL, "iflag_timeadvance must be 2, 3 or 4");
    }
    // the residual norms of every step, as the reference's loop forms them
    // (its output interval is 1, main.f90:341,575-603); acfd_residual reads
    // them back
    if (!rc) rc = launch_residual_all(ctx) ? ACFD_EDEVICE : 0;
    if (!rc) ctx->pri_current = true;
    return rc;
}

// Record one step into a graph (every launch of one_step on the context's
// stream, no host synchronisation inside) and instantiate it.
int capture_step(acfd_ctx *ctx)
{
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the graph
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    ctx->graph_geo_cached = false;
    ctx->graph_needs_pri = false;
    ctx->graph_sl_in[0] = ctx->sl_block;
    ctx->graph_sl_in[1] = ctx->sl_shape;
    const bool stale = ctx->rhsv_stale;
    ctx->rhsv_stale = false;
    const int rc = one_step(ctx);
    ctx->graph_rhsv = ctx->rhsv_stale;
    ctx->rhsv_stale = stale;  // nothing ran yet
    ctx->capturing = false;
    ctx->graph_sl_out[0] = ctx->sl_block;
    ctx->graph_sl_out[1] = ctx->sl_shape;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = timing;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess || !g) return fail(ACFD_EDEVICE, "acfd_step: stream capture failed");
    const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->graph_exec = nullptr;
        return fail(ACFD_EDEVICE, "acfd_step: graph instantiation failed");
    }
    return 0;
}

void drop_graph(acfd_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
    ctx->graph_geo_cached = false;
    ctx->graph_needs_pri = false;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char *acfd_last_error(void) { return g_err.c_str(); }

int acfd_create(const acfd_config *cfg, acfd_ctx **out)
{
    if (!cfg || !out) return fail(ACFD_EINVAL, "acfd_create: null argument");
    if (cfg->ma <= 0.0) return fail(ACFD_EINVAL, "acfd_create: ma must be positive");
    if (cfg->iflag_timeadvance < 2 || cfg->iflag_timeadvance > 4)
        return fail(ACFD_EINVAL,
                    "acfd_create: iflag_timeadvance must be 2 (1st implicit), 3 (2nd crank) or 4 (explicit RK3)");
    if (cfg->inviscid_scheme < 0 || cfg->inviscid_scheme > 2)
        return fail(ACFD_EINVAL, "acfd_create: inviscid_scheme must be 0 (WENO3 + Roe), 1 (Steger-Warming + WENO5) "
                                 "or 2 (WCNS)");
    if (cfg->lusgs_kernel < 0 || cfg->lusgs_kernel > 2)
        return fail(ACFD_EINVAL, "acfd_create: lusgs_kernel must be 0 (hyperplanes), 1 (tile sweep per block) or 2 "
                                 "(tile sweeps of all blocks in one launch)");
    if (cfg->iflag_timeadvance == 4 && cfg->iflag_turbulence == 2 && cfg->iflag_solver == 0)
        return fail(ACFD_EINVAL, "acfd_create: the explicit RK3 advance is for laminar or Euler flow");
    if (cfg->iflag_turbulence != 0 && cfg->iflag_turbulence != 2)
        return fail(ACFD_EINVAL, "acfd_create: wrong turbulence model type");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(ACFD_EDEVICE, "acfd_create: no HIP device");
    acfd_ctx *ctx = new acfd_ctx();
    ctx->cfg = *cfg;
    ctx->c = consts_of(*cfg);
    auto bad = [&](const std::string &m) {
        acfd_destroy(ctx);
        return fail(ACFD_EDEVICE, m);
    };
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    if (hipMalloc((void **)&ctx->d_cfl, sizeof(double)) != hipSuccess) return bad("alloc");
    if (hipMalloc((void **)&ctx->d_stats, 2 * sizeof(unsigned long long)) != hipSuccess) return bad("alloc");
    if (hipMalloc((void **)&ctx->d_nneg, sizeof(unsigned int)) != hipSuccess) return bad("alloc");
    if (hipMalloc((void **)&ctx->d_res, 8 * sizeof(double)) != hipSuccess) return bad("alloc");
    double cfl = cfg->cfl;
    if (hipMemcpy(ctx->d_cfl, &cfl, sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return bad("copy");
    if (hipMemsetAsync(ctx->d_nneg, 0, sizeof(unsigned int), ctx->stream) != hipSuccess) return bad("memset");
    if (hipMemsetAsync(ctx->d_res, 0, 8 * sizeof(double), ctx->stream) != hipSuccess) return bad("memset");
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return bad("sync");
    unsigned long long st[2] = {0ull, ~0ull};
    if (hipMemcpy(ctx->d_stats, st, sizeof(st), hipMemcpyHostToDevice) != hipSuccess) return bad("copy");
    *out = ctx;
    return 0;
}

int acfd_destroy(acfd_ctx *ctx)
{
    if (!ctx) return 0;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->arena) (void)hipFree(ctx->arena);
    if (ctx->geo) (void)hipFree(ctx->geo);
    for (void *p : {(void *)ctx->d_wcells, (void *)ctx->d_geo_off, (void *)ctx->d_fwork})
        if (p) (void)hipFree(p);
    if (ctx->lu.ex) uc_release(ctx->lu.ex);
    if (ctx->lu.sl) (void)hipFree(ctx->lu.sl);
    if (ctx->lu.err) (void)hipFree(ctx->lu.err);
    for (auto &b : ctx->lub) lusgs_batch_free(&b);
    for (auto &b : ctx->sab) lusgs_batch_free(&b);
    if (ctx->lu_ctr) (void)hipFree(ctx->lu_ctr);
    if (ctx->d_cfl_blk) (void)hipFree(ctx->d_cfl_blk);
    if (ctx->d_blocks) (void)hipFree(ctx->d_blocks);
    for (BatchGrid *g : {&ctx->bg_getpri, &ctx->bg_inv[0], &ctx->bg_inv[1], &ctx->bg_inv[2], &ctx->bg_inv[3],
                         &ctx->bg_visc, &ctx->bg_dt, &ctx->bg_prep})
        free_grid(*g);
    for (size_t r = 0; r < ctx->bg_bc.size(); r++) {
        if (ctx->d_bc_round[r]) (void)hipFree(ctx->d_bc_round[r]);
        free_grid(ctx->bg_bc[r]);
    }
    void *sp[] = {ctx->s_w, ctx->s_m, ctx->s_v, ctx->s_grad, ctx->s_loss, ctx->s_part, ctx->s_lpart, ctx->s_net,
                  ctx->s_X, ctx->s_Y, ctx->s_stats, ctx->s_dpart};
    for (void *p : sp)
        if (p) (void)hipFree(p);
    for (unsigned char *p : ctx->s_mask)
        if (p) (void)hipFree(p);
    void *ptrs[] = {ctx->d_cfl, ctx->d_stats, ctx->d_nneg, ctx->d_res, ctx->d_partial,
                    ctx->d_stage, ctx->d_qtab, ctx->d_nuttab, ctx->d_ntab, ctx->d_h_sb, ctx->d_h_db,
                    ctx->d_h_sn, ctx->d_h_dn, ctx->d_a_sb, ctx->d_a_gb, ctx->d_a_gs, ctx->d_a_sn,
                    ctx->d_a_gd, ctx->d_a_buf};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_rfork) (void)hipEventDestroy(ctx->ev_rfork);
    if (ctx->ev_rjoin) (void)hipEventDestroy(ctx->ev_rjoin);
    if (ctx->sa_ctr) (void)hipFree(ctx->sa_ctr);
    if (ctx->sa_ex) uc_release(ctx->sa_ex);
    if (ctx->sa_sl) (void)hipFree(ctx->sa_sl);
    free_grid(ctx->bg_sa);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

int acfd_set_topology(acfd_ctx *ctx, int nblk, const int *dims, const int *nsub, const acfd_subface *subs)
{
    if (!ctx || nblk <= 0 || !dims || !nsub) return fail(ACFD_EINVAL, "acfd_set_topology: bad arguments");
    if (ctx->topo) return fail(ACFD_ESTATE, "acfd_set_topology: topology already set");
    ctx->blk.resize(nblk);
    long off = 0;
    ctx->totalnodes = 0;
    for (int m = 0; m < nblk; m++) {
        int ni = dims[3 * m], nj = dims[3 * m + 1], nk = dims[3 * m + 2];
        if (ni < 4 || nj < 4 || nk < 4) return fail(ACFD_EINVAL, "acfd_set_topology: blocks need >= 4 nodes per direction");
        Block &b = ctx->blk[m];
        b.sub.assign(subs + off, subs + off + nsub[m]);
        off += nsub[m];
        block_ranges(ni, nj, nk, b.sub.data(), nsub[m], b.e);
        ctx->totalnodes += (long)ni * nj * nk;
    }
    // validate interfaces and faces (main.f90:655-690: no face mixes
    // physical and interface subfaces)
    for (int m = 0; m < nblk; m++) {
        int phys[7] = {0}, intf[7] = {0};
        const Ext &e = ctx->blk[m].e;
        for (auto &s : ctx->blk[m].sub) {
            if (s.face < 1 || s.face > 6) return fail(ACFD_EINVAL, "acfd_set_topology: face must be 1..6");
            // the subface must lie on its face inside the physical range
            bool ok = s.is >= e.is0 && s.is <= s.ie && s.ie <= e.ie0 && s.js >= e.js0 && s.js <= s.je &&
                      s.je <= e.je0 && s.ks >= e.ks0 && s.ks <= s.ke && s.ke <= e.ke0;
            switch (s.face) {
            case 1: ok = ok && s.is == e.is0 && s.ie == e.is0; break;
            case 2: ok = ok && s.is == e.ie0 && s.ie == e.ie0; break;
            case 3: ok = ok && s.js == e.js0 && s.je == e.js0; break;
            case 4: ok = ok && s.js == e.je0 && s.je == e.je0; break;
            case 5: ok = ok && s.ks == e.ks0 && s.ke == e.ks0; break;
            case 6: ok = ok && s.ks == e.ke0 && s.ke == e.ke0; break;
            }
            if (!ok)
                return fail(ACFD_EINVAL, "acfd_set_topology: subface " + std::to_string(s.f_no) + " of block " +
                                             std::to_string(m + 1) + " does not lie on its face");
            if (s.blk_t < 0) phys[s.face] = 1;
            if (s.blk_t > 0) {
                intf[s.face] = 1;
                if (s.blk_t > nblk || s.subface_t < 1 || s.subface_t > (int)ctx->blk[s.blk_t - 1].sub.size())
                    return fail(ACFD_EINVAL, "acfd_set_topology: interface names a missing block/subface");
            }
        }
        for (int f = 1; f <= 6; f++)
            if (phys[f] && intf[f])
                return fail(ACFD_EINVAL, "acfd_set_topology: block " + std::to_string(m + 1) +
                                             " mixes physical and interface subfaces on one face");
    }
    {
        int rc = alloc_blocks(ctx);
        if (rc) return rc;
    }
    // halo copy list: every (sender inner layer n, receiver halo layer n)
    // node pair, n = 1, 2 (sub_bc_updatebuffer.f90:57-135, :398-539)
    std::vector<PairNode> pn;
    ctx->h_sb.clear(); ctx->h_db.clear(); ctx->h_sn.clear(); ctx->h_dn.clear(); ctx->h_sub.clear();
    std::vector<int> a_sb;
    std::vector<long> a_sn;
    std::vector<long> a_key;  // (block, node) of the receiving face node
    std::vector<int> a_blk;
    int gsub = 0;
    for (int m = 0; m < nblk; m++) {
        for (auto &s : ctx->blk[m].sub) {
            int this_sub = gsub++;
            if (s.blk_t <= 0) continue;
            if (!pair_nodes(ctx->blk, s, pn))
                return fail(ACFD_EINVAL, "acfd_set_topology: interface faces of block " + std::to_string(m + 1) +
                                             " do not match in size for their ori");
            const Block &sb = ctx->blk[s.blk_t - 1];
            const acfd_subface &t = sb.sub[s.subface_t - 1];
            int sdi, sdj, sdk, rdi, rdj, rdk;
            inward(t.face, sdi, sdj, sdk);
            inward(s.face, rdi, rdj, rdk);
            const Ext &re = ctx->blk[m].e, &se = sb.e;
            for (auto &p : pn) {
                for (int nbuf = 1; nbuf <= 2; nbuf++) {
                    ctx->h_sb.push_back(s.blk_t - 1);
                    ctx->h_db.push_back(m);
                    ctx->h_sn.push_back(se.ix(p.si + nbuf * sdi, p.sj + nbuf * sdj, p.sk + nbuf * sdk));
                    ctx->h_dn.push_back(re.ix(p.ri - nbuf * rdi, p.rj - nbuf * rdj, p.rk - nbuf * rdk));
                    ctx->h_sub.push_back(this_sub);
                }
                a_sb.push_back(s.blk_t - 1);
                a_sn.push_back(se.ix(p.si, p.sj, p.sk));
                a_blk.push_back(m);
                a_key.push_back(re.ix(p.ri, p.rj, p.rk));
            }
        }
    }
    ctx->halo_n = (int)ctx->h_sb.size();
    int rc = 0;
    if ((rc = dev_upload(&ctx->d_h_sb, ctx->h_sb)) || (rc = dev_upload(&ctx->d_h_db, ctx->h_db)) ||
        (rc = dev_upload(&ctx->d_h_sn, ctx->h_sn)) || (rc = dev_upload(&ctx->d_h_dn, ctx->h_dn)))
        return rc;
    // average list grouped by receiving node; entries keep subface order
    std::vector<int> order(a_sb.size());
    for (size_t x = 0; x < order.size(); x++) order[x] = (int)x;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (a_blk[a] != a_blk[b]) return a_blk[a] < a_blk[b];
        return a_key[a] < a_key[b];
    });
    std::vector<int> e_sb, g_blk, g_start;
    std::vector<long> e_sn, g_dst;
    for (size_t x = 0; x < order.size(); x++) {
        int o = order[x];
        if (x == 0 || a_blk[o] != g_blk.back() || a_key[o] != g_dst.back()) {
            g_blk.push_back(a_blk[o]);
            g_dst.push_back(a_key[o]);
            g_start.push_back((int)x);
        }
        e_sb.push_back(a_sb[o]);
        e_sn.push_back(a_sn[o]);
    }
    g_start.push_back((int)order.size());
    if ((rc = dev_upload(&ctx->d_a_sb, e_sb)) || (rc = dev_upload(&ctx->d_a_sn, e_sn)) ||
        (rc = dev_upload(&ctx->d_a_gb, g_blk)) || (rc = dev_upload(&ctx->d_a_gd, g_dst)) ||
        (rc = dev_upload(&ctx->d_a_gs, g_start)))
        return rc;
    if (!order.empty()) HIP_TRY(hipMalloc((void **)&ctx->d_a_buf, order.size() * 6 * sizeof(double)));
    ctx->avg.count = (int)order.size();
    ctx->avg.src_blk = ctx->d_a_sb;
    ctx->avg.src = ctx->d_a_sn;
    ctx->avg.ngroups = (int)g_blk.size();
    ctx->avg.grp_blk = ctx->d_a_gb;
    ctx->avg.grp_dst = ctx->d_a_gd;
    ctx->avg.grp_start = ctx->d_a_gs;
    // physical boundary tasks in block/subface order (sub_bc.f90:26-62);
    // farfield runs with the block's js in place of the subface's js
    // (sub_bc.f90:57-59)
    ctx->bc.clear();
    for (int m = 0; m < nblk; m++)
        for (auto &s : ctx->blk[m].sub) {
            if (s.blk_t >= 0) continue;
            if (s.blk_t != -10 && s.blk_t != -20 && s.blk_t != -30 && s.blk_t != -40 && s.blk_t != -50) continue;
            BcTask t{m, s.face, s.blk_t, s.is, s.ie, s.js, s.je, s.ks, s.ke};
            if (s.blk_t == -10) t.js = ctx->blk[m].e.js;
            ctx->bc.push_back(t);
        }
    // pointer tables
    std::vector<double *> qt(nblk), nt(nblk);
    std::vector<long> ntab(nblk);
    long maxn = 0;
    int parts = 0;
    for (int m = 0; m < nblk; m++) {
        qt[m] = ctx->blk[m].f.q;
        nt[m] = ctx->blk[m].f.nut;
        ntab[m] = ctx->blk[m].e.n;
        maxn = std::max(maxn, ctx->blk[m].e.n);
        parts += residual_partials(ctx->blk[m].e);
    }
    if ((rc = dev_upload(&ctx->d_qtab, qt)) || (rc = dev_upload(&ctx->d_nuttab, nt)) ||
        (rc = dev_upload(&ctx->d_ntab, ntab)))
        return rc;
    HIP_TRY(hipMalloc((void **)&ctx->d_partial, (size_t)parts * 6 * sizeof(double)));
    rc = alloc_lu_ws(ctx);
    if (!rc) rc = build_batch(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->topo = true;
    return 0;
}

int acfd_block_ranges(acfd_ctx *ctx, int m, int *r)
{
    if (!ctx || !ctx->topo || m < 1 || m > (int)ctx->blk.size() || !r)
        return fail(ACFD_EINVAL, "acfd_block_ranges: bad block");
    const Ext &e = ctx->blk[m - 1].e;
    int v[18] = {e.is, e.ie, e.js, e.je, e.ks, e.ke, e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0,
                 e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    std::memcpy(r, v, sizeof(v));
    return 0;
}

int acfd_upload(acfd_ctx *ctx, int m, int field, const double *host)
{
    if (!ctx || !ctx->topo || m < 1 || m > (int)ctx->blk.size() || !host || field < 1 || field > kLastField)
        return fail(ACFD_EINVAL, "acfd_upload: bad arguments");
    ctx->pri_current = false;
    ctx->res_current = false;
    if (field == ACFD_RHSV) {
        int rc = rhsv_refresh(ctx);  // the other blocks' rhsv
        if (rc) return rc;
    }
    if (field >= ACFD_X) {
        int rc = ensure_geo(ctx);
        if (rc) return rc;
    }
    Block &b = ctx->blk[m - 1];
    double *dst = field_ptr(b, field);
    if (!dst) return fail(ACFD_EINVAL, "acfd_upload: field not allocated for this configuration");
    if (field == ACFD_DXIDX || field == ACFD_INV_J) sl_invalidate(ctx);
    if (field == ACFD_DXIDX || field == ACFD_INV_J || field == ACFD_AALPHA || field == ACFD_BBETA ||
        field == ACFD_GGAMMA) {
        ctx->abg_derived = false;
        drop_graph(ctx);
    }
    if (field == ACFD_DIST || field == ACFD_SPACING || field == ACFD_LENGTH) {
        ctx->sa_length_ready = false;
        drop_graph(ctx);
    }
    ctx->lu_prep_ready = false;
    if (field == ACFD_RHSI) {
        ctx->rhsi_clean = false;
        drop_graph(ctx);
    }
    int nc = kNcomp[field];
    long tot = b.e.n * nc;
    int rc = ensure_stage(ctx, tot);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->d_stage, host, tot * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    KRN_TRY(launch_aos_to_soa(ctx->stream, ctx->d_stage, dst, b.e.n, nc));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_download(acfd_ctx *ctx, int m, int field, double *host)
{
    if (!ctx || !ctx->topo || m < 1 || m > (int)ctx->blk.size() || !host || field < 1 || field > kLastField)
        return fail(ACFD_EINVAL, "acfd_download: bad arguments");
    Block &b = ctx->blk[m - 1];
    double *src = field_ptr(b, field);
    if (!src) return fail(ACFD_EINVAL, "acfd_download: field not allocated for this configuration");
    int nc = kNcomp[field];
    long tot = b.e.n * nc;
    int rc = field == ACFD_RHSV ? rhsv_refresh(ctx) : 0;
    if (!rc) rc = ensure_stage(ctx, tot);
    if (rc) return rc;
    KRN_TRY(launch_soa_to_aos(ctx->stream, src, ctx->d_stage, b.e.n, nc));
    HIP_TRY(hipMemcpyAsync(host, ctx->d_stage, tot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_run_stage(acfd_ctx *ctx, int stage)
{
    if (!ctx || !ctx->topo) return fail(ACFD_ESTATE, "acfd_run_stage: no topology");
    ctx->pri_current = false;
    ctx->res_current = false;
    ctx->lu_prep_ready = false;
    const int nb = (int)ctx->blk.size();
    int rc = rhsv_refresh(ctx);  // the stages read rhsv or reuse w_flux
    if (rc) return rc;
    switch (stage) {
    case ACFD_STAGE_GETPRI:
        for (int m = 0; m < nb && !rc; m++) rc = st_getpri(ctx, m);
        if (!rc) rc = check_nneg(ctx);
        break;
    case ACFD_STAGE_INVISCID:
        for (int m = 0; m < nb && !rc; m++) rc = st_inviscid(ctx, m);
        break;
    case ACFD_STAGE_VISCOUS:
        for (int m = 0; m < nb && !rc; m++) rc = st_viscous(ctx, m, 2);
        break;
    case ACFD_STAGE_FIND_DT:
        for (int m = 0; m < nb && !rc; m++) rc = st_find_dt(ctx, m);
        break;
    case ACFD_STAGE_IMPLICIT1ST:
        rc = st_implicit_all(ctx, nullptr);
        break;
    case ACFD_STAGE_SA:
        if (!(ctx->c.turb == 2 && ctx->c.solver == 0)) return fail(ACFD_EINVAL, "SA stage needs NS + SA");
        for (int m = 0; m < nb && !rc; m++) rc = st_sa(ctx, m, ctx->d_cfl);
        break;
    case ACFD_STAGE_PHYSICAL_BC: rc = st_physical_bc(ctx); break;
    case ACFD_STAGE_UPDATE_BUFFER: rc = st_update_buffer(ctx); break;
    case ACFD_STAGE_AVERAGE: rc = st_average(ctx); break;
    case ACFD_STAGE_CRANK2ND:
        if (ctx->cfg.iflag_timeadvance != 3) return fail(ACFD_EINVAL, "crank stage needs iflag_timeadvance = 3");
        for (int m = 0; m < nb && !rc; m++) rc = st_crank(ctx, m);
        break;
    case ACFD_STAGE_UPDATE_Q0:
        if (ctx->cfg.iflag_timeadvance != 3) return fail(ACFD_EINVAL, "q0 stage needs iflag_timeadvance = 3");
        for (int m = 0; m < nb && !rc; m++) rc = st_update_q0(ctx, m);
        break;
    default: return fail(ACFD_EINVAL, "acfd_run_stage: unknown stage");
    }
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_prepare(acfd_ctx *ctx)
{
    if (!ctx || !ctx->topo) return fail(ACFD_ESTATE, "acfd_prepare: no topology");
    ctx->pri_current = false;
    ctx->res_current = false;
    int rc = refresh(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_step(acfd_ctx *ctx, int nsteps)
{
    if (!ctx || !ctx->topo) return fail(ACFD_ESTATE, "acfd_step: no topology");
    if (nsteps > 0) ctx->res_current = false;
    int s = 0;
    // graph mode: the first step runs eagerly (it also tunes the tile
    // sweeps), the next one is captured, and every later step replays it;
    // audited sensor steps read results on the host, sensor steps that reuse
    // decisions alternate between two kernels, and steps under the
    // per-stage timers record events between launches: all stay eager
    const bool graph = ctx->cfg.use_graph && !ctx->timing && !(ctx->s_on && (ctx->s_audit > 0 || ctx->s_refresh > 1));
    if (ctx->graph_exec && ctx->graph_geo_cached &&
        (ctx->sl_block != ctx->graph_sl_in[0] || ctx->sl_shape != ctx->graph_sl_in[1]))
        drop_graph(ctx);  // the sweep layout holds other grid operands than at capture
    if (ctx->graph_exec && ctx->graph_needs_pri && !ctx->pri_current)
        drop_graph(ctx);  // q or pri_v was written since the last step
    if (graph && !ctx->graph_exec && nsteps > 0) {
        int rc = one_step(ctx);
        if (rc) return rc;
        s = 1;
        if (s < nsteps) {
            rc = capture_step(ctx);
            if (rc) return rc;
        }
    }
    for (; s < nsteps; s++) {
        if (graph && ctx->graph_exec) {
            HIP_TRY(hipGraphLaunch(ctx->graph_exec, ctx->stream));
            ctx->sl_block = ctx->graph_sl_out[0];
            ctx->sl_shape = ctx->graph_sl_out[1];
            ctx->pri_current = true;
            if (ctx->graph_rhsv) ctx->rhsv_stale = true;
            continue;
        }
        int rc = one_step(ctx);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (nsteps > 0) ctx->res_current = true;
    int rc = tile_error(ctx);
    if (rc) return rc;
    return check_nneg(ctx);
}

int acfd_residual(acfd_ctx *ctx, double *out)
{
    if (!ctx || !ctx->topo || !out) return fail(ACFD_EINVAL, "acfd_residual: bad arguments");
    // the last step's norms stand until a call changes the fields they
    // come from; otherwise formed now
    int rc = ctx->res_current ? 0 : launch_residual_all(ctx);
    if (rc) return rc;
    double s[6];
    HIP_TRY(hipMemcpyAsync(s, ctx->d_res, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int v = 0; v < 6; v++) out[v] = std::sqrt(s[v] / (double)ctx->totalnodes);
    out[6] = 0.0;
    return 0;
}

int acfd_stats(acfd_ctx *ctx, double *out)
{
    if (!ctx || !out) return fail(ACFD_EINVAL, "acfd_stats: bad arguments");
    unsigned long long st[2];
    double cfl;
    HIP_TRY(hipMemcpyAsync(st, ctx->d_stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&cfl, ctx->d_cfl, sizeof(cfl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double mx, mn;
    std::memcpy(&mx, &st[0], 8);
    std::memcpy(&mn, &st[1], 8);
    if (ctx->c.itime == 0) {
        out[0] = mx; out[1] = st[1] == ~0ull ? 1.0e10 : mn; out[2] = 0.0; out[3] = 1.0e10;
    } else {
        out[0] = 0.0; out[1] = 1.0e10; out[2] = cfl; out[3] = cfl;
    }
    out[4] = cfl;
    return 0;
}

int acfd_sync(acfd_ctx *ctx)
{
    if (!ctx) return fail(ACFD_EINVAL, "acfd_sync: null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_timing(acfd_ctx *ctx, int enable, double *ms, long *calls)
{
    if (!ctx) return fail(ACFD_EINVAL, "acfd_timing: null context");
    timing_flush(ctx);
    if (ms)
        for (int s = 0; s < kTimeSlots; s++) ms[s] = ctx->stage_ms[s];
    if (calls)
        for (int s = 0; s < kTimeSlots; s++) calls[s] = ctx->stage_calls[s];
    if (enable >= 0) {
        ctx->timing = enable != 0;
        for (int s = 0; s < kTimeSlots; s++) {
            ctx->stage_ms[s] = 0;
            ctx->stage_calls[s] = 0;
        }
    }
    return 0;
}

int acfd_grid_metrics(acfd_ctx *ctx, double *const *xyz)
{
    if (!ctx || !ctx->topo || !xyz) return fail(ACFD_EINVAL, "acfd_grid_metrics: bad arguments");
    ctx->res_current = false;
    {
        const int rc = rhsv_refresh(ctx);  // w_flux is scratch below
        if (rc) return rc;
    }
    const int nb = (int)ctx->blk.size();
    for (int m = 0; m < nb; m++)
        if (!xyz[m]) return fail(ACFD_EINVAL, "acfd_grid_metrics: missing coordinates of a block");
    {
        int rc = ensure_geo(ctx);
        if (rc) return rc;
    }
    const bool sa = ctx->c.turb == 2 && ctx->c.solver == 0;
    sl_invalidate(ctx);
    for (int m = 0; m < nb; m++) {
        Block &b = ctx->blk[m];
        const long n = b.e.n;
        HIP_TRY(hipMemcpyAsync(b.geo, xyz[m], 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        const double *x = b.geo, *y = b.geo + n, *z = b.geo + 2 * n;
        KRN_TRY(launch_grid_spacing(ctx->stream, b.e, x, y, z, b.geo + 3 * n, sa ? b.f.spacing : nullptr));
        KRN_TRY(launch_grid_metrics(ctx->stream, b.e, x, y, z, b.f.w_flux, b.f.inv_j, b.f.dxidx,
                                    sa ? b.f.aal : nullptr, sa ? b.f.bbe : nullptr, sa ? b.f.gga : nullptr));
    }
    ctx->abg_derived = sa;
    ctx->sa_length_ready = false;  // new dist and spacing
    drop_graph(ctx);
    if (sa) {
        // wall points: the face nodes of every wall subface of every block
        // (sub_init_readbcinfo.f90:294-400), SoA x | y | z
        std::vector<double> wx, wy, wz;
        for (int m = 0; m < nb; m++) {
            const Block &b = ctx->blk[m];
            const Ext &e = b.e;
            for (const acfd_subface &f : b.sub) {
                if (f.blk_t != -20) continue;
                for (int k = f.ks; k <= f.ke; k++)
                    for (int j = f.js; j <= f.je; j++)
                        for (int i = f.is; i <= f.ie; i++) {
                            const int ii = f.face == 1 ? e.is0 : f.face == 2 ? e.ie0 : i;
                            const int jj = f.face == 3 ? e.js0 : f.face == 4 ? e.je0 : j;
                            const int kk = f.face == 5 ? e.ks0 : f.face == 6 ? e.ke0 : k;
                            const long n = e.ix(ii, jj, kk);
                            wx.push_back(xyz[m][n]);
                            wy.push_back(xyz[m][e.n + n]);
                            wz.push_back(xyz[m][2 * e.n + n]);
                        }
            }
        }
        const long nw = (long)wx.size();
        double *dw = nullptr;
        std::vector<double> w(wx);  // kept until the stream is synchronised below
        w.insert(w.end(), wy.begin(), wy.end());
        w.insert(w.end(), wz.begin(), wz.end());
        // bounding box of every chunk of kWallChunk consecutive wall points
        const int nchunk = (int)((nw + kWallChunk - 1) / kWallChunk);
        std::vector<double> cbox(6 * (size_t)nchunk);
        for (int c = 0; c < nchunk; c++) {
            double *bx = &cbox[6 * (size_t)c];
            bx[0] = bx[1] = bx[2] = 1.0e300;
            bx[3] = bx[4] = bx[5] = -1.0e300;
            for (long p = (long)c * kWallChunk; p < std::min(nw, (long)(c + 1) * kWallChunk); p++) {
                const double v[3] = {wx[p], wy[p], wz[p]};
                for (int d = 0; d < 3; d++) {
                    bx[d] = std::min(bx[d], v[d]);
                    bx[3 + d] = std::max(bx[3 + d], v[d]);
                }
            }
        }
        if (nw > 0) {
            w.insert(w.end(), cbox.begin(), cbox.end());
            HIP_TRY(hipMalloc((void **)&dw, w.size() * sizeof(double)));
            HIP_TRY(hipMemcpyAsync(dw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        int rc = 0;
        for (int m = 0; m < nb && !rc; m++) {
            const Block &b = ctx->blk[m];
            const long n = b.e.n;
            rc = launch_walldist(ctx->stream, b.e, b.geo, b.geo + n, b.geo + 2 * n, dw, nw, dw ? dw + 3 * nw : nullptr,
                                 nchunk, b.f.dist);
        }
        const hipError_t es = hipStreamSynchronize(ctx->stream);
        if (dw) (void)hipFree(dw);
        if (rc) return fail(ACFD_EDEVICE, "acfd_grid_metrics: wall distance launch failed");
        if (es != hipSuccess) return fail(ACFD_EDEVICE, std::string("acfd_grid_metrics: ") + hipGetErrorString(es));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_forces(acfd_ctx *ctx, double sref, double *out)
{
    if (!ctx || !ctx->topo || !out) return fail(ACFD_EINVAL, "acfd_forces: bad arguments");
    if (!ctx->geo) return fail(ACFD_ESTATE, "acfd_forces: no edge lengths (acfd_grid_metrics or upload ACFD_SPCI..K)");
    if (ctx->n_wcells < 0) {
        // wall cells in the reference's accumulation order: blocks, subfaces,
        // outer then inner tangential index
        std::vector<WallCell> cells;
        for (int m = 0; m < (int)ctx->blk.size(); m++)
            for (const acfd_subface &f : ctx->blk[m].sub) {
                if (f.blk_t != -20) continue;
                const int d = (f.face - 1) / 2;
                const int r[3][2] = {{f.is, f.ie}, {f.js, f.je}, {f.ks, f.ke}};
                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                const int wall = (f.face & 1) ? r[d][0] : r[d][1];
                for (int v = r[t2][0]; v <= r[t2][1] - 1; v++)
                    for (int u = r[t1][0]; u <= r[t1][1] - 1; u++) cells.push_back(WallCell{m, f.face, wall, u, v, 0});
            }
        std::vector<long> goff;
        for (auto &b : ctx->blk) goff.push_back((long)(b.geo - ctx->geo));
        int rc = ensure_block_table(ctx);
        if (!rc && !cells.empty()) rc = dev_upload(&ctx->d_wcells, cells);
        if (!rc) rc = dev_upload(&ctx->d_geo_off, goff);
        if (rc) return rc;
        HIP_TRY(hipMalloc((void **)&ctx->d_fwork, (6 * cells.size() + 8) * sizeof(double)));
        ctx->n_wcells = (int)cells.size();
    }
    double *out6 = ctx->d_fwork + 6L * ctx->n_wcells;
    KRN_TRY(launch_wall_forces(ctx->stream, ctx->d_blocks, ctx->d_wcells, ctx->n_wcells, ctx->arena, ctx->geo,
                               ctx->d_geo_off, ctx->c.p0, ctx->c.re, ctx->d_fwork, out6));
    double f[6];
    HIP_TRY(hipMemcpyAsync(f, out6, sizeof(f), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // lift and drag (sub_output_forceintegration.f90:455-470), unit freestream
    const double div = 1.0 / (0.5 * ctx->c.d0 * ctx->c.u0 * ctx->c.u0 * sref);
    for (int c = 0; c < 6; c++) out[c] = f[c];
    out[6] = ((f[1] + f[4]) * std::cos(ctx->c.aoa) - (f[0] + f[3]) * std::sin(ctx->c.aoa)) * div;
    out[7] = ((f[1] + f[4]) * std::sin(ctx->c.aoa) + (f[0] + f[3]) * std::cos(ctx->c.aoa)) * div;
    return 0;
}

int acfd_exchange_coordinates(acfd_ctx *ctx, double **xyz, const double *shift)
{
    if (!ctx || !ctx->topo || !xyz) return fail(ACFD_EINVAL, "acfd_exchange_coordinates: bad arguments");
    // gather first (every pack precedes every unpack), then scatter
    std::vector<double> val(ctx->h_sb.size() * 3);
    for (size_t x = 0; x < ctx->h_sb.size(); x++) {
        const Block &sb = ctx->blk[ctx->h_sb[x]];
        const double *src = xyz[ctx->h_sb[x]];
        for (int c = 0; c < 3; c++) val[3 * x + c] = src[c * sb.e.n + ctx->h_sn[x]];
    }
    for (size_t x = 0; x < ctx->h_db.size(); x++) {
        const Block &db = ctx->blk[ctx->h_db[x]];
        double *dst = xyz[ctx->h_db[x]];
        for (int c = 0; c < 3; c++) {
            double s = shift ? shift[3 * ctx->h_sub[x] + c] : 0.0;
            dst[c * db.e.n + ctx->h_dn[x]] = val[3 * x + c] + s;
        }
    }
    return 0;
}


// ------------------------------------------------------- learned sensor
}  // extern "C"
namespace {

int sensor_ready(acfd_ctx *ctx)
{
    if (ctx->s_w) return 0;
    const size_t pb = kSensorParams * sizeof(float);
    HIP_TRY(hipMalloc((void **)&ctx->s_w, pb));
    HIP_TRY(hipMalloc((void **)&ctx->s_m, pb));
    HIP_TRY(hipMalloc((void **)&ctx->s_v, pb));
    HIP_TRY(hipMalloc((void **)&ctx->s_grad, pb));
    HIP_TRY(hipMalloc((void **)&ctx->s_loss, sizeof(float)));
    HIP_TRY(hipMalloc((void **)&ctx->s_net, sensor_net_bytes()));
    HIP_TRY(hipMalloc((void **)&ctx->s_stats, kSensorStatBytes));
    HIP_TRY(hipMemsetAsync(ctx->s_stats, 0, kSensorStatBytes, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->s_m, 0, pb, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->s_v, 0, pb, ctx->stream));
    // He initialisation from a fixed generator (64-bit LCG + Box-Muller)
    std::vector<float> w(kSensorParams, 0.f);
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    const int nf = 12, h = 32;
    for (int e = 0; e < nf * h; e++) w[e] = (float)(gauss() * std::sqrt(2.0 / nf));
    for (int e = 0; e < h * h; e++) w[nf * h + h + e] = (float)(gauss() * std::sqrt(2.0 / h));
    for (int e = 0; e < h; e++) w[nf * h + h + h * h + h + e] = (float)(gauss() * std::sqrt(1.0 / h));
    HIP_TRY(hipMemcpy(ctx->s_w, w.data(), pb, hipMemcpyHostToDevice));
    KRN_TRY(launch_sensor_pack(ctx->stream, ctx->s_w, ctx->s_net));
    ctx->s_calls.assign(ctx->blk.size(), 0);
    ctx->s_off.assign(ctx->blk.size(), 0);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int ensure_mlp_partials(acfd_ctx *ctx, long count)
{
    const long need = mlp_partials_needed(count);
    if (ctx->s_part_cap >= need) return 0;
    if (ctx->s_part) (void)hipFree(ctx->s_part);
    if (ctx->s_lpart) (void)hipFree(ctx->s_lpart);
    ctx->s_part = ctx->s_lpart = nullptr;
    HIP_TRY(hipMalloc((void **)&ctx->s_part, need * sizeof(float)));
    HIP_TRY(hipMalloc((void **)&ctx->s_lpart, (need / kSensorParams + 1) * sizeof(float)));
    ctx->s_part_cap = need;
    return 0;
}

}  // namespace
extern "C" {

int acfd_sensor_weights(acfd_ctx *ctx, float *w, int set)
{
    if (!ctx || !ctx->topo || !w) return fail(ACFD_EINVAL, "acfd_sensor_weights: bad arguments");
    int rc = sensor_ready(ctx);
    if (rc) return rc;
    const size_t pb = kSensorParams * sizeof(float);
    if (set) {
        HIP_TRY(hipMemcpy(ctx->s_w, w, pb, hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(ctx->s_m, 0, pb, ctx->stream));
        HIP_TRY(hipMemsetAsync(ctx->s_v, 0, pb, ctx->stream));
        ctx->s_adam_t = 0;
        KRN_TRY(launch_sensor_pack(ctx->stream, ctx->s_w, ctx->s_net));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } else {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(w, ctx->s_w, pb, hipMemcpyDeviceToHost));
    }
    return 0;
}

int acfd_sensor_collect(acfd_ctx *ctx, int stride, long *count)
{
    if (!ctx || !ctx->topo || stride < 1) return fail(ACFD_EINVAL, "acfd_sensor_collect: bad arguments");
    int rc = sensor_ready(ctx);
    if (rc) return rc;
    long total = 0;
    for (auto &b : ctx->blk)
        for (int ax = 0; ax < 3; ax++) total += sensor_samples(b.e, ax, stride);
    if (total > ctx->s_cap) {
        if (ctx->s_X) (void)hipFree(ctx->s_X);
        if (ctx->s_Y) (void)hipFree(ctx->s_Y);
        ctx->s_X = ctx->s_Y = nullptr;
        HIP_TRY(hipMalloc((void **)&ctx->s_X, (size_t)std::max(total, 1L) * kSensorFeatures * sizeof(float)));
        HIP_TRY(hipMalloc((void **)&ctx->s_Y, (size_t)std::max(total, 1L) * sizeof(float)));
        ctx->s_cap = total;
    }
    long off = 0;
    for (auto &b : ctx->blk)
        for (int ax = 0; ax < 3; ax++) {
            KRN_TRY(launch_sensor_collect(ctx->stream, b.e, b.f, ctx->c.gamma, ax, stride,
                                          ctx->s_X + off * kSensorFeatures, ctx->s_Y + off));
            off += sensor_samples(b.e, ax, stride);
        }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->s_n = total;
    if (count) *count = total;
    return 0;
}

int acfd_sensor_train(acfd_ctx *ctx, int iters, int batch, float lr, unsigned seed, float *losses)
{
    if (!ctx || !ctx->topo || iters < 0 || batch < 1) return fail(ACFD_EINVAL, "acfd_sensor_train: bad arguments");
    int rc = sensor_ready(ctx);
    if (rc) return rc;
    const long n = ctx->s_n;
    if (n <= 0) return fail(ACFD_ESTATE, "acfd_sensor_train: no samples (call acfd_sensor_collect)");
    const long count = std::min<long>(batch, n);
    rc = ensure_mlp_partials(ctx, count);
    if (rc) return rc;
    // a fixed pseudo-random visiting order: k -> (k * stride + offset) mod n
    const long primes[] = {1000003, 999983, 999979, 999961, 999959, 999953};
    long stride = 1;
    for (long p : primes)
        if (n % p != 0) {
            stride = p % n == 0 ? 1 : p;
            break;
        }
    const long offset = (long)(seed % (unsigned long)n);
    float *hist = nullptr;
    HIP_TRY(hipMalloc((void **)&hist, std::max(iters, 1) * sizeof(float)));
    for (int it = 0; it < iters; it++) {
        KRN_TRY(launch_mlp_grad(ctx->stream, ctx->s_w, ctx->s_X, ctx->s_Y, n, (long)it * count, count, stride, offset,
                                ctx->s_part, ctx->s_lpart, ctx->s_grad, hist + it));
        KRN_TRY(launch_adam(ctx->stream, ctx->s_w, ctx->s_grad, ctx->s_m, ctx->s_v, lr, ++ctx->s_adam_t));
    }
    KRN_TRY(launch_sensor_pack(ctx->stream, ctx->s_w, ctx->s_net));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (losses && iters > 0) HIP_TRY(hipMemcpy(losses, hist, iters * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(hist);
    return 0;
}

int acfd_sensor_evaluate(acfd_ctx *ctx, float threshold, double *out)
{
    if (!ctx || !ctx->topo || !out) return fail(ACFD_EINVAL, "acfd_sensor_evaluate: bad arguments");
    int rc = sensor_ready(ctx);
    if (rc) return rc;
    const long n = ctx->s_n;
    if (n <= 0) return fail(ACFD_ESTATE, "acfd_sensor_evaluate: no samples");
    float *prob = nullptr;
    HIP_TRY(hipMalloc((void **)&prob, n * sizeof(float)));
    KRN_TRY(launch_sensor_infer(ctx->stream, ctx->s_net, ctx->s_X, n, prob));
    std::vector<float> p(n), y(n);
    HIP_TRY(hipMemcpyAsync(p.data(), prob, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(y.data(), ctx->s_Y, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    (void)hipFree(prob);
    long correct = 0, lin = 0, false_lin = 0, pos = 0;
    for (long k = 0; k < n; k++) {
        const bool need = y[k] > 0.5f, pred = p[k] >= threshold;
        correct += need == pred;
        lin += !pred;
        pos += need;
        false_lin += need && !pred;
    }
    out[0] = (double)correct / n;            // accuracy
    out[1] = (double)lin / n;                // faces given linear weights
    out[2] = (double)false_lin / n;          // nonlinear faces given linear weights
    out[3] = (double)pos / n;                // faces whose label is nonlinear
    return 0;
}

int acfd_sensor_enable(acfd_ctx *ctx, int on, float threshold, int audit_interval, double audit_bound)
{
    if (!ctx || !ctx->topo) return fail(ACFD_EINVAL, "acfd_sensor_enable: no topology");
    if (on && ctx->cfg.inviscid_scheme != 2)
        return fail(ACFD_EINVAL, "acfd_sensor_enable: the sensor drives the WCNS scheme (inviscid_scheme = 2)");
    int rc = sensor_ready(ctx);
    if (rc) return rc;
    drop_graph(ctx);  // the captured step used the other inviscid kernels
    ctx->s_on = on != 0;
    ctx->s_thr = threshold;
    ctx->s_audit = std::max(audit_interval, 0);
    ctx->s_bound = audit_bound;
    ctx->s_calls.assign(ctx->blk.size(), 0);
    ctx->s_off.assign(ctx->blk.size(), 0);
    ctx->s_audits = 0;
    ctx->s_audit_max = 0.0;
    HIP_TRY(hipMemsetAsync(ctx->s_stats, 0, kSensorStatBytes, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int acfd_sensor_refresh(acfd_ctx *ctx, int interval)
{
    if (!ctx || !ctx->topo || interval < 1) return fail(ACFD_EINVAL, "acfd_sensor_refresh: bad arguments");
    drop_graph(ctx);
    if (interval > 1) {
        ctx->s_mask.resize(ctx->blk.size(), nullptr);
        for (size_t m = 0; m < ctx->blk.size(); m++)
            if (!ctx->s_mask[m]) HIP_TRY(hipMalloc((void **)&ctx->s_mask[m], 3 * (size_t)ctx->blk[m].e.n));
    }
    ctx->s_refresh = interval;
    // the next evaluation of every block runs the network
    ctx->s_calls.assign(ctx->blk.size(), 0);
    return 0;
}

int acfd_sensor_stats(acfd_ctx *ctx, double *out)
{
    if (!ctx || !out) return fail(ACFD_EINVAL, "acfd_sensor_stats: bad arguments");
    unsigned long long st[2] = {0, 0};
    if (ctx->s_stats) {
        std::vector<unsigned long long> h(kSensorStatBytes / sizeof(unsigned long long));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(h.data(), ctx->s_stats, kSensorStatBytes, hipMemcpyDeviceToHost));
        for (int k = 0; k < kSensorStatSlots; k++) {
            st[0] += h[(size_t)k * kSensorStatStride];
            st[1] += h[(size_t)k * kSensorStatStride + 1];
        }
    }
    long off = 0;
    for (char c : ctx->s_off) off += c;
    out[0] = (double)st[0];
    out[1] = (double)st[1];
    out[2] = ctx->s_audits;
    out[3] = ctx->s_audit_max;
    out[4] = (double)off;
    out[5] = 0.0;
    return 0;
}

int acfd_mlp_grad(const float *w, const float *X, const float *Y, long n, float *loss, float *grad)
{
    if (!w || !X || !Y || n <= 0 || !loss || !grad) return fail(ACFD_EINVAL, "acfd_mlp_grad: bad arguments");
    float *dw = nullptr, *dx = nullptr, *dy = nullptr, *dg = nullptr, *dl = nullptr, *dp = nullptr, *dlp = nullptr;
    const long np = mlp_partials_needed(n);
    int rc = 0;
    if (hipMalloc((void **)&dw, kSensorParams * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dx, n * kSensorFeatures * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dy, n * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dg, kSensorParams * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dl, sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dp, np * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dlp, (np / kSensorParams + 1) * sizeof(float)) != hipSuccess)
        rc = fail(ACFD_EDEVICE, "acfd_mlp_grad: device memory");
    if (!rc && (hipMemcpy(dw, w, kSensorParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dx, X, n * kSensorFeatures * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dy, Y, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(ACFD_EDEVICE, "acfd_mlp_grad: copy");
    if (!rc && launch_mlp_grad(nullptr, dw, dx, dy, n, 0, n, 1, 0, dp, dlp, dg, dl))
        rc = fail(ACFD_EDEVICE, "acfd_mlp_grad: launch");
    if (!rc && (hipMemcpy(grad, dg, kSensorParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
        rc = fail(ACFD_EDEVICE, "acfd_mlp_grad: copy back");
    void *all[] = {dw, dx, dy, dg, dl, dp, dlp};
    for (void *p : all)
        if (p) (void)hipFree(p);
    return rc;
}

int acfd_sensor_infer(const float *w, const float *X, long n, float *prob)
{
    if (!w || !X || n <= 0 || !prob) return fail(ACFD_EINVAL, "acfd_sensor_infer: bad arguments");
    float *dw = nullptr, *dx = nullptr, *dp = nullptr;
    SensorNet *net = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&dw, kSensorParams * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dx, n * kSensorFeatures * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dp, n * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&net, sensor_net_bytes()) != hipSuccess)
        rc = fail(ACFD_EDEVICE, "acfd_sensor_infer: device memory");
    if (!rc && (hipMemcpy(dw, w, kSensorParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dx, X, n * kSensorFeatures * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(ACFD_EDEVICE, "acfd_sensor_infer: copy");
    if (!rc && (launch_sensor_pack(nullptr, dw, net) || launch_sensor_infer(nullptr, net, dx, n, dp)))
        rc = fail(ACFD_EDEVICE, "acfd_sensor_infer: launch");
    if (!rc && hipMemcpy(prob, dp, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(ACFD_EDEVICE, "acfd_sensor_infer: copy back");
    void *all[] = {dw, dx, dp, net};
    for (void *p : all)
        if (p) (void)hipFree(p);
    return rc;
}

}  // extern "C"

// ======================================================= stateless routines
namespace {

// single-block context with explicit ranges and no interfaces
// The Ext of one block from its 18 ranges (is..ke, is0..ke0, is1..ke1).
Ext ext_of_ranges(const int *r)
{
    Ext e;
    e.is = r[0]; e.ie = r[1]; e.js = r[2]; e.je = r[3]; e.ks = r[4]; e.ke = r[5];
    e.is0 = r[6]; e.ie0 = r[7]; e.js0 = r[8]; e.je0 = r[9]; e.ks0 = r[10]; e.ke0 = r[11];
    e.is1 = r[12]; e.ie1 = r[13]; e.js1 = r[14]; e.je1 = r[15]; e.ks1 = r[16]; e.ke1 = r[17];
    e.ni = e.ie - e.is + 1; e.nj = e.je - e.js + 1; e.nk = e.ke - e.ks + 1;
    e.n = (long)e.ni * e.nj * e.nk;
    return e;
}

int make_single(const acfd_config *cfg, const int *r, acfd_ctx **out)
{
    acfd_config c = cfg ? *cfg : acfd_config{};
    if (!cfg) {
        c.ma = 0.5; c.re = 1e6; c.tinf = 288.15; c.iflag_timeadvance = 2; c.iflag_time = 1;
    }
    if (c.iflag_timeadvance != 3) c.iflag_timeadvance = 2;
    int rc = acfd_create(&c, out);
    if (rc) return rc;
    acfd_ctx *ctx = *out;
    ctx->blk.resize(1);
    Ext &e = ctx->blk[0].e;
    e = ext_of_ranges(r);
    if (e.ni < 3 || e.nj < 3 || e.nk < 3) {
        acfd_destroy(ctx);
        return fail(ACFD_EINVAL, "ranges: block too small");
    }
    rc = alloc_blocks(ctx);
    if (rc) {
        acfd_destroy(ctx);
        return rc;
    }
    std::vector<double *> qt{ctx->blk[0].f.q}, nt{ctx->blk[0].f.nut};
    std::vector<long> nn{e.n};
    dev_upload(&ctx->d_qtab, qt);
    dev_upload(&ctx->d_nuttab, nt);
    dev_upload(&ctx->d_ntab, nn);
    HIP_TRY(hipMalloc((void **)&ctx->d_partial, (size_t)residual_partials(e) * 6 * sizeof(double)));
    rc = alloc_lu_ws(ctx);
    if (rc) return rc;
    ctx->totalnodes = (long)(e.ie0 - e.is0 + 1) * (e.je0 - e.js0 + 1) * (e.ke0 - e.ks0 + 1);
    ctx->topo = true;
    return 0;
}

struct Single {
    acfd_ctx *ctx = nullptr;
    ~Single() { acfd_destroy(ctx); }
};

#define ST_TRY(x)            \
    do {                     \
        int r__ = (x);       \
        if (r__) return r__; \
    } while (0)

int put(acfd_ctx *c, int f, const double *h)
{
    return h ? acfd_upload(c, 1, f, h) : 0;
}

int set_cfl(acfd_ctx *c, double v)
{
    HIP_TRY(hipMemcpy(c->d_cfl, &v, sizeof(v), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

extern "C" {

int acfd_getpri(const int *r, double gamma, double cv, const double *q, double *pri_v)
{
    if (!r || !q || !pri_v) return fail(ACFD_EINVAL, "acfd_getpri: null argument");
    (void)gamma;
    acfd_config c{};
    c.ma = std::sqrt(1.0 / (1.4 * 0.4 * cv));  // cv = 1/(gamma (gamma-1) ma^2)
    c.re = 1e6; c.tinf = 288.15; c.iflag_timeadvance = 2; c.iflag_time = 1; c.iflag_solver = 1;
    Single s;
    ST_TRY(make_single(&c, r, &s.ctx));
    s.ctx->c.cv = cv;
    ST_TRY(put(s.ctx, ACFD_Q, q));
    int rc = acfd_run_stage(s.ctx, ACFD_STAGE_GETPRI);
    if (rc && rc != ACFD_ENUMERIC) return rc;
    ST_TRY(acfd_download(s.ctx, 1, ACFD_PRI_V, pri_v));
    return rc;
}

int acfd_inviscid_rhs(const acfd_config *cfg, const int *r, const double *q, const double *dxidx,
                      const double *inv_j, double *rhsi)
{
    if (!r || !q || !dxidx || !inv_j || !rhsi) return fail(ACFD_EINVAL, "acfd_inviscid_rhs: null argument");
    Single s;
    ST_TRY(make_single(cfg, r, &s.ctx));
    ST_TRY(put(s.ctx, ACFD_Q, q));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_INV_J, inv_j));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_INVISCID));
    return acfd_download(s.ctx, 1, ACFD_RHSI, rhsi);
}

int acfd_viscous_rhs(const acfd_config *cfg, const int *r, const double *pri_v, double *amu,
                     const double *dxidx, const double *inv_j, double *rhsv, double *vor, double *dudx,
                     double *tao)
{
    if (!r || !pri_v || !amu || !dxidx || !inv_j || !rhsv) return fail(ACFD_EINVAL, "acfd_viscous_rhs: null argument");
    Single s;
    ST_TRY(make_single(cfg, r, &s.ctx));
    s.ctx->c.solver = 0;
    ST_TRY(put(s.ctx, ACFD_PRI_V, pri_v));
    ST_TRY(put(s.ctx, ACFD_AMU, amu));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_INV_J, inv_j));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_VISCOUS));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_RHSV, rhsv));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_AMU, amu));
    if (vor) ST_TRY(acfd_download(s.ctx, 1, ACFD_VOR, vor));
    if (dudx) ST_TRY(acfd_download(s.ctx, 1, ACFD_DUDX, dudx));
    if (tao) ST_TRY(acfd_download(s.ctx, 1, ACFD_TAO, tao));
    return 0;
}

int acfd_find_dt(const acfd_config *cfg, const int *r, const double *pri_v, const double *amu,
                 const double *dxidx, double *dt, double *cfl)
{
    if (!cfg || !r || !pri_v || !amu || !dxidx || !dt || !cfl) return fail(ACFD_EINVAL, "acfd_find_dt: null argument");
    Single s;
    ST_TRY(make_single(cfg, r, &s.ctx));
    ST_TRY(put(s.ctx, ACFD_PRI_V, pri_v));
    ST_TRY(put(s.ctx, ACFD_AMU, amu));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_DT, dt));
    ST_TRY(set_cfl(s.ctx, *cfl));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_FIND_DT));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_DT, dt));
    HIP_TRY(hipMemcpy(cfl, s.ctx->d_cfl, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int acfd_lusgs(const acfd_config *cfg, const int *r, double *q, const double *dt, const double *pri_v,
               const double *rhsi, const double *rhsv, const double *dxidx, const double *inv_j, double cfl,
               double *rhs)
{
    if (!cfg || !r || !q || !dt || !pri_v || !rhsi || !rhsv || !dxidx || !inv_j)
        return fail(ACFD_EINVAL, "acfd_lusgs: null argument");
    Single s;
    ST_TRY(make_single(cfg, r, &s.ctx));
    ST_TRY(put(s.ctx, ACFD_Q, q));
    ST_TRY(put(s.ctx, ACFD_DT, dt));
    ST_TRY(put(s.ctx, ACFD_PRI_V, pri_v));
    ST_TRY(put(s.ctx, ACFD_RHSI, rhsi));
    ST_TRY(put(s.ctx, ACFD_RHSV, rhsv));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_INV_J, inv_j));
    ST_TRY(set_cfl(s.ctx, cfl));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_IMPLICIT1ST));
    {
        int rc = tile_error(s.ctx);
        if (rc) return rc;
    }
    ST_TRY(acfd_download(s.ctx, 1, ACFD_Q, q));
    if (rhs) ST_TRY(acfd_download(s.ctx, 1, ACFD_RHS, rhs));
    return 0;
}

int acfd_bdf2(const acfd_config *cfg, const int *r, double *q, const double *q0, const double *dt,
              const double *pri_v, const double *rhsi, const double *rhsv, const double *dxidx,
              const double *inv_j, double *rhs)
{
    if (!cfg || !r || !q || !q0 || !dt || !pri_v || !rhsi || !rhsv || !dxidx || !inv_j)
        return fail(ACFD_EINVAL, "acfd_bdf2: null argument");
    acfd_config c = *cfg;
    c.iflag_timeadvance = 3;
    Single s;
    ST_TRY(make_single(&c, r, &s.ctx));
    ST_TRY(put(s.ctx, ACFD_Q, q));
    ST_TRY(put(s.ctx, ACFD_Q0, q0));
    ST_TRY(put(s.ctx, ACFD_DT, dt));
    ST_TRY(put(s.ctx, ACFD_PRI_V, pri_v));
    ST_TRY(put(s.ctx, ACFD_RHSI, rhsi));
    ST_TRY(put(s.ctx, ACFD_RHSV, rhsv));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_INV_J, inv_j));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_CRANK2ND));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_Q, q));
    if (rhs) ST_TRY(acfd_download(s.ctx, 1, ACFD_RHS, rhs));
    return 0;
}

int acfd_sa_step(const acfd_config *cfg, const int *r, double *nut, double *sa_rhs, const double *dt,
                 const double *pri_v, const double *amu, const double *dudx, const double *vor,
                 const double *dxidx, const double *inv_j, const double *aalpha, const double *bbeta,
                 const double *ggamma, double *length, const double *dist, const double *spacing, double cfl)
{
    if (!cfg || !r || !nut || !sa_rhs || !dt || !pri_v || !amu || !dudx || !vor || !dxidx || !inv_j || !aalpha ||
        !bbeta || !ggamma || !length || !dist || !spacing)
        return fail(ACFD_EINVAL, "acfd_sa_step: null argument");
    acfd_config c = *cfg;
    c.iflag_solver = 0;
    c.iflag_turbulence = 2;
    Single s;
    ST_TRY(make_single(&c, r, &s.ctx));
    ST_TRY(put(s.ctx, ACFD_NUT, nut));
    ST_TRY(put(s.ctx, ACFD_DT, dt));
    ST_TRY(put(s.ctx, ACFD_PRI_V, pri_v));
    ST_TRY(put(s.ctx, ACFD_AMU, amu));
    ST_TRY(put(s.ctx, ACFD_DUDX, dudx));
    ST_TRY(put(s.ctx, ACFD_VOR, vor));
    ST_TRY(put(s.ctx, ACFD_DXIDX, dxidx));
    ST_TRY(put(s.ctx, ACFD_INV_J, inv_j));
    ST_TRY(put(s.ctx, ACFD_AALPHA, aalpha));
    ST_TRY(put(s.ctx, ACFD_BBETA, bbeta));
    ST_TRY(put(s.ctx, ACFD_GGAMMA, ggamma));
    ST_TRY(put(s.ctx, ACFD_DIST, dist));
    ST_TRY(put(s.ctx, ACFD_SPACING, spacing));
    ST_TRY(set_cfl(s.ctx, cfl));
    ST_TRY(acfd_run_stage(s.ctx, ACFD_STAGE_SA));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_NUT, nut));
    ST_TRY(acfd_download(s.ctx, 1, ACFD_SA_RHS, sa_rhs));
    return acfd_download(s.ctx, 1, ACFD_LENGTH, length);
}

}  // extern "C"

// ------------------------------------------------- no-allocation viscous_flux
// The viscous node fluxes (12 doubles per node) of acfd_viscous_rhs_noalloc
// live in one file-scope device array instead of an allocation.  Each GPU
// that loads the library holds its own copy (48 MiB) until the process
// ends.  Host code never uses the array's name as an address: the address
// on each device comes from hipGetSymbolAddress, once per device, with that
// device current.
namespace {

constexpr long kNoAllocMaxNodes = 1L << 19;  // extended nodes of one call
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[12 * kNoAllocMaxNodes];

double *g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mutex;

int scratch_on_current_device(double **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ACFD_EDEVICE, "acfd_viscous_rhs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ACFD_EDEVICE, "acfd_viscous_rhs_noalloc: no address for the scratch array");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int acfd_viscous_rhs_noalloc(const acfd_config *cfg, const int *r, const double *pri_v, double *amu,
                             const double *dxidx, const double *inv_j, double *rhsv, double *vor, double *dudx,
                             double *tao, void *stream)
{
    if (!cfg || !r || !pri_v || !amu || !dxidx || !inv_j || !rhsv)
        return fail(ACFD_EINVAL, "acfd_viscous_rhs_noalloc: null argument");
    if ((dudx || tao) && !(vor && dudx && tao))
        return fail(ACFD_EINVAL, "acfd_viscous_rhs_noalloc: dudx and tao need vor, dudx and tao together");
    if (cfg->precision != 0 && cfg->precision != 1)
        return fail(ACFD_EINVAL, "acfd_viscous_rhs_noalloc: precision must be 0 or 1");
    const Ext e = ext_of_ranges(r);
    if (e.ni < 3 || e.nj < 3 || e.nk < 3 || e.is0 < e.is || e.ie0 > e.ie || e.js0 < e.js || e.je0 > e.je ||
        e.ks0 < e.ks || e.ke0 > e.ke || e.ie0 - e.is0 < 1 || e.je0 - e.js0 < 1 || e.ke0 - e.ks0 < 1)
        return fail(ACFD_EINVAL, "acfd_viscous_rhs_noalloc: bad ranges");
    if (e.n > kNoAllocMaxNodes)
        return fail(ACFD_EINVAL, "acfd_viscous_rhs_noalloc: block of " + std::to_string(e.n) +
                                     " extended nodes exceeds the scratch limit of " +
                                     std::to_string(kNoAllocMaxNodes));
    double *w = nullptr;
    int rc = scratch_on_current_device(&w);
    if (rc) return rc;
    Consts c = consts_of(*cfg);
    c.solver = 0;
    Fields f{};
    f.pri = const_cast<double *>(pri_v);
    f.amu = amu;
    f.dxidx = const_cast<double *>(dxidx);
    f.inv_j = const_cast<double *>(inv_j);
    f.rhsv = rhsv;
    f.vor = vor;
    f.dudx = dudx;
    f.tao = tao;
    f.w_flux = w;
    const int store_aux = dudx ? 2 : vor ? 1 : 0;
    KRN_TRY(launch_viscous((hipStream_t)stream, e, f, c, cfg->precision, store_aux));
    return 0;
}

}  // extern "C"
