// First-order implicit update: the LU-SGS sweeps of implicit1st.
//
// Reference: sub_timeadvancement_jac.f90:7-63 (calc_jm) and
// sub_timeadvancement_1stimplicit.f90:7-585 (implicit1st).
//
// LU-SGS is a lexicographic Gauss-Seidel sweep: node (i,j,k) needs dq* of
// (i-1,j,k), (i,j-1,k) and (i,j,k-1) and nothing else of the current sweep.
// Nodes on one hyperplane i+j+k = l are independent, and any schedule that
// respects these three dependencies computes every node from the same
// operands as the sequential loop.  Two schedules:
//
//   hyperplanes  one launch per plane l (2 (Li+Lj+Lk-2) launches per step),
//                operands read in place.
//
//   tiles        one persistent launch per sweep.  Each workgroup owns a
//                TJ x TK tile of (j,k) lines and marches along i on a skewed
//                clock: thread (tj,tk) handles the node at sweep position
//                s - tj - tk at step s.
//                * Sweep layout.  Before the forward sweep the operands of
//                  every node (metric rows, 1/J, q, sound speed, the time
//                  term phi/dt, rhs: 22 values; the sweeps form the spectral
//                  radii and the diagonal from them) are gathered into an
//                  array ordered by (tile, step, field pair, lane), so at
//                  every step a tile reads 11 contiguous rows of TJ*TK
//                  16-byte pairs -- fully coalesced, where
//                  the natural layout would have every lane on its own cache
//                  line.  The sweeps write dq* and dq over the rhs slots and a
//                  final pass adds dq to q.
//                * Push form.  A node, once solved, forms the three
//                  off-diagonal products 1/2 inv_j (A(q; grad xi_d) +- r_d I) dq
//                  that its successors need, with its own operands: the i
//                  product stays in registers, the j and k products go to the
//                  next lanes through LDS, and the tile's last row/column
//                  hand them to the neighbouring tile through exchange slots.
//                  Each product is the same expression of the same operands
//                  as in the pull form, so both schedules give the same bits.
//                * Exchange.  Slots live in uncached memory and hold a
//                  sentinel bit pattern (a signalling NaN no arithmetic
//                  produces) until the producer stores the value; the
//                  consumer spins on the value itself and writes the sentinel
//                  back.  No counters, no fences, no waits for store
//                  acknowledgements on the per-step path; operands and
//                  exchange values are loaded one step ahead, and a slot
//                  found late is polled for all its late components at once.
//
// The flux Jacobian products use the closed form of A(q) v for the Euler
// flux in generalised coordinates (one reciprocal of rho instead of the ~25
// divisions of forming calc_jm's matrix); the values agree with the matrix
// form to rounding.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "kernels.hpp"

// No implicit contraction in this file: the sweeps' arithmetic is spelled
// out (explicit fma) so that every schedule produces the same bits.
#pragma clang fp contract(off)

namespace acfd {

// 0.5 * (A(q; m) v +- r v) * inv_j for the operands of one node
// (sub_timeadvancement_1stimplicit.f90:283-301 forward, :459-477 backward).
// A v = [ m.v_mom,
//         U v_mom + u theta + m dp,
//         a1 theta + U (dp + v5) ]
// with u = mom/rho, U = m.u, theta = m.v_mom - U v1,
// dp = (gamma-1)(|u|^2/2 v1 - u.v_mom + v5), a1 = gamma E/rho - (gamma-1)|u|^2/2.
template <int SGN>
__device__ inline void jac_mv_r(const double *qv, double mx, double my, double mz, double ij, double rr,
                                double gamma, const double *v, double *out)
{
    const double rho = qv[0], mu = qv[1], mv = qv[2], mw = qv[3], E = qv[4];
    const double ri = 1.0 / rho;
    const double u = mu * ri, w2 = mv * ri, w3 = mw * ri;
    const double U = fma(mz, w3, fma(my, w2, mx * u));
    const double q2 = fma(w3, w3, fma(w2, w2, u * u));
    const double g1 = gamma - 1.0;
    const double a1 = fma(gamma * E, ri, -0.5 * g1 * q2);
    const double mvm = fma(mz, v[3], fma(my, v[2], mx * v[1]));
    const double theta = fma(-U, v[0], mvm);
    const double uv = fma(w3, v[3], fma(w2, v[2], u * v[1]));
    const double dp = g1 * (fma(0.5 * q2, v[0], -uv) + v[4]);
    const double r = SGN > 0 ? rr : -rr;
    const double s = 0.5 * ij;
    out[0] = s * fma(r, v[0], mvm);
    out[1] = s * fma(r, v[1], fma(mx, dp, fma(u, theta, U * v[1])));
    out[2] = s * fma(r, v[2], fma(my, dp, fma(w2, theta, U * v[2])));
    out[3] = s * fma(r, v[3], fma(mz, dp, fma(w3, theta, U * v[3])));
    out[4] = s * fma(r, v[4], fma(U, dp + v[4], a1 * theta));
}

// the same product with the operands of node n read from the SoA fields
template <int SGN>
__device__ inline void jac_mv(const double *__restrict__ q, const double *__restrict__ dxidx,
                              const double *__restrict__ inv_j, const double *__restrict__ rr, long N, long n,
                              int c0, double gamma, const double *v, double *out)
{
    const double qv[5] = {q[n], q[N + n], q[2 * N + n], q[3 * N + n], q[4 * N + n]};
    jac_mv_r<SGN>(qv, dxidx[c0 * N + n], dxidx[(c0 + 1) * N + n], dxidx[(c0 + 2) * N + n], inv_j[n], rr[n],
                  gamma, v, out);
}

// forward node (:345-365): dq* = (rhs + A+ dq*_i-1 + B+ dq*_j-1 + C+ dq*_k-1)/diag/inv_j
__device__ inline void lu_fwd_combine(const double *rhs, const double *ta, const double *tb, const double *tc,
                                      double dg, double ij, double *res)
{
#pragma unroll
    for (int x = 0; x < 5; x++) res[x] = (rhs[x] + (ta[x] + tb[x] + tc[x])) / dg / ij;
}

// backward node (:521-542): dq = dq* - (A- dq_i+1 + B- dq_j+1 + C- dq_k+1)/inv_j/diag
__device__ inline void lu_bwd_combine(const double *dqs, const double *ta, const double *tb, const double *tc,
                                      double dg, double ij, double *res)
{
#pragma unroll
    for (int x = 0; x < 5; x++) res[x] = dqs[x] - (ta[x] + tb[x] + tc[x]) / ij / dg;
}

struct LuArgs {
    Ext e;
    const double *q, *dxidx, *inv_j, *rr, *diag, *rhs;
    double *dqs, *dq;
    double gamma;
};

__device__ inline void lu_forward(const LuArgs &a, long n, const double *vi, const double *vj,
                                  const double *vk, double *res)
{
    const long N = a.e.n, sj = a.e.ni, sk = (long)a.e.ni * a.e.nj;
    double ta[5], tb[5], tc[5], rhs[5];
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr, N, n - 1, 0, a.gamma, vi, ta);
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr + N, N, n - sj, 3, a.gamma, vj, tb);
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr + 2 * N, N, n - sk, 6, a.gamma, vk, tc);
#pragma unroll
    for (int x = 0; x < 5; x++) rhs[x] = a.rhs[x * N + n];
    lu_fwd_combine(rhs, ta, tb, tc, a.diag[n], a.inv_j[n], res);
}

__device__ inline void lu_backward(const LuArgs &a, long n, const double *vi, const double *vj,
                                   const double *vk, double *res)
{
    const long N = a.e.n, sj = a.e.ni, sk = (long)a.e.ni * a.e.nj;
    double ta[5], tb[5], tc[5], dqs[5];
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr, N, n + 1, 0, a.gamma, vi, ta);
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr + N, N, n + sj, 3, a.gamma, vj, tb);
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr + 2 * N, N, n + sk, 6, a.gamma, vk, tc);
#pragma unroll
    for (int x = 0; x < 5; x++) dqs[x] = a.dqs[x * N + n];
    lu_bwd_combine(dqs, ta, tb, tc, a.diag[n], a.inv_j[n], res);
}

// rhs = rhsi + rhsv on the computed range (:72-78).  For the hyperplane
// sweeps (tiles = 0): spectral radii ra, rb, rc and diag on the ring
// is1-1..ie1+1 (:98-156), and dq* and dq cleared on the ring so that
// out-of-range neighbours read 0 (:105-106).  For the tile sweeps, which
// form radii and diagonal themselves: only the time term phi*1/dt into diag.
__device__ inline void lu_prep_body(const Ext &e, long wg, const double *__restrict__ pri,
                                    const double *__restrict__ dxidx, const double *__restrict__ dt,
                                    const double *__restrict__ rhsi, const double *__restrict__ rhsv,
                                    double *__restrict__ rhs, double *__restrict__ rr, double *__restrict__ diag,
                                    double *__restrict__ dqs, double *__restrict__ dq, const double *cfl,
                                    int tiles)
{
    const Box ring{e.is1 - 1, e.ie1 + 1, e.js1 - 1, e.je1 + 1, e.ks1 - 1, e.ke1 + 1};
    int i, j, k;
    long n;
    if (!lin_node_at(e, ring, wg, i, j, k, n)) return;
    const long N = e.n;
    const double phi = *cfl >= 1.0e3 ? 0.0 : 1.0;  // :98-102
    if (tiles) diag[n] = phi * 1.0 / dt[n];
    else diag[n] = lu_radii_diag(pri, dxidx, N, n, dt[n], phi, rr);
    bool inner = i >= e.is1 && i <= e.ie1 && j >= e.js1 && j <= e.je1 && k >= e.ks1 && k <= e.ke1;
#pragma unroll
    for (int x = 0; x < 5; x++) {
        if (!tiles) {
            dqs[x * N + n] = 0.0;
            dq[x * N + n] = 0.0;
        }
        if (inner) rhs[x * N + n] = rhsi[x * N + n] + rhsv[x * N + n];
    }
}

__global__ void __launch_bounds__(256) k_lu_prep(Ext e, const double *__restrict__ pri,
                                                 const double *__restrict__ dxidx, const double *__restrict__ dt,
                                                 const double *__restrict__ rhsi, const double *__restrict__ rhsv,
                                                 double *__restrict__ rhs, double *__restrict__ rr,
                                                 double *__restrict__ diag, double *__restrict__ dqs,
                                                 double *__restrict__ dq, const double *cfl, int tiles)
{
    lu_prep_body(e, xcd_slab(blockIdx.x, gridDim.x), pri, dxidx, dt, rhsi, rhsv, rhs, rr, diag, dqs, dq, cfl,
                 tiles);
}

// every block, for the batched tile sweeps (rhs and the time term; dq*
// and dq are never touched); block m reads cfl[m * cfl_stride]
__global__ void __launch_bounds__(256) k_lu_prep_batch(const BlockRec *__restrict__ blks,
                                                       const int *__restrict__ map, const KGrid *__restrict__ kg,
                                                       double *arena, const double *cfl, int cfl_stride)
{
    int bx, by, bz;
    const int m = batch_wg(map, kg, bx, by, bz);
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    lu_prep_body(e, bx, arena + b.o.pri, arena + b.o.dxidx, arena + b.o.dt, arena + b.o.rhsi,
                 arena + b.o.rhsv, arena + b.o.rhs, arena + b.o.w_r, arena + b.o.w_diag, arena + b.o.w_dqs,
                 arena + b.o.w_dq, cfl + (long)m * cfl_stride, 1);
}

// linear tiles of 256 nodes over the block's storage (the ring of the
// computed range is masked inside), 64 x 4 threads
dim3 lu_prep_grid(const Ext &e) { return lin_grid(e); }

int launch_lu_prep_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const double *cfl,
                         int cfl_stride)
{
    if (bg.total <= 0) return 0;
    hipLaunchKernelGGL(k_lu_prep_batch, dim3((unsigned)bg.total), dim3(64, 4, 1), 0, s, blks, bg.map, bg.kg, arena,
                       cfl, cfl_stride);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// One hyperplane l of the forward (DIR = +1) or backward (DIR = -1) sweep.
// Threads cover (lj, lk) from (ja, ka); li = l - lj - lk; the backward sweep
// counts from the far corner.
template <int DIR>
__global__ void __launch_bounds__(256) k_lu_plane(LuArgs a, int l, int ja, int ka)
{
    const Ext &e = a.e;
    int lj = ja + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int lk = ka + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    int li = l - lj - lk;
    if (lj >= Lj || lk >= Lk || li < 0 || li >= Li) return;
    const long N = e.n, sj = e.ni, sk = (long)e.ni * e.nj;
    long n = DIR > 0 ? e.ix(e.is1 + li, e.js1 + lj, e.ks1 + lk) : e.ix(e.ie1 - li, e.je1 - lj, e.ke1 - lk);
    double *out = DIR > 0 ? a.dqs : a.dq;
    double vi[5], vj[5], vk[5], res[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        vi[x] = out[x * N + n - DIR];
        vj[x] = out[x * N + n - DIR * sj];
        vk[x] = out[x * N + n - DIR * sk];
    }
    if (DIR > 0) lu_forward(a, n, vi, vj, vk, res);
    else lu_backward(a, n, vi, vj, vk, res);
#pragma unroll
    for (int x = 0; x < 5; x++) out[x * N + n] = res[x];
}

// q += dq on the computed range (:568-574)
__global__ void __launch_bounds__(256) k_add5(Ext e, Box b, double *__restrict__ q,
                                              const double *__restrict__ dq)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
#pragma unroll
    for (int x = 0; x < 5; x++) q[x * N + n] = q[x * N + n] + dq[x * N + n];
}

// ------------------------------------------------------------ tile sweeps
// Sweep-layout fields: dxidx 0-8, inv_j 9, q 10-14, sound speed 15, the
// time term phi/dt 16, rhs 17-21 (overwritten by dq* in the forward and dq
// in the backward sweep).  The sweeps form the spectral radii and the
// diagonal from these (lu_radius, lu_diag), with the velocity q[1..3]/q[0]
// as get_primitivevariables forms it, so the results are those of radii
// and diagonal read from the preparation pass.  The grid's operands come
// first, as whole field pairs: while the sweep layout holds the same block
// and tiling and the metrics are unchanged, a gather can leave those pairs
// in place and move only the other 12 fields.
constexpr int kSlF = 22;
constexpr int kSlQ = 10;
constexpr int kSlC = 15;
constexpr int kSlT = 16;
constexpr int kSlOut = 17;
static_assert(kSlOut % 2 == 1 && kSlOut + 5 == kSlF && kSlT == kSlOut - 1,
              "the results are the upper half of the time term's pair and the two pairs after it");
constexpr int kSlGeoPairs = 5;
// exchange sentinel: a negative signalling NaN with a payload; arithmetic
// only ever produces quiet NaNs
constexpr unsigned long long kSentinel = 0xFFF5A5A5A5A5A5A5ull;

// geometry of one tiling of a block's computed range
struct SlGeom {
    int TJ, TK, TT, nbj, nbk, ntiles, nsteps, Li, Lj, Lk;
    int lgTJ;  // TJ is a power of two
    int refill;  // consumers leave exchange slots as they are; the region is refilled after each sweep
    long exj_tile, exk_tile;  // doubles per tile of the j / k exchange regions
    // fields are stored in pairs, so a lane's operands of one step are 12
    // 16-byte loads; double index of field f of lane l at row t
    __host__ __device__ long slot(int tile, int t, int f, int l) const {
        return ((((long)tile * nsteps + t) * (kSlF / 2) + (f >> 1)) * TT + l) * 2 + (f & 1);
    }
};

// tile shapes the autotuner chooses from (TJ, TK)
const int kTileShapes[][2] = {{16, 16}, {16, 8}, {8, 8}, {16, 4}};
const int kNumTileShapes = 4;

// Exchange slots are reset to the sentinel by one fill kernel after each
// flow sweep rather than by their consumers, one store each: the per-step
// stores were on the sweep's critical path.
static int ex_refill()
{
    static int r = -1;
    if (r < 0) {
        const char *v = getenv("ACFD_EX_REFILL");
        r = v && atoi(v) == 0 ? 0 : 1;
    }
    return r;
}

static SlGeom sl_geom(const Ext &e, int shape)
{
    SlGeom g;
    g.refill = ex_refill();
    g.TJ = kTileShapes[shape][0];
    g.TK = kTileShapes[shape][1];
    g.TT = g.TJ * g.TK;
    g.lgTJ = 0;
    while ((1 << g.lgTJ) < g.TJ) g.lgTJ++;
    g.Li = e.ie1 - e.is1 + 1;
    g.Lj = e.je1 - e.js1 + 1;
    g.Lk = e.ke1 - e.ks1 + 1;
    g.nbj = (g.Lj + g.TJ - 1) / g.TJ;
    g.nbk = (g.Lk + g.TK - 1) / g.TK;
    g.ntiles = g.nbj * g.nbk;
    g.nsteps = g.Li + g.TJ + g.TK - 2;
    g.exj_tile = (long)(g.Li + g.TK - 1) * g.TK * 5;
    g.exk_tile = (long)(g.Li + g.TJ - 1) * g.TJ * 5;
    return g;
}

__device__ inline unsigned long long xload_bits(const double *p)
{
    return __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore_bits(double *p, unsigned long long v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Take the five values of one exchange slot (component x at slot + x *
// stride): `pre` holds what was loaded a step earlier; components still at the sentinel are polled until the
// producer's store arrives, each poll reloading the later components too,
// so that a slot that was late costs one round trip more, not one per
// component.  The wait is bounded: running out sets *err and every later
// wait of the launch gives up at once, so the sweep always ends (the host
// then reports the failure).  With `reset` the slot is set back to the
// sentinel for the next sweep (otherwise the host refills the region).
__device__ inline void ex_take(double *slot, int stride, const unsigned long long *pre, double *v, int *err,
                               bool &failed, bool reset)
{
    unsigned long long cur[5];
#pragma unroll
    for (int x = 0; x < 5; x++) cur[x] = pre[x];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        unsigned long long b = cur[x];
        if (b == kSentinel) {
            for (long spins = 0; !failed; spins++) {
                b = xload_bits(slot + x * stride);
#pragma unroll
                for (int y = x + 1; y < 5; y++)
                    if (cur[y] == kSentinel) cur[y] = xload_bits(slot + y * stride);
                if (b != kSentinel) break;
                if ((spins & 1023) == 1023 &&
                    (spins > (1L << 21) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    atomicOr(err, 1);
                    failed = true;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (failed) b = 0;
        }
        v[x] = __longlong_as_double((long long)b);
        if (reset) xstore_bits(slot + x * stride, kSentinel);
    }
}

// LDS barrier that does not wait for outstanding global loads (the
// prefetched operands stay in flight across it)
__device__ inline void lds_barrier()
{
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The layout passes move data between the natural layout (i fastest) and
// the sweep layout through LDS.  One workgroup per tile walks the tile's
// steps in chunks of CT steps, one field pair at a time through all the
// chunks, so consecutive chunks of a line are adjacent in memory and the
// cache lines a chunk boundary splits are completed by the same workgroup
// one pair-chunk (2 CT TT doubles) later: near enough for the device-wide
// caches to return them (all pairs per chunk: 5.24 against 4.80 ms for the
// gather on the 40-block channel).  In a chunk starting at
// step t0, lane l's line contributes the CT nodes li = t0 + ii - tj - tk,
// ii = 0..CT-1, which are exactly rows t0 + ii of the sweep layout.
// Natural side: thread x handles ii = x % CT of lines x / CT + m TT/CT,
// so CT consecutive lanes touch CT consecutive doubles of one line
// (CT = min(32, 4096/TT)).
// Sweep side: thread x is lane x.  LDS holds [CT][TT] doubles per field
// with the line index XOR-swizzled by ii (no bank conflicts either way).
// steps a layout kernel transposes at a time: CT * TT <= kSlElems
constexpr int kSlElems = 4096;
constexpr int sl_ct_of(int tt) { return kSlElems / tt < 32 ? kSlElems / tt : 32; }
// the SA return pass (one value per node) takes at most 16: measured on
// 64-lane tiles, 0.86 against 1.25 ms (the two gathers and the flow add are
// faster at 32)
constexpr int sa_ret_ct_of(int tt) { return sl_ct_of(tt) < 16 ? sl_ct_of(tt) : 16; }
// dynamic LDS of the flow layout kernels: [2][CT][TT] doubles
constexpr size_t sl_lds_bytes(int tt) { return 2 * (size_t)sl_ct_of(tt) * tt * sizeof(double); }

// node index of (tile, line, li), li in range
__device__ inline long sl_nat(const Ext &e, const SlGeom &g, int tile, int line, int li)
{
    const int tj = line & (g.TJ - 1), tk = line >> g.lgTJ;
    return e.ix(e.is1 + li, e.js1 + (tile % g.nbj) * g.TJ + tj, e.ks1 + (tile / g.nbj) * g.TK + tk);
}
__device__ inline bool sl_line_ok(const SlGeom &g, int tile, int line)
{
    return (tile % g.nbj) * g.TJ + (line & (g.TJ - 1)) < g.Lj && (tile / g.nbj) * g.TK + (line >> g.lgTJ) < g.Lk;
}

// csnd: the sound speed (component 7 of pri_v), tterm: the time term
__device__ inline const double *gather_src(int f, long N, const double *q, const double *dxidx,
                                           const double *inv_j, const double *csnd, const double *tterm,
                                           const double *rhs)
{
    return f < 9 ? dxidx + f * N : f == 9 ? inv_j : f < kSlC ? q + (f - kSlQ) * N
         : f == kSlC ? csnd : f == kSlT ? tterm : rhs + (f - kSlOut) * N;
}

// Gather the operands of every node into the sweep layout, a field pair at
// a time (each lane then stores whole 16-byte pairs), from pair fp0 on (0:
// all kSlF fields, kSlGeoPairs: all but the grid's).  TT is a template
// parameter so that a thread's CT loads are issued back to back.
template <int TT>
__device__ inline void sl_gather_tile(const Ext &e, const SlGeom &g, int tile, const double *__restrict__ q,
                                      const double *__restrict__ dxidx, const double *__restrict__ inv_j,
                                      const double *__restrict__ csnd, const double *__restrict__ tterm,
                                      const double *__restrict__ rhs, double *__restrict__ sl, int fp0,
                                      double *buf /* [2][CT][TT] LDS, swizzled */)
{
    constexpr int CT = sl_ct_of(TT), lstep = TT / CT;
    const int x = threadIdx.x;
    const int ii_n = x % CT;
    const int tj = x & (g.TJ - 1), tk = x >> g.lgTJ;
    const bool me_ok = sl_line_ok(g, tile, x);
    const long N = e.n;
    double2 *sl2 = reinterpret_cast<double2 *>(sl);
    for (int fp = fp0; fp < kSlF / 2; fp++) {
        for (int t0 = 0; t0 < g.nsteps; t0 += CT) {
            // natural-side node offsets of this thread's CT elements (-1: none)
            int nat[CT];  // node offsets fit 32 bits (blocks below 2^31 nodes)
#pragma unroll
            for (int m = 0; m < CT; m++) {
                const int line = m * lstep + x / CT;
                const int li = t0 + ii_n - (line & (g.TJ - 1)) - (line >> g.lgTJ);
                nat[m] = (t0 + ii_n < g.nsteps && li >= 0 && li < g.Li && sl_line_ok(g, tile, line))
                             ? (int)sl_nat(e, g, tile, line, li) : -1;
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double *src = gather_src(2 * fp + h, N, q, dxidx, inv_j, csnd, tterm, rhs);
                double v[CT];
#pragma unroll
                for (int m = 0; m < CT; m++) v[m] = nat[m] >= 0 ? src[nat[m]] : 0.0;
#pragma unroll
                for (int m = 0; m < CT; m++) {
                    const int line = m * lstep + x / CT;
                    buf[(h * CT + ii_n) * TT + (line ^ ii_n)] = v[m];
                }
            }
            __syncthreads();
            if (me_ok) {
#pragma unroll
                for (int ii = 0; ii < CT; ii++) {
                    const int t = t0 + ii, li = t - tj - tk;
                    if (t < g.nsteps && li >= 0 && li < g.Li)
                        sl2[g.slot(tile, t, 2 * fp, x) >> 1] =
                            make_double2(buf[ii * TT + (x ^ ii)], buf[(CT + ii) * TT + (x ^ ii)]);
                }
            }
            __syncthreads();
        }
    }
}

template <int TT>
__global__ void __launch_bounds__(TT) k_sl_gather(Ext e, SlGeom g, const double *__restrict__ q,
                                                  const double *__restrict__ dxidx,
                                                  const double *__restrict__ inv_j,
                                                  const double *__restrict__ csnd,
                                                  const double *__restrict__ tterm,
                                                  const double *__restrict__ rhs, double *__restrict__ sl, int fp0)
{
    extern __shared__ double buf[];
    sl_gather_tile<TT>(e, g, (int)blockIdx.x, q, dxidx, inv_j, csnd, tterm, rhs, sl, fp0, buf);
}

// q += dq (:568-574): the backward sweep left q + dq in the result slots of
// the sweep layout; this pass moves it to q in the natural layout, same walk
// as the gather, one field pair at a time.
template <int TT>
__device__ inline void sl_add_tile(const Ext &e, const SlGeom &g, int tile, const double *__restrict__ sl,
                                   double *__restrict__ q, double *buf /* [2][CT][TT] LDS, swizzled */)
{
    constexpr int CT = sl_ct_of(TT), lstep = TT / CT;
    const int x = threadIdx.x;
    const int ii_n = x % CT;
    const int tj = x & (g.TJ - 1), tk = x >> g.lgTJ;
    const bool me_ok = sl_line_ok(g, tile, x);
    const long N = e.n;
    // the results are fields 19-23: the upper half of pair 9 (whose lower
    // half, diag, is not written) and pairs 10 and 11, read as whole
    // 16-byte pairs
    for (int fp = kSlOut / 2; fp < kSlF / 2; fp++) {
        for (int t0 = 0; t0 < g.nsteps; t0 += CT) {
            int nat[CT];  // node offsets fit 32 bits (blocks below 2^31 nodes)
#pragma unroll
            for (int m = 0; m < CT; m++) {
                const int line = m * lstep + x / CT;
                const int li = t0 + ii_n - (line & (g.TJ - 1)) - (line >> g.lgTJ);
                nat[m] = (t0 + ii_n < g.nsteps && li >= 0 && li < g.Li && sl_line_ok(g, tile, line))
                             ? (int)sl_nat(e, g, tile, line, li) : -1;
            }
            const int v0 = 2 * fp - kSlOut;  // component of the pair's lower field (-1: diag)
            if (me_ok) {
                const double2 *sl2 = reinterpret_cast<const double2 *>(sl);
                double2 d[CT];
#pragma unroll
                for (int ii = 0; ii < CT; ii++) {
                    const int t = t0 + ii, li = t - tj - tk;
                    d[ii] = (t < g.nsteps && li >= 0 && li < g.Li) ? sl2[g.slot(tile, t, 2 * fp, x) >> 1]
                                                                   : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int ii = 0; ii < CT; ii++) {
                    buf[ii * TT + (x ^ ii)] = d[ii].x;
                    buf[(CT + ii) * TT + (x ^ ii)] = d[ii].y;
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < CT; m++) {
                const int line = m * lstep + x / CT;
                if (nat[m] >= 0) {
                    if (v0 >= 0) q[v0 * N + nat[m]] = buf[ii_n * TT + (line ^ ii_n)];
                    q[(v0 + 1) * N + nat[m]] = buf[(CT + ii_n) * TT + (line ^ ii_n)];
                }
            }
            __syncthreads();
        }
    }
}

template <int TT>
__global__ void __launch_bounds__(TT) k_sl_add(Ext e, SlGeom g, const double *__restrict__ sl,
                                               double *__restrict__ q)
{
    extern __shared__ double buf[];
    sl_add_tile<TT>(e, g, (int)blockIdx.x, sl, q, buf);
}

template <int TT>
static void launch_layout(hipStream_t s, int which, const Ext &e, const SlGeom &g, const Fields &f,
                          const LuWs &ws, int fp0)
{
    const size_t lds = sl_lds_bytes(TT);
    if (which == 0)
        hipLaunchKernelGGL(k_sl_gather<TT>, dim3((unsigned)g.ntiles), dim3(TT), lds, s, e, g, f.q, f.dxidx,
                           f.inv_j, f.pri + 6 * e.n, f.w_diag, f.rhs, ws.sl, fp0);
    else
        hipLaunchKernelGGL(k_sl_add<TT>, dim3((unsigned)g.ntiles), dim3(TT), lds, s, e, g, ws.sl, f.q);
}

static void launch_layout_tt(hipStream_t s, int which, const Ext &e, const SlGeom &g, const Fields &f,
                             const LuWs &ws, int fp0 = 0)
{
    if (g.TT == 256) launch_layout<256>(s, which, e, g, f, ws, fp0);
    else if (g.TT == 128) launch_layout<128>(s, which, e, g, f, ws, fp0);
    else launch_layout<64>(s, which, e, g, f, ws, fp0);
}

// Fill exchange slots with the sentinel.
__global__ void __launch_bounds__(256) k_fill_sentinel(double *p, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xstore_bits(p + i, kSentinel);
}

// The node's spectral radii and diagonal from its operands: velocity as
// get_primitivevariables forms it (sub_getpri_v.f90:39-41)
__device__ inline double lu_node_radii(const double *op, double *r)
{
    const double d = op[kSlQ];
    const double u = op[kSlQ + 1] / d, v = op[kSlQ + 2] / d, w = op[kSlQ + 3] / d;
#pragma unroll
    for (int x = 0; x < 3; x++) r[x] = lu_radius(op[3 * x], op[3 * x + 1], op[3 * x + 2], u, v, w, op[kSlC]);
    return lu_diag(op[kSlT], r);
}

// One step of a lane: combine, solve, push.  op = the node's operands, dg
// and r its diagonal and spectral radii.
template <int DIR>
__device__ inline void lu_node(const double *op, double dg, const double *r, const double *ta, const double *tb,
                               const double *tc, double gamma, double *res, double *pa, double *pb, double *pc)
{
    if (DIR > 0) lu_fwd_combine(op + kSlOut, ta, tb, tc, dg, op[9], res);
    else lu_bwd_combine(op + kSlOut, ta, tb, tc, dg, op[9], res);
    jac_mv_r<DIR>(op + kSlQ, op[0], op[1], op[2], op[9], r[0], gamma, res, pa);
    jac_mv_r<DIR>(op + kSlQ, op[3], op[4], op[5], op[9], r[1], gamma, res, pb);
    jac_mv_r<DIR>(op + kSlQ, op[6], op[7], op[8], op[9], r[2], gamma, res, pc);
}

// Persistent tile sweep over the sweep layout.  The forward sweep numbers
// tiles bj + nbj*bk and waits only on lower numbers; the backward sweep runs
// the same tiles mirrored (lane (tj,tk) schedules as (TJ-1-tj, TK-1-tk)) and
// is dispatched in reverse order.  Either way every wait is on a tile that
// was dispatched earlier, so it is running or done.
// With q (backward sweep only) the update q += dq is fused: each node's dq
// is added to its q operand and stored into the natural layout (e: the
// block's ranges), in place of writing dq to the sweep layout.
template <int DIR, int TJ, int TK>
__device__ inline void lu_tiles_body(const SlGeom &g, int tile, double *__restrict__ sl, double gamma, double *ex,
                                     int *err, double (*pj)[5][TJ * TK], double (*pk)[5][TJ * TK],
                                     const Ext *e = nullptr, double *__restrict__ q = nullptr)
{
    constexpr int TT = TJ * TK;
    const int bj = tile % g.nbj, bk = tile / g.nbj;
    const int l = threadIdx.x, tj = l % TJ, tk = l / TJ;
    const int lj = bj * TJ + tj, lk = bk * TK + tk;
    const bool active = lj < g.Lj && lk < g.Lk;
    const bool fuse_add = DIR < 0 && q != nullptr;
    const long qN = fuse_add ? e->n : 0;
    const long nat0 = fuse_add && active ? e->ix(e->is1, e->js1 + lj, e->ks1 + lk) : 0;
    // schedule coordinates (mirrored for the backward sweep)
    const int sj = DIR > 0 ? tj : TJ - 1 - tj, sk = DIR > 0 ? tk : TK - 1 - tk;
    // the j / k predecessor: outside the computed range (0), in the
    // neighbouring tile (exchange slot) or the previous lane (LDS)
    const bool j_ring = DIR > 0 ? lj == 0 : lj + 1 >= g.Lj;
    const bool k_ring = DIR > 0 ? lk == 0 : lk + 1 >= g.Lk;
    const bool j_in = !j_ring && sj == 0, k_in = !k_ring && sk == 0;
    const int lpj = l - DIR, lpk = l - DIR * TJ;
    // this lane's products leave the tile when it is last in schedule and a
    // neighbouring tile exists on that side
    const bool j_out = sj == TJ - 1 && (DIR > 0 ? bj + 1 < g.nbj : bj > 0);
    const bool k_out = sk == TK - 1 && (DIR > 0 ? bk + 1 < g.nbk : bk > 0);
    double *exj = ex, *exk = ex + (long)g.ntiles * g.exj_tile;
    // slots: j region [tile][d = pos + sk][5][tk], k region [tile][d = pos + sj][5][tj]:
    // the lanes of a tile face store and load one component of their slots
    // as one contiguous run (TK or TJ doubles)
    double *jin = exj + (long)(DIR > 0 ? tile - 1 : tile + 1) * g.exj_tile + tk;
    double *kin = exk + (long)(DIR > 0 ? tile - g.nbj : tile + g.nbj) * g.exk_tile + tj;
    double *jout = exj + (long)tile * g.exj_tile + tk;
    double *kout = exk + (long)tile * g.exk_tile + tj;
    bool failed = false;
    double ta[5] = {0, 0, 0, 0, 0};
    double opA[kSlF], opB[kSlF];
    unsigned long long xjA[5], xkA[5], xjB[5], xkB[5];

    auto row = [&](int s) { return DIR > 0 ? s : g.nsteps - 1 - s; };
    auto on_at = [&](int s) {
        const int pos = s - sj - sk;
        return active && pos >= 0 && pos < g.Li;
    };
    // Issue the loads of step s (operands and exchange values).  The loads
    // are unconditional -- lanes with nothing to do at step s read a valid
    // slot and ignore it -- so the compiler can count outstanding loads
    // exactly and the prefetch stays in flight across the previous step.
    const int xj_last = g.Li + TK - 2, xk_last = g.Li + TJ - 2;
    // In a tile with a neighbour on that side every lane loads from the
    // face lanes' slots of the step (so one load instruction touches one
    // contiguous run); otherwise from the start of the region.  Only the
    // face lanes use what they load.
    const bool j_pred = DIR > 0 ? bj > 0 : bj + 1 < g.nbj;
    const bool k_pred = DIR > 0 ? bk > 0 : bk + 1 < g.nbk;
    const double *jsrc = j_pred ? jin : ex;
    const double *ksrc = k_pred ? kin : ex;
    auto fetch = [&](int s, double *op, unsigned long long *xj, unsigned long long *xk) {
        const int sc = min(s, g.nsteps - 1);
        const double2 *src = reinterpret_cast<const double2 *>(sl + g.slot(tile, row(sc), 0, l));
#pragma unroll
        for (int f = 0; f < kSlF / 2; f++) {
            const double2 v = src[(long)f * TT];
            op[2 * f] = v.x;
            op[2 * f + 1] = v.y;
        }
        // a face lane (sj == 0 / sk == 0) takes slot pos + sk = s / pos + sj = s
        const int dj = j_pred ? min(max(s, 0), xj_last) : 0;
        const int dk = k_pred ? min(max(s, 0), xk_last) : 0;
#pragma unroll
        for (int x = 0; x < 5; x++) xj[x] = xload_bits(jsrc + (long)dj * TK * 5 + x * TK);
#pragma unroll
        for (int x = 0; x < 5; x++) xk[x] = xload_bits(ksrc + (long)dk * TJ * 5 + x * TJ);
    };
    auto step = [&](int s, const double *op, const unsigned long long *xj, const unsigned long long *xk) {
        const int p = s & 1;
        if (on_at(s)) {
            const int pos = s - sj - sk;
            // from the node's own operands, ahead of the waits on its
            // neighbours' values
            double rad[3];
            const double dg = lu_node_radii(op, rad);
            double tb[5], tc[5], res[5], pa[5], pb[5], pc[5];
            if (j_ring) {
#pragma unroll
                for (int x = 0; x < 5; x++) tb[x] = 0.0;
            } else if (j_in) {
                ex_take(jin + (long)(pos + sk) * TK * 5, TK, xj, tb, err, failed, !g.refill);
            } else {
#pragma unroll
                for (int x = 0; x < 5; x++) tb[x] = pj[p ^ 1][x][lpj];
            }
            if (k_ring) {
#pragma unroll
                for (int x = 0; x < 5; x++) tc[x] = 0.0;
            } else if (k_in) {
                ex_take(kin + (long)(pos + sj) * TJ * 5, TJ, xk, tc, err, failed, !g.refill);
            } else {
#pragma unroll
                for (int x = 0; x < 5; x++) tc[x] = pk[p ^ 1][x][lpk];
            }
            lu_node<DIR>(op, dg, rad, ta, tb, tc, gamma, res, pa, pb, pc);
            if (fuse_add) {
                const long n = nat0 + (g.Li - 1 - pos);
#pragma unroll
                for (int x = 0; x < 5; x++) q[x * qN + n] = op[kSlQ + x] + res[x];
            } else {
                // the backward sweep leaves q + dq (its q operand plus the
                // result, the update's one addition), the forward sweep dq*;
                // stored as whole 16-byte pairs, the first one with the
                // diagonal the lane read (its value unchanged)
                double o[5];
#pragma unroll
                for (int x = 0; x < 5; x++) o[x] = DIR < 0 ? op[kSlQ + x] + res[x] : res[x];
                double2 *dst = reinterpret_cast<double2 *>(sl + g.slot(tile, row(s), kSlT, l));
                dst[0] = make_double2(op[kSlT], o[0]);
                dst[TT] = make_double2(o[1], o[2]);
                dst[2 * TT] = make_double2(o[3], o[4]);
            }
#pragma unroll
            for (int x = 0; x < 5; x++) {
                ta[x] = pa[x];
                pj[p][x][l] = pb[x];
                pk[p][x][l] = pc[x];
            }
            if (j_out) {
#pragma unroll
                for (int x = 0; x < 5; x++) xstore(jout + (long)(pos + sk) * TK * 5 + x * TK, pb[x]);
            }
            if (k_out) {
#pragma unroll
                for (int x = 0; x < 5; x++) xstore(kout + (long)(pos + sj) * TJ * 5 + x * TJ, pc[x]);
            }
        }
        lds_barrier();
    };

    // operands are loaded one step ahead (two register sets in rotation):
    // measured faster than two steps ahead with three sets (256^3 block,
    // 16x8 tiles: 1.83 against 2.05 ms per forward sweep); steps past the
    // end are no-ops that keep the barriers uniform
    fetch(0, opA, xjA, xkA);
    for (int s = 0; s < g.nsteps; s += 2) {
        fetch(s + 1, opB, xjB, xkB);
        step(s, opA, xjA, xkA);
        fetch(s + 2, opA, xjA, xkA);
        step(s + 1, opB, xjB, xkB);
    }
}

template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_lu_tiles(SlGeom g, double *__restrict__ sl, double gamma,
                                                      double *ex, int *err, Ext e, double *q)
{
    __shared__ double pj[2][5][TJ * TK], pk[2][5][TJ * TK];
    const int tile = DIR > 0 ? (int)blockIdx.x : g.ntiles - 1 - (int)blockIdx.x;
    lu_tiles_body<DIR, TJ, TK>(g, tile, sl, gamma, ex, err, pj, pk, &e, q);
}

// ------------------------------------------------------------ batched tiles
// Every block of a context in one launch per phase.  A table holds one
// entry per block (ranges, tiling, operand fields, its part of the
// workspace) and the launch's flat tile numbers run over the blocks one
// after another, tile0 being a block's first.  Fields are offsets from the
// base of the blocks' arena and workspace parts offsets from the
// workspace's bases, all passed as kernel arguments: pointers derived from
// kernel arguments are known to be global memory, so the accesses compile
// to global loads and stores (a pointer read from a table would compile to
// flat ones, which the LDS barriers of the sweep would also wait for).
struct LuBatchEnt {
    Ext e;
    SlGeom g;
    long q, dxidx, inv_j, csnd, tterm, rhs;
    long sl, ex;
    int tile0;
};

template <int TT>
__global__ void __launch_bounds__(TT) k_sl_gather_batch(const LuBatchEnt *__restrict__ ents,
                                                        const int *__restrict__ tile_ent, double *arena,
                                                        double *slb, int fp0)
{
    extern __shared__ double buf[];
    const int flat = (int)blockIdx.x;
    const LuBatchEnt &b = ents[tile_ent[flat]];
    const Ext e = b.e;
    const SlGeom g = b.g;
    sl_gather_tile<TT>(e, g, flat - b.tile0, arena + b.q, arena + b.dxidx, arena + b.inv_j, arena + b.csnd,
                       arena + b.tterm, arena + b.rhs, slb + b.sl, fp0, buf);
}

template <int TT>
__global__ void __launch_bounds__(TT) k_sl_add_batch(const LuBatchEnt *__restrict__ ents,
                                                     const int *__restrict__ tile_ent, double *arena,
                                                     const double *slb)
{
    extern __shared__ double buf[];
    const int flat = (int)blockIdx.x;
    const LuBatchEnt &b = ents[tile_ent[flat]];
    const Ext e = b.e;
    const SlGeom g = b.g;
    sl_add_tile<TT>(e, g, flat - b.tile0, slb + b.sl, arena + b.q, buf);
}

// Flat tile numbers.  When every workgroup of the launch fits on the device
// at once (`resident`), tile = workgroup index: all tiles run together and
// every wait ends.  Otherwise the tiles are handed out in the order
// workgroups start (a ticket counter): a tile then only ever waits on tiles
// whose workgroups started before it, so the sweep progresses however the
// launch is dispatched; the last workgroup to finish resets both counters
// for the next launch.  (The first form is the faster one where it applies:
// the hardware's placement of consecutive workgroups suits the sweep.)
template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_lu_tiles_batch(const LuBatchEnt *__restrict__ ents,
                                                            const int *__restrict__ tile_ent, int total,
                                                            int resident, double *slb, double *exb,
                                                            double gamma, int *err, int *ctr, double *arena,
                                                            int fuse_add)
{
    __shared__ double pj[2][5][TJ * TK], pk[2][5][TJ * TK];
    __shared__ int ticket;
    if (!resident) {
        if (threadIdx.x == 0)
            ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    const int flat = resident ? (int)blockIdx.x : __builtin_amdgcn_readfirstlane(ticket);
    const LuBatchEnt &b = ents[tile_ent[flat]];
    const SlGeom g = b.g;
    const int local = flat - b.tile0;
    const Ext e = b.e;
    lu_tiles_body<DIR, TJ, TK>(g, DIR > 0 ? local : g.ntiles - 1 - local, slb + b.sl, gamma, exb + b.ex, err, pj,
                               pk, &e, fuse_add ? arena + b.q : nullptr);
    if (!resident && threadIdx.x == 0 &&
        __hip_atomic_fetch_add(ctr + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1) {
        __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ctr + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ------------------------------------------------ scalar sweeps (SA model)
// The LU-SGS of the SA working variable (sub_tur_sa.f90:755-812) has the
// same dependencies with one unknown per node:
//   forward   dn* = (rhs + a+ dn*_i-1 + b+ dn*_j-1 + c+ dn*_k-1) / diag
//   backward  dn  = dn* - (a- dn_i+1 + b- dn_j+1 + c- dn_k+1) / diag
// with the node's own a+- = (a +- |a|)/2 (a, b, c the contravariant
// velocities), so a neighbour only supplies its value.  Same tiles and
// skewed clock as the flow sweep; the operands (a, b, c, diag and rhs or
// dn*) are read in place, one i-line per lane (a cache line serves 16
// consecutive steps), and one value per exchange slot crosses a tile edge.
struct SaGeom {
    int TJ, TK, nbj, nbk, ntiles, nsteps, Li, Lj, Lk;
    long exj_tile, exk_tile;  // doubles per tile of the j / k exchange regions
};

static SaGeom sa_geom(const Ext &e, int shape)
{
    const SlGeom s = sl_geom(e, shape);
    SaGeom g;
    g.TJ = s.TJ;
    g.TK = s.TK;
    g.nbj = s.nbj;
    g.nbk = s.nbk;
    g.ntiles = s.ntiles;
    g.nsteps = s.nsteps;
    g.Li = s.Li;
    g.Lj = s.Lj;
    g.Lk = s.Lk;
    g.exj_tile = (long)(g.Li + g.TK - 1) * g.TK;
    g.exk_tile = (long)(g.Li + g.TJ - 1) * g.TJ;
    return g;
}

__device__ inline double ex_take1(double *slot, unsigned long long pre, int *err, bool &failed)
{
    unsigned long long b = pre;
    if (b == kSentinel) {
        for (long spins = 0; !failed; spins++) {
            b = xload_bits(slot);
            if (b != kSentinel) break;
            if ((spins & 1023) == 1023 &&
                (spins > (1L << 21) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                atomicOr(err, 1);
                failed = true;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (failed) b = 0;
    }
    xstore_bits(slot, kSentinel);
    return __longlong_as_double((long long)b);
}

// cu = a, b, c, diag (component stride N); fwd reads rhs and writes dns,
// bwd reads dns and writes dn
template <int DIR, int TJ, int TK>
__device__ inline void sa_tiles_body(const Ext &e, const SaGeom &g, int tile, const double *__restrict__ cu,
                                     const double *__restrict__ rhs, double *__restrict__ dns,
                                     double *__restrict__ dn, double *ex, int *err, double (*pj)[TJ * TK],
                                     double (*pk)[TJ * TK])
{
    const int bj = tile % g.nbj, bk = tile / g.nbj;
    const int l = threadIdx.x, tj = l % TJ, tk = l / TJ;
    const int lj = bj * TJ + tj, lk = bk * TK + tk;
    const bool active = lj < g.Lj && lk < g.Lk;
    const int sj = DIR > 0 ? tj : TJ - 1 - tj, sk = DIR > 0 ? tk : TK - 1 - tk;
    const bool j_ring = DIR > 0 ? lj == 0 : lj + 1 >= g.Lj;
    const bool k_ring = DIR > 0 ? lk == 0 : lk + 1 >= g.Lk;
    const bool j_in = !j_ring && sj == 0, k_in = !k_ring && sk == 0;
    const int lpj = l - DIR, lpk = l - DIR * TJ;
    const bool j_out = sj == TJ - 1 && (DIR > 0 ? bj + 1 < g.nbj : bj > 0);
    const bool k_out = sk == TK - 1 && (DIR > 0 ? bk + 1 < g.nbk : bk > 0);
    double *exj = ex, *exk = ex + (long)g.ntiles * g.exj_tile;
    double *jin = exj + (long)(DIR > 0 ? tile - 1 : tile + 1) * g.exj_tile + tk;
    double *kin = exk + (long)(DIR > 0 ? tile - g.nbj : tile + g.nbj) * g.exk_tile + tj;
    double *jout = exj + (long)tile * g.exj_tile + tk;
    double *kout = exk + (long)tile * g.exk_tile + tj;
    // every lane of a tile with a neighbour on that side loads the face
    // lanes' slot run of the step; only the face lanes use it
    const bool j_pred = DIR > 0 ? bj > 0 : bj + 1 < g.nbj;
    const bool k_pred = DIR > 0 ? bk > 0 : bk + 1 < g.nbk;
    const double *jsrc = j_pred ? jin : ex;
    const double *ksrc = k_pred ? kin : ex;
    const int xj_last = g.Li + TK - 2, xk_last = g.Li + TJ - 2;
    const long N = e.n;
    // first node of the lane's line in sweep order (inactive lanes read a
    // valid node and ignore it)
    const int cj = min(lj, g.Lj - 1), ck = min(lk, g.Lk - 1);
    const long n0 = e.ix(e.is1, e.js1 + cj, e.ks1 + ck);
    const double *src = DIR > 0 ? rhs : dns;
    double *out = DIR > 0 ? dns : dn;
    bool failed = false;
    double vi = 0.0;
    // operands of step s: a, b, c, diag, rhs / dn*, and the exchange bits
    auto fetch = [&](int s, double *op, unsigned long long &xj, unsigned long long &xk) {
        const int pos = min(max(s - sj - sk, 0), g.Li - 1);
        const long n = n0 + (DIR > 0 ? pos : g.Li - 1 - pos);
        op[0] = cu[n];
        op[1] = cu[N + n];
        op[2] = cu[2 * N + n];
        op[3] = cu[3 * N + n];
        op[4] = src[n];
        const int dj = j_pred ? min(max(s, 0), xj_last) : 0;
        const int dk = k_pred ? min(max(s, 0), xk_last) : 0;
        xj = xload_bits(jsrc + (long)dj * TK);
        xk = xload_bits(ksrc + (long)dk * TJ);
    };
    auto step = [&](int s, const double *op, unsigned long long xj, unsigned long long xk) {
        const int par = s & 1;
        const int pos = s - sj - sk;
        if (active && pos >= 0 && pos < g.Li) {
            const long n = n0 + (DIR > 0 ? pos : g.Li - 1 - pos);
            const double vj = j_ring ? 0.0 : j_in ? ex_take1(jin + (long)(pos + sk) * TK, xj, err, failed)
                                                  : pj[par ^ 1][lpj];
            const double vk = k_ring ? 0.0 : k_in ? ex_take1(kin + (long)(pos + sj) * TJ, xk, err, failed)
                                                  : pk[par ^ 1][lpk];
            double res;
            if (DIR > 0) {
                const double ta = 0.5 * (op[0] + fabs(op[0])) * vi;
                const double tb = 0.5 * (op[1] + fabs(op[1])) * vj;
                const double tc = 0.5 * (op[2] + fabs(op[2])) * vk;
                res = (op[4] + (ta + tb + tc)) / op[3];
            } else {
                const double ta = 0.5 * (op[0] - fabs(op[0])) * vi;
                const double tb = 0.5 * (op[1] - fabs(op[1])) * vj;
                const double tc = 0.5 * (op[2] - fabs(op[2])) * vk;
                res = op[4] - (ta + tb + tc) / op[3];
            }
            out[n] = res;
            vi = res;
            pj[par][l] = res;
            pk[par][l] = res;
            if (j_out) xstore(jout + (long)(pos + sk) * TK, res);
            if (k_out) xstore(kout + (long)(pos + sj) * TJ, res);
        }
        lds_barrier();
    };
    double opA[5], opB[5];
    unsigned long long xjA, xkA, xjB, xkB;
    fetch(0, opA, xjA, xkA);
    for (int s = 0; s < g.nsteps; s += 2) {
        fetch(s + 1, opB, xjB, xkB);
        step(s, opA, xjA, xkA);
        fetch(s + 2, opA, xjA, xkA);
        step(s + 1, opB, xjB, xkB);
    }
}

template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_sa_tiles(Ext e, SaGeom g, const double *__restrict__ cu,
                                                      const double *__restrict__ rhs, double *__restrict__ dns,
                                                      double *__restrict__ dn, double *ex, int *err)
{
    __shared__ double pj[2][TJ * TK], pk[2][TJ * TK];
    const int tile = DIR > 0 ? (int)blockIdx.x : g.ntiles - 1 - (int)blockIdx.x;
    sa_tiles_body<DIR, TJ, TK>(e, g, tile, cu, rhs, dns, dn, ex, err, pj, pk);
}

struct SaBatchEnt {
    Ext e;
    SaGeom g;
    long cu, rhs, dns, dn, ex;
    long sl;  // the block's part of the SA sweep layout (offset from its base)
    long nut;
    int tile0;
};

template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_sa_tiles_batch(const SaBatchEnt *__restrict__ ents,
                                                            const int *__restrict__ tile_ent, int total,
                                                            int resident, double *arena, double *exb, int *err,
                                                            int *ctr)
{
    __shared__ double pj[2][TJ * TK], pk[2][TJ * TK];
    __shared__ int ticket;
    if (!resident) {
        if (threadIdx.x == 0)
            ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    const int flat = resident ? (int)blockIdx.x : __builtin_amdgcn_readfirstlane(ticket);
    const SaBatchEnt &b = ents[tile_ent[flat]];
    const Ext e = b.e;
    const SaGeom g = b.g;
    const int local = flat - b.tile0;
    sa_tiles_body<DIR, TJ, TK>(e, g, DIR > 0 ? local : g.ntiles - 1 - local, arena + b.cu, arena + b.rhs,
                               arena + b.dns, arena + b.dn, exb + b.ex, err, pj, pk);
    if (!resident && threadIdx.x == 0 &&
        __hip_atomic_fetch_add(ctr + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1) {
        __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ctr + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- scalar sweeps over an SA sweep layout (every block in one launch)
// The operands a, b, c, diag and the right-hand side of every computed node
// are moved into a sweep layout ordered (tile, step, operand, lane): at
// step t, lane l of a tile reads its five operands as five rows of TT
// consecutive doubles, whole lines for the wavefront, where the natural
// layout has each lane on a line of its own.  Row t of lane (tj, tk) holds
// the node li = t - tj - tk of the lane's i-line; the backward sweep's step
// s reads row nsteps-1-s for every lane (its skewed clock runs the same
// rows in reverse).  The forward sweep leaves dnut* in the rhs operand's
// slot and the backward sweep dnut there; a last pass adds it to nut.
constexpr int kSaOps = 5;

__host__ __device__ inline long sa_slot(const SaGeom &g, int tt, int tile, int t, int f, int l)
{
    return (((long)tile * g.nsteps + t) * kSaOps + f) * tt + l;
}

// node of (tile, line, li) of a block and whether the line exists
__device__ inline long sa_nat(const Ext &e, const SaGeom &g, int tile, int line, int li)
{
    const int tj = line % g.TJ, tk = line / g.TJ;
    return e.ix(e.is1 + li, e.js1 + (tile % g.nbj) * g.TJ + tj, e.ks1 + (tile / g.nbj) * g.TK + tk);
}
__device__ inline bool sa_line_ok(const SaGeom &g, int tile, int line)
{
    return (tile % g.nbj) * g.TJ + line % g.TJ < g.Lj && (tile / g.nbj) * g.TK + line / g.TJ < g.Lk;
}

// Natural layout <-> SA sweep layout for one tile (one workgroup of TT
// threads), the walk of the flow sweep's layout passes: chunks of CT steps,
// a natural-side thread handling CT consecutive nodes of a line, an
// XOR-swizzled [CT][TT] LDS buffer in between.  TO_SL: cu (4 operands,
// component stride N) and rhs into the layout; otherwise nut += the layout's
// last operand.
template <int TT, bool TO_SL>
__device__ inline void sa_layout_tile(const Ext &e, const SaGeom &g, int tile, const double *__restrict__ cu,
                                      const double *__restrict__ rhs, double *__restrict__ nut, double *__restrict__ L,
                                      double *buf /* [CT][TT] */)
{
    constexpr int CT = TO_SL ? sl_ct_of(TT) : sa_ret_ct_of(TT), lstep = TT / CT;
    const int x = threadIdx.x;
    const int ii_n = x % CT;
    const int tj = x % g.TJ, tk = x / g.TJ;
    const bool me_ok = sa_line_ok(g, tile, x);
    const long N = e.n;
    for (int t0 = 0; t0 < g.nsteps; t0 += CT) {
        int nat[CT];  // node offsets fit 32 bits (blocks below 2^31 nodes)
#pragma unroll
        for (int m = 0; m < CT; m++) {
            const int line = m * lstep + x / CT;
            const int li = t0 + ii_n - line % g.TJ - line / g.TJ;
            nat[m] = (t0 + ii_n < g.nsteps && li >= 0 && li < g.Li && sa_line_ok(g, tile, line))
                         ? (int)sa_nat(e, g, tile, line, li) : -1;
        }
        if (TO_SL) {
            for (int f = 0; f < kSaOps; f++) {
                const double *src = f < 4 ? cu + f * N : rhs;
                double v[CT];
#pragma unroll
                for (int m = 0; m < CT; m++) v[m] = nat[m] >= 0 ? src[nat[m]] : 0.0;
#pragma unroll
                for (int m = 0; m < CT; m++) {
                    const int line = m * lstep + x / CT;
                    buf[ii_n * TT + (line ^ ii_n)] = v[m];
                }
                __syncthreads();
                if (me_ok) {
#pragma unroll
                    for (int ii = 0; ii < CT; ii++) {
                        const int t = t0 + ii, li = t - tj - tk;
                        if (t < g.nsteps && li >= 0 && li < g.Li) L[sa_slot(g, TT, tile, t, f, x)] = buf[ii * TT + (x ^ ii)];
                    }
                }
                __syncthreads();
            }
        } else {
            if (me_ok) {
#pragma unroll
                for (int ii = 0; ii < CT; ii++) {
                    const int t = t0 + ii, li = t - tj - tk;
                    buf[ii * TT + (x ^ ii)] =
                        (t < g.nsteps && li >= 0 && li < g.Li) ? L[sa_slot(g, TT, tile, t, kSaOps - 1, x)] : 0.0;
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < CT; m++) {
                const int line = m * lstep + x / CT;
                if (nat[m] >= 0) nut[nat[m]] = nut[nat[m]] + buf[ii_n * TT + (line ^ ii_n)];
            }
            __syncthreads();
        }
    }
}

template <int TT, bool TO_SL>
__global__ void __launch_bounds__(TT) k_sa_layout_batch(const SaBatchEnt *__restrict__ ents,
                                                        const int *__restrict__ tile_ent, double *arena,
                                                        double *slbase)
{
    __shared__ double buf[(TO_SL ? sl_ct_of(TT) : sa_ret_ct_of(TT)) * TT];
    const int flat = (int)xcd_slab(blockIdx.x, gridDim.x);
    const SaBatchEnt &b = ents[tile_ent[flat]];
    const Ext e = b.e;
    const SaGeom g = b.g;
    sa_layout_tile<TT, TO_SL>(e, g, flat - b.tile0, arena + b.cu, arena + b.rhs, arena + b.nut, slbase + b.sl, buf);
}

// The scalar sweep of sa_tiles_body with its operands in the SA sweep
// layout (the same expressions, the same exchange and skewed clock).
template <int DIR, int TJ, int TK>
__device__ inline void sa_sl_body(const SaGeom &g, int tile, double *__restrict__ L, double *ex, int *err,
                                  double (*pj)[TJ * TK], double (*pk)[TJ * TK])
{
    constexpr int TT = TJ * TK;
    const int bj = tile % g.nbj, bk = tile / g.nbj;
    const int l = threadIdx.x, tj = l % TJ, tk = l / TJ;
    const int lj = bj * TJ + tj, lk = bk * TK + tk;
    const bool active = lj < g.Lj && lk < g.Lk;
    const int sj = DIR > 0 ? tj : TJ - 1 - tj, sk = DIR > 0 ? tk : TK - 1 - tk;
    const bool j_ring = DIR > 0 ? lj == 0 : lj + 1 >= g.Lj;
    const bool k_ring = DIR > 0 ? lk == 0 : lk + 1 >= g.Lk;
    const bool j_in = !j_ring && sj == 0, k_in = !k_ring && sk == 0;
    const int lpj = l - DIR, lpk = l - DIR * TJ;
    const bool j_out = sj == TJ - 1 && (DIR > 0 ? bj + 1 < g.nbj : bj > 0);
    const bool k_out = sk == TK - 1 && (DIR > 0 ? bk + 1 < g.nbk : bk > 0);
    double *exj = ex, *exk = ex + (long)g.ntiles * g.exj_tile;
    double *jin = exj + (long)(DIR > 0 ? tile - 1 : tile + 1) * g.exj_tile + tk;
    double *kin = exk + (long)(DIR > 0 ? tile - g.nbj : tile + g.nbj) * g.exk_tile + tj;
    double *jout = exj + (long)tile * g.exj_tile + tk;
    double *kout = exk + (long)tile * g.exk_tile + tj;
    // every lane of a tile with a neighbour on that side loads the face
    // lanes' slot run of the step (one contiguous run per load); only the
    // face lanes use it
    const bool j_pred = DIR > 0 ? bj > 0 : bj + 1 < g.nbj;
    const bool k_pred = DIR > 0 ? bk > 0 : bk + 1 < g.nbk;
    const double *jsrc = j_pred ? jin : ex;
    const double *ksrc = k_pred ? kin : ex;
    const int xj_last = g.Li + TK - 2, xk_last = g.Li + TJ - 2;
    bool failed = false;
    double vi = 0.0;
    auto row_of = [&](int s) -> long {
        const int sc = min(s, g.nsteps - 1);
        return ((long)tile * g.nsteps + (DIR > 0 ? sc : g.nsteps - 1 - sc)) * kSaOps * TT + l;
    };
    auto fetch = [&](int s, double *op, unsigned long long &xj, unsigned long long &xk) {
        const double *p = L + row_of(s);
#pragma unroll
        for (int f = 0; f < kSaOps; f++) op[f] = p[f * TT];
        const int dj = j_pred ? min(max(s, 0), xj_last) : 0;
        const int dk = k_pred ? min(max(s, 0), xk_last) : 0;
        xj = xload_bits(jsrc + (long)dj * TK);
        xk = xload_bits(ksrc + (long)dk * TJ);
    };
    auto step = [&](int s, const double *op, unsigned long long xj, unsigned long long xk) {
        const int par = s & 1;
        const int pos = s - sj - sk;
        if (active && pos >= 0 && pos < g.Li) {
            const double vj = j_ring ? 0.0 : j_in ? ex_take1(jin + (long)(pos + sk) * TK, xj, err, failed)
                                                  : pj[par ^ 1][lpj];
            const double vk = k_ring ? 0.0 : k_in ? ex_take1(kin + (long)(pos + sj) * TJ, xk, err, failed)
                                                  : pk[par ^ 1][lpk];
            double res;
            if (DIR > 0) {
                const double ta = 0.5 * (op[0] + fabs(op[0])) * vi;
                const double tb = 0.5 * (op[1] + fabs(op[1])) * vj;
                const double tc = 0.5 * (op[2] + fabs(op[2])) * vk;
                res = (op[4] + (ta + tb + tc)) / op[3];
            } else {
                const double ta = 0.5 * (op[0] - fabs(op[0])) * vi;
                const double tb = 0.5 * (op[1] - fabs(op[1])) * vj;
                const double tc = 0.5 * (op[2] - fabs(op[2])) * vk;
                res = op[4] - (ta + tb + tc) / op[3];
            }
            L[row_of(s) + (kSaOps - 1) * TT] = res;
            vi = res;
            pj[par][l] = res;
            pk[par][l] = res;
            if (j_out) xstore(jout + (long)(pos + sk) * TK, res);
            if (k_out) xstore(kout + (long)(pos + sj) * TJ, res);
        }
        lds_barrier();
    };
    double opA[kSaOps], opB[kSaOps];
    unsigned long long xjA, xkA, xjB, xkB;
    fetch(0, opA, xjA, xkA);
    for (int s = 0; s < g.nsteps; s += 2) {
        fetch(s + 1, opB, xjB, xkB);
        step(s, opA, xjA, xkA);
        fetch(s + 2, opA, xjA, xkA);
        step(s + 1, opB, xjB, xkB);
    }
}

template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_sa_sl_batch(const SaBatchEnt *__restrict__ ents,
                                                         const int *__restrict__ tile_ent, int total, int resident,
                                                         double *slbase, double *exb, int *err, int *ctr)
{
    __shared__ double pj[2][TJ * TK], pk[2][TJ * TK];
    __shared__ int ticket;
    if (!resident) {
        if (threadIdx.x == 0)
            ticket = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    const int flat = resident ? (int)blockIdx.x : __builtin_amdgcn_readfirstlane(ticket);
    const SaBatchEnt &b = ents[tile_ent[flat]];
    const SaGeom g = b.g;
    const int local = flat - b.tile0;
    sa_sl_body<DIR, TJ, TK>(g, DIR > 0 ? local : g.ntiles - 1 - local, slbase + b.sl, exb + b.ex, err, pj, pk);
    if (!resident && threadIdx.x == 0 &&
        __hip_atomic_fetch_add(ctr + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1) {
        __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ctr + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

int lusgs_tile_shapes() { return kNumTileShapes; }

long lusgs_tiles(const Ext &e, int shape)
{
    SlGeom g = sl_geom(e, shape);
    return g.Li > 0 && g.Lj > 0 && g.Lk > 0 ? g.ntiles : 0;
}

void lusgs_ws_sizes(const Ext &e, long *sl_n, long *ex_n)
{
    long a = 0, b = 0;
    for (int s = 0; s < kNumTileShapes; s++) {
        SlGeom g = sl_geom(e, s);
        if (g.Li <= 0 || g.Lj <= 0 || g.Lk <= 0) continue;
        a = std::max(a, (long)g.ntiles * g.nsteps * kSlF * g.TT);
        b = std::max(b, (long)g.ntiles * (g.exj_tile + g.exk_tile));
    }
    *sl_n = a;
    *ex_n = b;
}

// Whether the backward sweep adds dq to q itself (and the add pass is
// skipped).  Its stores then go to the natural layout, one line per lane.
// On one block's launch that is free up to a few million computed nodes
// (128^3: 0.644 against 0.642 ms per sweep, 0.1 ms of add pass saved) and
// dearer on larger ones (256^3: +1.0 ms in the sweep against 0.64 ms
// saved).  In the batched sweeps it costs more than the pass it saves
// (flat plate, 12 blocks: +0.25 ms against 0.10 ms; 128^3 in 64 blocks:
// +0.16 ms against 0.12 ms), so batched launches never fuse.
// ACFD_LU_FUSED_ADD=0/1 forces either form.
constexpr long kFuseAddMaxNodes = 1L << 22;
static bool fused_add(long launch_nodes, bool batched = false)
{
    const char *v = std::getenv("ACFD_LU_FUSED_ADD");
    if (v) return std::atoi(v) != 0;
    return !batched && launch_nodes < kFuseAddMaxNodes;
}
static long computed_nodes(const Ext &e)
{
    return (long)(e.ie1 - e.is1 + 1) * (e.je1 - e.js1 + 1) * (e.ke1 - e.ks1 + 1);
}

static long round_up32(long n) { return (n + 31) / 32 * 32; }
static void batch_residency(LuBatch *b);

// blocks of the batch in launch order: longest sweep first
static std::vector<int> batch_order(int nblk, const Ext *e, int shape)
{
    std::vector<int> order;
    for (int m = 0; m < nblk; m++)
        if (lusgs_tiles(e[m], shape) > 0) order.push_back(m);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const SlGeom ga = sl_geom(e[a], shape), gb = sl_geom(e[b], shape);
        return ga.nsteps + ga.nbj + ga.nbk > gb.nsteps + gb.nbj + gb.nbk;
    });
    return order;
}

void lusgs_batch_ws(int nblk, const Ext *e, int shape, long *sl_n, long *ex_n)
{
    long a = 0, b = 0;
    for (int m : batch_order(nblk, e, shape)) {
        const SlGeom g = sl_geom(e[m], shape);
        a += round_up32((long)g.ntiles * g.nsteps * kSlF * g.TT);
        b += round_up32((long)g.ntiles * (g.exj_tile + g.exk_tile));
    }
    *sl_n = a;
    *ex_n = b;
}

int lusgs_batch_make(int nblk, const Ext *e, const Fields *f, double *arena, int shape, double *sl, double *ex,
                     LuBatch *out)
{
    std::vector<LuBatchEnt> ents;
    std::vector<int> te;
    long so = 0, xo = 0;
    for (int m : batch_order(nblk, e, shape)) {
        const SlGeom g = sl_geom(e[m], shape);
        LuBatchEnt b;
        b.e = e[m];
        b.g = g;
        b.q = f[m].q - arena;
        b.dxidx = f[m].dxidx - arena;
        b.inv_j = f[m].inv_j - arena;
        b.csnd = f[m].pri + 6 * e[m].n - arena;
        b.tterm = f[m].w_diag - arena;
        b.rhs = f[m].rhs - arena;
        b.sl = so;
        b.ex = xo;
        b.tile0 = (int)te.size();
        so += round_up32((long)g.ntiles * g.nsteps * kSlF * g.TT);
        xo += round_up32((long)g.ntiles * (g.exj_tile + g.exk_tile));
        for (int t = 0; t < g.ntiles; t++) te.push_back((int)ents.size());
        ents.push_back(b);
    }
    out->shape = shape;
    out->arena = arena;
    out->sl = sl;
    out->ex = ex;
    out->nent = (int)ents.size();
    out->total = (int)te.size();
    out->ex_n = xo;
    out->refill = ex_refill();
    if (ents.empty()) return 0;
    batch_residency(out);
    long total = 0;
    for (const LuBatchEnt &b : ents) total += computed_nodes(b.e);
    out->fuse_add = fused_add(total, true) ? 1 : 0;
    if (hipMalloc(&out->ents, ents.size() * sizeof(LuBatchEnt)) != hipSuccess) return 1;
    if (hipMalloc((void **)&out->tile_ent, te.size() * sizeof(int)) != hipSuccess) return 1;
    if (hipMemcpy(out->ents, ents.data(), ents.size() * sizeof(LuBatchEnt), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(out->tile_ent, te.data(), te.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return 1;
    return 0;
}

void lusgs_batch_free(LuBatch *b)
{
    if (b->ents) (void)hipFree(b->ents);
    if (b->tile_ent) (void)hipFree(b->tile_ent);
    b->ents = nullptr;
    b->tile_ent = nullptr;
    b->nent = b->total = 0;
}

// whether all `total` workgroups of the sweep kernel fit on the device at once
template <int DIR, int TJ, int TK>
static bool fits_resident(int total)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lu_tiles_batch<DIR, TJ, TK>, TJ * TK, 0) !=
            hipSuccess)
        return false;
    return (long)cus * per_cu >= total;
}

template <int DIR, int TJ, int TK>
static void launch_tiles_batch(hipStream_t s, const LuBatch &b, double gamma, int *err, int *ctr)
{
    hipLaunchKernelGGL((k_lu_tiles_batch<DIR, TJ, TK>), dim3((unsigned)b.total), dim3(TJ * TK), 0, s,
                       (const LuBatchEnt *)b.ents, b.tile_ent, b.total, b.resident[DIR > 0 ? 0 : 1], b.sl, b.ex,
                       gamma, err, ctr, b.arena, DIR < 0 && b.fuse_add ? 1 : 0);
}

static void batch_residency(LuBatch *b)
{
    bool f = false, r = false;
    switch (b->shape) {
    case 0: f = fits_resident<1, 16, 16>(b->total); r = fits_resident<-1, 16, 16>(b->total); break;
    case 1: f = fits_resident<1, 16, 8>(b->total); r = fits_resident<-1, 16, 8>(b->total); break;
    case 2: f = fits_resident<1, 8, 8>(b->total); r = fits_resident<-1, 8, 8>(b->total); break;
    default: f = fits_resident<1, 16, 4>(b->total); r = fits_resident<-1, 16, 4>(b->total); break;
    }
    // ACFD_LU_TICKETS=1 hands tiles out by ticket even when all would fit
    const char *t = std::getenv("ACFD_LU_TICKETS");
    const bool tickets = t && std::atoi(t) == 1;
    b->resident[0] = f && !tickets ? 1 : 0;
    b->resident[1] = r && !tickets ? 1 : 0;
}

template <int DIR>
static void launch_tiles_batch_shape(hipStream_t s, const LuBatch &b, double gamma, int *err, int *ctr)
{
    switch (b.shape) {
    case 0: launch_tiles_batch<DIR, 16, 16>(s, b, gamma, err, ctr); break;
    case 1: launch_tiles_batch<DIR, 16, 8>(s, b, gamma, err, ctr); break;
    case 2: launch_tiles_batch<DIR, 8, 8>(s, b, gamma, err, ctr); break;
    default: launch_tiles_batch<DIR, 16, 4>(s, b, gamma, err, ctr); break;
    }
}

template <int TT>
static void launch_layout_batch(hipStream_t s, int which, const LuBatch &b, int fp0)
{
    const size_t lds = sl_lds_bytes(TT);
    const LuBatchEnt *ents = (const LuBatchEnt *)b.ents;
    if (which == 0)
        hipLaunchKernelGGL(k_sl_gather_batch<TT>, dim3((unsigned)b.total), dim3(TT), lds, s, ents, b.tile_ent,
                           b.arena, b.sl, fp0);
    else
        hipLaunchKernelGGL(k_sl_add_batch<TT>, dim3((unsigned)b.total), dim3(TT), lds, s, ents, b.tile_ent,
                           b.arena, (const double *)b.sl);
}

// Phases 1-4 of the LU-SGS update for every block of the batch at once
// (phase 0, the per-node preparation, stays per block).  geo_cached: the
// sweep layout already holds this batch's grid operands (phase 1 then
// gathers the other fields only).
int launch_lusgs_batch(hipStream_t s, const LuBatch &b, int phase, double gamma, int *err, int *ctr, int geo_cached)
{
    if (b.total <= 0 || phase < 1 || phase > 4) return 0;
    const int tt = kTileShapes[b.shape][0] * kTileShapes[b.shape][1];
    if (phase == 2 || phase == 3) {
        if (phase == 2) launch_tiles_batch_shape<1>(s, b, gamma, err, ctr);
        else launch_tiles_batch_shape<-1>(s, b, gamma, err, ctr);
        if (b.refill)
            hipLaunchKernelGGL(k_fill_sentinel, dim3((unsigned)((b.ex_n + 255) / 256)), dim3(256), 0, s, b.ex,
                               b.ex_n);
    } else if (phase == 4 && b.fuse_add) return 0;
    else {
        const int which = phase == 1 ? 0 : 1, fp0 = geo_cached ? kSlGeoPairs : 0;
        if (tt == 256) launch_layout_batch<256>(s, which, b, fp0);
        else if (tt == 128) launch_layout_batch<128>(s, which, b, fp0);
        else launch_layout_batch<64>(s, which, b, fp0);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ---- scalar sweeps: host side
template <int DIR, int TJ, int TK>
static bool sa_fits_resident(int total)
{
    int dev = 0, cus = 0, per_cu = 0, per_cu_sl = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sa_tiles_batch<DIR, TJ, TK>, TJ * TK, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_sl, k_sa_sl_batch<DIR, TJ, TK>, TJ * TK, 0) !=
            hipSuccess)
        return false;
    return (long)cus * std::min(per_cu, per_cu_sl) >= total;
}

template <int TJ, int TK>
static void sa_sweeps_t(hipStream_t s, const Ext &e, const SaGeom &g, const double *cu, const double *rhs,
                        double *dns, double *dn, double *ex, int *err)
{
    hipLaunchKernelGGL((k_sa_tiles<1, TJ, TK>), dim3((unsigned)g.ntiles), dim3(TJ * TK), 0, s, e, g, cu, rhs, dns,
                       dn, ex, err);
    hipLaunchKernelGGL((k_sa_tiles<-1, TJ, TK>), dim3((unsigned)g.ntiles), dim3(TJ * TK), 0, s, e, g, cu, rhs,
                       dns, dn, ex, err);
}

// both SA sweeps of one block (ex: sentinel-filled, at least
// ntiles * (exj_tile + exk_tile) doubles -- less than the flow sweep's)
int launch_sa_sweeps(hipStream_t s, const Ext &e, int shape, const double *cu, const double *rhs, double *dns,
                     double *dn, double *ex, int *err)
{
    const SaGeom g = sa_geom(e, shape);
    if (g.Li <= 0 || g.Lj <= 0 || g.Lk <= 0) return 0;
    switch (shape) {
    case 0: sa_sweeps_t<16, 16>(s, e, g, cu, rhs, dns, dn, ex, err); break;
    case 1: sa_sweeps_t<16, 8>(s, e, g, cu, rhs, dns, dn, ex, err); break;
    case 2: sa_sweeps_t<8, 8>(s, e, g, cu, rhs, dns, dn, ex, err); break;
    default: sa_sweeps_t<16, 4>(s, e, g, cu, rhs, dns, dn, ex, err); break;
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

void sa_batch_sl(int nblk, const Ext *e, int shape, long *sl_n)
{
    long so = 0;
    for (int m = 0; m < nblk; m++) {
        const SaGeom g = sa_geom(e[m], shape);
        if (g.Li <= 0 || g.Lj <= 0 || g.Lk <= 0) continue;
        so += round_up32((long)g.ntiles * g.nsteps * kSaOps * g.TJ * g.TK);
    }
    *sl_n = so;
}

int sa_batch_make(int nblk, const Ext *e, const Fields *f, double *arena, int shape, double *ex, double *sl,
                  LuBatch *out)
{
    std::vector<SaBatchEnt> ents;
    std::vector<int> te;
    long xo = 0, so = 0;
    for (int m : batch_order(nblk, e, shape)) {
        const SaGeom g = sa_geom(e[m], shape);
        if (!f[m].w_sa) continue;
        SaBatchEnt b;
        b.e = e[m];
        b.g = g;
        b.cu = f[m].w_dqs - arena;
        b.rhs = f[m].sa_rhs - arena;
        b.dns = f[m].w_sa + 8 * e[m].n - arena;
        b.dn = f[m].w_dq - arena;
        b.nut = f[m].nut - arena;
        b.ex = xo;
        b.sl = so;
        b.tile0 = (int)te.size();
        xo += round_up32((long)g.ntiles * (g.exj_tile + g.exk_tile));
        so += round_up32((long)g.ntiles * g.nsteps * kSaOps * g.TJ * g.TK);
        for (int t = 0; t < g.ntiles; t++) te.push_back((int)ents.size());
        ents.push_back(b);
    }
    out->shape = shape;
    out->arena = arena;
    out->ex = ex;
    out->sl = sl;
    out->nent = (int)ents.size();
    out->total = (int)te.size();
    if (ents.empty()) return 0;
    const char *t = std::getenv("ACFD_LU_TICKETS");
    const bool tickets = t && std::atoi(t) == 1;
    bool fwd = false, bwd = false;
    switch (shape) {
    case 0: fwd = sa_fits_resident<1, 16, 16>(out->total); bwd = sa_fits_resident<-1, 16, 16>(out->total); break;
    case 1: fwd = sa_fits_resident<1, 16, 8>(out->total); bwd = sa_fits_resident<-1, 16, 8>(out->total); break;
    case 2: fwd = sa_fits_resident<1, 8, 8>(out->total); bwd = sa_fits_resident<-1, 8, 8>(out->total); break;
    default: fwd = sa_fits_resident<1, 16, 4>(out->total); bwd = sa_fits_resident<-1, 16, 4>(out->total); break;
    }
    out->resident[0] = fwd && !tickets ? 1 : 0;
    out->resident[1] = bwd && !tickets ? 1 : 0;
    if (hipMalloc(&out->ents, ents.size() * sizeof(SaBatchEnt)) != hipSuccess) return 1;
    if (hipMalloc((void **)&out->tile_ent, te.size() * sizeof(int)) != hipSuccess) return 1;
    if (hipMemcpy(out->ents, ents.data(), ents.size() * sizeof(SaBatchEnt), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(out->tile_ent, te.data(), te.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return 1;
    return 0;
}

template <int TJ, int TK>
static void sa_batch_t(hipStream_t s, const LuBatch &b, int *err, int *ctr, int dir)
{
    const SaBatchEnt *ents = (const SaBatchEnt *)b.ents;
    if (dir >= 0)
        hipLaunchKernelGGL((k_sa_tiles_batch<1, TJ, TK>), dim3((unsigned)b.total), dim3(TJ * TK), 0, s, ents,
                           b.tile_ent, b.total, b.resident[0], b.arena, b.ex, err, ctr);
    if (dir <= 0)
        hipLaunchKernelGGL((k_sa_tiles_batch<-1, TJ, TK>), dim3((unsigned)b.total), dim3(TJ * TK), 0, s, ents,
                           b.tile_ent, b.total, b.resident[1], b.arena, b.ex, err, ctr);
}

template <int TJ, int TK>
static void sa_sl_batch_t(hipStream_t s, const LuBatch &b, int *err, int *ctr, int dir)
{
    const SaBatchEnt *ents = (const SaBatchEnt *)b.ents;
    if (dir >= 0)
        hipLaunchKernelGGL((k_sa_sl_batch<1, TJ, TK>), dim3((unsigned)b.total), dim3(TJ * TK), 0, s, ents,
                           b.tile_ent, b.total, b.resident[0], b.sl, b.ex, err, ctr);
    if (dir <= 0)
        hipLaunchKernelGGL((k_sa_sl_batch<-1, TJ, TK>), dim3((unsigned)b.total), dim3(TJ * TK), 0, s, ents,
                           b.tile_ent, b.total, b.resident[1], b.sl, b.ex, err, ctr);
}

template <int TT>
static void sa_layout_batch_t(hipStream_t s, const LuBatch &b, bool to_sl)
{
    const SaBatchEnt *ents = (const SaBatchEnt *)b.ents;
    if (to_sl)
        hipLaunchKernelGGL((k_sa_layout_batch<TT, true>), dim3((unsigned)b.total), dim3(TT), 0, s, ents, b.tile_ent,
                           b.arena, b.sl);
    else
        hipLaunchKernelGGL((k_sa_layout_batch<TT, false>), dim3((unsigned)b.total), dim3(TT), 0, s, ents,
                           b.tile_ent, b.arena, b.sl);
}

// SA sweeps over the SA sweep layout: phase 0 moves a, b, c, diag and rhs
// into the layout, 1 / -1 run the forward / backward sweep, 2 adds dnut
// to nut
int launch_sa_sl_batch(hipStream_t s, const LuBatch &b, int *err, int *ctr, int phase)
{
    if (b.total <= 0 || !b.sl) return 0;
    const int tt = kTileShapes[b.shape][0] * kTileShapes[b.shape][1];
    if (phase == 0 || phase == 2) {
        if (tt == 256) sa_layout_batch_t<256>(s, b, phase == 0);
        else if (tt == 128) sa_layout_batch_t<128>(s, b, phase == 0);
        else sa_layout_batch_t<64>(s, b, phase == 0);
    } else {
        switch (b.shape) {
        case 0: sa_sl_batch_t<16, 16>(s, b, err, ctr, phase); break;
        case 1: sa_sl_batch_t<16, 8>(s, b, err, ctr, phase); break;
        case 2: sa_sl_batch_t<8, 8>(s, b, err, ctr, phase); break;
        default: sa_sl_batch_t<16, 4>(s, b, err, ctr, phase); break;
        }
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// the SA sweeps of every block of the batch (dir 0: both, in order)
int launch_sa_batch(hipStream_t s, const LuBatch &b, int *err, int *ctr, int dir)
{
    if (b.total <= 0) return 0;
    switch (b.shape) {
    case 0: sa_batch_t<16, 16>(s, b, err, ctr, dir); break;
    case 1: sa_batch_t<16, 8>(s, b, err, ctr, dir); break;
    case 2: sa_batch_t<8, 8>(s, b, err, ctr, dir); break;
    default: sa_batch_t<16, 4>(s, b, err, ctr, dir); break;
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_fill_sentinel(hipStream_t s, double *p, long n)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_sentinel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    ACFD_CHECK_LAUNCH();
    return 0;
}

template <int DIR, int TJ, int TK>
static void launch_tiles(hipStream_t s, const SlGeom &g, const LuWs &ws, double gamma, const Ext &e, double *q)
{
    hipLaunchKernelGGL((k_lu_tiles<DIR, TJ, TK>), dim3((unsigned)g.ntiles), dim3(TJ * TK), 0, s, g, ws.sl, gamma,
                       ws.ex, ws.err, e, q);
}

template <int DIR>
static void launch_tiles_shape(hipStream_t s, int shape, const SlGeom &g, const LuWs &ws, double gamma,
                               const Ext &e, double *q)
{
    switch (shape) {
    case 0: launch_tiles<DIR, 16, 16>(s, g, ws, gamma, e, q); break;
    case 1: launch_tiles<DIR, 16, 8>(s, g, ws, gamma, e, q); break;
    case 2: launch_tiles<DIR, 8, 8>(s, g, ws, gamma, e, q); break;
    default: launch_tiles<DIR, 16, 4>(s, g, ws, gamma, e, q); break;
    }
}

// Phases of one LU-SGS update:
//   0 rhs and, for hyperplanes, spectral radii and diag (and dq* and dq
//     cleared), for tiles the time term phi/dt,
//   1 tiles: gather the operands into the sweep layout (with geo_cached,
//     which says that it already holds this block's grid operands in this
//     tiling, all but those),
//   2 forward sweep, 3 backward sweep,
//   4 q += dq (tiles: from the sweep layout, or fused into the backward
//     sweep, which then stores q + dq itself).
// variant 0 = hyperplanes, 1 = tiles of the given shape index.
int launch_lusgs_phase(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                       int variant, const LuWs *ws, int phase, int shape, int geo_cached)
{
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    const bool tiles = variant == 1 && ws && ws->sl && ws->ex;
    const bool empty = Li <= 0 || Lj <= 0 || Lk <= 0;
    if (phase == 0) {
        hipLaunchKernelGGL(k_lu_prep, lu_prep_grid(e), dim3(64, 4, 1), 0, s, e, f.pri, f.dxidx, f.dt, f.rhsi,
                           f.rhsv, f.rhs, f.w_r, f.w_diag, f.w_dqs, f.w_dq, cfl, tiles ? 1 : 0);
    } else if (empty) {
        return 0;
    } else if (tiles) {
        const SlGeom g = sl_geom(e, shape);
        if (phase == 1)
            launch_layout_tt(s, 0, e, g, f, *ws, geo_cached ? kSlGeoPairs : 0);
        else if (phase == 2 || phase == 3) {
            if (phase == 2) launch_tiles_shape<1>(s, shape, g, *ws, c.gamma, e, nullptr);
            else launch_tiles_shape<-1>(s, shape, g, *ws, c.gamma, e, fused_add(computed_nodes(e)) ? f.q : nullptr);
            if (g.refill) {
                const long n = (long)g.ntiles * (g.exj_tile + g.exk_tile);
                hipLaunchKernelGGL(k_fill_sentinel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ws->ex, n);
            }
        } else if (!fused_add(computed_nodes(e)))
            launch_layout_tt(s, 1, e, g, f, *ws);
    } else if (phase == 4) {
        Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
        hipLaunchKernelGGL(k_add5, box_grid(b), dim3(64, 4, 1), 0, s, e, b, f.q, f.w_dq);
    } else if (phase == 2 || phase == 3) {
        LuArgs a{e, f.q, f.dxidx, f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma};
        const bool fwd = phase == 2;
        int L = Li + Lj + Lk - 2;
        for (int l = 0; l < L; l++) {
            int ja = max(0, l - (Li - 1) - (Lk - 1)), jb = min(Lj - 1, l);
            int ka = max(0, l - (Li - 1) - (Lj - 1)), kb = min(Lk - 1, l);
            dim3 gr((jb - ja + 64) / 64, (kb - ka + 4) / 4, 1);
            if (fwd) hipLaunchKernelGGL((k_lu_plane<1>), gr, dim3(64, 4, 1), 0, s, a, l, ja, ka);
            else hipLaunchKernelGGL((k_lu_plane<-1>), gr, dim3(64, 4, 1), 0, s, a, l, ja, ka);
        }
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_lusgs(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                 int variant, const LuWs *ws, int shape)
{
    for (int ph = 0; ph < kLusgsPhases; ph++) {
        int rc = launch_lusgs_phase(s, e, f, c, cfl, variant, ws, ph, shape);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace acfd
