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
//   hyperplanes: one launch per plane l (2 (Li+Lj+Lk-2) launches per step);
//   tiles:       one persistent launch per sweep.  Each workgroup owns a
//                TJ x TK tile of (j,k) lines and marches along i on a skewed
//                clock (thread (tj,tk) handles i = t - tj - tk at step t):
//                the i neighbour stays in registers, the j and k neighbours
//                inside the tile come through LDS, and the tile's first
//                row/column read the neighbouring tile's output from global
//                memory once that tile has published its step counter.
//                dq*, dq and the counters live in uncached device memory, so
//                a producer only has to wait for its stores to complete
//                before publishing, and consumers read memory directly; no
//                cache writeback or invalidation is needed across XCDs.
//
// The flux Jacobian products use the closed form of A(q) v for the Euler
// flux in generalised coordinates (one reciprocal of rho instead of the ~25
// divisions of forming calc_jm's matrix); the values agree with the matrix
// form to rounding.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

// No implicit contraction in this file: the sweeps' arithmetic is spelled
// out (explicit fma) so that every schedule produces the same bits.
#pragma clang fp contract(off)

namespace acfd {

// 0.5 * (A(q_n; m) v +- r_n v) * inv_j_n, with m the metric row c0 of node n
// (sub_timeadvancement_1stimplicit.f90:283-301 forward, :459-477 backward).
// A v = [ m.v_mom,
//         U v_mom + u theta + m dp,
//         a1 theta + U (dp + v5) ]
// with u = mom/rho, U = m.u, theta = m.v_mom - U v1,
// dp = (gamma-1)(|u|^2/2 v1 - u.v_mom + v5), a1 = gamma E/rho - (gamma-1)|u|^2/2.
template <int SGN>
__device__ inline void jac_mv(const double *__restrict__ q, const double *__restrict__ dxidx,
                              const double *__restrict__ inv_j, const double *__restrict__ rr, long N, long n,
                              int c0, double gamma, const double *v, double *out)
{
    // every fused multiply-add is written out (the file is compiled without
    // contraction), so both schedules round identically
    const double rho = q[n], mu = q[N + n], mv = q[2 * N + n], mw = q[3 * N + n], E = q[4 * N + n];
    const double mx = dxidx[c0 * N + n], my = dxidx[(c0 + 1) * N + n], mz = dxidx[(c0 + 2) * N + n];
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
    const double r = SGN > 0 ? rr[n] : -rr[n];
    const double s = 0.5 * inv_j[n];
    out[0] = s * fma(r, v[0], mvm);
    out[1] = s * fma(r, v[1], fma(mx, dp, fma(u, theta, U * v[1])));
    out[2] = s * fma(r, v[2], fma(my, dp, fma(w2, theta, U * v[2])));
    out[3] = s * fma(r, v[3], fma(mz, dp, fma(w3, theta, U * v[3])));
    out[4] = s * fma(r, v[4], fma(U, dp + v[4], a1 * theta));
}

struct LuArgs {
    Ext e;
    const double *q, *dxidx, *inv_j, *rr, *diag, *rhs;
    double *dqs, *dq;
    double gamma;
};

// forward node (:345-365): dq* = (rhs + A+ dq*_i-1 + B+ dq*_j-1 + C+ dq*_k-1)/diag/inv_j
__device__ inline void lu_forward(const LuArgs &a, long n, const double *vi, const double *vj,
                                  const double *vk, double *res)
{
    const long N = a.e.n, sj = a.e.ni, sk = (long)a.e.ni * a.e.nj;
    double ta[5], tb[5], tc[5];
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr, N, n - 1, 0, a.gamma, vi, ta);
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr + N, N, n - sj, 3, a.gamma, vj, tb);
    jac_mv<1>(a.q, a.dxidx, a.inv_j, a.rr + 2 * N, N, n - sk, 6, a.gamma, vk, tc);
    const double dg = a.diag[n], ij = a.inv_j[n];
#pragma unroll
    for (int x = 0; x < 5; x++) res[x] = (a.rhs[x * N + n] + (ta[x] + tb[x] + tc[x])) / dg / ij;
}

// backward node (:521-542): dq = dq* - (A- dq_i+1 + B- dq_j+1 + C- dq_k+1)/inv_j/diag
__device__ inline void lu_backward(const LuArgs &a, long n, const double *vi, const double *vj,
                                   const double *vk, double *res)
{
    const long N = a.e.n, sj = a.e.ni, sk = (long)a.e.ni * a.e.nj;
    double ta[5], tb[5], tc[5];
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr, N, n + 1, 0, a.gamma, vi, ta);
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr + N, N, n + sj, 3, a.gamma, vj, tb);
    jac_mv<-1>(a.q, a.dxidx, a.inv_j, a.rr + 2 * N, N, n + sk, 6, a.gamma, vk, tc);
    const double dg = a.diag[n], ij = a.inv_j[n];
#pragma unroll
    for (int x = 0; x < 5; x++) res[x] = a.dqs[x * N + n] - (ta[x] + tb[x] + tc[x]) / ij / dg;
}

// rhs = rhsi + rhsv on the computed range; spectral radii ra, rb, rc and
// diag on the ring is1-1..ie1+1 (:72-156); dq* and dq cleared on the ring so
// that out-of-range neighbours read 0 (:105-106).
__global__ void __launch_bounds__(256) k_lu_prep(Ext e, Box ring, const double *__restrict__ pri,
                                                 const double *__restrict__ dxidx,
                                                 const double *__restrict__ dt,
                                                 const double *__restrict__ rhsi,
                                                 const double *__restrict__ rhsv,
                                                 double *__restrict__ rhs, double *__restrict__ rr,
                                                 double *__restrict__ diag, double *__restrict__ dqs,
                                                 double *__restrict__ dq, const double *cfl)
{
    int i, j, k;
    if (!box_node(ring, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
    double phi = *cfl >= 1.0e3 ? 0.0 : 1.0;  // :98-102
    double u = pri[N + n], v = pri[2 * N + n], w = pri[3 * N + n], c = pri[6 * N + n];
    double r3[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double m0 = dxidx[(3 * d) * N + n], m1 = dxidx[(3 * d + 1) * N + n], m2 = dxidx[(3 * d + 2) * N + n];
        double temp = m0 * m0 + m1 * m1 + m2 * m2;
        temp = sqrt(temp);
        double cu = m0 * u + m1 * v + m2 * w;
        r3[d] = fabs(cu) + fabs(c * temp);
        rr[d * N + n] = r3[d];
    }
    diag[n] = phi * 1.0 / dt[n] + r3[0] + r3[1] + r3[2];
    bool inner = i >= e.is1 && i <= e.ie1 && j >= e.js1 && j <= e.je1 && k >= e.ks1 && k <= e.ke1;
#pragma unroll
    for (int x = 0; x < 5; x++) {
        dqs[x * N + n] = 0.0;
        dq[x * N + n] = 0.0;
        if (inner) rhs[x * N + n] = rhsi[x * N + n] + rhsv[x * N + n];
    }
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

// Wait (thread 0 only) until the counter at f reaches need.  The wait is
// bounded: running out sets *err and stops further waiting, so the sweep
// always ends (its results are then flagged as invalid by the host).
__device__ inline void wait_progress(const int *f, int need, int &known, bool &failed, int *err)
{
    if (!f || failed || known >= need) return;
    int v = 0;
    for (long spins = 0;; spins++) {
        v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v >= need) break;
        if (spins > (1L << 20)) {
            atomicOr(err, 1);
            failed = true;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    known = v;
}

// Exchange values between tiles are read and written at agent scope: the
// accesses bypass the CU's vector L1, and the array lives in uncached memory
// so no XCD's L2 holds a stale copy.  No acquire/release fence is needed (a
// fence would write back and invalidate the whole L2 at every step).
__device__ inline double xload(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Persistent tile sweep; flags[tile] = epoch + completed steps of the tile.
// Tiles are numbered bj + nbj*bk and only wait on lower numbers, which are
// dispatched first, so every wait is on a tile that is running or done.
template <int DIR, int TJ, int TK>
__global__ void __launch_bounds__(TJ *TK) k_lu_tiles(LuArgs a, int *flags, int epoch, int *err)
{
    __shared__ double sv[2][5][TK][TJ + 1];
    const Ext &e = a.e;
    const int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    const int nbj = (Lj + TJ - 1) / TJ;
    const int tile = blockIdx.x;
    const int bj = tile % nbj, bk = tile / nbj;
    const int tj = threadIdx.x % TJ, tk = threadIdx.x / TJ;
    const int lj = bj * TJ + tj, lk = bk * TK + tk;
    const bool active = lj < Lj && lk < Lk;
    const long N = e.n, sj = e.ni, sk = (long)e.ni * e.nj;
    const int nsteps = Li + TJ + TK - 2;
    double *out = DIR > 0 ? a.dqs : a.dq;
    int *myflag = flags + tile;
    const int *fj = bj > 0 ? flags + (tile - 1) : nullptr;
    const int *fk = bk > 0 ? flags + (tile - nbj) : nullptr;
    int known_j = 0, known_k = 0;
    bool failed = false;
    double vi[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < nsteps; t++) {
        if (threadIdx.x == 0) {
            // the neighbour tile's thread on the shared row/column handles
            // this plane TJ-1 (TK-1) steps later on its own clock
            wait_progress(fj, epoch + min(t + TJ, nsteps), known_j, failed, err);
            wait_progress(fk, epoch + min(t + TK, nsteps), known_k, failed, err);
        }
        __syncthreads();
        const int li = t - tj - tk;
        if (active && li >= 0 && li < Li) {
            int i = DIR > 0 ? e.is1 + li : e.ie1 - li;
            int j = DIR > 0 ? e.js1 + lj : e.je1 - lj;
            int k = DIR > 0 ? e.ks1 + lk : e.ke1 - lk;
            long n = e.ix(i, j, k);
            const int p = t & 1;
            double vj[5], vk[5], res[5];
#pragma unroll
            for (int x = 0; x < 5; x++) {
                vj[x] = tj > 0 ? sv[p ^ 1][x][tk][tj - 1] : xload(&out[x * N + n - DIR * sj]);
                vk[x] = tk > 0 ? sv[p ^ 1][x][tk - 1][tj] : xload(&out[x * N + n - DIR * sk]);
                if (li == 0) vi[x] = xload(&out[x * N + n - DIR]);
            }
            if (DIR > 0) lu_forward(a, n, vi, vj, vk, res);
            else lu_backward(a, n, vi, vj, vk, res);
#pragma unroll
            for (int x = 0; x < 5; x++) {
                xstore(&out[x * N + n], res[x]);
                vi[x] = res[x];
                sv[p][x][tk][tj] = res[x];
            }
        }
        // every store of this step is acknowledged before the barrier, so
        // the counter cannot overtake the values it announces
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(myflag, epoch + t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
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

// tile shapes the autotuner chooses from
const int kTileShapes[][2] = {{16, 16}, {8, 8}, {16, 8}, {8, 4}};
const int kNumTileShapes = 4;

int lusgs_tile_shapes() { return kNumTileShapes; }

long lusgs_tiles(const Ext &e, int shape)
{
    int Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    int TJ = kTileShapes[shape][0], TK = kTileShapes[shape][1];
    return (long)((Lj + TJ - 1) / TJ) * ((Lk + TK - 1) / TK);
}

// one error word (flags[0]) plus per-tile progress counters (flags[1..]),
// for any tile shape
long lusgs_tile_flags_needed(const Ext &e)
{
    long mx = 0;
    for (int s = 0; s < kNumTileShapes; s++) mx = std::max(mx, lusgs_tiles(e, s));
    return mx + 1;
}

static int lu_epoch = 0;

template <int DIR, int TJ, int TK>
static void launch_tiles(hipStream_t s, const LuArgs &a, int *flags, long ntiles, long nflags)
{
    const int Li = a.e.ie1 - a.e.is1 + 1;
    const int nsteps = Li + TJ + TK - 2;
    // counters only grow: each sweep starts above every value written before
    (void)nflags;
    hipLaunchKernelGGL((k_lu_tiles<DIR, TJ, TK>), dim3((unsigned)ntiles), dim3(TJ * TK), 0, s, a, flags + 1, lu_epoch,
                       flags);
    lu_epoch += nsteps;
}

// phase 0: rhs, spectral radii, diag, cleared dq*/dq; 1: forward sweep;
// 2: backward sweep; 3: q += dq.  variant 0 = hyperplanes, 1 = tiles of the
// given shape index.
int launch_lusgs_phase(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                       int variant, int *tile_flags, int phase, int shape)
{
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    if (phase == 0) {
        Box ring{e.is1 - 1, e.ie1 + 1, e.js1 - 1, e.je1 + 1, e.ks1 - 1, e.ke1 + 1};
        hipLaunchKernelGGL(k_lu_prep, box_grid(ring), dim3(64, 4, 1), 0, s, e, ring, f.pri, f.dxidx, f.dt,
                           f.rhsi, f.rhsv, f.rhs, f.w_r, f.w_diag, f.w_dqs, f.w_dq, cfl);
    } else if (phase == 3) {
        Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
        if (Li > 0 && Lj > 0 && Lk > 0)
            hipLaunchKernelGGL(k_add5, box_grid(b), dim3(64, 4, 1), 0, s, e, b, f.q, f.w_dq);
    } else {
        if (Li <= 0 || Lj <= 0 || Lk <= 0) return 0;
        LuArgs a{e, f.q, f.dxidx, f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma};
        const bool fwd = phase == 1;
        if (variant == 1 && tile_flags) {
            long nt = lusgs_tiles(e, shape), nf = lusgs_tile_flags_needed(e);
            switch (shape) {
            case 0:
                if (fwd) launch_tiles<1, 16, 16>(s, a, tile_flags, nt, nf);
                else launch_tiles<-1, 16, 16>(s, a, tile_flags, nt, nf);
                break;
            case 1:
                if (fwd) launch_tiles<1, 8, 8>(s, a, tile_flags, nt, nf);
                else launch_tiles<-1, 8, 8>(s, a, tile_flags, nt, nf);
                break;
            case 2:
                if (fwd) launch_tiles<1, 16, 8>(s, a, tile_flags, nt, nf);
                else launch_tiles<-1, 16, 8>(s, a, tile_flags, nt, nf);
                break;
            default:
                if (fwd) launch_tiles<1, 8, 4>(s, a, tile_flags, nt, nf);
                else launch_tiles<-1, 8, 4>(s, a, tile_flags, nt, nf);
                break;
            }
        } else {
            int L = Li + Lj + Lk - 2;
            for (int l = 0; l < L; l++) {
                int ja = max(0, l - (Li - 1) - (Lk - 1)), jb = min(Lj - 1, l);
                int ka = max(0, l - (Li - 1) - (Lj - 1)), kb = min(Lk - 1, l);
                dim3 g((jb - ja + 64) / 64, (kb - ka + 4) / 4, 1);
                if (fwd) hipLaunchKernelGGL((k_lu_plane<1>), g, dim3(64, 4, 1), 0, s, a, l, ja, ka);
                else hipLaunchKernelGGL((k_lu_plane<-1>), g, dim3(64, 4, 1), 0, s, a, l, ja, ka);
            }
        }
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_lusgs(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                 int variant, int *tile_flags)
{
    for (int ph = 0; ph < 4; ph++) {
        int rc = launch_lusgs_phase(s, e, f, c, cfl, variant, tile_flags, ph, 0);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace acfd
