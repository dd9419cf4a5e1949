// Implicit update kernels: LU-SGS first-order implicit step and the BDF2
// line-sweep sub-iteration.
//
// Reference routines: sub_timeadvancement_jac.f90:7-63 (calc_jm),
// sub_timeadvancement_1stimplicit.f90:7-585 (implicit1st),
// sub_timeadvancement_2ndcrank.f90:7-467 (crank2nd), main.f90:692-711
// (update_q0).
//
// LU-SGS is a lexicographic Gauss-Seidel sweep: node (i,j,k) needs dq* of
// (i-1,j,k), (i,j-1,k), (i,j,k-1) and nothing else of the current sweep.
// Nodes on one hyperplane i+j+k = l are therefore independent, and sweeping
// the hyperplanes in order computes every node from exactly the same
// operands as the sequential loop.  Two schedules are provided:
//   variant 0: one launch per hyperplane;
//   variant 1: a persistent launch in which each workgroup owns a 16x16 tile
//              of (j,k) lines and marches along i on a skewed schedule
//              (thread (tj,tk) handles i = t - tj - tk at step t), exchanging
//              the j/k neighbours through LDS inside the tile and through
//              global memory plus per-tile progress counters across tiles.
#include "common.hpp"
#include "kernels.hpp"

namespace acfd {

// calc_jm (sub_timeadvancement_jac.f90:7-63) at node n with metric row c0,
// then (a +- r I) * 0.5 * inv_j (sub_timeadvancement_1stimplicit.f90:283-301,
// :459-477) and its product with v: out[r] = sum_c a(r,c) v(c).
template <int SGN>
__device__ inline void split_jac_mv(const double *__restrict__ q, const double *__restrict__ dxidx,
                                    const double *__restrict__ inv_j, const double *__restrict__ rr,
                                    long N, long n, int c0, double gamma, const double *v, double *out)
{
    double q1 = q[n], q2 = q[N + n], q3 = q[2 * N + n], q4 = q[3 * N + n], q5 = q[4 * N + n];
    double dxidx0 = dxidx[c0 * N + n], dxidy0 = dxidx[(c0 + 1) * N + n], dxidz0 = dxidx[(c0 + 2) * N + n];
    double a2 = gamma - 1.0, a3 = gamma - 2.0;
    double rho = q1, mu0 = q2, mv0 = q3, mw0 = q4, e0 = q5;
    double uu = (dxidx0 * mu0 + dxidy0 * mv0 + dxidz0 * mw0) / rho;
    double phi = 0.5 * a2 * (mu0 * mu0 + mv0 * mv0 + mw0 * mw0) / (rho * rho);
    double a1 = gamma * e0 / rho - phi;
    double a[5][5];
    a[0][0] = 0.0; a[0][1] = dxidx0; a[0][2] = dxidy0; a[0][3] = dxidz0; a[0][4] = 0.0;
    a[1][0] = phi * dxidx0 - uu * mu0 / rho;
    a[1][1] = uu - a3 * dxidx0 * mu0 / rho;
    a[1][2] = dxidy0 * mu0 / rho - a2 * dxidx0 * mv0 / rho;
    a[1][3] = dxidz0 * mu0 / rho - a2 * dxidx0 * mw0 / rho;
    a[1][4] = a2 * dxidx0;
    a[2][0] = phi * dxidy0 - uu * mv0 / rho;
    a[2][1] = dxidx0 * mv0 / rho - a2 * dxidy0 * mu0 / rho;
    a[2][2] = uu - a3 * dxidy0 * mv0 / rho;
    a[2][3] = dxidz0 * mv0 / rho - a2 * dxidy0 * mw0 / rho;
    a[2][4] = a2 * dxidy0;
    a[3][0] = phi * dxidz0 - uu * mw0 / rho;
    a[3][1] = dxidx0 * mw0 / rho - a2 * dxidz0 * mu0 / rho;
    a[3][2] = dxidy0 * mw0 / rho - a2 * dxidz0 * mv0 / rho;
    a[3][3] = uu - a3 * dxidz0 * mw0 / rho;
    a[3][4] = a2 * dxidz0;
    a[4][0] = (phi - a1) * uu;
    a[4][1] = a1 * dxidx0 - a2 * uu * mu0 / rho;
    a[4][2] = a1 * dxidy0 - a2 * uu * mv0 / rho;
    a[4][3] = a1 * dxidz0 - a2 * uu * mw0 / rho;
    a[4][4] = gamma * uu;
    double r = rr[n], ij = inv_j[n];
#pragma unroll
    for (int d = 0; d < 5; d++) a[d][d] = SGN > 0 ? a[d][d] + r : a[d][d] - r;
#pragma unroll
    for (int x = 0; x < 5; x++) {
        double s = 0.0;
#pragma unroll
        for (int y = 0; y < 5; y++) {
            double m = 0.5 * a[x][y] * ij;
            s = y == 0 ? m * v[0] : s + m * v[y];
        }
        out[x] = s;
    }
}

// rhs = rhsi + rhsv on the computed range; spectral radii ra, rb, rc and
// diag on the ring is1-1..ie1+1 (sub_timeadvancement_1stimplicit.f90:72-156);
// dq* and dq cleared on the ring so that out-of-range neighbours read 0.
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
    double phi = *cfl >= 1.0e3 ? 0.0 : 1.0;
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

// Forward update of one node (sub_timeadvancement_1stimplicit.f90:279-368).
__device__ inline void lu_forward_node(const Ext &e, long n, const double *__restrict__ q,
                                       const double *__restrict__ dxidx,
                                       const double *__restrict__ inv_j,
                                       const double *__restrict__ rr, const double *__restrict__ diag,
                                       const double *__restrict__ rhs, double *__restrict__ dqs,
                                       double gamma, const double *vi, const double *vj,
                                       const double *vk)
{
    const long N = e.n;
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    double ta[5], tb[5], tc[5];
    split_jac_mv<1>(q, dxidx, inv_j, rr, N, n - 1, 0, gamma, vi, ta);
    split_jac_mv<1>(q, dxidx, inv_j, rr + N, N, n - sj, 3, gamma, vj, tb);
    split_jac_mv<1>(q, dxidx, inv_j, rr + 2 * N, N, n - sk, 6, gamma, vk, tc);
    double dg = diag[n], ij = inv_j[n];
#pragma unroll
    for (int x = 0; x < 5; x++) dqs[x * N + n] = (rhs[x * N + n] + (ta[x] + tb[x] + tc[x])) / dg / ij;
}

// Backward update of one node (:455-546).
__device__ inline void lu_backward_node(const Ext &e, long n, const double *__restrict__ q,
                                        const double *__restrict__ dxidx,
                                        const double *__restrict__ inv_j,
                                        const double *__restrict__ rr,
                                        const double *__restrict__ diag,
                                        const double *__restrict__ dqs, double *__restrict__ dq,
                                        double gamma, const double *vi, const double *vj,
                                        const double *vk)
{
    const long N = e.n;
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    double ta[5], tb[5], tc[5];
    split_jac_mv<-1>(q, dxidx, inv_j, rr, N, n + 1, 0, gamma, vi, ta);
    split_jac_mv<-1>(q, dxidx, inv_j, rr + N, N, n + sj, 3, gamma, vj, tb);
    split_jac_mv<-1>(q, dxidx, inv_j, rr + 2 * N, N, n + sk, 6, gamma, vk, tc);
    double dg = diag[n], ij = inv_j[n];
#pragma unroll
    for (int x = 0; x < 5; x++) dq[x * N + n] = dqs[x * N + n] - (ta[x] + tb[x] + tc[x]) / ij / dg;
}

// One hyperplane l of the forward (DIR = +1) or backward (DIR = -1) sweep.
// Threads cover (lj, lk) in [ja, ja+gridDim.x*64) x [ka, ...); li = l-lj-lk.
template <int DIR>
__global__ void __launch_bounds__(256) k_lu_plane(Ext e, int l, int ja, int ka, const double *__restrict__ q,
                                                  const double *__restrict__ dxidx,
                                                  const double *__restrict__ inv_j,
                                                  const double *__restrict__ rr,
                                                  const double *__restrict__ diag,
                                                  const double *__restrict__ rhs,
                                                  double *__restrict__ dqs, double *__restrict__ dq,
                                                  double gamma)
{
    int lj = ja + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int lk = ka + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    int li = l - lj - lk;
    if (lj >= Lj || lk >= Lk || li < 0 || li >= Li) return;
    const long N = e.n;
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    if (DIR > 0) {
        long n = e.ix(e.is1 + li, e.js1 + lj, e.ks1 + lk);
        double vi[5], vj[5], vk[5];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            vi[x] = dqs[x * N + n - 1];
            vj[x] = dqs[x * N + n - sj];
            vk[x] = dqs[x * N + n - sk];
        }
        lu_forward_node(e, n, q, dxidx, inv_j, rr, diag, rhs, dqs, gamma, vi, vj, vk);
    } else {
        // backward planes count from the far corner
        long n = e.ix(e.ie1 - li, e.je1 - lj, e.ke1 - lk);
        double vi[5], vj[5], vk[5];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            vi[x] = dq[x * N + n + 1];
            vj[x] = dq[x * N + n + sj];
            vk[x] = dq[x * N + n + sk];
        }
        lu_backward_node(e, n, q, dxidx, inv_j, rr, diag, dqs, dq, gamma, vi, vj, vk);
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

// ------------------------------------------------------- pipelined sweep
// Tile of TJ x TK (j,k) lines per workgroup; thread (tj,tk) processes
// li = t - tj - tk at local step t.  Cross-tile neighbours are read from
// global memory after the producing tile has published its progress.
constexpr int TJ = 16, TK = 16;

__device__ inline int ld_acquire(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_release(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// flags[tile] = number of completed local steps of that tile.  Grid: one
// workgroup per tile, tiles numbered bj + nbj*bk; every tile only waits on
// tiles with a smaller number, and workgroups are dispatched in increasing
// order, so the waits always terminate once earlier tiles are resident.
template <int DIR>
__global__ void __launch_bounds__(TJ *TK) k_lu_tiles(Ext e, const double *__restrict__ q,
                                                     const double *__restrict__ dxidx,
                                                     const double *__restrict__ inv_j,
                                                     const double *__restrict__ rr,
                                                     const double *__restrict__ diag,
                                                     const double *__restrict__ rhs,
                                                     double *__restrict__ dqs, double *__restrict__ dq,
                                                     double gamma, int *flags, int epoch, int *err)
{
    __shared__ double sv[2][5][TK][TJ + 1];  // double-buffered dq of the previous step
    const int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    const int nbj = (Lj + TJ - 1) / TJ;
    const int tile = blockIdx.x;
    const int bj = tile % nbj, bk = tile / nbj;
    const int tj = threadIdx.x % TJ, tk = threadIdx.x / TJ;
    const int lj = bj * TJ + tj, lk = bk * TK + tk;
    const bool active = lj < Lj && lk < Lk;
    const long N = e.n;
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    const int nsteps = Li + TJ + TK - 2;
    double *out = DIR > 0 ? dqs : dq;
    int *myflag = flags + tile;
    const int *fj = bj > 0 ? flags + (tile - 1) : nullptr;
    const int *fk = bk > 0 ? flags + (tile - nbj) : nullptr;
    double vi[5] = {0, 0, 0, 0, 0};  // dq of (i -+ 1) kept in registers
    for (int t = 0; t < nsteps; t++) {
        int li = t - tj - tk;
        bool work = active && li >= 0 && li < Li;
        // cross-tile dependencies: neighbour tile must have finished the
        // step that produced the plane before ours
        if (threadIdx.x == 0) {
            // bounded waits: a wait that runs out flags the sweep as failed
            // (the host turns it into an error) instead of hanging the device
            if (fj) {
                int need = epoch + min(t + TJ, nsteps);
                long spins = 0;
                while (ld_acquire(fj) < need && spins < (1L << 20)) {
                    __builtin_amdgcn_s_sleep(1);
                    spins++;
                }
                if (spins >= (1L << 20)) atomicOr(err, 1);
            }
            if (fk) {
                int need = epoch + min(t + TK, nsteps);
                long spins = 0;
                while (ld_acquire(fk) < need && spins < (1L << 20)) {
                    __builtin_amdgcn_s_sleep(1);
                    spins++;
                }
                if (spins >= (1L << 20)) atomicOr(err, 1);
            }
        }
        __syncthreads();
        if (work) {
            // sweep-local coordinates; backward counts from the far corner
            int i = DIR > 0 ? e.is1 + li : e.ie1 - li;
            int j = DIR > 0 ? e.js1 + lj : e.je1 - lj;
            int k = DIR > 0 ? e.ks1 + lk : e.ke1 - lk;
            long n = e.ix(i, j, k);
            const int p = t & 1;
            double vj[5], vk[5];
#pragma unroll
            for (int x = 0; x < 5; x++) {
                if (tj > 0) vj[x] = sv[p ^ 1][x][tk][tj - 1];
                else vj[x] = (DIR > 0 ? dqs : dq)[x * N + n - DIR * sj];
                if (tk > 0) vk[x] = sv[p ^ 1][x][tk - 1][tj];
                else vk[x] = (DIR > 0 ? dqs : dq)[x * N + n - DIR * sk];
            }
            if (li == 0) {
#pragma unroll
                for (int x = 0; x < 5; x++) vi[x] = (DIR > 0 ? dqs : dq)[x * N + n - DIR];
            }
            if (DIR > 0)
                lu_forward_node(e, n, q, dxidx, inv_j, rr, diag, rhs, dqs, gamma, vi, vj, vk);
            else
                lu_backward_node(e, n, q, dxidx, inv_j, rr, diag, dqs, dq, gamma, vi, vj, vk);
#pragma unroll
            for (int x = 0; x < 5; x++) {
                double r = out[x * N + n];
                vi[x] = r;
                sv[p][x][tk][tj] = r;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            st_release(myflag, epoch + t + 1);
        }
    }
}

// per-tile progress counters plus one error word
long lusgs_tile_flags_needed(const Ext &e)
{
    int Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    return (long)((Lj + TJ - 1) / TJ) * ((Lk + TK - 1) / TK) + 1;
}

static int lu_epoch = 0;

// phase 0: rhs, spectral radii, diag, cleared dq*/dq; 1: forward sweep;
// 2: backward sweep; 3: q += dq.
int launch_lusgs_phase(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
                       int variant, int *tile_flags, int phase)
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
        const int dir = phase == 1 ? 1 : -1;
        if (variant == 1 && tile_flags) {
            long ntiles = lusgs_tile_flags_needed(e) - 1;
            int nsteps = Li + TJ + TK - 2;
            // flags grow monotonically across sweeps: each sweep starts at an
            // epoch above every value written before, so no reset is needed
            int ep = lu_epoch;
            if (dir > 0)
                hipLaunchKernelGGL((k_lu_tiles<1>), dim3((unsigned)ntiles), dim3(TJ * TK), 0, s, e, f.q, f.dxidx,
                                   f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma, tile_flags, ep,
                                   tile_flags + ntiles);
            else
                hipLaunchKernelGGL((k_lu_tiles<-1>), dim3((unsigned)ntiles), dim3(TJ * TK), 0, s, e, f.q, f.dxidx,
                                   f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma, tile_flags, ep,
                                   tile_flags + ntiles);
            lu_epoch = ep + nsteps;
        } else {
            int L = Li + Lj + Lk - 2;
            for (int l = 0; l < L; l++) {
                int ja = max(0, l - (Li - 1) - (Lk - 1)), jb = min(Lj - 1, l);
                int ka = max(0, l - (Li - 1) - (Lj - 1)), kb = min(Lk - 1, l);
                dim3 g((jb - ja + 64) / 64, (kb - ka + 4) / 4, 1);
                if (dir > 0)
                    hipLaunchKernelGGL((k_lu_plane<1>), g, dim3(64, 4, 1), 0, s, e, l, ja, ka, f.q, f.dxidx,
                                       f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma);
                else
                    hipLaunchKernelGGL((k_lu_plane<-1>), g, dim3(64, 4, 1), 0, s, e, l, ja, ka, f.q, f.dxidx,
                                       f.inv_j, f.w_r, f.w_diag, f.rhs, f.w_dqs, f.w_dq, c.gamma);
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
        int rc = launch_lusgs_phase(s, e, f, c, cfl, variant, tile_flags, ph);
        if (rc) return rc;
    }
    return 0;
}

// ------------------------------------------------------------------ BDF2
// rhs of sub_timeadvancement_2ndcrank.f90:61-82 on the computed range and
// dq cleared on the ring.
__global__ void __launch_bounds__(256) k_crank_prep(Ext e, Box ring, const double *__restrict__ q,
                                                    const double *__restrict__ q0,
                                                    const double *__restrict__ dt,
                                                    const double *__restrict__ inv_j,
                                                    const double *__restrict__ rhsi,
                                                    const double *__restrict__ rhsv,
                                                    double *__restrict__ rhs, double *__restrict__ dq)
{
    int i, j, k;
    if (!box_node(ring, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
    bool inner = i >= e.is1 && i <= e.ie1 && j >= e.js1 && j <= e.je1 && k >= e.ks1 && k <= e.ke1;
    const double alph = 0.5, alph_i = 1.0 / (1.0 + alph);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        dq[v * N + n] = 0.0;
        if (inner) {
            double r = rhsi[v * N + n] + rhsv[v * N + n];
            double qm1 = q0[(2 * v) * N + n], qz = q0[(2 * v + 1) * N + n];
            rhs[v * N + n] = alph_i * (r * dt[n] - ((1.0 + alph) * q[v * N + n] - (1.0 + 2.0 * alph) * qz
                                                    + (alph)*qm1) * inv_j[n]);
        }
    }
}

__device__ inline double crank_rou(const double *pri, const double *dxidx, long N, long n, int c0)
{
    double u = pri[N + n], v = pri[2 * N + n], w = pri[3 * N + n], c = pri[6 * N + n];
    double m0 = dxidx[c0 * N + n], m1 = dxidx[(c0 + 1) * N + n], m2 = dxidx[(c0 + 2) * N + n];
    double temp = m0 * m0 + m1 * m1 + m2 * m2;
    temp = sqrt(temp);
    double cu = u * m0 + v * m1 + w * m2;
    return fabs(cu) + fabs(c * temp);
}

// a1..a5 of the neighbour nb (sub_timeadvancement_2ndcrank.f90:129-160)
__device__ inline void crank_terms(double *a, const double *pri, const double *dxidx, double r,
                                   const double *dq, long N, long nb, int c0, double hdv, int sgn,
                                   double gamma)
{
    double u = pri[N + nb], v = pri[2 * N + nb], w = pri[3 * N + nb], c = pri[6 * N + nb];
    double m0 = dxidx[c0 * N + nb], m1 = dxidx[(c0 + 1) * N + nb], m2 = dxidx[(c0 + 2) * N + nb];
    double q2 = 0.5 * (u * u + v * v + w * w);
    double h = c * c / (gamma - 1.0) + q2;
    double cu = m0 * u + m1 * v + m2 * w;
    double s0 = dq[nb], s1 = dq[N + nb], s2 = dq[2 * N + nb], s3 = dq[3 * N + nb], s4 = dq[4 * N + nb];
    double b1 = hdv * (-cu * s0 + m0 * s1 + m1 * s2 + m2 * s3);
    double b2 = hdv * (q2 * s0 - u * s1 - v * s2 - w * s3 + s4) * (gamma - 1.0);
    double b3 = sgn > 0 ? hdv * (cu + r) : hdv * (cu - r);
    a[0] = b1 + b3 * s0;
    a[1] = b1 * u + b2 * m0 + b3 * s1;
    a[2] = b1 * v + b2 * m1 + b3 * s2;
    a[3] = b1 * w + b2 * m2 + b3 * s3;
    a[4] = b1 * h + b2 * cu + b3 * s4;
}

// One family of line recurrences (AXIS 0: xi lines over (j,k); 1: eta over
// (i,k); 2: zeta over (i,j)), one thread per line.
template <int AXIS>
__global__ void __launch_bounds__(256) k_crank_lines(Ext e, const double *__restrict__ pri,
                                                     const double *__restrict__ dxidx,
                                                     const double *__restrict__ dt,
                                                     const double *__restrict__ rhs,
                                                     double *__restrict__ rou, double *__restrict__ dq,
                                                     double gamma)
{
    const int ax1 = AXIS == 0 ? 1 : 0, ax2 = AXIS == 2 ? 1 : 2;
    int a = e.lo1(ax1) + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int b = e.lo1(ax2) + (int)blockIdx.y;
    if (a > e.hi1(ax1) || b > e.hi1(ax2)) return;
    const long N = e.n;
    const long st = e.stride(AXIS);
    const int c0 = 3 * AXIS;
    int i0 = AXIS == 0 ? 0 : a, j0 = AXIS == 1 ? 0 : (AXIS == 0 ? a : b), k0 = AXIS == 2 ? 0 : b;
    // node of line position x along AXIS
    auto node = [&](int x) -> long {
        return e.ix(AXIS == 0 ? x : i0, AXIS == 1 ? x : j0, AXIS == 2 ? x : k0);
    };
    int lo = e.lo1(AXIS), hi = e.hi1(AXIS);
    for (int x = lo - 1; x <= hi + 1; x++) {
        long n = node(x);
        rou[n] = crank_rou(pri, dxidx, N, n, c0);
    }
    const double alph = 0.5;
    double av[5];
    // forward: xi and eta from lo, zeta from lo+1 (:128, :244, :360)
    int f0 = AXIS == 2 ? lo + 1 : lo;
    for (int x = f0; x <= hi; x++) {
        long n = node(x), nb = n - st;
        double dts = dt[n] / (1.0 + alph);
        crank_terms(av, pri, dxidx, rou[nb], dq, N, nb, c0, 0.5 * dts, 1, gamma);
        double div = 1.0 / (1.0 + dts * rou[n]);
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double base = AXIS == 0 ? rhs[v * N + n] : dq[v * N + n];
            dq[v * N + n] = div * (base + av[v]);
        }
    }
    // backward: xi and eta from hi, zeta from hi-1 (:172, :288, :404)
    int b0 = AXIS == 2 ? hi - 1 : hi;
    for (int x = b0; x >= lo; x--) {
        long n = node(x), nb = n + st;
        double dts = dt[n] / (1.0 + alph);
        crank_terms(av, pri, dxidx, rou[nb], dq, N, nb, c0, 0.5 * dts, -1, gamma);
        double div = 1.0 / (1.0 + dts * rou[n]);
#pragma unroll
        for (int v = 0; v < 5; v++) dq[v * N + n] = dq[v * N + n] - div * av[v];
    }
}

// q += dq / inv_j (:455-461)
__global__ void __launch_bounds__(256) k_crank_update(Ext e, Box b, double *__restrict__ q,
                                                      const double *__restrict__ dq,
                                                      const double *__restrict__ inv_j)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
#pragma unroll
    for (int x = 0; x < 5; x++) q[x * N + n] = q[x * N + n] + dq[x * N + n] / inv_j[n];
}

int launch_crank(hipStream_t s, const Ext &e, const Fields &f, const Consts &c)
{
    Box ring{e.is1 - 1, e.ie1 + 1, e.js1 - 1, e.je1 + 1, e.ks1 - 1, e.ke1 + 1};
    hipLaunchKernelGGL(k_crank_prep, box_grid(ring), dim3(64, 4, 1), 0, s, e, ring, f.q, f.q0, f.dt,
                       f.inv_j, f.rhsi, f.rhsv, f.rhs, f.w_dq);
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    hipLaunchKernelGGL((k_crank_lines<0>), dim3((Lj + 63) / 64, Lk), dim3(64), 0, s, e, f.pri, f.dxidx,
                       f.dt, f.rhs, f.w_r, f.w_dq, c.gamma);
    hipLaunchKernelGGL((k_crank_lines<1>), dim3((Li + 63) / 64, Lk), dim3(64), 0, s, e, f.pri, f.dxidx,
                       f.dt, f.rhs, f.w_r + e.n, f.w_dq, c.gamma);
    hipLaunchKernelGGL((k_crank_lines<2>), dim3((Li + 63) / 64, Lj), dim3(64), 0, s, e, f.pri, f.dxidx,
                       f.dt, f.rhs, f.w_r + 2 * e.n, f.w_dq, c.gamma);
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    hipLaunchKernelGGL(k_crank_update, box_grid(b), dim3(64, 4, 1), 0, s, e, b, f.q, f.w_dq, f.inv_j);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// update_q0 on the physical range (main.f90:701-708)
__global__ void __launch_bounds__(256) k_update_q0(Ext e, Box b, double *__restrict__ q0,
                                                   const double *__restrict__ q)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        q0[(2 * v) * N + n] = q0[(2 * v + 1) * N + n];
        q0[(2 * v + 1) * N + n] = q[v * N + n];
    }
}

int launch_update_q0(hipStream_t s, const Ext &e, const Fields &f)
{
    Box b{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0};
    hipLaunchKernelGGL(k_update_q0, box_grid(b), dim3(64, 4, 1), 0, s, e, b, f.q0, f.q);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
