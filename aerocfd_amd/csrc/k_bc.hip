// Boundary kernels: physical boundary conditions, the SA clamp, the
// inter-block halo fill, overlap averaging, the residual norm, and the
// AoS <-> SoA layout transposition used at the host boundary.
//
// Reference: sub_bc.f90:9-90, sub_bc_-10_farfield.f90, sub_bc_-20_solidwall.f90,
// sub_bc_-30_symm.f90, sub_bc_-40_outflow.f90, sub_bc_-50_inflow.f90,
// sub_bc_updatebuffer.f90:10-548, sub_bc_updatebufferaverage.f90:10-562,
// sub_output_residual.f90:8-69.
#include "batch.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace acfd {

constexpr double kNutInf = 1.341946;  // module_tursa.f90:10

// One thread per face node of a physical subface.  The six faces of every
// condition file differ only in the axis and the side, so they share one
// kernel: boundary index lo (odd face) or hi (even face) along the axis,
// interior neighbour one step inwards.
__device__ inline void bc_node(const Ext &e, const Consts &c, const BcTask &t, long x, double *__restrict__ q,
                               double *__restrict__ nut, const double *__restrict__ dxidx)
{
    const int axis = (t.face - 1) / 2;
    const bool minus = (t.face & 1) == 1;
    int lo[3] = {t.is, t.js, t.ks}, hi[3] = {t.ie, t.je, t.ke};
    int fix = minus ? lo[axis] : hi[axis];
    lo[axis] = hi[axis] = fix;
    int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
    if (x >= (long)n0 * n1 * n2) return;
    int i = lo[0] + (int)(x % n0);
    int j = lo[1] + (int)((x / n0) % n1);
    int k = lo[2] + (int)(x / ((long)n0 * n1));
    const long N = e.n;
    const long st = e.stride(axis) * (minus ? 1 : -1);
    long nb = e.ix(i, j, k), np = nb + st;
    const int turb = c.turb;
    if (t.kind == -50) {  // inflow, sub_bc_-50_inflow.f90:29-37
        q[nb] = c.d0;
        q[N + nb] = c.d0 * c.u0 * cos(c.aoa);
        q[2 * N + nb] = c.d0 * c.u0 * sin(c.aoa);
        q[3 * N + nb] = 0.0;
        q[4 * N + nb] = c.p0 / (c.gamma - 1.0) + 0.5 * (c.u0 * c.u0);
        if (turb >= 1) nut[nb] = kNutInf;
        return;
    }
    if (t.kind == -40) {  // outflow, sub_bc_-40_outflow.f90:24-32
#pragma unroll
        for (int v = 0; v < 5; v++) q[v * N + nb] = q[v * N + np];
        if (turb >= 1) nut[nb] = nut[np];
        return;
    }
    if (t.kind == -20 && c.solver == 0) {  // no-slip wall, sub_bc_-20_solidwall.f90:228-235
        long n2 = np + st, n3 = np + 2 * st;
        q[nb] = (11.0 * q[np] - 7.0 * q[n2] + 2.0 * q[n3]) / 6.0;
        q[N + nb] = 0.0;
        q[2 * N + nb] = 0.0;
        q[3 * N + nb] = 0.0;
        q[4 * N + nb] = (11.0 * q[4 * N + np] - 7.0 * q[4 * N + n2] + 2.0 * q[4 * N + n3]) / 6.0;
        if (turb >= 1) nut[nb] = 0.0;
        return;
    }
    // unit outward normal from the metric row of the axis (e.g.
    // sub_bc_-10_farfield.f90:43-52)
    const int row = 3 * axis;
    double mx = dxidx[row * N + nb], my = dxidx[(row + 1) * N + nb], mz = dxidx[(row + 2) * N + nb];
    double temp = mx * mx + my * my + mz * mz;
    temp = sqrt(temp);
    double nx, ny, nz;
    if (minus) {
        nx = -mx / temp; ny = -my / temp; nz = -mz / temp;
    } else {
        nx = mx / temp; ny = my / temp; nz = mz / temp;
    }
    if (t.kind == -30 || t.kind == -20) {  // symmetry / slip wall, sub_bc_-30_symm.f90:39-49
        double q1 = q[np], q2 = q[N + np], q3 = q[2 * N + np], q4 = q[3 * N + np], q5 = q[4 * N + np];
        double cu = q2 * nx + q3 * ny + q4 * nz;
        q[nb] = q1;
        q[N + nb] = q2 - cu * nx;
        q[2 * N + nb] = q3 - cu * ny;
        q[3 * N + nb] = q4 - cu * nz;
        q[4 * N + nb] = q5;
        if (turb >= 1) nut[nb] = nut[np];
        return;
    }
    if (t.kind == -10) {  // farfield, sub_bc_-10_farfield.f90:29-104
        const double gamma = c.gamma;
        double u_inf = c.u0 * cos(c.aoa), v_inf = c.u0 * sin(c.aoa), w_inf = 0.0;
        double s_inf = c.p0 / pow(c.d0, gamma), c_inf = c.c0;
        double d_p = q[np];
        double u_p = q[N + np] / d_p, v_p = q[2 * N + np] / d_p, w_p = q[3 * N + np] / d_p;
        double p_p = (q[4 * N + np] - 0.5 * d_p * (u_p * u_p + v_p * v_p + w_p * w_p)) * (gamma - 1.0);
        double c_p = sqrt(gamma * p_p / d_p);
        double s_p = p_p / pow(d_p, gamma);
        double vn_inf = u_inf * nx + v_inf * ny + w_inf * nz;
        double vn_p = u_p * nx + v_p * ny + w_p * nz;
        double rm = vn_inf - 2.0 * c_inf / (gamma - 1.0);
        double rp = vn_p + 2.0 * c_p / (gamma - 1.0);
        double vn_b = 0.5 * (rp + rm);
        double s_b = 0, u_ref = 0, v_ref = 0, w_ref = 0, turval = 0;
        if (vn_b < 0.0) {
            s_b = s_inf; u_ref = u_inf; v_ref = v_inf; w_ref = w_inf; turval = kNutInf;
        }
        if (vn_b >= 0.0) {
            s_b = s_p; u_ref = u_p; v_ref = v_p; w_ref = w_p; turval = nut ? nut[np] : 0.0;
        }
        double u_ref_t = u_ref * nx + v_ref * ny + w_ref * nz;
        double u_b = u_ref + (vn_b - u_ref_t) * nx;
        double v_b = v_ref + (vn_b - u_ref_t) * ny;
        double w_b = w_ref + (vn_b - u_ref_t) * nz;
        // density from the interior sound speed c_p (:93)
        double d_b = pow(c_p * c_p / (gamma * s_b), 1.0 / (gamma - 1.0));
        double p_b = s_b * pow(d_b, gamma);
        q[nb] = d_b;
        q[N + nb] = d_b * u_b;
        q[2 * N + nb] = d_b * v_b;
        q[3 * N + nb] = d_b * w_b;
        q[4 * N + nb] = p_b / (gamma - 1.0) + 0.5 * d_b * (u_b * u_b + v_b * v_b + w_b * w_b);
        if (turb >= 1) nut[nb] = turval;
    }
}

__global__ void __launch_bounds__(256) k_bc(Ext e, Consts c, BcTask t, double *__restrict__ q,
                                            double *__restrict__ nut, const double *__restrict__ dxidx)
{
    bc_node(e, c, t, (long)blockIdx.x * blockDim.x + threadIdx.x, q, nut, dxidx);
}

// One subface of each of several blocks in one launch (tasks[m] for the
// workgroups of entry m of the grid).  Blocks touch only their own arrays,
// so the subfaces of different blocks are independent; the subfaces of one
// block go to successive launches, in the reference's order.
__global__ void __launch_bounds__(256) k_bc_batch(const BcTask *__restrict__ tasks, const int *__restrict__ map,
                                                  const KGrid *__restrict__ kg, const BlockRec *__restrict__ blks,
                                                  double *arena, Consts c)
{
    int bx, by, bz;
    const BcTask t = tasks[batch_wg(map, kg, bx, by, bz)];
    const BlockRec &b = blks[t.block];
    bc_node(b.e, c, t, (long)bx * blockDim.x + threadIdx.x, arena + b.o.q, b.o.nut >= 0 ? arena + b.o.nut : nullptr,
            arena + b.o.dxidx);
}

long bc_nodes(const BcTask &t)
{
    const int axis = (t.face - 1) / 2;
    int lo[3] = {t.is, t.js, t.ks}, hi[3] = {t.ie, t.je, t.ke};
    lo[axis] = hi[axis] = 0;
    return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

int launch_bc(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const BcTask &t)
{
    const long cnt = bc_nodes(t);
    if (cnt <= 0) return 0;
    int bs = 256;
    hipLaunchKernelGGL(k_bc, dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0, s, e, c, t, f.q, f.nut, f.dxidx);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_bc_batch(hipStream_t s, const BcTask *tasks, const BatchGrid &bg, const BlockRec *blks, double *arena,
                    const Consts &c)
{
    if (bg.total <= 0) return 0;
    hipLaunchKernelGGL(k_bc_batch, dim3((unsigned)bg.total), dim3(256), 0, s, tasks, bg.map, bg.kg, blks, arena, c);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// SA clamp of sub_bc.f90:65-87 over the whole extended block: nut keeps
// min(nut, 1e6) (the max(nut,1e-4) of :80 is overwritten by :81) and
// mu_t = nut * fv1 * rho with rho from pri_v.
__device__ inline void sa_clamp_node(const Ext &e, long n, double *__restrict__ nut, double *__restrict__ amu,
                                     const double *__restrict__ pri)
{
    if (n >= e.n) return;
    const long N = e.n;
    const double cv1 = 7.1;
    double d = pri[n];
    double nu = nut[n];
    double leix = nu / amu[n] * d;
    leix = fmax(leix, 0.001);
    double l3 = leix * leix * leix;
    double fv1 = l3 / (l3 + cv1 * cv1 * cv1);
    double nn = fmin(nu, 1.0e6);
    if (!(nn == nu)) nut[n] = nn;  // stored only where it changes (NaN included)
    amu[N + n] = nn * fv1 * d;
}

__global__ void __launch_bounds__(256) k_sa_clamp(Ext e, double *__restrict__ nut, double *__restrict__ amu,
                                                  const double *__restrict__ pri)
{
    sa_clamp_node(e, (long)blockIdx.x * blockDim.x + threadIdx.x, nut, amu, pri);
}

// every block (1-D grid over each block's extended nodes)
__global__ void __launch_bounds__(256) k_sa_clamp_batch(const BlockRec *__restrict__ blks,
                                                        const int *__restrict__ map, const KGrid *__restrict__ kg,
                                                        double *arena)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    sa_clamp_node(b.e, (long)bx * blockDim.x + threadIdx.x, arena + b.o.nut, arena + b.o.amu, arena + b.o.pri);
}

int launch_sa_clamp_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena)
{
    if (bg.total <= 0) return 0;
    hipLaunchKernelGGL(k_sa_clamp_batch, dim3((unsigned)bg.total), dim3(256), 0, s, blks, bg.map, bg.kg, arena);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_sa_clamp(hipStream_t s, const Ext &e, const Fields &f)
{
    int bs = 256;
    hipLaunchKernelGGL(k_sa_clamp, dim3((unsigned)((e.n + bs - 1) / bs)), dim3(bs), 0, s, e, f.nut, f.amu, f.pri);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// Halo fill: a pure gather over a precomputed list of (source node,
// destination halo node) pairs covering every interface, both layers.
__global__ void k_halo_copy(int count, const int *__restrict__ sb, const long *__restrict__ sn,
                            const int *__restrict__ db, const long *__restrict__ dn,
                            double *const *__restrict__ qtab, double *const *__restrict__ nuttab,
                            const long *__restrict__ ntab, int with_nut)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count) return;
    int b0 = sb[x], b1 = db[x];
    long s = sn[x], d = dn[x];
    const double *qs = qtab[b0];
    double *qd = qtab[b1];
    long Ns = ntab[b0], Nd = ntab[b1];
#pragma unroll
    for (int v = 0; v < 5; v++) qd[v * Nd + d] = qs[v * Ns + s];
    if (with_nut) nuttab[b1][d] = nuttab[b0][s];
}

int launch_halo_copy(hipStream_t s, const CopyList &cl, double *const *q_tab, double *const *nut_tab,
                     const long *n_tab, int with_nut)
{
    if (cl.count <= 0) return 0;
    int bs = 256;
    hipLaunchKernelGGL(k_halo_copy, dim3((cl.count + bs - 1) / bs), dim3(bs), 0, s, cl.count, cl.src_blk, cl.src,
                       cl.dst_blk, cl.dst, q_tab, nut_tab, n_tab, with_nut);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// average1: gather the neighbour's face node of every pairing entry.
__global__ void k_avg_gather(int count, const int *__restrict__ sb, const long *__restrict__ sn,
                             double *const *__restrict__ qtab, double *const *__restrict__ nuttab,
                             const long *__restrict__ ntab, double *__restrict__ buf, int with_nut)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= count) return;
    int b0 = sb[x];
    long s = sn[x];
    const double *qs = qtab[b0];
    long Ns = ntab[b0];
#pragma unroll
    for (int v = 0; v < 5; v++) buf[(long)6 * x + v] = qs[v * Ns + s];
    buf[(long)6 * x + 5] = with_nut ? nuttab[b0][s] : 0.0;
}

// average2: for each destination node, apply q = (q + buf)*0.5 for its
// entries in subface order (a node shared by several subfaces is averaged
// once per subface, as the sequential loop does).
__global__ void k_avg_apply(int ngroups, const int *__restrict__ gb, const long *__restrict__ gd,
                            const int *__restrict__ gs, double *const *__restrict__ qtab,
                            double *const *__restrict__ nuttab, const long *__restrict__ ntab,
                            const double *__restrict__ buf, int with_nut)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int b = gb[g];
    long d = gd[g];
    double *qd = qtab[b];
    long Nd = ntab[b];
    double v0 = qd[d], v1 = qd[Nd + d], v2 = qd[2 * Nd + d], v3 = qd[3 * Nd + d], v4 = qd[4 * Nd + d];
    double vn = with_nut ? nuttab[b][d] : 0.0;
    for (int x = gs[g]; x < gs[g + 1]; x++) {
        const double *p = buf + (long)6 * x;
        v0 = (v0 + p[0]) * 0.5;
        v1 = (v1 + p[1]) * 0.5;
        v2 = (v2 + p[2]) * 0.5;
        v3 = (v3 + p[3]) * 0.5;
        v4 = (v4 + p[4]) * 0.5;
        vn = (vn + p[5]) * 0.5;
    }
    qd[d] = v0;
    qd[Nd + d] = v1;
    qd[2 * Nd + d] = v2;
    qd[3 * Nd + d] = v3;
    qd[4 * Nd + d] = v4;
    if (with_nut) nuttab[b][d] = vn;
}

int launch_average(hipStream_t s, const AvgList &al, double *const *q_tab, double *const *nut_tab,
                   const long *n_tab, double *buf, int with_nut)
{
    if (al.count <= 0) return 0;
    int bs = 256;
    hipLaunchKernelGGL(k_avg_gather, dim3((al.count + bs - 1) / bs), dim3(bs), 0, s, al.count, al.src_blk, al.src,
                       q_tab, nut_tab, n_tab, buf, with_nut);
    hipLaunchKernelGGL(k_avg_apply, dim3((al.ngroups + bs - 1) / bs), dim3(bs), 0, s, al.ngroups, al.grp_blk,
                       al.grp_dst, al.grp_start, q_tab, nut_tab, n_tab, buf, with_nut);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// Residual partial sums over the computed range (sub_output_residual.f90:32-44):
// 6 sums per workgroup, reduced in a fixed order.
constexpr int kResBlocks = 512;

__device__ inline void residual_body(const Ext &e, int bx, const double *__restrict__ rhs,
                                     const double *__restrict__ dt, const double *__restrict__ inv_j,
                                     const double *__restrict__ sa_rhs, double *__restrict__ partial, int with_sa)
{
    __shared__ double red[6][256];
    const long N = e.n;
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    long total = (Li > 0 && Lj > 0 && Lk > 0) ? (long)Li * Lj * Lk : 0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long x = (long)bx * blockDim.x + threadIdx.x; x < total; x += (long)kResBlocks * blockDim.x) {
        int i = e.is1 + (int)(x % Li);
        int j = e.js1 + (int)((x / Li) % Lj);
        int k = e.ks1 + (int)(x / ((long)Li * Lj));
        long n = e.ix(i, j, k);
        double d = dt[n], ij = inv_j[n];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double y = rhs[v * N + n] * d / ij;
            acc[v] += y * y;
        }
        double y = (with_sa ? sa_rhs[n] : 0.0) * d;
        acc[5] += y * y;
    }
#pragma unroll
    for (int v = 0; v < 6; v++) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int v = 0; v < 6; v++) red[v][threadIdx.x] += red[v][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * bx + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_residual(Ext e, const double *__restrict__ rhs,
                                                  const double *__restrict__ dt,
                                                  const double *__restrict__ inv_j,
                                                  const double *__restrict__ sa_rhs,
                                                  double *__restrict__ partial, int with_sa)
{
    residual_body(e, (int)blockIdx.x, rhs, dt, inv_j, sa_rhs, partial, with_sa);
}

// every block: workgroup x = m * kResBlocks + b writes block m's partial b
// where the per-block launches would have put it
__global__ void __launch_bounds__(256) k_residual_batch(const BlockRec *__restrict__ blks, double *arena,
                                                        double *__restrict__ partial, int with_sa)
{
    const int m = (int)blockIdx.x / kResBlocks, bx = (int)blockIdx.x % kResBlocks;
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    residual_body(e, bx, arena + b.o.rhs, arena + b.o.dt, arena + b.o.inv_j, arena + b.o.sa_rhs,
                  partial + 6L * kResBlocks * m, with_sa);
}

int residual_partials(const Ext &) { return kResBlocks; }

int launch_residual(hipStream_t s, const Ext &e, const Fields &f, double *partial, int with_sa)
{
    hipLaunchKernelGGL(k_residual, dim3(kResBlocks), dim3(256), 0, s, e, f.rhs, f.dt, f.inv_j, f.sa_rhs, partial,
                       with_sa);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// One workgroup of kResFinal threads.  A fixed order, so the sums are
// reproducible and the same for every launch form that lays the partials
// out alike: thread t adds partials t, t + kResFinal, ... in turn, then a
// pairwise tree over the threads in LDS.
constexpr int kResFinal = 256;
__global__ void __launch_bounds__(kResFinal) k_residual_final(const double *__restrict__ partial, int count,
                                                              double *__restrict__ out)
{
    __shared__ double red[6][kResFinal];
    const int t = threadIdx.x;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int x = t; x < count; x += kResFinal) {
#pragma unroll
        for (int v = 0; v < 6; v++) s[v] += partial[6L * x + v];
    }
#pragma unroll
    for (int v = 0; v < 6; v++) red[v][t] = s[v];
    __syncthreads();
    for (int h = kResFinal / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int v = 0; v < 6; v++) red[v][t] += red[v][t + h];
        }
        __syncthreads();
    }
    if (t < 6) out[t] = red[t][0];
}

int launch_residual_batch(hipStream_t s, const BlockRec *blks, int nblk, double *arena, double *partial, int with_sa)
{
    hipLaunchKernelGGL(k_residual_batch, dim3((unsigned)(nblk * kResBlocks)), dim3(256), 0, s, blks, arena, partial,
                       with_sa);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_residual_final(hipStream_t s, const double *partial, int count, double *out)
{
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kResFinal), 0, s, partial, count, out);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ----------------------------------------------------------- transposition
__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ soa, long n, int nc)
{
    long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n * nc) return;
    long node = x / nc;
    int c = (int)(x % nc);
    soa[c * n + node] = aos[x];
}

__global__ void k_soa_to_aos(const double *__restrict__ soa, double *__restrict__ aos, long n, int nc)
{
    long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n * nc) return;
    long node = x / nc;
    int c = (int)(x % nc);
    aos[x] = soa[c * n + node];
}

int launch_aos_to_soa(hipStream_t s, const double *aos, double *soa, long n, int ncomp)
{
    long tot = n * ncomp;
    int bs = 256;
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((tot + bs - 1) / bs)), dim3(bs), 0, s, aos, soa, n, ncomp);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_soa_to_aos(hipStream_t s, const double *soa, double *aos, long n, int ncomp)
{
    long tot = n * ncomp;
    int bs = 256;
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((tot + bs - 1) / bs)), dim3(bs), 0, s, soa, aos, n, ncomp);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
