// Grid pre-processing on the device: metrics, edge lengths and wall
// distance, the path's geometric inputs.
//
// Reference: sub_init_gridtransformation.f90:27-874 (3-D branch),
// sub_init_spacing.f90:16-110, sub_init_walldist.f90:9-85.
//
// The reference evaluates the coordinate derivatives, 1/J and the metric
// terms on the block box and its six face halo slabs (the nodes with at
// most one coordinate outside the physical range), with central
// differences, one-sided at the ends of the line: the extended line when
// the node's two other coordinates are physical, the physical range on a
// halo slab.  The edge and corner halo nodes are left to uninitialised
// temporaries there; here they get zero (no stencil of the path reads
// them).
#include "common.hpp"
#include "kernels.hpp"

// Bitwise agreement with the reference: every expression in its order, no
// contraction.
#pragma clang fp contract(off)

namespace acfd {

__device__ inline bool outside(const Ext &e, int a, int p)
{
    return p < e.lo0(a) || p > e.hi0(a);
}

// Derivative along axis a of the scalar field v at node (i,j,k) under the
// reference's region rule; false where it is not defined.
__device__ inline bool grid_deriv(const Ext &e, const double *__restrict__ v, int i, int j, int k, int a, double &out)
{
    const int p[3] = {i, j, k};
    const int others = (a != 0 && outside(e, 0, i)) + (a != 1 && outside(e, 1, j)) + (a != 2 && outside(e, 2, k));
    int lo, hi;
    if (others == 0) {
        lo = e.lo(a);
        hi = e.hi(a);
    } else if (others == 1) {
        lo = e.lo0(a);
        hi = e.hi0(a);
        if (p[a] < lo || p[a] > hi) return false;
    } else {
        return false;
    }
    const long n = e.ix(i, j, k), st = e.stride(a);
    if (p[a] == lo) out = v[n + st] - v[n];
    else if (p[a] == hi) out = v[n] - v[n - st];
    else out = (v[n + st] - v[n - st]) / 2.0;
    return true;
}

__device__ inline int halo_count(const Ext &e, int i, int j, int k)
{
    return outside(e, 0, i) + outside(e, 1, j) + outside(e, 2, k);
}

// Pass 1: dx/dxi, dx/deta, dx/dzeta (d[3a + c] = d x_c / d xi_a, 9 fields of
// stride n) and 1/J (:103-529).
__global__ void __launch_bounds__(256) k_grid_d1(Ext e, const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, double *__restrict__ d,
                                                 double *__restrict__ inv_j)
{
    const Box b{e.is, e.ie, e.js, e.je, e.ks, e.ke};
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n, n = e.ix(i, j, k);
    const double *xyz[3] = {x, y, z};
    double g[9];
    const bool ok = halo_count(e, i, j, k) <= 1;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double v = 0.0;
            if (ok) grid_deriv(e, xyz[c], i, j, k, a, v);
            g[3 * a + c] = v;
            d[(3 * a + c) * N + n] = v;
        }
    const double dxdxi0 = g[0], dydxi0 = g[1], dzdxi0 = g[2];
    const double dxdeta0 = g[3], dydeta0 = g[4], dzdeta0 = g[5];
    const double dxdzeta0 = g[6], dydzeta0 = g[7], dzdzeta0 = g[8];
    inv_j[n] = ok ? dxdxi0 * dydeta0 * dzdzeta0 + dxdeta0 * dydzeta0 * dzdxi0 + dxdzeta0 * dydxi0 * dzdeta0
                        - dxdzeta0 * dydeta0 * dzdxi0 - dxdxi0 * dydzeta0 * dzdeta0 - dxdeta0 * dydxi0 * dzdzeta0
                  : 0.0;
}

// The temporaries of :560-585 at node n: tx (tempxi), te (tempeta), tz
// (tempzeta), 6 values each.
__device__ inline void grid_temps(const double *__restrict__ d, long N, long n, double X, double Y, double Z,
                                  int which, double *t)
{
    // which = 0: tempxi (dzeta, deta); 1: tempeta (dzeta, dxi); 2: tempzeta (deta, dxi)
    const int A = which == 2 ? 1 : 2;  // first derivative set (zeta or eta)
    const int B = which == 0 ? 1 : 0;  // second (eta or xi)
    t[0] = d[(3 * A + 1) * N + n] * Z;
    t[1] = d[(3 * A + 2) * N + n] * X;
    t[2] = d[(3 * A + 0) * N + n] * Y;
    t[3] = d[(3 * B + 1) * N + n] * Z;
    t[4] = d[(3 * B + 2) * N + n] * X;
    t[5] = d[(3 * B + 0) * N + n] * Y;
}

// Derivative along axis a of temporary set `which` (a == which) at node
// (i,j,k) under the region rule (:587-807).
__device__ inline void grid_temp_deriv(const Ext &e, const double *__restrict__ x, const double *__restrict__ y,
                                       const double *__restrict__ z, const double *__restrict__ d, int i, int j,
                                       int k, int a, double *out)
{
    const int p[3] = {i, j, k};
    const int others = (a != 0 && outside(e, 0, i)) + (a != 1 && outside(e, 1, j)) + (a != 2 && outside(e, 2, k));
    const int lo = others == 0 ? e.lo(a) : e.lo0(a);
    const int hi = others == 0 ? e.hi(a) : e.hi0(a);
    const long N = e.n, n = e.ix(i, j, k), st = e.stride(a);
    double t0[6], t1[6];
    if (p[a] == lo) {
        grid_temps(d, N, n + st, x[n + st], y[n + st], z[n + st], a, t1);
        grid_temps(d, N, n, x[n], y[n], z[n], a, t0);
#pragma unroll
        for (int c = 0; c < 6; c++) out[c] = t1[c] - t0[c];
    } else if (p[a] == hi) {
        grid_temps(d, N, n, x[n], y[n], z[n], a, t1);
        grid_temps(d, N, n - st, x[n - st], y[n - st], z[n - st], a, t0);
#pragma unroll
        for (int c = 0; c < 6; c++) out[c] = t1[c] - t0[c];
    } else {
        grid_temps(d, N, n + st, x[n + st], y[n + st], z[n + st], a, t1);
        grid_temps(d, N, n - st, x[n - st], y[n - st], z[n - st], a, t0);
#pragma unroll
        for (int c = 0; c < 6; c++) out[c] = (t1[c] - t0[c]) / 2.0;
    }
}

// Pass 2: dxidx and alpha/beta/gamma (:809-866).
__global__ void __launch_bounds__(256) k_grid_d2(Ext e, const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const double *__restrict__ d,
                                                 const double *__restrict__ inv_j, double *__restrict__ dxidx,
                                                 double *__restrict__ aal, double *__restrict__ bbe,
                                                 double *__restrict__ gga)
{
    const Box b{e.is, e.ie, e.js, e.je, e.ks, e.ke};
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n, n = e.ix(i, j, k);
    double m[9];
    const double ij = inv_j[n];
    if (halo_count(e, i, j, k) <= 1) {
        double txx[6], tee[6], tzz[6];
        grid_temp_deriv(e, x, y, z, d, i, j, k, 0, txx);
        grid_temp_deriv(e, x, y, z, d, i, j, k, 1, tee);
        grid_temp_deriv(e, x, y, z, d, i, j, k, 2, tzz);
        m[0] = (tzz[0] - tee[0]) / ij;
        m[1] = (tzz[1] - tee[1]) / ij;
        m[2] = (tzz[2] - tee[2]) / ij;
        m[3] = (txx[0] - tzz[3]) / ij;
        m[4] = (txx[1] - tzz[4]) / ij;
        m[5] = (txx[2] - tzz[5]) / ij;
        m[6] = (tee[3] - txx[3]) / ij;
        m[7] = (tee[4] - txx[4]) / ij;
        m[8] = (tee[5] - txx[5]) / ij;
    } else {
#pragma unroll
        for (int c = 0; c < 9; c++) m[c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < 9; c++) dxidx[c * N + n] = m[c];
    if (aal) {
        double p[9];
        sa_metric_products(m, ij, p);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            aal[c * N + n] = p[c];
            bbe[c * N + n] = p[3 + c];
            gga[c * N + n] = p[6 + c];
        }
    }
}

// sub_init_spacing.f90:37-104 on the physical range: edge lengths (the last
// node along a line takes its neighbour's) and their maximum.
__global__ void __launch_bounds__(256) k_grid_spacing(Ext e, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      double *__restrict__ spc, double *__restrict__ spacing)
{
    const Box b{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0};
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    const int p[3] = {i, j, k};
    double s[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a == 2 && !(e.ke0 > e.ks0)) {
            s[a] = 0.0;
            continue;
        }
        const long st = e.stride(a);
        const long n = e.ix(i, j, k) - (p[a] == e.hi0(a) ? st : 0);
        const double dx = x[n + st] - x[n], dy = y[n + st] - y[n], dz = z[n + st] - z[n];
        s[a] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const long n = e.ix(i, j, k);
#pragma unroll
    for (int a = 0; a < 3; a++) spc[a * N + n] = s[a];
    if (spacing) {
        double m = s[0];
        if (s[1] > m) m = s[1];
        if (s[2] > m) m = s[2];
        spacing[n] = m;
    }
}

// sub_init_walldist.f90:30-85: the nearest of nw wall points (SoA x, y, z
// of stride nw), brute force with the wall points staged through LDS.  The
// minimum of the squared distances is taken and rooted once: the square
// root is monotone and correctly rounded, so this equals the minimum of the
// rooted distances the reference compares.
constexpr int kWallTile = kWallChunk;

// workgroup-wide min (MAXOP false) or max of v, every thread gets the result
template <bool MAXOP>
__device__ inline double wg_reduce(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kWallTile / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double a = red[threadIdx.x], c = red[threadIdx.x + o];
            red[threadIdx.x] = MAXOP ? (c > a ? c : a) : (c < a ? c : a);
        }
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// squared distance between two axis-aligned boxes (0 where they overlap)
__device__ inline double box_gap2(const double *a, const double *b)
{
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double g = fmax(0.0, fmax(b[d] - a[3 + d], a[d] - b[3 + d]));
        s += g * g;
    }
    return s;
}

// Wall points are taken in chunks of kWallTile (cbox: each chunk's bounding
// box, lo xyz then hi xyz).  A workgroup bounds its own nodes by a box,
// takes first the chunk nearest that box, then every chunk whose box could
// hold a point nearer than the largest distance its nodes have found; a
// chunk it skips holds no point nearer to any of its nodes than what they
// hold already (the test keeps a relative margin of 1e-12 over the rounding
// of the box gap), so the minimum is the one over all points.
__global__ void __launch_bounds__(kWallTile) k_walldist(Ext e, const double *__restrict__ x,
                                                        const double *__restrict__ y, const double *__restrict__ z,
                                                        const double *__restrict__ wall, long nw,
                                                        const double *__restrict__ cbox, int nchunk,
                                                        double *__restrict__ dist)
{
    __shared__ double wx[kWallTile], wy[kWallTile], wz[kWallTile], red[kWallTile];
    const long L = (long)(e.ie0 - e.is0 + 1) * (e.je0 - e.js0 + 1) * (e.ke0 - e.ks0 + 1);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = t < L;
    long n = 0;
    double xx = 0.0, yy = 0.0, zz = 0.0;
    if (ok) {
        const int li = e.ie0 - e.is0 + 1, lj = e.je0 - e.js0 + 1;
        const int i = e.is0 + (int)(t % li), j = e.js0 + (int)((t / li) % lj), k = e.ks0 + (int)(t / ((long)li * lj));
        n = e.ix(i, j, k);
        xx = x[n];
        yy = y[n];
        zz = z[n];
    }
    double best = 1.0e18;  // (1e9)^2: the reference's distance with no wall
    if (nw > 0) {
        // the workgroup's box
        double wb[6];
        wb[0] = wg_reduce<false>(ok ? xx : 1.0e300, red);
        wb[1] = wg_reduce<false>(ok ? yy : 1.0e300, red);
        wb[2] = wg_reduce<false>(ok ? zz : 1.0e300, red);
        wb[3] = wg_reduce<true>(ok ? xx : -1.0e300, red);
        wb[4] = wg_reduce<true>(ok ? yy : -1.0e300, red);
        wb[5] = wg_reduce<true>(ok ? zz : -1.0e300, red);
        // the nearest chunk (smallest box gap, lowest index among equals)
        double g0 = 1.0e308;
        int c0 = 0;
        for (int c = threadIdx.x; c < nchunk; c += kWallTile) {
            const double g = box_gap2(wb, cbox + 6L * c);
            if (g < g0) {
                g0 = g;
                c0 = c;
            }
        }
        const double gmin = wg_reduce<false>(g0, red);
        const double cand = g0 == gmin ? (double)c0 : 1.0e300;
        const int first = (int)wg_reduce<false>(cand, red);
        auto take = [&](int c) {
            const long p0 = (long)c * kWallTile;
            const long p = p0 + threadIdx.x;
            if (p < nw) {
                wx[threadIdx.x] = wall[p];
                wy[threadIdx.x] = wall[nw + p];
                wz[threadIdx.x] = wall[2 * nw + p];
            }
            __syncthreads();
            const int cnt = (int)min((long)kWallTile, nw - p0);
            for (int q = 0; q < cnt; q++) {
                const double dx = xx - wx[q], dy = yy - wy[q], dz = zz - wz[q];
                const double d2 = dx * dx + dy * dy + dz * dz;
                best = d2 < best ? d2 : best;
            }
            __syncthreads();
        };
        take(first);
        double bound = wg_reduce<true>(ok ? best : 0.0, red);
        int taken = 0;
        for (int c = 0; c < nchunk; c++) {
            if (c == first || box_gap2(wb, cbox + 6L * c) > bound * (1.0 + 1.0e-12)) continue;
            take(c);
            if (++taken % 16 == 0) bound = wg_reduce<true>(ok ? best : 0.0, red);
        }
    }
    if (ok) dist[n] = nw > 0 ? sqrt(best) : 1.0e9;
}

int launch_grid_metrics(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z, double *scratch9,
                        double *inv_j, double *dxidx, double *aal, double *bbe, double *gga)
{
    const dim3 g = box_grid(Box{e.is, e.ie, e.js, e.je, e.ks, e.ke}), t(64, 4, 1);
    hipLaunchKernelGGL(k_grid_d1, g, t, 0, s, e, x, y, z, scratch9, inv_j);
    hipLaunchKernelGGL(k_grid_d2, g, t, 0, s, e, x, y, z, scratch9, inv_j, dxidx, aal, bbe, gga);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_grid_spacing(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z, double *spc,
                        double *spacing)
{
    hipLaunchKernelGGL(k_grid_spacing, box_grid(Box{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0}), dim3(64, 4, 1), 0,
                       s, e, x, y, z, spc, spacing);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_walldist(hipStream_t s, const Ext &e, const double *x, const double *y, const double *z,
                    const double *wall, long nw, const double *cbox, int nchunk, double *dist)
{
    const long L = (long)(e.ie0 - e.is0 + 1) * (e.je0 - e.js0 + 1) * (e.ke0 - e.ks0 + 1);
    if (L <= 0) return 0;
    hipLaunchKernelGGL(k_walldist, dim3((unsigned)((L + kWallTile - 1) / kWallTile)), dim3(kWallTile), 0, s, e, x, y,
                       z, wall, nw, cbox, nchunk, dist);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd

namespace acfd {

// ------------------------------------------------------------ wall forces
// clcd_output (sub_output_forceintegration.f90:182-509, 3-D).  One thread
// per wall cell evaluates the cell's pressure and viscous force exactly as
// the reference does (velocity from the first interior layer, pressure and
// viscosity on the wall, each averaged over the cell's four nodes; the
// face-normal vector averaged from the unit metric of the four interior
// nodes with the reference's mix of components: the second pair of nodes
// takes the next component).  Six threads then add the cells up in the
// reference's order, so the sums are the reference's to the last bit.
__global__ void __launch_bounds__(256) k_wall_forces(const BlockRec *__restrict__ blks, const WallCell *__restrict__ cells,
                                                     int ncell, double *arena, double *geo,
                                                     const long *__restrict__ geo_off, double pinf, double re,
                                                     double *__restrict__ f)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (x >= ncell) return;
    const WallCell cl = cells[x];
    const BlockRec &b = blks[cl.m];
    const Ext e = b.e;
    const long N = e.n;
    const double *pri = arena + b.o.pri, *dxidx = arena + b.o.dxidx, *amu = arena + b.o.amu;
    const double *spc = geo + geo_off[cl.m] + 3 * N;  // spci, spcj, spck
    const int d = (cl.face - 1) / 2, c0 = 3 * d;
    const double sg = (cl.face & 1) ? -1.0 : 1.0;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int in = (cl.face & 1) ? cl.wall + 1 : cl.wall - 1;
    long W[4], I[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int p[3];
        p[d] = cl.wall;
        p[t1] = cl.u + (c & 1);
        p[t2] = cl.v + (c >> 1);
        W[c] = e.ix(p[0], p[1], p[2]);
        p[d] = in;
        I[c] = e.ix(p[0], p[1], p[2]);
    }
    const double uc = (pri[N + I[0]] + pri[N + I[1]] + pri[N + I[2]] + pri[N + I[3]]) * 0.25;
    const double vc = (pri[2 * N + I[0]] + pri[2 * N + I[1]] + pri[2 * N + I[2]] + pri[2 * N + I[3]]) * 0.25;
    const double wc = (pri[3 * N + I[0]] + pri[3 * N + I[1]] + pri[3 * N + I[2]] + pri[3 * N + I[3]]) * 0.25;
    const double pc = (pri[4 * N + W[0]] + pri[4 * N + W[1]] + pri[4 * N + W[2]] + pri[4 * N + W[3]]) * 0.25;
    double t[4], M[4][3];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int k = 0; k < 3; k++) M[c][k] = dxidx[(c0 + k) * N + I[c]];
        t[c] = 1.0 / sqrt(M[c][0] * M[c][0] + M[c][1] * M[c][1] + M[c][2] * M[c][2]);
    }
    const double nx = sg * (M[0][0] * t[0] + M[1][0] * t[1] + M[2][1] * t[2] + M[3][1] * t[3]) * 0.25;
    const double ny = sg * (M[0][2] * t[0] + M[1][2] * t[1] + M[2][0] * t[2] + M[3][0] * t[3]) * 0.25;
    const double nz = sg * (M[0][1] * t[0] + M[1][1] * t[1] + M[2][2] * t[2] + M[3][2] * t[3]) * 0.25;
    const double miuc = (amu[2 * N + W[0]] + amu[2 * N + W[1]] + amu[2 * N + W[2]] + amu[2 * N + W[3]]) * 0.25;
    const long w0 = W[0];
    const double si = spc[w0], sj = spc[N + w0], sk = spc[2 * N + w0];
    const double area = d == 0 ? sj * sk : d == 1 ? si * sk : si * sj;
    const double dn = d == 0 ? si : d == 1 ? sj : sk;
    // force_compute (:481-509)
    const double un = uc * nx + vc * ny + wc * nz;
    double *o = f + 6L * x;
    o[0] = (pc - pinf) * area * nx;
    o[1] = (pc - pinf) * area * ny;
    o[2] = (pc - pinf) * area * nz;
    o[3] = miuc * (uc - un * nx) / dn * area / re;
    o[4] = miuc * (vc - un * ny) / dn * area / re;
    o[5] = miuc * (wc - un * nz) / dn * area / re;
}

// out[c] = sum over cells of f[6 x + c], in cell order
__global__ void k_wall_sum(const double *__restrict__ f, int ncell, double *__restrict__ out)
{
    const int c = (int)threadIdx.x;
    if (c >= 6) return;
    double s = 0.0;
#pragma unroll 16
    for (int x = 0; x < ncell; x++) s = s + f[6L * x + c];
    out[c] = s;
}

int launch_wall_forces(hipStream_t s, const BlockRec *blks, const WallCell *cells, int ncell, double *arena,
                       double *geo, const long *geo_off, double pinf, double re, double *work, double *out6)
{
    if (ncell > 0)
        hipLaunchKernelGGL(k_wall_forces, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, blks, cells, ncell,
                           arena, geo, geo_off, pinf, re, work);
    hipLaunchKernelGGL(k_wall_sum, dim3(1), dim3(64), 0, s, work, ncell, out6);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
