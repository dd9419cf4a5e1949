// Residual kernels: primitive variables, WENO3 + Roe inviscid flux,
// second-order central viscous flux, local time step.
//
// Reference routines: sub_getpri_v.f90:19-63, sub_inviscid1.f90:8-189,
// sub_inviscid_upwind.f90:1-180, sub_inviscid_splitting.f90:1-172,
// sub_viscous.f90:9-378, sub_timeadvancement_finddt.f90:24-130.
//
// The reconstruction and splitting are templated on the arithmetic type R:
// R = double reproduces the reference in FP64, R = float is the FP32-flux
// variant (inputs and the accumulated right-hand side stay FP64).
#include "batch.hpp"
#include "common.hpp"
#include "flux_common.hpp"
#include "kernels.hpp"

namespace acfd {

// ------------------------------------------------------------------ getpri
// sub_getpri_v.f90:33-59 over the whole extended block.  Counts nodes with
// negative density into *nneg.
__device__ inline void getpri_body(const Ext &e, int bx, const double *__restrict__ q, double *__restrict__ pri,
                                   double gamma, double cv, unsigned int *nneg)
{
    long n = (long)bx * blockDim.x + threadIdx.x;
    if (n >= e.n) return;
    const long N = e.n;
    double d = q[n];
    double u = q[N + n] / d, v = q[2 * N + n] / d, w = q[3 * N + n] / d;
    double q2 = 0.5 * (u * u + v * v + w * w);
    double en = q[4 * N + n] / d;
    double t = (en - q2) / cv;
    double p = (q[4 * N + n] - q2 * d) * (gamma - 1.0);
    double c = sqrt(gamma * p / d);
    pri[n] = d;
    pri[N + n] = u;
    pri[2 * N + n] = v;
    pri[3 * N + n] = w;
    pri[4 * N + n] = p;
    pri[5 * N + n] = t;
    pri[6 * N + n] = c;
    if (d < 0.0) atomicAdd(nneg, 1u);
}

__global__ void __launch_bounds__(256) k_getpri(Ext e, const double *__restrict__ q, double *__restrict__ pri,
                                                double gamma, double cv, unsigned int *nneg)
{
    getpri_body(e, (int)blockIdx.x, q, pri, gamma, cv, nneg);
}

__global__ void __launch_bounds__(256) k_getpri_batch(const BlockRec *__restrict__ blks,
                                                      const int *__restrict__ map, const KGrid *__restrict__ kg,
                                                      double *arena, double gamma, double cv, unsigned int *nneg)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const Ext e = b.e;
    getpri_body(e, bx, arena + b.o.q, arena + b.o.pri, gamma, cv, nneg);
}

dim3 grid_1d(long n) { return dim3((unsigned)((n + 255) / 256)); }

int launch_getpri(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, unsigned int *nneg)
{
    hipLaunchKernelGGL(k_getpri, grid_1d(e.n), dim3(256), 0, s, e, f.q, f.pri, c.gamma, c.cv, nneg);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_getpri_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                        unsigned int *nneg)
{
    if (bg.total <= 0) return 0;
    hipLaunchKernelGGL(k_getpri_batch, dim3((unsigned)bg.total), dim3(256), 0, s, blks, bg.map, bg.kg, arena,
                       c.gamma, c.cv, nneg);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------ fill helper
__global__ void k_fill(double *p, long n, double v)
{
    long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) p[x] = v;
}

// rhsi = 0 on every block (1-D grid over 5n per block)
__global__ void __launch_bounds__(256) k_zero_rhsi_batch(const BlockRec *__restrict__ blks,
                                                         const int *__restrict__ map,
                                                         const KGrid *__restrict__ kg, double *arena)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const long x = (long)bx * blockDim.x + threadIdx.x;
    if (x < 5 * b.e.n) arena[b.o.rhsi + x] = 0.0;
}

int launch_fill(hipStream_t s, double *p, long n, double v)
{
    if (n <= 0) return 0;
    int bs = 256;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, p, n, v);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------ WENO3 + Roe
// weno3, roe_flux and cons2prim are in flux_common.hpp.

// Flux through the face between nodes a and a+1 of one line from the four
// values w[0..3] = q at a-1, a, a+1, a+2 of every component (alo / ahi are
// the extended range ends: first order there) and the face metric.
template <typename R>
__device__ inline void face_flux_w(const R (*w)[4], int a, int alo, int ahi, R mx, R my, R mz, R gamma,
                                   R *F)
{
    R ul[5], ur[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        ul[v] = a == alo ? w[v][1] : weno3<R>(w[v][0], w[v][1], w[v][2]);
        ur[v] = a == ahi - 1 ? w[v][2] : weno3<R>(w[v][3], w[v][2], w[v][1]);
    }
    R pl[5], pr[5];
    cons2prim<R>(ul, pl, gamma);
    cons2prim<R>(ur, pr, gamma);
    roe_flux<R>(pl, pr, mx, my, mz, gamma, F);
}

// face metric 0.5 (grad xi_d / J at a + at a+1), sub_inviscid_splitting.f90:65-67
__device__ inline void face_metric(const double *__restrict__ dxidx, const double *__restrict__ inv_j, long N,
                                   long n, long st, int c0, double *m)
{
    const double j0 = inv_j[n], j1 = inv_j[n + st];
#pragma unroll
    for (int c = 0; c < 3; c++) m[c] = 0.5 * (dxidx[(c0 + c) * N + n] * j0 + dxidx[(c0 + c) * N + n + st] * j1);
}

// The i direction (sub_inviscid1.f90:113-145): a wavefront of 64 lanes along
// i evaluates the faces a+1/2, a = i0-1 .. i0+62, once each; lane L > 0 takes
// F(a-1/2) from lane L-1, so a wave produces 63 nodes.  rhsi = -(F+ - F-).
constexpr int kInvI = 63;
template <typename R>
__device__ inline void inviscid_i_body(const Ext &e, int bx, int by, int bz, const double *__restrict__ q,
                                       const double *__restrict__ dxidx, const double *__restrict__ inv_j,
                                       double *__restrict__ rhsi, double gamma)
{
    const int lane = threadIdx.x;
    const int j = e.js1 + (int)(by * blockDim.y + threadIdx.y);
    const int k = e.ks1 + bz;
    const int a_raw = e.is1 - 1 + bx * kInvI + lane;
    const bool row = j <= e.je1;
    // lanes past the range evaluate a valid face and discard it, so that
    // every lane takes part in the exchange
    const int a = min(a_raw, e.ie1);
    const long N = e.n;
    const long n = e.ix(a, row ? j : e.je1, k);
    // the outer values of the window are not used at the extended range
    // ends (first-order states there) and would lie outside the block
    const long nm = a > e.is ? n - 1 : n, np = a + 2 <= e.ie ? n + 2 : n + 1;
    R w[5][4];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double *qv = q + v * N;
        w[v][0] = (R)qv[nm];
        w[v][1] = (R)qv[n];
        w[v][2] = (R)qv[n + 1];
        w[v][3] = (R)qv[np];
    }
    double m[3];
    face_metric(dxidx, inv_j, N, n, 1, 0, m);
    R F[5];
    face_flux_w<R>(w, a, e.is, e.ie, (R)m[0], (R)m[1], (R)m[2], (R)gamma, F);
    double Fm[5];
#pragma unroll
    for (int v = 0; v < 5; v++) Fm[v] = __shfl_up((double)F[v], 1, 64);
    if (row && lane > 0 && a_raw <= e.ie1) {
#pragma unroll
        for (int v = 0; v < 5; v++) rhsi[v * N + n] = -((double)F[v] - Fm[v]);
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_inviscid_i(Ext e, const double *__restrict__ q,
                                                   const double *__restrict__ dxidx,
                                                   const double *__restrict__ inv_j,
                                                   double *__restrict__ rhsi, double gamma)
{
    inviscid_i_body<R>(e, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, q, dxidx, inv_j, rhsi, gamma);
}

// The j and k directions (:160-243): each thread marches along the axis
// over kInvRun nodes, keeping the four-value window and the previous face
// flux in registers, so each face is evaluated once.  rhsi -= (F+ - F-).
constexpr int kInvRun = 16;
// Waves per SIMD asked of the k kernel: FP64 it needs 176 VGPRs (two
// waves); held to 168 with a few values in scratch it runs three, and on
// the 40-block channel takes 3.84 ms instead of 4.32.  The same bound on
// the j kernel (four waves: 22 values in scratch) or four on the k kernel
// is slower by a third or more.
constexpr int kInvKWaves = 3;
template <typename R, int AXIS>
__device__ inline void inviscid_jk_body(const Ext &e, int bx, int by, int bz, const double *__restrict__ q,
                                        const double *__restrict__ dxidx, const double *__restrict__ inv_j,
                                        double *__restrict__ rhsi, double gamma)
{
    const int a0 = e.lo1(AXIS) + bz * kInvRun;
    const int a1 = min(a0 + kInvRun - 1, e.hi1(AXIS));
    int i, j, k;
    if (AXIS == 1) {
        i = e.is1 + (int)(bx * blockDim.x + threadIdx.x);
        j = a0;
        k = e.ks1 + (int)(by * blockDim.y + threadIdx.y);
        if (i > e.ie1 || k > e.ke1) return;
    } else {
        // a k-plane is contiguous in storage: the 256 lanes of workgroup bx
        // take 256 consecutive positions of it (whole cache lines whatever
        // the row length); positions outside the computed range idle
        const int p = bx * 256 + (int)(threadIdx.x + blockDim.x * threadIdx.y);
        i = e.is + p % e.ni;
        j = e.js + p / e.ni;
        k = a0;
        if (i < e.is1 || i > e.ie1 || j < e.js1 || j > e.je1) return;
    }
    const long N = e.n, st = e.stride(AXIS);
    const int alo = e.lo(AXIS), ahi = e.hi(AXIS);
    const int c0 = 3 * AXIS;
    long n = e.ix(i, j, k) - st;  // node a0-1: first face is (a0-1)+1/2
    // window values outside the extended range are never used (first-order
    // states at its ends); their loads are redirected inside the block
    const long back = a0 - 2 >= alo ? st : 0;
    R w[5][4];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double *qv = q + v * N + n;
        w[v][0] = (R)qv[-back];
        w[v][1] = (R)qv[0];
        w[v][2] = (R)qv[st];
        w[v][3] = (R)qv[2 * st];
    }
    R Fm[5], Fp[5];
    {
        double m[3];
        face_metric(dxidx, inv_j, N, n, st, c0, m);
        face_flux_w<R>(w, a0 - 1, alo, ahi, (R)m[0], (R)m[1], (R)m[2], (R)gamma, Fm);
    }
    for (int a = a0; a <= a1; a++) {
        n += st;
        const long ahead = a + 2 <= ahi ? 2 * st : st;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            w[v][0] = w[v][1];
            w[v][1] = w[v][2];
            w[v][2] = w[v][3];
            w[v][3] = (R)q[v * N + n + ahead];
        }
        double m[3];
        face_metric(dxidx, inv_j, N, n, st, c0, m);
        face_flux_w<R>(w, a, alo, ahi, (R)m[0], (R)m[1], (R)m[2], (R)gamma, Fp);
#pragma unroll
        for (int v = 0; v < 5; v++) {
            rhsi[v * N + n] = rhsi[v * N + n] - ((double)Fp[v] - (double)Fm[v]);
            Fm[v] = Fp[v];
        }
    }
}

template <typename R, int AXIS>
__global__ void __launch_bounds__(256, AXIS == 2 ? kInvKWaves : 1) k_inviscid_jk(Ext e, const double *__restrict__ q,
                                                    const double *__restrict__ dxidx,
                                                    const double *__restrict__ inv_j,
                                                    double *__restrict__ rhsi, double gamma)
{
    inviscid_jk_body<R, AXIS>(e, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, q, dxidx, inv_j, rhsi, gamma);
}

// AXIS 0 = the i kernel, 1 / 2 = the j / k kernels.  The k kernel runs at
// three waves per SIMD (kInvKWaves).
template <typename R, int AXIS>
__global__ void __launch_bounds__(256, AXIS == 2 ? kInvKWaves : 1) k_inviscid_batch(const BlockRec *__restrict__ blks,
                                                        const int *__restrict__ map,
                                                        const KGrid *__restrict__ kg, double *arena, double gamma)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const Ext e = b.e;
    if (AXIS == 0)
        inviscid_i_body<R>(e, bx, by, bz, arena + b.o.q, arena + b.o.dxidx, arena + b.o.inv_j, arena + b.o.rhsi,
                           gamma);
    else
        inviscid_jk_body<R, AXIS>(e, bx, by, bz, arena + b.o.q, arena + b.o.dxidx, arena + b.o.inv_j,
                                  arena + b.o.rhsi, gamma);
}

// per-block grids of the three inviscid kernels (64 x 4 threads)
dim3 inviscid_grid(const Ext &e, int axis)
{
    const int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    if (Li <= 0 || Lj <= 0 || Lk <= 0) return dim3(0, 0, 0);
    if (axis == 0) return dim3((Li + kInvI - 1) / kInvI, (Lj + 3) / 4, Lk);
    if (axis == 1) return dim3((Li + 63) / 64, (Lk + 3) / 4, (Lj + kInvRun - 1) / kInvRun);
    // k: linear positions over the whole (extended) i-j plane
    return dim3((unsigned)(((long)e.ni * e.nj + 255) / 256), 1, (Lk + kInvRun - 1) / kInvRun);
}

template <typename R>
static void launch_inviscid_t(hipStream_t s, const Ext &e, const Fields &f, double gamma)
{
    dim3 t(64, 4, 1);
    hipLaunchKernelGGL((k_inviscid_i<R>), inviscid_grid(e, 0), t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi, gamma);
    hipLaunchKernelGGL((k_inviscid_jk<R, 1>), inviscid_grid(e, 1), t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi,
                       gamma);
    hipLaunchKernelGGL((k_inviscid_jk<R, 2>), inviscid_grid(e, 2), t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi,
                       gamma);
}

// WENO3 + Roe residual of every block: bg[0] = rhsi zeroing (1-D over 5n),
// bg[1..3] = the i, j, k kernels
template <typename R>
static void inviscid_batch_dirs(hipStream_t s, const BlockRec *blks, const BatchGrid *bg, double *arena, double gamma,
                                int part)
{
    const dim3 t(64, 4, 1);
    if (part < 0 || part == 1)
        hipLaunchKernelGGL((k_inviscid_batch<R, 0>), dim3((unsigned)bg[1].total), t, 0, s, blks, bg[1].map,
                           bg[1].kg, arena, gamma);
    if (part < 0 || part == 2)
        hipLaunchKernelGGL((k_inviscid_batch<R, 1>), dim3((unsigned)bg[2].total), t, 0, s, blks, bg[2].map,
                           bg[2].kg, arena, gamma);
    if (part < 0 || part == 3)
        hipLaunchKernelGGL((k_inviscid_batch<R, 2>), dim3((unsigned)bg[3].total), t, 0, s, blks, bg[3].map,
                           bg[3].kg, arena, gamma);
}

int launch_inviscid_batch(hipStream_t s, const BlockRec *blks, const BatchGrid *bg, double *arena,
                          const Consts &c, int precision, int part)
{
    if (bg[0].total > 0 && (part == -1 || part == 0))
        hipLaunchKernelGGL(k_zero_rhsi_batch, dim3((unsigned)bg[0].total), dim3(256), 0, s, blks, bg[0].map,
                           bg[0].kg, arena);
    if (bg[1].total <= 0 || part == 0) return 0;
    const int dirs = part < 0 ? -1 : part;
    if (precision == 1) inviscid_batch_dirs<float>(s, blks, bg, arena, c.gamma, dirs);
    else inviscid_batch_dirs<double>(s, blks, bg, arena, c.gamma, dirs);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_inviscid(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision)
{
    // rhsi is zero outside the computed range (sub_inviscid1.f90:102)
    int rc = launch_fill(s, f.rhsi, 5 * e.n, 0.0);
    if (rc) return rc;
    if (e.ie1 < e.is1 || e.je1 < e.js1 || e.ke1 < e.ks1) return 0;
    if (precision == 1) launch_inviscid_t<float>(s, e, f, c.gamma);
    else launch_inviscid_t<double>(s, e, f, c.gamma);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------------ viscous
// Node difference of sub_viscous.f90:107-178: one-sided at the physical
// range ends, central inside; in FP64 whatever the storage type S.
template <typename S>
__device__ inline double ndiff(const S *a, long n, long st, int x, int x0, int x1)
{
    if (x == x0) return (double)a[n + st] - (double)a[n];
    if (x == x1) return (double)a[n] - (double)a[n - st];
    return ((double)a[n + st] - (double)a[n - st]) / 2.0;
}

// Stage 1 (sub_viscous.f90:107-293) at node (i,j,k) of is0..ie0: velocity
// and temperature gradients, Sutherland viscosity, stress, heat flux and
// the contravariant node fluxes fv/gv/hv (components 2..5) into w[12]
// (fv 0-3, gv 4-7, hv 8-11); am1 / am3 are the laminar / total viscosity,
// du the velocity gradient, ta the stress.
template <typename R>
__device__ inline void viscous_node_eval(const Ext &e, int i, int j, int k, const Consts &c,
                                         const double *__restrict__ pri, const double *__restrict__ amu,
                                         const double *__restrict__ dxidx, const double *__restrict__ inv_j,
                                         double *w, R &am1, R &am3, R *du, R *ta)
{
    const long N = e.n;
    long n = e.ix(i, j, k);
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    // gradients along xi, eta, zeta of t,u,v,w (pri components 6,2,3,4)
    R g[4][3];
    const int comp[4] = {5, 1, 2, 3};
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double *a = pri + comp[f] * N;
        g[f][0] = (R)ndiff(a, n, 1, i, e.is0, e.ie0);
        g[f][1] = (R)ndiff(a, n, sj, j, e.js0, e.je0);
        g[f][2] = (R)ndiff(a, n, sk, k, e.ks0, e.ke0);
    }
    R m[9];
#pragma unroll
    for (int x = 0; x < 9; x++) m[x] = (R)dxidx[x * N + n];
    R u0 = (R)pri[N + n], v0 = (R)pri[2 * N + n], w0 = (R)pri[3 * N + n], t0 = (R)pri[5 * N + n];
    // Sutherland; t^1.5 as t sqrt(t) (the reference's t**1.5 to an ulp)
    am1 = (R)((1.0 + 124.0 / c.tinf) / ((double)t0 + 124.0 / c.tinf) * ((double)t0 * sqrt((double)t0)));
    R am2 = (R)ld_once(amu + N + n);
    am3 = am1 + am2;
    R ki = (R)(c.gamma * c.cv) * (am1 / (R)c.prl + am2 / (R)c.prt);
    R tx = g[0][0] * m[0] + g[0][1] * m[3] + g[0][2] * m[6];
    R ty = g[0][0] * m[1] + g[0][1] * m[4] + g[0][2] * m[7];
    R tz = g[0][0] * m[2] + g[0][1] * m[5] + g[0][2] * m[8];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        du[3 * f + 0] = g[f + 1][0] * m[0] + g[f + 1][1] * m[3] + g[f + 1][2] * m[6];
        du[3 * f + 1] = g[f + 1][0] * m[1] + g[f + 1][1] * m[4] + g[f + 1][2] * m[7];
        du[3 * f + 2] = g[f + 1][0] * m[2] + g[f + 1][1] * m[5] + g[f + 1][2] * m[8];
    }
    const R c43 = (R)(4.0 / 3.0), c23 = (R)(2.0 / 3.0);
    ta[0] = am3 * (c43 * du[0] - c23 * (du[4] + du[8]));
    ta[1] = am3 * (du[1] + du[3]);
    ta[2] = am3 * (du[6] + du[2]);
    ta[3] = am3 * (c43 * du[4] - c23 * (du[0] + du[8]));
    ta[4] = am3 * (du[7] + du[5]);
    ta[5] = am3 * (c43 * du[8] - c23 * (du[0] + du[4]));
    R bx = u0 * ta[0] + v0 * ta[1] + w0 * ta[2] + ki * tx;
    R by = u0 * ta[1] + v0 * ta[3] + w0 * ta[4] + ki * ty;
    R bz = u0 * ta[2] + v0 * ta[4] + w0 * ta[5] + ki * tz;
    R ij = (R)ld_once(inv_j + n);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const R *r = m + 3 * d;
        w[4 * d + 0] = (double)((r[0] * ta[0] + r[1] * ta[1] + r[2] * ta[2]) * ij);
        w[4 * d + 1] = (double)((r[0] * ta[1] + r[1] * ta[3] + r[2] * ta[4]) * ij);
        w[4 * d + 2] = (double)((r[0] * ta[2] + r[1] * ta[4] + r[2] * ta[5]) * ij);
        w[4 * d + 3] = (double)((r[0] * bx + r[1] * by + r[2] * bz) * ij);
    }
}

// the node's own outputs besides the fluxes: viscosities, and with
// store_aux the vorticity magnitude (1) plus gradient and stress (2)
template <typename R>
__device__ inline void viscous_node_store(long N, long n, R am1, R am3, const R *du, const R *ta,
                                          double *__restrict__ amu, double *__restrict__ vor,
                                          double *__restrict__ dudx, double *__restrict__ tao, int store_aux)
{
    st_out(amu + n, (double)am1);
    st_out(amu + 2 * N + n, (double)am3);
    if (store_aux) {
        st_out(vor + n, (double)sqrt((du[5] - du[7]) * (du[5] - du[7]) + (du[2] - du[6]) * (du[2] - du[6])
                              + (du[1] - du[3]) * (du[1] - du[3])));
        if (store_aux > 1) {
#pragma unroll
            for (int x = 0; x < 9; x++) st_out(dudx + x * N + n, (double)du[x]);
#pragma unroll
            for (int x = 0; x < 6; x++) st_out(tao + x * N + n, (double)ta[x]);
        }
    }
}

// The node fluxes are stored in the arithmetic type R (FP32 in the FP32-flux
// variant, where they are FP32 values already): w_flux holds 12 components
// of R at component stride n.
template <typename R>
__device__ inline void viscous_node_body(const Ext &e, long wg, const Consts &c,
                                         const double *__restrict__ pri, double *__restrict__ amu,
                                         const double *__restrict__ dxidx, const double *__restrict__ inv_j,
                                         R *__restrict__ w, double *__restrict__ vor, double *__restrict__ dudx,
                                         double *__restrict__ tao, int store_aux)
{
    const Box b{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0};
    int i, j, k;
    long n;
    if (!lin_node_at(e, b, wg, i, j, k, n)) return;
    const long N = e.n;
    double w12[12];
    R am1, am3, du[9], ta[6];
    viscous_node_eval<R>(e, i, j, k, c, pri, amu, dxidx, inv_j, w12, am1, am3, du, ta);
    viscous_node_store<R>(N, n, am1, am3, du, ta, amu, vor, dudx, tao, store_aux);
#pragma unroll
    for (int x = 0; x < 12; x++) w[x * N + n] = (R)w12[x];
}

template <typename R>
__global__ void __launch_bounds__(256) k_viscous_node(Ext e, Consts c, const double *__restrict__ pri,
                                                      double *__restrict__ amu, const double *__restrict__ dxidx,
                                                      const double *__restrict__ inv_j, double *__restrict__ w,
                                                      double *__restrict__ vor, double *__restrict__ dudx,
                                                      double *__restrict__ tao, int store_aux)
{
    viscous_node_body<R>(e, xcd_slab(blockIdx.x, gridDim.x), c, pri, amu, dxidx, inv_j, reinterpret_cast<R *>(w),
                         vor, dudx, tao, store_aux);
}

// Stage 2 (sub_viscous.f90:296-365): rhsv = (d/dxi fv + d/deta gv + d/dzeta hv)/re;
// with rhsi given, also rhs = rhsi + rhsv on the computed range (the
// first line of implicit1st, sub_timeadvancement_1stimplicit.f90:72-78),
// and with rhsv null only rhs
template <typename S>
__device__ inline void viscous_diff_body(const Ext &e, long wg, const S *__restrict__ w,
                                         double *__restrict__ rhsv, double re, const double *__restrict__ rhsi = nullptr,
                                         double *__restrict__ rhs = nullptr)
{
    const Box b{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0};
    int i, j, k;
    long n;
    if (!lin_node_at(e, b, wg, i, j, k, n)) return;
    const long N = e.n;
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    const bool inner = rhs && i >= e.is1 && i <= e.ie1 && j >= e.js1 && j <= e.je1 && k >= e.ks1 && k <= e.ke1;
    if (!rhsv && !inner) return;
    const double v0 = (0.0 + 0.0 + 0.0) / re;
    if (rhsv) st_out(rhsv + n, v0);
    if (inner) st_out(rhs + n, ld_once(rhsi + n) + v0);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double a = ndiff(w + (0 + c) * N, n, 1, i, e.is0, e.ie0);
        double bb = ndiff(w + (4 + c) * N, n, sj, j, e.js0, e.je0);
        double h = ndiff(w + (8 + c) * N, n, sk, k, e.ks0, e.ke0);
        const double v = (a + bb + h) / re;
        if (rhsv) st_out(rhsv + (c + 1) * N + n, v);
        if (inner) st_out(rhs + (c + 1) * N + n, ld_once(rhsi + (c + 1) * N + n) + v);
    }
}

// S: the storage type of the node fluxes
template <typename S>
__global__ void __launch_bounds__(256) k_viscous_diff(Ext e, const double *__restrict__ w,
                                                      double *__restrict__ rhsv, double re)
{
    viscous_diff_body<S>(e, xcd_slab(blockIdx.x, gridDim.x), reinterpret_cast<const S *>(w), rhsv, re);
}

template <typename R>
__global__ void __launch_bounds__(256) k_viscous_node_batch(const BlockRec *__restrict__ blks,
                                                            const int *__restrict__ map,
                                                            const KGrid *__restrict__ kg, double *arena, Consts c,
                                                            int store_aux)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const Ext e = b.e;
    viscous_node_body<R>(e, bx, c, arena + b.o.pri, arena + b.o.amu, arena + b.o.dxidx, arena + b.o.inv_j,
                         reinterpret_cast<R *>(arena + b.o.w_flux), arena + b.o.vor, arena + b.o.dudx, arena + b.o.tao,
                         store_aux);
}

// with_rhs 1: also rhs = rhsi + rhsv on the computed range; 2: rhs only
// (rhsv not stored); S: the storage type of the node fluxes
template <typename S>
__global__ void __launch_bounds__(256) k_viscous_diff_batch(const BlockRec *__restrict__ blks,
                                                            const int *__restrict__ map,
                                                            const KGrid *__restrict__ kg, double *arena, double re,
                                                            int with_rhs = 0)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const Ext e = b.e;
    const S *w = reinterpret_cast<const S *>(arena + b.o.w_flux);
    if (with_rhs)
        viscous_diff_body<S>(e, bx, w, with_rhs == 2 ? nullptr : arena + b.o.rhsv, re, arena + b.o.rhsi,
                             arena + b.o.rhs);
    else viscous_diff_body<S>(e, bx, w, arena + b.o.rhsv, re);
}

// per-block grid of both viscous kernels: linear tiles of 256 nodes over the
// block's storage (the physical range is masked inside), 64 x 4 threads
dim3 viscous_grid(const Ext &e) { return lin_grid(e); }

int launch_viscous(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision,
                   int store_aux)
{
    dim3 g = viscous_grid(e), t(64, 4, 1);
    if (precision == 1)
        hipLaunchKernelGGL((k_viscous_node<float>), g, t, 0, s, e, c, f.pri, f.amu, f.dxidx, f.inv_j, f.w_flux,
                           f.vor, f.dudx, f.tao, store_aux);
    else
        hipLaunchKernelGGL((k_viscous_node<double>), g, t, 0, s, e, c, f.pri, f.amu, f.dxidx, f.inv_j, f.w_flux,
                           f.vor, f.dudx, f.tao, store_aux);
    if (precision == 1) hipLaunchKernelGGL(k_viscous_diff<float>, g, t, 0, s, e, f.w_flux, f.rhsv, c.re);
    else hipLaunchKernelGGL(k_viscous_diff<double>, g, t, 0, s, e, f.w_flux, f.rhsv, c.re);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_viscous_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                         int precision, int store_aux, int part)
{
    if (bg.total <= 0) return 0;
    dim3 g((unsigned)bg.total), t(64, 4, 1);
    if (part < 0 || part == 1) {
        if (precision == 1)
            hipLaunchKernelGGL((k_viscous_node_batch<float>), g, t, 0, s, blks, bg.map, bg.kg, arena, c, store_aux);
        else
            hipLaunchKernelGGL((k_viscous_node_batch<double>), g, t, 0, s, blks, bg.map, bg.kg, arena, c, store_aux);
    }
    if (part < 0 || part == 2) {
        if (precision == 1)
            hipLaunchKernelGGL(k_viscous_diff_batch<float>, g, t, 0, s, blks, bg.map, bg.kg, arena, c.re, 0);
        else
            hipLaunchKernelGGL(k_viscous_diff_batch<double>, g, t, 0, s, blks, bg.map, bg.kg, arena, c.re, 0);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------------ find_dt
// sub_timeadvancement_finddt.f90:52-125.  Steady: dt on the whole extended
// block (min/max folded into stats[0..1] as ordered bit patterns of the
// positive values).  Unsteady: the local CFL; the reference keeps the value
// of the last node of the block in the global cfl (:116), written to *cfl.
__device__ inline double local_rate_a3(const Ext &e, long n, const double *pri, double a3, const double *dxidx,
                                       double re);

__device__ inline double local_rate(const Ext &e, long n, const double *pri, const double *amu,
                                    const double *dxidx, double re)
{
    return local_rate_a3(e, n, pri, amu[2 * e.n + n], dxidx, re);
}

// the rate with the total viscosity a3 of the node given
__device__ inline double local_rate_a3(const Ext &e, long n, const double *pri, double a3, const double *dxidx,
                                       double re)
{
    const long N = e.n;
    double m[9];
#pragma unroll
    for (int x = 0; x < 9; x++) m[x] = ld_once(dxidx + x * N + n);
    double txi = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double teta = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
    double tzeta = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
    double d0 = ld_once(pri + n), u0 = pri[N + n], v0 = pri[2 * N + n], w0 = pri[3 * N + n],
           c0 = ld_once(pri + 6 * N + n);
    double uu = (m[0] * u0 + m[1] * v0 + m[2] * w0) / txi;
    double vv = (m[3] * u0 + m[4] * v0 + m[5] * w0) / teta;
    double ww = (m[6] * u0 + m[7] * v0 + m[8] * w0) / tzeta;
    // 2 |grad xi| a3 4/3 / re / d0 of sub_timeadvancement_finddt.f90:75-77
    // with one division
    const double kv = a3 * (8.0 / 3.0) / (re * d0);
    double t1 = fabs(uu) + c0 + txi * kv;
    double t2 = fabs(vv) + c0 + teta * kv;
    double t3 = fabs(ww) + c0 + tzeta * kv;
    return txi * t1 + teta * t2 + tzeta * t3;
}

// block-wide min/max of finite positive dt values via ordered bit patterns
__device__ inline void dt_stats(bool ok, double v, unsigned long long *stats);

__device__ inline void find_dt_steady_body(const Ext &e, int bx, const double *__restrict__ pri,
                                           const double *__restrict__ amu, const double *__restrict__ dxidx,
                                           double *__restrict__ dt, const double *cfl, double re,
                                           unsigned long long *stats)
{
    long n = (long)bx * blockDim.x + threadIdx.x;
    double v = 0.0;
    bool ok = n < e.n;
    if (ok) {
        v = *cfl / local_rate(e, n, pri, amu, dxidx, re);
        dt[n] = v;
    }
    dt_stats(ok, v, stats);
}

__device__ inline void dt_stats(bool ok, double v, unsigned long long *stats)
{
    unsigned long long bits = ok && v > 0.0 && isfinite(v) ? (unsigned long long)__double_as_longlong(v) : 0ull;
    unsigned long long mx = bits, mn = bits ? bits : ~0ull;
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long a = __shfl_xor(mx, o), c = __shfl_xor(mn, o);
        mx = a > mx ? a : mx;
        mn = c < mn ? c : mn;
    }
    if ((threadIdx.x & 63) == 0 && stats) {
        atomicMax(stats, mx);
        atomicMin(stats + 1, mn);
    }
}

__global__ void __launch_bounds__(256) k_find_dt_steady(Ext e, const double *__restrict__ pri,
                                                        const double *__restrict__ amu,
                                                        const double *__restrict__ dxidx, double *__restrict__ dt,
                                                        const double *cfl, double re, unsigned long long *stats)
{
    find_dt_steady_body(e, (int)blockIdx.x, pri, amu, dxidx, dt, cfl, re, stats);
}

// every block; the dt statistics come from the last block, as when the
// blocks run one after another with the statistics reset before each
__global__ void __launch_bounds__(256) k_find_dt_steady_batch(const BlockRec *__restrict__ blks,
                                                              const int *__restrict__ map,
                                                              const KGrid *__restrict__ kg, double *arena,
                                                              int nblk, const double *cfl, double re,
                                                              unsigned long long *stats)
{
    int bx, by, bz;
    const int m = batch_wg(map, kg, bx, by, bz);
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    find_dt_steady_body(e, bx, arena + b.o.pri, arena + b.o.amu, arena + b.o.dxidx, arena + b.o.dt, cfl, re,
                        m == nblk - 1 ? stats : nullptr);
}

__global__ void k_find_dt_unsteady(Ext e, const double *__restrict__ pri, const double *__restrict__ amu,
                                   const double *__restrict__ dxidx, const double *__restrict__ dt,
                                   double *cfl, double *cfl_blk, double re)
{
    long n = e.n - 1;  // node (ie, je, ke)
    const double v = dt[n] * local_rate(e, n, pri, amu, dxidx, re);
    *cfl = v;
    if (cfl_blk) *cfl_blk = v;
}

// unsteady, every block: thread m writes block m's cfl to cfl_blk[m]; the
// global cfl ends as the last block's, as after the blocks in order
__global__ void k_find_dt_unsteady_batch(const BlockRec *__restrict__ blks, double *arena, int nblk,
                                         double *cfl, double *cfl_blk, double re)
{
    const int m = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (m >= nblk) return;
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    const long n = e.n - 1;
    const double v = arena[b.o.dt + n] * local_rate(e, n, arena + b.o.pri, arena + b.o.amu, arena + b.o.dxidx, re);
    cfl_blk[m] = v;
    if (m == nblk - 1) *cfl = v;
}

__global__ void k_reset_stats(unsigned long long *st)
{
    st[0] = 0ull;
    st[1] = ~0ull;
}

// viscous_flux's node stage and the steady find_dt in one pass over every
// block's storage: a node of is0..ie0 evaluates its viscous node fluxes
// (k_viscous_node_batch), then dt from the total viscosity it has just
// formed; every other node of the extended block only dt (k_find_dt_steady
// reads the stored viscosity there).  The same expressions as the two
// kernels; the dt statistics from the last block.
//
// With PREP, the node also forms implicit1st's time term phi*1/dt
// (sub_timeadvancement_1stimplicit.f90:98-102,151, as lu_prep_body forms it
// for the tile sweeps, which add the spectral radii themselves) from its
// dt; the difference pass then forms rhs.
template <typename R, bool PREP>
__global__ void __launch_bounds__(256) k_viscous_node_dt_batch(const BlockRec *__restrict__ blks,
                                                               const int *__restrict__ map,
                                                               const KGrid *__restrict__ kg, double *arena,
                                                               Consts c, int store_aux, int nblk, const double *cfl,
                                                               unsigned long long *stats)
{
    int bx, by, bz;
    const int m = batch_wg(map, kg, bx, by, bz);
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    const long N = e.n;
    const long n = (long)bx * 256 + (long)(threadIdx.x + blockDim.x * threadIdx.y);
    const bool ok = n < N;
    double v = 0.0;
    if (ok) {
        const unsigned plane = (unsigned)e.ni * (unsigned)e.nj, un = (unsigned)n;
        const unsigned kk = un / plane, r = un - kk * plane, jj = r / (unsigned)e.ni;
        const int i = e.is + (int)(r - jj * (unsigned)e.ni), j = e.js + (int)jj, k = e.ks + (int)kk;
        const double *pri = arena + b.o.pri, *dxidx = arena + b.o.dxidx;
        double *amu = arena + b.o.amu;
        double a3;
        if (i >= e.is0 && i <= e.ie0 && j >= e.js0 && j <= e.je0 && k >= e.ks0 && k <= e.ke0) {
            double w12[12];
            R am1, am3, du[9], ta[6];
            viscous_node_eval<R>(e, i, j, k, c, pri, amu, dxidx, arena + b.o.inv_j, w12, am1, am3, du, ta);
            viscous_node_store<R>(N, n, am1, am3, du, ta, amu, arena + b.o.vor, arena + b.o.dudx, arena + b.o.tao,
                                  store_aux);
            R *w = reinterpret_cast<R *>(arena + b.o.w_flux);
#pragma unroll
            for (int x = 0; x < 12; x++) st_out(w + x * N + n, (R)w12[x]);
            a3 = (double)am3;
        } else {
            a3 = amu[2 * N + n];
        }
        v = *cfl / local_rate_a3(e, n, pri, a3, dxidx, c.re);
        st_out(arena + b.o.dt + n, v);
        if (PREP) {
            const double phi = *cfl >= 1.0e3 ? 0.0 : 1.0;  // implicit1st :98-102
            st_out(arena + b.o.w_diag + n, phi * 1.0 / v);
        }
    }
    dt_stats(ok, v, m == nblk - 1 ? stats : nullptr);
}

template <bool PREP>
static void viscous_dt_node(hipStream_t s, const BlockRec *blks, int nblk, const BatchGrid &bg, double *arena,
                            const Consts &c, int precision, int store_aux, const double *cfl, unsigned long long *stats)
{
    const dim3 g((unsigned)bg.total), t(64, 4, 1);
    if (precision == 1)
        hipLaunchKernelGGL((k_viscous_node_dt_batch<float, PREP>), g, t, 0, s, blks, bg.map, bg.kg, arena, c,
                           store_aux, nblk, cfl, stats);
    else
        hipLaunchKernelGGL((k_viscous_node_dt_batch<double, PREP>), g, t, 0, s, blks, bg.map, bg.kg, arena, c,
                           store_aux, nblk, cfl, stats);
}

int launch_viscous_dt_batch(hipStream_t s, const BlockRec *blks, int nblk, const BatchGrid &bg, double *arena,
                            const Consts &c, int precision, int store_aux, const double *cfl,
                            unsigned long long *stats, int part, int prep)
{
    if (bg.total <= 0) return 0;
    const dim3 g((unsigned)bg.total), t(64, 4, 1);
    if (part < 0 || part == 1) {
        if (prep) viscous_dt_node<true>(s, blks, nblk, bg, arena, c, precision, store_aux, cfl, stats);
        else viscous_dt_node<false>(s, blks, nblk, bg, arena, c, precision, store_aux, cfl, stats);
    }
    if (part < 0 || part == 2) {
        if (precision == 1)
            hipLaunchKernelGGL(k_viscous_diff_batch<float>, g, t, 0, s, blks, bg.map, bg.kg, arena, c.re, 2 * prep);
        else
            hipLaunchKernelGGL(k_viscous_diff_batch<double>, g, t, 0, s, blks, bg.map, bg.kg, arena, c.re, 2 * prep);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_reset_stats(hipStream_t s, unsigned long long *stats)
{
    hipLaunchKernelGGL(k_reset_stats, dim3(1), dim3(1), 0, s, stats);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_find_dt_batch(hipStream_t s, const BlockRec *blks, int nblk, const BatchGrid &bg, double *arena,
                         const Consts &c, double *cfl, double *cfl_blk, unsigned long long *stats)
{
    if (c.itime == 0) {
        if (bg.total > 0)
            hipLaunchKernelGGL(k_find_dt_steady_batch, dim3((unsigned)bg.total), dim3(256), 0, s, blks, bg.map, bg.kg,
                               arena, nblk, cfl, c.re, stats);
    } else {
        hipLaunchKernelGGL(k_find_dt_unsteady_batch, dim3((unsigned)((nblk + 63) / 64)), dim3(64), 0, s, blks,
                           arena, nblk, cfl, cfl_blk, c.re);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_find_dt(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, double *cfl, double *cfl_blk,
                   unsigned long long *stats)
{
    if (c.itime == 0) {
        int bs = 256;
        hipLaunchKernelGGL(k_find_dt_steady, dim3((unsigned)((e.n + bs - 1) / bs)), dim3(bs), 0, s, e,
                           f.pri, f.amu, f.dxidx, f.dt, cfl, c.re, stats);
    } else {
        hipLaunchKernelGGL(k_find_dt_unsteady, dim3(1), dim3(1), 0, s, e, f.pri, f.amu, f.dxidx, f.dt, cfl, cfl_blk,
                           c.re);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
