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
#include "common.hpp"
#include "kernels.hpp"

namespace acfd {

// ------------------------------------------------------------------ getpri
// sub_getpri_v.f90:33-59 over the whole extended block.  Counts nodes with
// negative density into *nneg.
__global__ void k_getpri(Ext e, const double *__restrict__ q, double *__restrict__ pri,
                         double gamma, double cv, unsigned int *nneg)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
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

int launch_getpri(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, unsigned int *nneg)
{
    int bs = 256;
    long nb = (e.n + bs - 1) / bs;
    hipLaunchKernelGGL(k_getpri, dim3((unsigned)nb), dim3(bs), 0, s, e, f.q, f.pri, c.gamma, c.cv, nneg);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------ fill helper
__global__ void k_fill(double *p, long n, double v)
{
    long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) p[x] = v;
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
// Candidate pairs of sub_inviscid_upwind.f90:32-41 (right state from
// u1,u2,u3 = f(a),f(a+1),f(a+2)) and :57-66 (left state from f(a-1),f(a),f(a+1)).
template <typename R>
__device__ inline R weno3_right(R u1, R u2, R u3)
{
    const R eps = (R)1.0e-6;
    R s1 = (u2 - u3) * (u2 - u3);
    R s2 = (u1 - u2) * (u1 - u2);
    R t1 = eps + s1, t2 = eps + s2;
    R a1 = (R)1.0 / ((R)3.0 * (t1 * t1));
    R a2 = (R)2.0 / ((R)3.0 * (t2 * t2));
    R w1 = a1 / (a1 + a2);
    R w2 = a2 / (a1 + a2);
    return w1 * (-u3 / (R)2.0 + (R)3.0 * u2 / (R)2.0) + w2 * (u2 / (R)2.0 + u1 / (R)2.0);
}

template <typename R>
__device__ inline R weno3_left(R u1, R u2, R u3)
{
    const R eps = (R)1.0e-6;
    R s1 = (u2 - u1) * (u2 - u1);
    R s2 = (u3 - u2) * (u3 - u2);
    R t1 = eps + s1, t2 = eps + s2;
    R a1 = (R)1.0 / ((R)3.0 * (t1 * t1));
    R a2 = (R)2.0 / ((R)3.0 * (t2 * t2));
    R w1 = a1 / (a1 + a2);
    R w2 = a2 / (a1 + a2);
    return w1 * (-u1 / (R)2.0 + (R)3.0 * u2 / (R)2.0) + w2 * (u2 / (R)2.0 + u3 / (R)2.0);
}

// Roe flux with Harten's entropy fix (delta = 0.06) on a face with metric
// (mx,my,mz) from primitive left/right states (d,u,v,w,p);
// sub_inviscid_splitting.f90:65-162.
template <typename R>
__device__ inline void roe_flux(const R *ql, const R *qr, R mx, R my, R mz, R gamma, R *out)
{
    const R delt = (R)0.06;
    R temp = sqrt(mx * mx + my * my + mz * mz);
    R nx = mx / temp, ny = my / temp, nz = mz / temp;
    R dl = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4];
    R dr = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4];
    R hl = gamma * pl / ((gamma - (R)1.0) * dl) + (R)0.5 * (ul * ul + vl * vl + wl * wl);
    R hr = gamma * pr / ((gamma - (R)1.0) * dr) + (R)0.5 * (ur * ur + vr * vr + wr * wr);
    R uul = ul * mx + vl * my + wl * mz;
    R uur = ur * mx + vr * my + wr * mz;
    R ubarl = uul / temp, ubarr = uur / temp;
    R tmpp = sqrt(dr / dl);
    R avd = sqrt(dl * dr);
    R avu = (ul + ur * tmpp) / ((R)1.0 + tmpp);
    R avv = (vl + vr * tmpp) / ((R)1.0 + tmpp);
    R avw = (wl + wr * tmpp) / ((R)1.0 + tmpp);
    R avh = (hl + hr * tmpp) / ((R)1.0 + tmpp);
    R ava = sqrt((gamma - (R)1.0) * (avh - (R)0.5 * (avu * avu + avv * avv + avw * avw)));
    R avuu = avu * nx + avv * ny + avw * nz;
    R l1 = fabs(avuu - ava), l2 = fabs(avuu), l5 = fabs(avuu + ava);
    // Harten: sign(1, delt - lamda) >= 0 selects (l^2 + d^2)/(2d)
    if (delt - l1 >= (R)0) l1 = (l1 * l1 + delt * delt) / ((R)2.0 * delt);
    if (delt - l2 >= (R)0) l2 = (l2 * l2 + delt * delt) / ((R)2.0 * delt);
    if (delt - l5 >= (R)0) l5 = (l5 * l5 + delt * delt) / ((R)2.0 * delt);
    R deltad = dr - dl, deltap = pr - pl;
    R deltau = ur - ul, deltav = vr - vl, deltaw = wr - wl;
    R deltauu = ubarr - ubarl;
    R ava2 = ava * ava;
    R alpha1 = l2 * (deltad - deltap / ava2);
    R alpha2 = l5 * (deltap + avd * ava * deltauu) / ((R)2.0 * ava2);
    R alpha3 = l1 * (deltap - avd * ava * deltauu) / ((R)2.0 * ava2);
    R alpha4 = alpha1 + alpha2 + alpha3;
    R alpha5 = ava * (alpha2 - alpha3);
    R alpha6 = l2 * (avd * deltau - avd * deltauu * nx);
    R alpha7 = l2 * (avd * deltav - avd * deltauu * ny);
    R alpha8 = l2 * (avd * deltaw - avd * deltauu * nz);
    R f0 = alpha4;
    R f1 = avu * alpha4 + alpha5 * nx + alpha6;
    R f2 = avv * alpha4 + alpha5 * ny + alpha7;
    R f3 = avw * alpha4 + alpha5 * nz + alpha8;
    R f4 = avh * alpha4 + avuu * alpha5 + avu * alpha6 + avv * alpha7 + avw * alpha8
         - ava2 * alpha1 / (gamma - (R)1.0);
    out[0] = (R)0.5 * (dl * uul + dr * uur - temp * f0);
    out[1] = (R)0.5 * (dl * ul * uul + mx * pl + (dr * ur * uur + mx * pr) - temp * f1);
    out[2] = (R)0.5 * (dl * vl * uul + my * pl + (dr * vr * uur + my * pr) - temp * f2);
    out[3] = (R)0.5 * (dl * wl * uul + mz * pl + (dr * wr * uur + mz * pr) - temp * f3);
    out[4] = (R)0.5 * (dl * hl * uul + dr * hr * uur - temp * f4);
}

// conservative -> (d,u,v,w,p), sub_inviscid1.f90:123-133
template <typename R>
__device__ inline void cons2prim(const R *u, R *o, R gamma)
{
    o[0] = u[0];
    o[1] = u[1] / u[0];
    o[2] = u[2] / u[0];
    o[3] = u[3] / u[0];
    o[4] = (u[4] - (R)0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]) * (gamma - (R)1.0);
}

// Flux through the face between node n (axis position a) and n + st, with
// WENO3 reconstruction (first order at the extended edges alo / ahi).
template <typename R>
__device__ inline void face_flux(const double *__restrict__ q, const double *__restrict__ dxidx,
                                 const double *__restrict__ inv_j, long N, long n, long st,
                                 int a, int alo, int ahi, int axis, R gamma, R *F)
{
    R ul[5], ur[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double *qv = q + v * N;
        R f0 = (R)qv[n], f1 = (R)qv[n + st];
        if (a == alo) ul[v] = f0;
        else ul[v] = weno3_left<R>((R)qv[n - st], f0, f1);
        if (a == ahi - 1) ur[v] = f1;
        else ur[v] = weno3_right<R>(f0, f1, (R)qv[n + 2 * st]);
    }
    R pl[5], pr[5];
    cons2prim<R>(ul, pl, gamma);
    cons2prim<R>(ur, pr, gamma);
    const int c0 = 3 * axis;
    double j0 = inv_j[n], j1 = inv_j[n + st];
    R mx = (R)(0.5 * (dxidx[c0 * N + n] * j0 + dxidx[c0 * N + n + st] * j1));
    R my = (R)(0.5 * (dxidx[(c0 + 1) * N + n] * j0 + dxidx[(c0 + 1) * N + n + st] * j1));
    R mz = (R)(0.5 * (dxidx[(c0 + 2) * N + n] * j0 + dxidx[(c0 + 2) * N + n + st] * j1));
    roe_flux<R>(pl, pr, mx, my, mz, gamma, F);
}

// One direction of inviscid_flux_1 (sub_inviscid1.f90:113-243) for the
// computed nodes: rhsi = -(F(a+1/2) - F(a-1/2)) for axis 0, rhsi -= ... else.
template <typename R, int AXIS>
__global__ void __launch_bounds__(256) k_inviscid_dir(Ext e, Box b, const double *__restrict__ q,
                                                      const double *__restrict__ dxidx,
                                                      const double *__restrict__ inv_j,
                                                      double *__restrict__ rhsi, double gamma)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
    long st = e.stride(AXIS);
    int a = AXIS == 0 ? i : AXIS == 1 ? j : k;
    int alo = e.lo(AXIS), ahi = e.hi(AXIS);
    R fm[5], fp[5];
    face_flux<R>(q, dxidx, inv_j, N, n - st, st, a - 1, alo, ahi, AXIS, (R)gamma, fm);
    face_flux<R>(q, dxidx, inv_j, N, n, st, a, alo, ahi, AXIS, (R)gamma, fp);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double dF = (double)fp[v] - (double)fm[v];
        if (AXIS == 0) rhsi[v * N + n] = -dF;
        else rhsi[v * N + n] = rhsi[v * N + n] - dF;
    }
}

int launch_inviscid(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision)
{
    // rhsi is zero outside the computed range (sub_inviscid1.f90:102)
    int rc = launch_fill(s, f.rhsi, 5 * e.n, 0.0);
    if (rc) return rc;
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    dim3 g = box_grid(b), t(64, 4, 1);
    if (g.x == 0) return 0;
    if (precision == 1) {
        hipLaunchKernelGGL((k_inviscid_dir<float, 0>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
        hipLaunchKernelGGL((k_inviscid_dir<float, 1>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
        hipLaunchKernelGGL((k_inviscid_dir<float, 2>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
    } else {
        hipLaunchKernelGGL((k_inviscid_dir<double, 0>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
        hipLaunchKernelGGL((k_inviscid_dir<double, 1>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
        hipLaunchKernelGGL((k_inviscid_dir<double, 2>), g, t, 0, s, e, b, f.q, f.dxidx, f.inv_j, f.rhsi, c.gamma);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------------ viscous
// Node difference of sub_viscous.f90:107-178: one-sided at the physical
// range ends, central inside.
__device__ inline double ndiff(const double *a, long n, long st, int x, int x0, int x1)
{
    if (x == x0) return a[n + st] - a[n];
    if (x == x1) return a[n] - a[n - st];
    return (a[n + st] - a[n - st]) / 2.0;
}

// Stage 1 (sub_viscous.f90:107-293) on nodes is0..ie0: velocity and
// temperature gradients, Sutherland viscosity, stress, heat flux and the
// contravariant node fluxes fv/gv/hv (components 2..5) into w (12 fields).
template <typename R>
__global__ void __launch_bounds__(256) k_viscous_node(Ext e, Box b, Consts c,
                                                      const double *__restrict__ pri,
                                                      double *__restrict__ amu,
                                                      const double *__restrict__ dxidx,
                                                      const double *__restrict__ inv_j,
                                                      double *__restrict__ w, double *__restrict__ vor,
                                                      double *__restrict__ dudx,
                                                      double *__restrict__ tao, int store_aux)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
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
    R am1 = (R)((1.0 + 124.0 / c.tinf) / ((double)t0 + 124.0 / c.tinf) * pow((double)t0, 1.5));
    R am2 = (R)amu[N + n];
    R am3 = am1 + am2;
    amu[n] = (double)am1;
    amu[2 * N + n] = (double)am3;
    R ki = (R)(c.gamma * c.cv) * (am1 / (R)c.prl + am2 / (R)c.prt);
    R tx = g[0][0] * m[0] + g[0][1] * m[3] + g[0][2] * m[6];
    R ty = g[0][0] * m[1] + g[0][1] * m[4] + g[0][2] * m[7];
    R tz = g[0][0] * m[2] + g[0][1] * m[5] + g[0][2] * m[8];
    R du[9];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        du[3 * f + 0] = g[f + 1][0] * m[0] + g[f + 1][1] * m[3] + g[f + 1][2] * m[6];
        du[3 * f + 1] = g[f + 1][0] * m[1] + g[f + 1][1] * m[4] + g[f + 1][2] * m[7];
        du[3 * f + 2] = g[f + 1][0] * m[2] + g[f + 1][1] * m[5] + g[f + 1][2] * m[8];
    }
    const R c43 = (R)(4.0 / 3.0), c23 = (R)(2.0 / 3.0);
    R ta[6];
    ta[0] = am3 * (c43 * du[0] - c23 * (du[4] + du[8]));
    ta[1] = am3 * (du[1] + du[3]);
    ta[2] = am3 * (du[6] + du[2]);
    ta[3] = am3 * (c43 * du[4] - c23 * (du[0] + du[8]));
    ta[4] = am3 * (du[7] + du[5]);
    ta[5] = am3 * (c43 * du[8] - c23 * (du[0] + du[4]));
    R bx = u0 * ta[0] + v0 * ta[1] + w0 * ta[2] + ki * tx;
    R by = u0 * ta[1] + v0 * ta[3] + w0 * ta[4] + ki * ty;
    R bz = u0 * ta[2] + v0 * ta[4] + w0 * ta[5] + ki * tz;
    R ij = (R)inv_j[n];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const R *r = m + 3 * d;
        w[(4 * d + 0) * N + n] = (double)((r[0] * ta[0] + r[1] * ta[1] + r[2] * ta[2]) * ij);
        w[(4 * d + 1) * N + n] = (double)((r[0] * ta[1] + r[1] * ta[3] + r[2] * ta[4]) * ij);
        w[(4 * d + 2) * N + n] = (double)((r[0] * ta[2] + r[1] * ta[4] + r[2] * ta[5]) * ij);
        w[(4 * d + 3) * N + n] = (double)((r[0] * bx + r[1] * by + r[2] * bz) * ij);
    }
    if (store_aux) {
        vor[n] = (double)sqrt((du[5] - du[7]) * (du[5] - du[7]) + (du[2] - du[6]) * (du[2] - du[6])
                              + (du[1] - du[3]) * (du[1] - du[3]));
        if (store_aux > 1) {
#pragma unroll
            for (int x = 0; x < 9; x++) dudx[x * N + n] = (double)du[x];
#pragma unroll
            for (int x = 0; x < 6; x++) tao[x * N + n] = (double)ta[x];
        }
    }
}

// Stage 2 (sub_viscous.f90:296-365): rhsv = (d/dxi fv + d/deta gv + d/dzeta hv)/re
__global__ void __launch_bounds__(256) k_viscous_diff(Ext e, Box b, const double *__restrict__ w,
                                                      double *__restrict__ rhsv, double re)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n;
    long n = e.ix(i, j, k);
    const long sj = e.ni, sk = (long)e.ni * e.nj;
    rhsv[n] = (0.0 + 0.0 + 0.0) / re;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double a = ndiff(w + (0 + c) * N, n, 1, i, e.is0, e.ie0);
        double bb = ndiff(w + (4 + c) * N, n, sj, j, e.js0, e.je0);
        double h = ndiff(w + (8 + c) * N, n, sk, k, e.ks0, e.ke0);
        rhsv[(c + 1) * N + n] = (a + bb + h) / re;
    }
}

int launch_viscous(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision,
                   int store_aux)
{
    Box b{e.is0, e.ie0, e.js0, e.je0, e.ks0, e.ke0};
    dim3 g = box_grid(b), t(64, 4, 1);
    if (precision == 1)
        hipLaunchKernelGGL((k_viscous_node<float>), g, t, 0, s, e, b, c, f.pri, f.amu, f.dxidx, f.inv_j,
                           f.w_flux, f.vor, f.dudx, f.tao, store_aux);
    else
        hipLaunchKernelGGL((k_viscous_node<double>), g, t, 0, s, e, b, c, f.pri, f.amu, f.dxidx, f.inv_j,
                           f.w_flux, f.vor, f.dudx, f.tao, store_aux);
    hipLaunchKernelGGL(k_viscous_diff, g, t, 0, s, e, b, f.w_flux, f.rhsv, c.re);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------------ find_dt
// sub_timeadvancement_finddt.f90:52-125.  Steady: dt on the whole extended
// block (min/max folded into stats[0..1] as ordered bit patterns of the
// positive values).  Unsteady: the local CFL; the reference keeps the value
// of the last node of the block in the global cfl (:116), written to *cfl.
__device__ inline double local_rate(const Ext &e, long n, const double *pri, const double *amu,
                                    const double *dxidx, double re)
{
    const long N = e.n;
    double m[9];
#pragma unroll
    for (int x = 0; x < 9; x++) m[x] = dxidx[x * N + n];
    double txi = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double teta = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
    double tzeta = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
    double d0 = pri[n], u0 = pri[N + n], v0 = pri[2 * N + n], w0 = pri[3 * N + n], c0 = pri[6 * N + n];
    double uu = (m[0] * u0 + m[1] * v0 + m[2] * w0) / txi;
    double vv = (m[3] * u0 + m[4] * v0 + m[5] * w0) / teta;
    double ww = (m[6] * u0 + m[7] * v0 + m[8] * w0) / tzeta;
    double a3 = amu[2 * N + n];
    double t1 = fabs(uu) + c0 + 2.0 * txi * a3 * 4.0 / 3.0 / re / d0;
    double t2 = fabs(vv) + c0 + 2.0 * teta * a3 * 4.0 / 3.0 / re / d0;
    double t3 = fabs(ww) + c0 + 2.0 * tzeta * a3 * 4.0 / 3.0 / re / d0;
    return txi * t1 + teta * t2 + tzeta * t3;
}

__global__ void k_find_dt_steady(Ext e, const double *__restrict__ pri, const double *__restrict__ amu,
                                 const double *__restrict__ dxidx, double *__restrict__ dt,
                                 const double *cfl, double re, unsigned long long *stats)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    bool ok = n < e.n;
    if (ok) {
        v = *cfl / local_rate(e, n, pri, amu, dxidx, re);
        dt[n] = v;
    }
    // block-wide min/max of finite positive values via ordered bit patterns
    unsigned long long bits = ok && v > 0.0 && isfinite(v) ? (unsigned long long)__double_as_longlong(v) : 0ull;
    unsigned long long mx = bits, mn = bits ? bits : ~0ull;
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long a = __shfl_xor(mx, o), c = __shfl_xor(mn, o);
        mx = a > mx ? a : mx;
        mn = c < mn ? c : mn;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(stats, mx);
        atomicMin(stats + 1, mn);
    }
}

__global__ void k_find_dt_unsteady(Ext e, const double *__restrict__ pri, const double *__restrict__ amu,
                                   const double *__restrict__ dxidx, const double *__restrict__ dt,
                                   double *cfl, double re)
{
    long n = e.n - 1;  // node (ie, je, ke)
    *cfl = dt[n] * local_rate(e, n, pri, amu, dxidx, re);
}

__global__ void k_reset_stats(unsigned long long *st)
{
    st[0] = 0ull;
    st[1] = ~0ull;
}

int launch_reset_stats(hipStream_t s, unsigned long long *stats)
{
    hipLaunchKernelGGL(k_reset_stats, dim3(1), dim3(1), 0, s, stats);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_find_dt(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, double *cfl,
                   unsigned long long *stats)
{
    if (c.itime == 0) {
        int bs = 256;
        hipLaunchKernelGGL(k_find_dt_steady, dim3((unsigned)((e.n + bs - 1) / bs)), dim3(bs), 0, s, e,
                           f.pri, f.amu, f.dxidx, f.dt, cfl, c.re, stats);
    } else {
        hipLaunchKernelGGL(k_find_dt_unsteady, dim3(1), dim3(1), 0, s, e, f.pri, f.amu, f.dxidx, f.dt, cfl, c.re);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
