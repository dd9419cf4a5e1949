// Inviscid residual schemes and time integration with no counterpart in the
// reference, from the published methods:
//   scheme 1  Steger-Warming flux-vector splitting (J. Comput. Phys. 40,
//             1981) with WENO5 flux reconstruction (Jiang & Shu, J. Comput.
//             Phys. 126, 1996): rhsi -= F(a+1/2) - F(a-1/2) per direction.
//   scheme 2  weighted compact nonlinear scheme (Deng & Zhang, J. Comput.
//             Phys. 165, 2000): WENO5 interpolation of the conservative
//             variables to the midpoints, Roe flux there, explicit
//             midpoint-to-node difference of order 6 / 4 / 2.
//   RK3       three-stage Shu-Osher Runge-Kutta (J. Comput. Phys. 77, 1988).
// Near the ends of a line's extended range the stencils drop to third
// order (WENO3 splitting / the reference's WENO3 states) and then first
// order, exactly as the oracle does.  Everything is templated on the flux
// arithmetic R (double, or float for the FP32-flux variant).
#include <type_traits>

#include "common.hpp"
#include "flux_common.hpp"
#include "kernels.hpp"
#include "sensor.hpp"

namespace acfd {

// Steger-Warming split fluxes through the node metric M = grad xi / J.
template <typename R>
__device__ inline void sw_split(const R *q, R mx, R my, R mz, R gamma, R *fp, R *fm)
{
    const R rho = q[0], ri = (R)1.0 / rho;
    const R u = q[1] * ri, v = q[2] * ri, w = q[3] * ri;
    const R v2 = u * u + v * v + w * w;
    const R p = (gamma - (R)1.0) * (q[4] - (R)0.5 * rho * v2);
    const R c = sqrt(gamma * p * ri);
    const R mm = sqrt(mx * mx + my * my + mz * mz);
    const R rm = (R)1.0 / mm;
    const R nx = mx * rm, ny = my * rm, nz = mz * rm;
    const R ut = u * nx + v * ny + w * nz;
    const R h = c * c / (gamma - (R)1.0) + (R)0.5 * v2;
    const R l[3] = {mm * ut, mm * (ut + c), mm * (ut - c)};
    const R s = rho / ((R)2.0 * gamma);
#pragma unroll
    for (int side = 0; side < 2; side++) {
        R lp[3];
#pragma unroll
        for (int a = 0; a < 3; a++) lp[a] = side == 0 ? (R)0.5 * (l[a] + fabs(l[a])) : (R)0.5 * (l[a] - fabs(l[a]));
        const R a1 = (R)2.0 * (gamma - (R)1.0) * lp[0];
        R *f = side == 0 ? fp : fm;
        f[0] = s * (a1 + lp[1] + lp[2]);
        f[1] = s * (a1 * u + lp[1] * (u + c * nx) + lp[2] * (u - c * nx));
        f[2] = s * (a1 * v + lp[1] * (v + c * ny) + lp[2] * (v - c * ny));
        f[3] = s * (a1 * w + lp[1] * (w + c * nz) + lp[2] * (w - c * nz));
        f[4] = s * (a1 * (R)0.5 * v2 + lp[1] * (h + c * ut) + lp[2] * (h - c * ut));
    }
}

// Jiang-Shu smoothness indicators of the three candidate stencils
template <typename R>
__device__ inline void js_beta(R a, R b, R c, R d, R e5, R &b0, R &b1, R &b2)
{
    const R k = (R)13.0 / (R)12.0;
    b0 = k * (a - (R)2.0 * b + c) * (a - (R)2.0 * b + c) + (R)0.25 * (a - (R)4.0 * b + (R)3.0 * c) * (a - (R)4.0 * b + (R)3.0 * c);
    b1 = k * (b - (R)2.0 * c + d) * (b - (R)2.0 * c + d) + (R)0.25 * (b - d) * (b - d);
    b2 = k * (c - (R)2.0 * d + e5) * (c - (R)2.0 * d + e5) + (R)0.25 * ((R)3.0 * c - (R)4.0 * d + e5) * ((R)3.0 * c - (R)4.0 * d + e5);
}

// sum_k (d_k / t_k) q_k / sum_k (d_k / t_k), t_k = (eps + beta_k)^2, with
// one division
template <typename R>
__device__ inline R weno_combine3(R q0, R q1, R q2, R b0, R b1, R b2, R d0, R d1, R d2)
{
    const R eps = (R)1.0e-6;
    const R t0 = (eps + b0) * (eps + b0), t1 = (eps + b1) * (eps + b1), t2 = (eps + b2) * (eps + b2);
    const R w0 = d0 * t1 * t2, w1 = d1 * t0 * t2, w2 = d2 * t0 * t1;
    return (w0 * q0 + w1 * q1 + w2 * q2) / (w0 + w1 + w2);
}

// WENO5 flux reconstruction at i+1/2 from f(i-2..i+2), left-biased
template <typename R>
__device__ inline R weno5_rec(R a, R b, R c, R d, R e5)
{
    const R q0 = a / (R)3.0 - (R)7.0 * b / (R)6.0 + (R)11.0 * c / (R)6.0;
    const R q1 = -b / (R)6.0 + (R)5.0 * c / (R)6.0 + d / (R)3.0;
    const R q2 = c / (R)3.0 + (R)5.0 * d / (R)6.0 - e5 / (R)6.0;
    R b0, b1, b2;
    js_beta(a, b, c, d, e5, b0, b1, b2);
    return weno_combine3(q0, q1, q2, b0, b1, b2, (R)0.1, (R)0.6, (R)0.3);
}

// WENO3 flux reconstruction at i+1/2 from f(i-1..i+1) (linear weights 1/3, 2/3)
template <typename R>
__device__ inline R weno3_rec(R a, R b, R c)
{
    const R eps = (R)1.0e-6;
    const R q0 = (R)(-0.5) * a + (R)1.5 * b, q1 = (R)0.5 * (b + c);
    const R t0 = (eps + (b - a) * (b - a)) * (eps + (b - a) * (b - a));
    const R t1 = (eps + (c - b) * (c - b)) * (eps + (c - b) * (c - b));
    return (q0 * t1 + (R)2.0 * q1 * t0) / (t1 + (R)2.0 * t0);
}

// WENO5 interpolation to the midpoint i+1/2 from u(i-2..i+2)
template <typename R>
__device__ inline R weno5_interp(R a, R b, R c, R d, R e5)
{
    const R q0 = ((R)3.0 * a - (R)10.0 * b + (R)15.0 * c) / (R)8.0;
    const R q1 = (-b + (R)6.0 * c + (R)3.0 * d) / (R)8.0;
    const R q2 = ((R)3.0 * c + (R)6.0 * d - e5) / (R)8.0;
    R b0, b1, b2;
    js_beta(a, b, c, d, e5, b0, b1, b2);
    return weno_combine3(q0, q1, q2, b0, b1, b2, (R)1.0 / (R)16.0, (R)10.0 / (R)16.0, (R)5.0 / (R)16.0);
}

__device__ inline int face_order(int a, int lo, int hi)
{
    if (a - 2 >= lo && a + 3 <= hi) return 5;
    if (a - 1 >= lo && a + 2 <= hi) return 3;
    return 1;
}

__device__ inline int wcns_order(int a, int lo, int hi)
{
    if (a - 3 >= lo && a + 2 <= hi - 1 && face_order(a - 3, lo, hi) == 5 && face_order(a + 2, lo, hi) == 5) return 6;
    if (a - 2 >= lo && a + 1 <= hi - 1 && face_order(a - 2, lo, hi) >= 3 && face_order(a + 1, lo, hi) >= 3) return 4;
    return 2;
}

// node quantity of a scheme at node n: split fluxes (scheme 1, 10 values)
// or the conservative variables (scheme 2, 5 values)
template <typename R, int SCHEME>
__device__ inline void node_quantity(const double *__restrict__ q, const double *__restrict__ dxidx,
                                     const double *__restrict__ inv_j, long N, long n, int c0, R gamma, R *out)
{
    R qv[5];
#pragma unroll
    for (int x = 0; x < 5; x++) qv[x] = (R)q[x * N + n];
    if (SCHEME == 1) {
        const double ij = inv_j[n];
        sw_split<R>(qv, (R)(dxidx[c0 * N + n] * ij), (R)(dxidx[(c0 + 1) * N + n] * ij),
                    (R)(dxidx[(c0 + 2) * N + n] * ij), gamma, out, out + 5);
    } else {
#pragma unroll
        for (int x = 0; x < 5; x++) out[x] = qv[x];
    }
}

// face flux at a+1/2 from node quantities W[0..5] at a-2..a+3 (order ord);
// scheme 2 also needs the face metric m
template <typename R, int SCHEME, int NQ>
__device__ inline void ext_face(const R (*W)[NQ], int ord, const R *m, R gamma, R *F, bool lin = false)
{
    if (SCHEME == 1) {
#pragma unroll
        for (int x = 0; x < 5; x++) {
            R p, mn;
            if (ord == 5) {
                p = weno5_rec<R>(W[0][x], W[1][x], W[2][x], W[3][x], W[4][x]);
                mn = weno5_rec<R>(W[5][5 + x], W[4][5 + x], W[3][5 + x], W[2][5 + x], W[1][5 + x]);
            } else if (ord == 3) {
                p = weno3_rec<R>(W[1][x], W[2][x], W[3][x]);
                mn = weno3_rec<R>(W[4][5 + x], W[3][5 + x], W[2][5 + x]);
            } else {
                p = W[2][x];
                mn = W[3][5 + x];
            }
            F[x] = p + mn;
        }
    } else {
        R ul[5], ur[5];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            if (ord == 5 && lin) {
                // linear weights: the fifth-order midpoint interpolation
                ul[x] = ((R)3.0 * W[0][x] - (R)20.0 * W[1][x] + (R)90.0 * W[2][x] + (R)60.0 * W[3][x] -
                         (R)5.0 * W[4][x]) / (R)128.0;
                ur[x] = ((R)3.0 * W[5][x] - (R)20.0 * W[4][x] + (R)90.0 * W[3][x] + (R)60.0 * W[2][x] -
                         (R)5.0 * W[1][x]) / (R)128.0;
            } else if (ord == 5) {
                ul[x] = weno5_interp<R>(W[0][x], W[1][x], W[2][x], W[3][x], W[4][x]);
                ur[x] = weno5_interp<R>(W[5][x], W[4][x], W[3][x], W[2][x], W[1][x]);
            } else if (ord == 3) {
                ul[x] = weno3<R>(W[1][x], W[2][x], W[3][x]);
                ur[x] = weno3<R>(W[4][x], W[3][x], W[2][x]);
            } else {
                ul[x] = W[2][x];
                ur[x] = W[3][x];
            }
        }
        R pl[5], pr[5];
        cons2prim<R>(ul, pl, gamma);
        cons2prim<R>(ur, pr, gamma);
        roe_flux<R>(pl, pr, m[0], m[1], m[2], gamma, F);
    }
}

template <typename R>
__device__ inline void face_metric_r(const double *__restrict__ dxidx, const double *__restrict__ inv_j, long N,
                                     long n, long st, int c0, R *m)
{
    const double j0 = inv_j[n], j1 = inv_j[n + st];
#pragma unroll
    for (int c = 0; c < 3; c++) m[c] = (R)(0.5 * (dxidx[(c0 + c) * N + n] * j0 + dxidx[(c0 + c) * N + n + st] * j1));
}

// node difference at node a of the face fluxes in FR, where FR[IA] is the
// face a+1/2 (and FR[IA-1] the face a-1/2, FR[IA+-2], FR[IA+-3] the wider
// faces of the WCNS difference)
template <typename R, int SCHEME, int IA>
__device__ inline double ext_diff(const R (*FR)[5], int x, int dord)
{
    const double d1 = (double)FR[IA][x] - (double)FR[IA - 1][x];
    if (SCHEME == 1 || dord == 2) return d1;
    const double d3 = (double)FR[IA + 1][x] - (double)FR[IA - 2][x];
    if (dord == 4) return 9.0 / 8.0 * d1 - 1.0 / 24.0 * d3;
    const double d5 = (double)FR[IA + 2][x] - (double)FR[IA - 3][x];
    return 75.0 / 64.0 * d1 - 25.0 / 384.0 * d3 + 3.0 / 640.0 * d5;
}

// j and k directions: each thread marches along the axis over kExtRun nodes
// with rolling windows of node quantities and face fluxes.
constexpr int kExtRun = 32;
// one wavefront's face counts into its slot (called by one lane)
__device__ inline void sensor_count(unsigned long long *stats, unsigned long long ne, unsigned long long nl)
{
    const unsigned slot = (blockIdx.x * 31u + blockIdx.y * 17u + blockIdx.z * 7u + threadIdx.y) % kSensorStatSlots;
    atomicAdd(&stats[slot * kSensorStatStride], ne);
    atomicAdd(&stats[slot * kSensorStatStride + 1], nl);
}

// SENS: 0 plain; 1 the network decides per face (and records the decisions
// in sa.mask when given); 2 the decisions are read from sa.mask
template <typename R, int AXIS, int SCHEME, int SENS>
__global__ void __launch_bounds__(256) k_ext_jk(Ext e, const double *__restrict__ q,
                                               const double *__restrict__ dxidx,
                                               const double *__restrict__ inv_j, double *__restrict__ rhsi,
                                               double gamma_d, SensorArgs sa)
{
    constexpr int NQ = SCHEME == 1 ? 10 : 5;
    const int ir = e.is1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int orr = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int a0 = e.lo1(AXIS) + (int)blockIdx.z * kExtRun;
    const int a1 = min(a0 + kExtRun - 1, e.hi1(AXIS));
    const bool active = ir <= e.ie1 && (AXIS == 1 ? e.ks1 + orr <= e.ke1 : e.js1 + orr <= e.je1);
    // with the sensor every lane takes part (in the wavefront-wide network,
    // and in the wavefront's sum of its face counts); lanes outside the
    // range work on a valid line and store nothing
    __shared__ typename std::conditional<SENS == 1, SensorLds, char>::type slds;
    if constexpr (SENS == 1) {
        sensor_load_net(sa.net, slds);
        __syncthreads();
    } else if constexpr (SENS == 0) {
        if (!active) return;
    }
    const int i = min(ir, e.ie1);
    const int o = min(orr, AXIS == 1 ? e.ke1 - e.ks1 : e.je1 - e.js1);
    const int j = AXIS == 1 ? e.lo(1) : e.js1 + o;
    const int k = AXIS == 2 ? e.lo(2) : e.ks1 + o;
    unsigned n_eval = 0, n_lin = 0;
    const R gamma = (R)gamma_d;
    const long N = e.n, st = e.stride(AXIS);
    const int lo = e.lo(AXIS), hi = e.hi(AXIS);
    const int c0 = 3 * AXIS;
    const long line0 = e.ix(i, j, k);  // node at position lo of this line
    auto nat = [&](int pos) { return line0 + (long)(pos - lo) * st; };
    const int bfirst = SCHEME == 1 ? a0 - 1 : a0 - 3;
    const int blast = SCHEME == 1 ? a1 : a1 + 2;
    R W[6][NQ];
#pragma unroll
    for (int w = 0; w < 6; w++) {
        const int pos = bfirst - 2 + w;
        if (pos >= lo && pos <= hi) node_quantity<R, SCHEME>(q, dxidx, inv_j, N, nat(pos), c0, gamma, W[w]);
        else
#pragma unroll
            for (int x = 0; x < NQ; x++) W[w][x] = (R)0;
    }
    R FR[6][5];
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
        for (int x = 0; x < 5; x++) FR[f][x] = (R)0;
    for (int b = bfirst; b <= blast; b++) {
        if (b > bfirst) {
#pragma unroll
            for (int w = 0; w < 5; w++)
#pragma unroll
                for (int x = 0; x < NQ; x++) W[w][x] = W[w + 1][x];
            const int pos = b + 3;
            if (pos <= hi) node_quantity<R, SCHEME>(q, dxidx, inv_j, N, nat(pos), c0, gamma, W[5]);
        }
#pragma unroll
        for (int f = 0; f < 5; f++)
#pragma unroll
            for (int x = 0; x < 5; x++) FR[f][x] = FR[f + 1][x];
        const int ord = face_order(b, lo, hi);
        const bool face = b >= lo && b + 1 <= hi;
        bool lin = false;
        if constexpr (SENS == 1) {
            float feat[kSensorNF];
            sensor_features<R>(W, gamma_d, feat);
            const float p = sensor_eval_wave(slds, threadIdx.y, feat);
            lin = p < sa.threshold;
            if (sa.mask && active && face) sa.mask[AXIS * N + nat(b)] = lin ? 1 : 0;
        } else if constexpr (SENS == 2) {
            lin = face && sa.mask[AXIS * N + nat(b)] != 0;
        }
        if constexpr (SENS != 0) {
            if (active && ord == 5) {
                n_eval++;
                n_lin += lin ? 1 : 0;
            }
        }
        if (face) {
            R m[3] = {(R)0, (R)0, (R)0};
            if (SCHEME == 2) face_metric_r<R>(dxidx, inv_j, N, nat(b), st, c0, m);
            ext_face<R, SCHEME, NQ>(W, ord, m, gamma, FR[5], lin);
        }
        const int a = SCHEME == 1 ? b : b - 2;
        if (active && a >= a0 && a <= a1) {
            const int dord = SCHEME == 2 ? wcns_order(a, lo, hi) : 2;
            const long n = nat(a);
#pragma unroll
            for (int x = 0; x < 5; x++)
                rhsi[x * N + n] = rhsi[x * N + n] - ext_diff<R, SCHEME, SCHEME == 1 ? 5 : 3>(FR, x, dord);
        }
    }
    if constexpr (SENS != 0) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            n_eval += __shfl_xor(n_eval, o);
            n_lin += __shfl_xor(n_lin, o);
        }
        if (threadIdx.x == 0 && n_eval) sensor_count(sa.stats, n_eval, n_lin);
    }
}

// i direction: a wavefront holds 64 consecutive positions; node quantities
// and face fluxes are exchanged through LDS.  Scheme 1 yields 58 nodes per
// wavefront, scheme 2 (wider difference) 54.
template <int SCHEME> struct ExtI {
    static constexpr int OFF = SCHEME == 1 ? 3 : 5;
    static constexpr int NPW = SCHEME == 1 ? 58 : 54;
};

template <typename R, int SCHEME, int SENS>
__global__ void __launch_bounds__(256) k_ext_i(Ext e, const double *__restrict__ q,
                                              const double *__restrict__ dxidx,
                                              const double *__restrict__ inv_j, double *__restrict__ rhsi,
                                              double gamma_d, SensorArgs sa)
{
    constexpr int NQ = SCHEME == 1 ? 10 : 5;
    __shared__ R nq[4][NQ][64];
    __shared__ R fq[4][5][64];
    __shared__ typename std::conditional<SENS == 1, SensorLds, char>::type slds;
    if constexpr (SENS == 1) sensor_load_net(sa.net, slds);
    const int lane = threadIdx.x, row = threadIdx.y;
    const int jr = e.js1 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const bool row_ok = jr <= e.je1;
    const int j = row_ok ? jr : e.je1;
    const int k = e.ks1 + (int)blockIdx.z;
    const int i0 = e.is1 + (int)blockIdx.x * ExtI<SCHEME>::NPW;
    const int p = i0 - ExtI<SCHEME>::OFF + lane;  // position of this lane
    const int lo = e.is, hi = e.ie;
    const R gamma = (R)gamma_d;
    const long N = e.n;
    const int pc = min(max(p, lo), hi);
    const long n = e.ix(pc, j, k);
    {
        R v[NQ];
        node_quantity<R, SCHEME>(q, dxidx, inv_j, N, n, 0, gamma, v);
#pragma unroll
        for (int x = 0; x < NQ; x++) nq[row][x][lane] = v[x];
    }
    __syncthreads();
    {
        R F[5] = {(R)0, (R)0, (R)0, (R)0, (R)0};
        const bool face_ok = lane >= 2 && lane <= 60 && p >= lo && p + 1 <= hi;
        R W[6][NQ];
#pragma unroll
        for (int w = 0; w < 6; w++)
#pragma unroll
            for (int x = 0; x < NQ; x++) W[w][x] = nq[row][x][min(max(lane - 2 + w, 0), 63)];
        const int ord = face_order(p, lo, hi);
        bool lin = false;
        if constexpr (SENS == 1) {
            float feat[kSensorNF];
            sensor_features<R>(W, gamma_d, feat);
            const float pr = sensor_eval_wave(slds, row, feat);
            lin = pr < sa.threshold;
            if (sa.mask && face_ok && row_ok) sa.mask[n] = lin ? 1 : 0;
        } else if constexpr (SENS == 2) {
            lin = face_ok && sa.mask[n] != 0;
        }
        if constexpr (SENS != 0) {
            const bool counted = face_ok && ord == 5 && row_ok && p >= e.is1 - 1 && p <= e.ie1;
            const unsigned long long ne = __popcll(__ballot(counted)), nl = __popcll(__ballot(counted && lin));
            if (lane == 0 && ne) sensor_count(sa.stats, ne, nl);
        }
        if (face_ok) {
            R m[3] = {(R)0, (R)0, (R)0};
            if (SCHEME == 2) face_metric_r<R>(dxidx, inv_j, N, n, 1, 0, m);
            ext_face<R, SCHEME, NQ>(W, ord, m, gamma, F, lin);
        }
#pragma unroll
        for (int x = 0; x < 5; x++) fq[row][x][lane] = F[x];
    }
    __syncthreads();
    if (row_ok && lane >= ExtI<SCHEME>::OFF && lane < ExtI<SCHEME>::OFF + ExtI<SCHEME>::NPW && p >= e.is1 &&
        p <= e.ie1) {
        R FR[6][5];
#pragma unroll
        for (int f = 0; f < 6; f++)
#pragma unroll
            for (int x = 0; x < 5; x++) {
                const int l = lane - 3 + f;
                FR[f][x] = (l >= 0 && l < 64) ? fq[row][x][l] : (R)0;
            }
        const int dord = SCHEME == 2 ? wcns_order(p, lo, hi) : 2;
#pragma unroll
        for (int x = 0; x < 5; x++) rhsi[x * N + n] = -ext_diff<R, SCHEME, 3>(FR, x, dord);
    }
}

template <typename R, int SCHEME, int SENS = 0>
static void launch_ext_t(hipStream_t s, const Ext &e, const Fields &f, double gamma, SensorArgs sa = {})
{
    const int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    dim3 t(64, 4, 1);
    dim3 gi((Li + ExtI<SCHEME>::NPW - 1) / ExtI<SCHEME>::NPW, (Lj + 3) / 4, Lk);
    hipLaunchKernelGGL((k_ext_i<R, SCHEME, SENS>), gi, t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi, gamma, sa);
    dim3 gj((Li + 63) / 64, (Lk + 3) / 4, (Lj + kExtRun - 1) / kExtRun);
    hipLaunchKernelGGL((k_ext_jk<R, 1, SCHEME, SENS>), gj, t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi, gamma, sa);
    dim3 gk((Li + 63) / 64, (Lj + 3) / 4, (Lk + kExtRun - 1) / kExtRun);
    hipLaunchKernelGGL((k_ext_jk<R, 2, SCHEME, SENS>), gk, t, 0, s, e, f.q, f.dxidx, f.inv_j, f.rhsi, gamma, sa);
}

int launch_inviscid_wcns_sensor(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision,
                                const SensorArgs &sa)
{
    int rc = launch_fill(s, f.rhsi, 5 * e.n, 0.0);
    if (rc) return rc;
    if (e.ie1 < e.is1 || e.je1 < e.js1 || e.ke1 < e.ks1) return 0;
    if (sa.replay) {
        if (precision == 1) launch_ext_t<float, 2, 2>(s, e, f, c.gamma, sa);
        else launch_ext_t<double, 2, 2>(s, e, f, c.gamma, sa);
    } else {
        if (precision == 1) launch_ext_t<float, 2, 1>(s, e, f, c.gamma, sa);
        else launch_ext_t<double, 2, 1>(s, e, f, c.gamma, sa);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_inviscid_ext(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, int precision, int scheme)
{
    // rhsi is zero outside the computed range
    int rc = launch_fill(s, f.rhsi, 5 * e.n, 0.0);
    if (rc) return rc;
    if (e.ie1 < e.is1 || e.je1 < e.js1 || e.ke1 < e.ks1) return 0;
    if (scheme == 1) {
        if (precision == 1) launch_ext_t<float, 1>(s, e, f, c.gamma);
        else launch_ext_t<double, 1>(s, e, f, c.gamma);
    } else {
        if (precision == 1) launch_ext_t<float, 2>(s, e, f, c.gamma);
        else launch_ext_t<double, 2>(s, e, f, c.gamma);
    }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// -------------------------------------------------------------------- RK3
// Stage s of the Shu-Osher scheme on the computed range:
// q = a_s qn + b_s (q + dt (rhsi + rhsv) / inv_j), (a,b) = (0,1), (3/4,1/4),
// (1/3,2/3); the first stage also stores rhs = rhsi + rhsv (the step's
// residual).
__global__ void __launch_bounds__(256) k_rk3_stage(Ext e, Box b, double *__restrict__ q,
                                                  const double *__restrict__ qn,
                                                  const double *__restrict__ rhsi,
                                                  const double *__restrict__ rhsv, double *__restrict__ rhs,
                                                  const double *__restrict__ dt, const double *__restrict__ inv_j,
                                                  double A, double B)
{
    int i, j, k;
    if (!box_node(b, i, j, k)) return;
    const long N = e.n, n = e.ix(i, j, k);
    const double dtn = dt[n], ij = inv_j[n];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        const double r = rhsv ? rhsi[x * N + n] + rhsv[x * N + n] : rhsi[x * N + n] + 0.0;
        if (rhs) rhs[x * N + n] = r;
        const double upd = q[x * N + n] + dtn * r / ij;
        q[x * N + n] = A * qn[x * N + n] + B * upd;
    }
}

__global__ void __launch_bounds__(256) k_copy(const double *__restrict__ a, double *__restrict__ b, long n)
{
    long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) b[x] = a[x];
}

int launch_copy(hipStream_t s, const double *a, double *b, long n)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_rk3_stage(hipStream_t s, const Ext &e, const Fields &f, const double *qn, int stage, bool viscous)
{
    static const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    if (b.i1 < b.i0 || b.j1 < b.j0 || b.k1 < b.k0) return 0;
    hipLaunchKernelGGL(k_rk3_stage, box_grid(b), dim3(64, 4, 1), 0, s, e, b, f.q, qn, f.rhsi,
                       viscous ? f.rhsv : nullptr, stage == 0 ? f.rhs : nullptr, f.dt, f.inv_j, A[stage],
                       B[stage]);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
