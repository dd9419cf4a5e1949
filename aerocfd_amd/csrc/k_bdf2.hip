// BDF2 sub-iteration kernels: the per-direction scalar-diagonal line
// recurrences of crank2nd and the history shift of update_q0.
//
// Reference: sub_timeadvancement_2ndcrank.f90:7-467, main.f90:692-711.
// Each line of a family (xi lines over (j,k), eta over (i,k), zeta over
// (i,j)) is an independent forward/backward recurrence, so one thread runs
// one line; the three families run in the reference's order.
#include "common.hpp"
#include "kernels.hpp"

namespace acfd {

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
