// Shared device-side definitions for the aerocfd_amd kernels (gfx950).
//
// Device layout: every block is resident in HBM as structure-of-arrays,
// field component c of node (i,j,k) at base[c*n + ix(i,j,k)], i fastest,
// over the block's extended extents (is..ie, js..je, ks..ke, 1-based as in
// the reference).  A wavefront of 64 lanes walks 64 consecutive i, so every
// field load is one coalesced 512-byte transaction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acfd {

// Index ranges of one block (module_block.f90:23-25) plus derived sizes.
struct Ext {
    int is, ie, js, je, ks, ke;
    int is0, ie0, js0, je0, ks0, ke0;
    int is1, ie1, js1, je1, ks1, ke1;
    int ni, nj, nk;
    long n;
    __host__ __device__ inline long ix(int i, int j, int k) const {
        return (long)(i - is) + (long)ni * ((long)(j - js) + (long)nj * (long)(k - ks));
    }
    __host__ __device__ inline long stride(int axis) const {
        return axis == 0 ? 1 : axis == 1 ? (long)ni : (long)ni * nj;
    }
    __host__ __device__ inline int lo(int axis) const { return axis == 0 ? is : axis == 1 ? js : ks; }
    __host__ __device__ inline int hi(int axis) const { return axis == 0 ? ie : axis == 1 ? je : ke; }
    __host__ __device__ inline int lo0(int axis) const { return axis == 0 ? is0 : axis == 1 ? js0 : ks0; }
    __host__ __device__ inline int hi0(int axis) const { return axis == 0 ? ie0 : axis == 1 ? je0 : ke0; }
    __host__ __device__ inline int lo1(int axis) const { return axis == 0 ? is1 : axis == 1 ? js1 : ks1; }
    __host__ __device__ inline int hi1(int axis) const { return axis == 0 ? ie1 : axis == 1 ? je1 : ke1; }
};

// Flow constants (sub_init_flowfield.f90:30-39, module_ns_consts.f90:31-33).
struct Consts {
    double gamma, cv, re, tinf, prl, prt;
    double d0, u0, p0, c0, aoa;
    int solver, itime, turb, des;
};

// Device pointers of one block.  Multi-component fields are SoA with
// component stride n.
struct Fields {
    double *q, *pri, *rhsi, *rhsv, *rhs;
    double *dxidx, *inv_j, *amu, *dt;
    double *dudx, *tao, *vor;
    double *aal, *bbe, *gga, *dist, *spacing, *length;
    double *nut, *sa_rhs, *q0;
    // work arrays
    double *w_flux;   // 5n face fluxes / 12n viscous node fluxes (max)
    double *w_dqs;    // 5n
    double *w_dq;     // 5n
    double *w_diag;   // n
    double *w_r;      // 3n spectral radii
    double *w_sa;     // 9n SA temporaries
};

// A box [i0,i1]x[j0,j1]x[k0,k1] of nodes covered by a (64,4,1) thread
// block tiling: blockIdx.x over i, y over j, z over k.
struct Box {
    int i0, i1, j0, j1, k0, k1;
};

inline dim3 box_grid(const Box &b, int bx = 64, int by = 4) {
    int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return dim3(0, 0, 0);
    return dim3((ni + bx - 1) / bx, (nj + by - 1) / by, nk);
}

__device__ inline bool box_node(const Box &b, int &i, int &j, int &k) {
    i = b.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    j = b.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    k = b.k0 + (int)blockIdx.z;
    return i <= b.i1 && j <= b.j1 && k <= b.k1;
}

// implicit1st's spectral radius of one direction at a node,
// |grad xi . u| + |c |grad xi|| (sub_timeadvancement_1stimplicit.f90:110-150),
// from the metric row m[0..2], the velocity and the sound speed, and the
// diagonal phi/dt + r_0 + r_1 + r_2 (:151) from the time term phi*1/dt.
// Never contracted, so every kernel that forms them (the preparation pass,
// the residual's viscous node pass, the tile sweeps) gives the same bits.
__device__ inline double lu_radius(double m0, double m1, double m2, double u, double v, double w, double c)
{
#pragma clang fp contract(off)
    double temp = m0 * m0 + m1 * m1 + m2 * m2;
    temp = sqrt(temp);
    const double cu = m0 * u + m1 * v + m2 * w;
    return fabs(cu) + fabs(c * temp);
}

__device__ inline double lu_diag(double tterm, const double *r)
{
#pragma clang fp contract(off)
    return tterm + r[0] + r[1] + r[2];
}

// the radii of node n into rr[d * N + n] and its diagonal, returned
__device__ inline double lu_radii_diag(const double *pri, const double *dxidx, long N, long n, double dt, double phi,
                                       double *rr)
{
    const double u = pri[N + n], v = pri[2 * N + n], w = pri[3 * N + n], c = pri[6 * N + n];
    double r3[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        r3[d] = lu_radius(dxidx[(3 * d) * N + n], dxidx[(3 * d + 1) * N + n], dxidx[(3 * d + 2) * N + n], u, v, w, c);
        rr[d * N + n] = r3[d];
    }
    return lu_diag(phi * 1.0 / dt, r3);
}

// The SA diffusion's metric products alpha, beta, gamma at a node
// (sub_init_gridtransformation.f90:809-866) from its nine dxidx values m
// and 1/J: p[0..2] alpha, p[3..5] beta, p[6..8] gamma.  Never contracted:
// the grid set-up stores them and the SA pass may form them again.
__device__ inline void sa_metric_products(const double *m, double ij, double *p)
{
#pragma clang fp contract(off)
    const double a0 = m[0], a1 = m[1], a2 = m[2], b0 = m[3], b1 = m[4], b2 = m[5];
    const double c0 = m[6], c1 = m[7], c2 = m[8];
    p[0] = (a0 * a0 + a1 * a1 + a2 * a2) * ij;
    p[1] = (a0 * b0 + a1 * b1 + a2 * b2) * ij;
    p[2] = (a0 * c0 + a1 * c1 + a2 * c2) * ij;
    p[3] = (b0 * a0 + b1 * a1 + b2 * a2) * ij;
    p[4] = (b0 * b0 + b1 * b1 + b2 * b2) * ij;
    p[5] = (b0 * c0 + b1 * c1 + b2 * c2) * ij;
    p[6] = (c0 * a0 + c1 * a1 + c2 * a2) * ij;
    p[7] = (c0 * b0 + c1 * b1 + c2 * b2) * ij;
    p[8] = (c0 * c0 + c1 * c1 + c2 * c2) * ij;
}

// Operands a pass reads once and results no later node of the pass reads:
// nontemporal (streaming) loads and stores, so that the L2 keeps the lines
// that neighbouring nodes' stencils read again.
template <typename T>
__device__ inline T ld_once(const T *p)
{
    return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ inline void st_out(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// 1/x for x of a positive normal range (in the inviscid flux WENO3's
// weight denominators, densities, 1 + sqrt(rho_r/rho_l), gamma - 1, 2 c^2;
// in SA's source the viscosity and the positive sums): the hardware's reciprocal estimate and
// two Newton-Raphson steps r <- r + r(1 - x r), within an ulp of the
// correctly rounded quotient (the inviscid kernels 6-7 % faster than with
// the division sequence); float keeps its division
template <typename R>
__device__ inline R rcp_pos(R x)
{
    return (R)1.0 / x;
}
template <>
__device__ inline double rcp_pos<double>(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

}  // namespace acfd

#define ACFD_CHECK_LAUNCH()                                                    \
    do {                                                                       \
        hipError_t e__ = hipGetLastError();                                    \
        if (e__ != hipSuccess) return (int)e__;                                \
    } while (0)
