// Spalart-Allmaras kernels: turbulent length scale, convection (first-order
// upwind), diffusion, production/destruction with the diagonal Jacobian, and
// the scalar LU-SGS update of nut.
//
// Reference: sub_tur_sa.f90:7-825, constants module_tursa.f90:4-11.
#include "batch.hpp"
#include "common.hpp"
#include "kernels.hpp"

// No implicit contraction: the hyperplane and tile sweeps evaluate the same
// expressions and must round alike.
#pragma clang fp contract(off)

namespace acfd {

namespace sa {
constexpr double dta = 2.0 / 3.0, kar = 0.41;
constexpr double cb1 = 0.1355, cb2 = 0.622;
constexpr double cw1 = cb1 / kar / kar + (1.0 + cb2) / dta;
constexpr double cw2 = 0.3, cw3 = 2.0;
constexpr double ct3 = 1.2, ct4 = 0.5;
constexpr double cv1 = 7.1;
constexpr double cdes = 0.65;
}  // namespace sa

// Work layout in f.w_sa (n each): 0-2 dnut/dxi,deta,dzeta; 3-5 alpha.grad,
// beta.grad, gamma.grad; 6 nu + (1+cb2) nut; 7 diasd; 8 dnut*.
// f.w_dqs holds contravariant velocities (0-2) and diag (3); f.w_dq dnut.

// length (sub_tur_sa.f90:50-100) of node n of the extended block
__device__ inline double sa_length_at(const Ext &e, const Consts &c, long n, const double *__restrict__ dist,
                                      const double *__restrict__ spac, const double *__restrict__ pri,
                                      const double *__restrict__ nut, const double *__restrict__ amu,
                                      const double *__restrict__ dudx)
{
    const long N = e.n;
    double dist0 = dist[n];
    if (c.des == 1) return fmin(dist0, sa::cdes * spac[n]);
    if (c.des == 2) {
        // dudx is declared dudx(3,is:ie,js:je,ks:ke,3) in sa_model
        // (sub_tur_sa.f90:29) while the array holds dudx(9,...): element
        // (a,n,b) of that view is flat element (a-1) + 3n + 3N(b-1) of the
        // reference layout, i.e. component F%9 of node F/9.
        auto du = [&](int a, int b) -> double {
            long F = (long)(a - 1) + 3 * n + 3 * N * (long)(b - 1);
            return dudx[(F % 9) * N + F / 9];
        };
        double uy = du(1, 2), uz = du(1, 3), vx = du(2, 1), vz = du(2, 3), wx = du(3, 1), wy = du(3, 2);
        double d0 = pri[n], nut0 = nut[n];
        double nu0 = amu[n] / d0;
        double spac0 = spac[n];
        double temp = uy * uy + uz * uz + vx * vx + vz * vz + wx * wx + wy * wy;
        temp = sqrt(temp);
        double rd = (nu0 + nut0) / (temp * c.re * (sa::kar * sa::kar) * (dist0 * dist0));
        double x = 8.0 * rd;
        double fd = 1.0 - tanh(x * x * x);
        return dist0 - fd * fmax(0.0, dist0 - sa::cdes * spac0);
    }
    return dist0;
}

__global__ void k_sa_length(Ext e, Consts c, const double *__restrict__ dist, const double *__restrict__ spac,
                            const double *__restrict__ pri, const double *__restrict__ nut,
                            const double *__restrict__ amu, const double *__restrict__ dudx,
                            double *__restrict__ length)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= e.n) return;
    length[n] = sa_length_at(e, c, n, dist, spac, pri, nut, amu, dudx);
}

// one-sided at the physical ends, central inside; the operand is zero
// outside the computed range when ZO (as the reference's zeroed temporaries)
template <bool ZO>
__device__ inline double sa_d(const Ext &e, const double *a, long n, long st, int x, int x0, int x1,
                              int c0, int c1)
{
    auto val = [&](long m, int xm) -> double {
        if (ZO && (xm < c0 || xm > c1)) return 0.0;
        return a[m];
    };
    if (x == x0) return val(n + st, x + 1) - a[n];
    if (x == x1) return a[n] - val(n - st, x - 1);
    return (val(n + st, x + 1) - val(n - st, x - 1)) / 2.0;
}

// stage A at node (i,j,k) = n of the computed range: convection
// (sub_tur_sa.f90:152-208) into sa_rhs, nut gradient (:294-338) into w[0..2]
// when W012, alpha/beta/gamma products (:340-356) into w[3..5] and
// nu + (1+cb2) nut (:405-413) into w[6].  With inv_j given, alpha, beta and
// gamma are formed from dxidx and 1/J as the grid set-up forms them
// (sa_metric_products) instead of read.
template <bool W012>
__device__ inline void sa_a_at(const Ext &e, int i, int j, int k, long n, const double *__restrict__ pri,
                               const double *__restrict__ dxidx, const double *__restrict__ nut,
                               const double *__restrict__ amu, const double *__restrict__ aal,
                               const double *__restrict__ bbe, const double *__restrict__ gga,
                               double *__restrict__ sa_rhs, double *__restrict__ w,
                               const double *__restrict__ inv_j = nullptr)
{
    const long N = e.n;
    const long st[3] = {1, (long)e.ni, (long)e.ni * e.nj};
    double u = pri[N + n], v = pri[2 * N + n], ww = pri[3 * N + n];
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double m0 = dxidx[(3 * d) * N + n], m1 = dxidx[(3 * d + 1) * N + n], m2 = dxidx[(3 * d + 2) * N + n];
        double uu = m0 * u + m1 * v + m2 * ww;
        double nutp = nut[n + st[d]], nutc = nut[n], nutm = nut[n - st[d]];
        r[d] = -(0.5 * (uu + fabs(uu)) * (nutc - nutm) + 0.5 * (uu - fabs(uu)) * (nutp - nutc));
    }
    sa_rhs[n] = r[0] + r[1] + r[2];
    double g0 = sa_d<false>(e, nut, n, st[0], i, e.is0, e.ie0, 0, 0);
    double g1 = sa_d<false>(e, nut, n, st[1], j, e.js0, e.je0, 0, 0);
    double g2 = sa_d<false>(e, nut, n, st[2], k, e.ks0, e.ke0, 0, 0);
    if (W012) {
        w[n] = g0;
        w[N + n] = g1;
        w[2 * N + n] = g2;
    }
    double p[9];
    if (inv_j) {
        double m[9];
#pragma unroll
        for (int x = 0; x < 9; x++) m[x] = dxidx[x * N + n];
        sa_metric_products(m, ld_once(inv_j + n), p);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            p[c] = aal[c * N + n];
            p[3 + c] = bbe[c * N + n];
            p[6 + c] = gga[c * N + n];
        }
    }
    w[3 * N + n] = p[0] * g0 + p[1] * g1 + p[2] * g2;
    w[4 * N + n] = p[3] * g0 + p[4] * g1 + p[5] * g2;
    w[5 * N + n] = p[6] * g0 + p[7] * g1 + p[8] * g2;
    double nu0 = amu[n] / pri[n];
    w[6 * N + n] = nu0 + (1.0 + sa::cb2) * nut[n];
}

__global__ void __launch_bounds__(256) k_sa_a(Ext e, Box b, const double *__restrict__ pri,
                                              const double *__restrict__ dxidx,
                                              const double *__restrict__ nut,
                                              const double *__restrict__ amu,
                                              const double *__restrict__ aal,
                                              const double *__restrict__ bbe,
                                              const double *__restrict__ gga,
                                              double *__restrict__ sa_rhs, double *__restrict__ w)
{
    int i, j, k;
    long n;
    if (!lin_node_at(e, b, xcd_slab(blockIdx.x, gridDim.x), i, j, k, n)) return;
    sa_a_at<true>(e, i, j, k, n, pri, dxidx, nut, amu, aal, bbe, gga, sa_rhs, w);
}

// stage B at node (i,j,k) = n of the computed range: diffusion
// (sub_tur_sa.f90:359-512) and source with the diagonal Jacobian (:561-623).
// sa_rhs is completed; the Jacobian term diasd (min(spn + sdn, 0)) is
// returned.  The nut gradient at n comes from w[0..2] (W012) or is formed
// again from nut (the same expressions as stage A).
template <bool W012>
__device__ inline double sa_b_at(const Ext &e, int i, int j, int k, long n, const Consts &c,
                                 const double *__restrict__ pri, const double *__restrict__ dxidx,
                                 const double *__restrict__ inv_j, const double *__restrict__ nut,
                                 const double *__restrict__ amu, const double *__restrict__ vor,
                                 const double *__restrict__ length, double *__restrict__ sa_rhs,
                                 const double *__restrict__ w)
{
    const long N = e.n;
    const long st[3] = {1, (long)e.ni, (long)e.ni * e.nj};
    const double re = c.re;
    double tempa = sa_d<true>(e, w + 3 * N, n, st[0], i, e.is0, e.ie0, e.is1, e.ie1);
    double tempb = sa_d<true>(e, w + 4 * N, n, st[1], j, e.js0, e.je0, e.js1, e.je1);
    double tempc = sa_d<true>(e, w + 5 * N, n, st[2], k, e.ks0, e.ke0, e.ks1, e.ke1);
    double txi = sa_d<true>(e, w + 6 * N, n, st[0], i, e.is0, e.ie0, e.is1, e.ie1);
    double teta = sa_d<true>(e, w + 6 * N, n, st[1], j, e.js0, e.je0, e.js1, e.je1);
    double tzeta = sa_d<true>(e, w + 6 * N, n, st[2], k, e.ks0, e.ke0, e.ks1, e.ke1);
    double m[9];
#pragma unroll
    for (int x = 0; x < 9; x++) m[x] = dxidx[x * N + n];
    double tx = txi * m[0] + teta * m[3] + tzeta * m[6];
    double ty = txi * m[1] + teta * m[4] + tzeta * m[7];
    double tz = txi * m[2] + teta * m[5] + tzeta * m[8];
    double nut0 = nut[n];
    double deltanut = (tempa + tempb + tempc) / ld_once(inv_j + n);
    double g0, g1, g2;
    if (W012) {
        g0 = w[n];
        g1 = w[N + n];
        g2 = w[2 * N + n];
    } else {
        g0 = sa_d<false>(e, nut, n, st[0], i, e.is0, e.ie0, 0, 0);
        g1 = sa_d<false>(e, nut, n, st[1], j, e.js0, e.je0, 0, 0);
        g2 = sa_d<false>(e, nut, n, st[2], k, e.ks0, e.ke0, 0, 0);
    }
    double dnutdx = g0 * m[0] + g1 * m[3] + g2 * m[6];
    double dnutdy = g0 * m[1] + g1 * m[4] + g2 * m[7];
    double dnutdz = g0 * m[2] + g1 * m[5] + g2 * m[8];
    double temp = w[6 * N + n];
    double diff = -(sa::cb2 / (re * sa::dta)) * nut0 * deltanut
                + (1.0 / (re * sa::dta)) * (temp * deltanut + tx * dnutdx + ty * dnutdy + tz * dnutdz);
    double rhs = ld_once(sa_rhs + n) + diff;

    // source (r = nut0 * div as in the reference, :579)
    const double kar2 = sa::kar * sa::kar;
    const double cv13 = sa::cv1 * sa::cv1 * sa::cv1;
    const double cw36 = sa::cw3 * sa::cw3 * sa::cw3 * sa::cw3 * sa::cw3 * sa::cw3;
    // one division per distinct denominator, multiplications by its
    // reciprocal elsewhere (the reference's expressions to a few ulps; pass
    // B 5.81 -> 5.68 ms on the channel)
    const double rkar2 = 1.0 / kar2, rre = rcp_pos(re);
    double d = ld_once(pri + n);
    const double rnu0 = d * rcp_pos(ld_once(amu + n));  // 1 / nu0
    double omega = fabs(ld_once(vor + n));
    double dist0 = ld_once(length + n);
    double leix = fmax(nut0 * rnu0, 0.001);
    double leix2 = leix * leix, leix3 = leix * leix * leix;
    const double den = leix3 + cv13, rden = rcp_pos(den);
    double fv1 = leix3 * rden;
    const double den2 = 1.0 + leix * fv1, rden2 = rcp_pos(den2);
    double fv2 = 1.0 - leix * rden2;
    double dist2 = dist0 * dist0;
    const double rkd = 1.0 / (re * kar2 * dist2), rd2 = 1.0 / dist2;
    double s = omega + nut0 * fv2 * rkd;
    s = fmax(s, 0.3 * omega);
    double div = s * re * kar2 * dist2;
    div = fmax(div, 1.0e-6);
    double r = nut0 * div;
    r = fmin(r, 10.0);
    double r6 = r * r * r * r * r * r;
    double g = r + sa::cw2 * (r6 - r);
    double gg6 = g * g * g * g * g * g;
    double fw = g * pow((1.0 + cw36) * rcp_pos(gg6 + cw36), 1.0 / 6.0);
    double ft2 = sa::ct3 * exp(-sa::ct4 * leix2);
    double aaa = sa::cb1 * ((1.0 - ft2) * fv2 + ft2) * rkar2;
    double bbb = sa::cw1 * fw;
    double ccc = nut0 * rd2;
    double sd1 = (aaa - bbb) * ccc * rre;
    double sd = sd1 * nut0;
    double sp = sa::cb1 * (1.0 - ft2) * omega * nut0;
    st_out(sa_rhs + n, rhs + sd + sp);
    double leixn = rnu0;
    double fv1n = (3.0 * leix2 * leixn * den - leix3 * (3.0 * leix2 * leixn)) * (rden * rden);
    double fv2n = (leixn * den2 - leix * (leixn * fv1 + leix * fv1n)) * (rden2 * rden2);
    double ft2n = -2.0 * sa::ct4 * leix * leixn * ft2;
    double sn = rkd * (fv2 + fv2n * nut0);
    div = (s * s) * (re * kar2 * dist2);
    div = fmax(div, 1.0e-6);
    double rn = (s - sn * nut0) * rcp_pos(div);
    double r5 = r * r * r * r * r;
    double gn = (1.0 + sa::cw2 * 6.0 * r5 - sa::cw2) * rn;
    // the reference's integer powers cw3**(-6), g**(-6), g**(-7) as
    // reciprocals of products (within a few ulps of pow, the same infinities
    // at g = 0; two pow evaluations less: 6.58 -> 5.82 ms on the channel)
    const double cw3m6 = 1.0 / cw36, r1cw = 1.0 / (1.0 + cw3m6);
    double fwn = -(1.0 / 6.0) * pow((1.0 / gg6 + cw3m6) * r1cw, -7.0 / 6.0);
    fwn = fwn * r1cw * (-6.0 * (1.0 / (gg6 * g))) * gn;
    double spn = sa::cb1 * (1.0 - ft2) * omega + sa::cb1 * (-ft2n) * omega * nut0;
    double sdn = rre * (sa::cb1 * (fv2 - ft2 * fv2 + ft2) * rkar2 - sa::cw1 * fw) * 2.0 * nut0 * rd2;
    sdn = rre * (sa::cb1 * (fv2n - ft2n * fv2 - ft2 * fv2n + ft2n) * rkar2 - sa::cw1 * fwn)
              * ((nut0 * nut0) * rd2) + sdn;
    return fmin(spn + sdn, 0.0);
}

__global__ void __launch_bounds__(256) k_sa_b(Ext e, Box b, Consts c, const double *__restrict__ pri,
                                              const double *__restrict__ dxidx,
                                              const double *__restrict__ inv_j,
                                              const double *__restrict__ nut,
                                              const double *__restrict__ amu,
                                              const double *__restrict__ vor,
                                              const double *__restrict__ length,
                                              double *__restrict__ sa_rhs, double *__restrict__ w)
{
    int i, j, k;
    long n;
    if (!lin_node_at(e, b, xcd_slab(blockIdx.x, gridDim.x), i, j, k, n)) return;
    w[7 * e.n + n] = sa_b_at<true>(e, i, j, k, n, c, pri, dxidx, inv_j, nut, amu, vor, length, sa_rhs, w);
}

// contravariant velocities a, b, c and diag of sa_lusgs (sub_tur_sa.f90:
// 719-753) at node n; ds = diasd at n (0 outside the computed range)
__device__ inline void sa_prep_at(const Ext &e, long n, const Consts &c, double phi, double ds,
                                  const double *__restrict__ pri, const double *__restrict__ dxidx,
                                  const double *__restrict__ nut, const double *__restrict__ dt, double *cu4)
{
    const long N = e.n;
    double u = ld_once(pri + N + n), v = ld_once(pri + 2 * N + n), ww = ld_once(pri + 3 * N + n);
    double m[9];
#pragma unroll
    for (int x = 0; x < 9; x++) m[x] = ld_once(dxidx + x * N + n);
    double x2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double c0 = m[0] * u + m[1] * v + m[2] * ww;
    x2 = x2 + m[3] * m[3] + m[4] * m[4] + m[5] * m[5];
    double c1 = m[3] * u + m[4] * v + m[5] * ww;
    x2 = x2 + m[6] * m[6] + m[7] * m[7] + m[8] * m[8];
    double c2 = m[6] * u + m[7] * v + m[8] * ww;
    double tm = 4.0 * x2 * nut[n] / c.re;
    cu4[0] = c0;
    cu4[1] = c1;
    cu4[2] = c2;
    cu4[3] = 1.0 / ld_once(dt + n) * phi + fabs(c0) + fabs(c1) + fabs(c2) + tm - ds;
}

__device__ inline double sa_phi(const Consts &c, const double *cfl)
{
    return (*cfl >= 1.0e3 && c.itime == 0) ? 0.0 : 1.0;
}

// on the ring; dnut* and dnut cleared there.
__global__ void __launch_bounds__(256) k_sa_prep(Ext e, Box ring, Consts c, const double *__restrict__ pri,
                                                 const double *__restrict__ dxidx,
                                                 const double *__restrict__ nut,
                                                 const double *__restrict__ dt, const double *cfl,
                                                 double *__restrict__ w, double *__restrict__ cu,
                                                 double *__restrict__ dnut)
{
    int i, j, k;
    long n;
    if (!lin_node_at(e, ring, xcd_slab(blockIdx.x, gridDim.x), i, j, k, n)) return;
    const long N = e.n;
    bool inner = i >= e.is1 && i <= e.ie1 && j >= e.js1 && j <= e.je1 && k >= e.ks1 && k <= e.ke1;
    double ds = inner ? w[7 * N + n] : 0.0;
    double cu4[4];
    sa_prep_at(e, n, c, sa_phi(c, cfl), ds, pri, dxidx, nut, dt, cu4);
    cu[n] = cu4[0];
    cu[N + n] = cu4[1];
    cu[2 * N + n] = cu4[2];
    cu[3 * N + n] = cu4[3];
    w[8 * N + n] = 0.0;
    dnut[n] = 0.0;
}

template <int DIR>
__global__ void __launch_bounds__(256) k_sa_plane(Ext e, int l, int ja, int ka, const double *__restrict__ cu,
                                                  const double *__restrict__ sa_rhs,
                                                  double *__restrict__ dn1, double *__restrict__ dn)
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
        double a = cu[n], b = cu[N + n], cc = cu[2 * N + n];
        double ta = 0.5 * (a + fabs(a)) * dn1[n - 1];
        double tb = 0.5 * (b + fabs(b)) * dn1[n - sj];
        double tc = 0.5 * (cc + fabs(cc)) * dn1[n - sk];
        dn1[n] = (sa_rhs[n] + (ta + tb + tc)) / cu[3 * N + n];
    } else {
        long n = e.ix(e.ie1 - li, e.je1 - lj, e.ke1 - lk);
        double a = cu[n], b = cu[N + n], cc = cu[2 * N + n];
        double ta = 0.5 * (a - fabs(a)) * dn[n + 1];
        double tb = 0.5 * (b - fabs(b)) * dn[n + sj];
        double tc = 0.5 * (cc - fabs(cc)) * dn[n + sk];
        dn[n] = dn1[n] - (ta + tb + tc) / cu[3 * N + n];
    }
}

__global__ void __launch_bounds__(256) k_add1(Ext e, Box b, double *__restrict__ x,
                                              const double *__restrict__ d)
{
    int i, j, k;
    long n;
    if (!lin_node_at(e, b, xcd_slab(blockIdx.x, gridDim.x), i, j, k, n)) return;
    x[n] = x[n] + d[n];
}

// ------------------------------------------------ every block in one launch
// Pass A over each block's whole storage (linear tiles of 256 nodes): the
// length scale on the extended block, stage A on the computed range
// (without the nut gradient, which pass B forms again).
template <bool DERIVE>
__global__ void __launch_bounds__(256) k_sa_a_batch(const BlockRec *__restrict__ blks, const int *__restrict__ map,
                                                    const KGrid *__restrict__ kg, double *arena, Consts c,
                                                    int length_on)
{
    int bx, by, bz;
    const BlockRec &b = blks[batch_wg(map, kg, bx, by, bz)];
    const Ext e = b.e;
    const Box all{e.is, e.ie, e.js, e.je, e.ks, e.ke};
    int i, j, k;
    long n;
    if (!lin_node_at(e, all, bx, i, j, k, n)) return;
    const double *pri = arena + b.o.pri, *nut = arena + b.o.nut, *amu = arena + b.o.amu;
    if (length_on)
        arena[b.o.length + n] = sa_length_at(e, c, n, arena + b.o.dist, arena + b.o.spacing, pri, nut, amu,
                                             arena + b.o.dudx);
    if (i < e.is1 || i > e.ie1 || j < e.js1 || j > e.je1 || k < e.ks1 || k > e.ke1) return;
    sa_a_at<false>(e, i, j, k, n, pri, arena + b.o.dxidx, nut, amu, arena + b.o.aal, arena + b.o.bbe,
                   arena + b.o.gga, arena + b.o.sa_rhs, arena + b.o.w_sa, DERIVE ? arena + b.o.inv_j : nullptr);
}

// Pass B on the computed range: stage B (sa_rhs completed) and the sweep
// operands a, b, c, diag of the node into w_dqs (components 0..3); block m
// reads its cfl at cfl[m * cfl_stride].  Five waves per SIMD: 86 VGPRs
// with nothing in scratch (126 unbounded, four waves), 3% faster.
__global__ void __launch_bounds__(256, 5) k_sa_b_batch(const BlockRec *__restrict__ blks, const int *__restrict__ map,
                                                    const KGrid *__restrict__ kg, double *arena, Consts c,
                                                    const double *cfl, int cfl_stride)
{
    int bx, by, bz;
    const int m = batch_wg(map, kg, bx, by, bz);
    const BlockRec &b = blks[m];
    const Ext e = b.e;
    const Box inner{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    int i, j, k;
    long n;
    if (!lin_node_at(e, inner, bx, i, j, k, n)) return;
    const double *pri = arena + b.o.pri, *dxidx = arena + b.o.dxidx, *nut = arena + b.o.nut;
    const double ds = sa_b_at<false>(e, i, j, k, n, c, pri, dxidx, arena + b.o.inv_j, nut, arena + b.o.amu,
                                     arena + b.o.vor, arena + b.o.length, arena + b.o.sa_rhs, arena + b.o.w_sa);
    double cu4[4];
    sa_prep_at(e, n, c, sa_phi(c, cfl + (long)m * cfl_stride), ds, pri, dxidx, nut, arena + b.o.dt, cu4);
    double *cu = arena + b.o.w_dqs;
    const long N = e.n;
    st_out(cu + n, cu4[0]);
    st_out(cu + N + n, cu4[1]);
    st_out(cu + 2 * N + n, cu4[2]);
    st_out(cu + 3 * N + n, cu4[3]);
}

int launch_sa_pre_batch(hipStream_t s, const BlockRec *blks, const BatchGrid &bg, double *arena, const Consts &c,
                        const double *cfl, int cfl_stride, int part, bool derive_abg, bool length_on)
{
    if (bg.total <= 0) return 0;
    const dim3 g((unsigned)bg.total), t(64, 4, 1);
    const int lo = length_on ? 1 : 0;
    if (part != 2) {
        if (derive_abg) hipLaunchKernelGGL(k_sa_a_batch<true>, g, t, 0, s, blks, bg.map, bg.kg, arena, c, lo);
        else hipLaunchKernelGGL(k_sa_a_batch<false>, g, t, 0, s, blks, bg.map, bg.kg, arena, c, lo);
    }
    if (part != 1) hipLaunchKernelGGL(k_sa_b_batch, g, t, 0, s, blks, bg.map, bg.kg, arena, c, cfl, cfl_stride);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// Everything before the sweeps: length scale, right-hand side, diagonal.
int launch_sa_pre(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl)
{
    int bs = 256;
    // The reference zeroes sa_rhs and its temporaries every call
    // (sub_tur_sa.f90:48, :276-291).  Here they live in the arena, which is
    // zeroed when allocated, and only their computed range is ever written
    // (w_sa[8] on the ring is reset by k_sa_prep below); sa_rhs outside the
    // computed range is never read.  So they hold those zeros already.
    hipLaunchKernelGGL(k_sa_length, dim3((unsigned)((e.n + bs - 1) / bs)), dim3(bs), 0, s, e, c, f.dist,
                       f.spacing, f.pri, f.nut, f.amu, f.dudx, f.length);
    // linear tiles of 256 nodes over the block's storage, the box masked
    // inside (whole cache lines per wavefront)
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    if (box_grid(b).x == 0) return 0;
    const dim3 g = lin_grid(e), t(64, 4, 1);
    hipLaunchKernelGGL(k_sa_a, g, t, 0, s, e, b, f.pri, f.dxidx, f.nut, f.amu, f.aal, f.bbe, f.gga, f.sa_rhs,
                       f.w_sa);
    hipLaunchKernelGGL(k_sa_b, g, t, 0, s, e, b, c, f.pri, f.dxidx, f.inv_j, f.nut, f.amu, f.vor, f.length,
                       f.sa_rhs, f.w_sa);
    Box ring{e.is1 - 1, e.ie1 + 1, e.js1 - 1, e.je1 + 1, e.ks1 - 1, e.ke1 + 1};
    hipLaunchKernelGGL(k_sa_prep, g, t, 0, s, e, ring, c, f.pri, f.dxidx, f.nut, f.dt, cfl, f.w_sa,
                       f.w_dqs, f.w_dq);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// The two sweeps as hyperplane launches.
int launch_sa_planes(hipStream_t s, const Ext &e, const Fields &f)
{
    dim3 t(64, 4, 1);
    int Li = e.ie1 - e.is1 + 1, Lj = e.je1 - e.js1 + 1, Lk = e.ke1 - e.ks1 + 1;
    if (Li <= 0 || Lj <= 0 || Lk <= 0) return 0;
    int L = Li + Lj + Lk - 2;
    double *dn1 = f.w_sa + 8 * e.n;
    for (int dir = 0; dir < 2; dir++)
        for (int l = 0; l < L; l++) {
            int ja = max(0, l - (Li - 1) - (Lk - 1)), jb = min(Lj - 1, l);
            int ka = max(0, l - (Li - 1) - (Lj - 1)), kb = min(Lk - 1, l);
            dim3 gg((jb - ja + 64) / 64, (kb - ka + 4) / 4, 1);
            if (dir == 0)
                hipLaunchKernelGGL((k_sa_plane<1>), gg, t, 0, s, e, l, ja, ka, f.w_dqs, f.sa_rhs, dn1, f.w_dq);
            else
                hipLaunchKernelGGL((k_sa_plane<-1>), gg, t, 0, s, e, l, ja, ka, f.w_dqs, f.sa_rhs, dn1, f.w_dq);
        }
    ACFD_CHECK_LAUNCH();
    return 0;
}

// nut += dnut on the computed range.
int launch_sa_post(hipStream_t s, const Ext &e, const Fields &f)
{
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    if (box_grid(b).x == 0) return 0;
    hipLaunchKernelGGL(k_add1, lin_grid(e), dim3(64, 4, 1), 0, s, e, b, f.nut, f.w_dq);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// sa_model of one block (sub_tur_sa.f90:7-122) with hyperplane sweeps, or
// with the tile sweeps when `sweep` is given (same results).
int launch_sa(hipStream_t s, const Ext &e, const Fields &f, const Consts &c, const double *cfl,
              const SaSweep *sweep)
{
    int rc = launch_sa_pre(s, e, f, c, cfl);
    if (!rc) {
        if (sweep)
            rc = launch_sa_sweeps(s, e, sweep->shape, f.w_dqs, f.sa_rhs, f.w_sa + 8 * e.n, f.w_dq, sweep->ex,
                                  sweep->err);
        else
            rc = launch_sa_planes(s, e, f);
    }
    if (!rc) rc = launch_sa_post(s, e, f);
    return rc;
}

}  // namespace acfd
