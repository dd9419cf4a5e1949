// Face-flux building blocks shared by the inviscid kernels (templated on
// the arithmetic type R: double, or float for the FP32-flux variant).
#pragma once
#include "common.hpp"

namespace acfd {

// The formulas are the reference's, rearranged to need fewer divisions
// (FP64 division is a ~10-instruction sequence on CDNA): the WENO3 weights
// a1/(a1+a2), a2/(a1+a2) with a1 = 1/(3 t1^2), a2 = 2/(3 t2^2) are
// t2^2/(t2^2 + 2 t1^2) and 2 t1^2/(t2^2 + 2 t1^2) (one division instead of
// four), and the Roe flux divides once per distinct denominator; the
// positive denominators (WENO3's, the densities, 1 + sqrt(rho_r/rho_l),
// gamma - 1, 2 c^2) go through rcp_pos, and sqrt(rho_r/rho_l) is
// sqrt(rho_l rho_r)/rho_l.  Results agree with the reference's operation order to
// rounding.
//

// WENO3 state at a face from three values; sub_inviscid_upwind.f90:32-41
// (right state: u1,u2,u3 = f(a),f(a+1),f(a+2)) and :57-66 (left state:
// f(a-1),f(a),f(a+1)).  `near` is the value next to the face on the side of
// the state, `far` the one beyond it, `other` the value across the face.
template <typename R>
__device__ inline R weno3(R far, R near, R other)
{
    const R eps = (R)1.0e-6;
    const R s1 = (near - far) * (near - far);
    const R s2 = (other - near) * (other - near);
    const R t1 = eps + s1, t2 = eps + s2;
    const R T1 = t1 * t1, T2 = t2 * t2;
    const R inv = rcp_pos<R>(T2 + (R)2.0 * T1);
    const R w1 = T2 * inv, w2 = (R)2.0 * T1 * inv;
    return w1 * ((R)1.5 * near - (R)0.5 * far) + w2 * ((R)0.5 * (near + other));
}

// Roe flux with Harten's entropy fix (delta = 0.06) on a face with metric
// (mx,my,mz) from primitive left/right states (d,u,v,w,p);
// sub_inviscid_splitting.f90:65-162.
template <typename R>
__device__ inline void roe_flux(const R *ql, const R *qr, R mx, R my, R mz, R gamma, R *out)
{
    const R delt = (R)0.06, rdelt2 = (R)1.0 / ((R)2.0 * (R)0.06);
    const R g1 = gamma - (R)1.0, rg1 = rcp_pos<R>(g1), ggg = gamma * rg1;
    const R temp = sqrt(mx * mx + my * my + mz * mz);
    const R rt = (R)1.0 / temp;
    const R nx = mx * rt, ny = my * rt, nz = mz * rt;
    const R dl = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4];
    const R dr = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4];
    const R rdl = rcp_pos<R>(dl);
    const R hl = ggg * pl * rdl + (R)0.5 * (ul * ul + vl * vl + wl * wl);
    const R hr = ggg * pr * rcp_pos<R>(dr) + (R)0.5 * (ur * ur + vr * vr + wr * wr);
    const R uul = ul * mx + vl * my + wl * mz;
    const R uur = ur * mx + vr * my + wr * mz;
    const R ubarl = uul * rt, ubarr = uur * rt;
    const R avd = sqrt(dl * dr);
    const R tmpp = avd * rdl;  // sqrt(dr / dl)
    const R r1t = rcp_pos<R>((R)1.0 + tmpp);
    const R avu = (ul + ur * tmpp) * r1t;
    const R avv = (vl + vr * tmpp) * r1t;
    const R avw = (wl + wr * tmpp) * r1t;
    const R avh = (hl + hr * tmpp) * r1t;
    const R ava = sqrt(g1 * (avh - (R)0.5 * (avu * avu + avv * avv + avw * avw)));
    const R avuu = avu * nx + avv * ny + avw * nz;
    R l1 = fabs(avuu - ava), l2 = fabs(avuu), l5 = fabs(avuu + ava);
    // Harten: sign(1, delt - lamda) >= 0 selects (l^2 + d^2)/(2d)
    if (delt - l1 >= (R)0) l1 = (l1 * l1 + delt * delt) * rdelt2;
    if (delt - l2 >= (R)0) l2 = (l2 * l2 + delt * delt) * rdelt2;
    if (delt - l5 >= (R)0) l5 = (l5 * l5 + delt * delt) * rdelt2;
    const R deltad = dr - dl, deltap = pr - pl;
    const R deltau = ur - ul, deltav = vr - vl, deltaw = wr - wl;
    const R deltauu = ubarr - ubarl;
    const R ava2 = ava * ava;
    const R rh = rcp_pos<R>((R)2.0 * ava2);
    const R alpha1 = l2 * (deltad - deltap * ((R)2.0 * rh));
    const R alpha2 = l5 * (deltap + avd * ava * deltauu) * rh;
    const R alpha3 = l1 * (deltap - avd * ava * deltauu) * rh;
    const R alpha4 = alpha1 + alpha2 + alpha3;
    const R alpha5 = ava * (alpha2 - alpha3);
    const R alpha6 = l2 * (avd * deltau - avd * deltauu * nx);
    const R alpha7 = l2 * (avd * deltav - avd * deltauu * ny);
    const R alpha8 = l2 * (avd * deltaw - avd * deltauu * nz);
    const R f0 = alpha4;
    const R f1 = avu * alpha4 + alpha5 * nx + alpha6;
    const R f2 = avv * alpha4 + alpha5 * ny + alpha7;
    const R f3 = avw * alpha4 + alpha5 * nz + alpha8;
    const R f4 = avh * alpha4 + avuu * alpha5 + avu * alpha6 + avv * alpha7 + avw * alpha8
               - ava2 * alpha1 * rg1;
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
    const R r0 = rcp_pos<R>(u[0]);
    o[0] = u[0];
    o[1] = u[1] * r0;
    o[2] = u[2] * r0;
    o[3] = u[3] * r0;
    o[4] = (u[4] - (R)0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * r0) * (gamma - (R)1.0);
}

}  // namespace acfd
