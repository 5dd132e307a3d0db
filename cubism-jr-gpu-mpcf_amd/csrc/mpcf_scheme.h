// mpcf_scheme.h -- device arithmetic of the stiffened-gas RHS.
//
// The scheme is the reference's (WENO5-JS reconstruction of primitives with
// clipping, Einfeldt/Batten wave speeds, HLLC fluxes, quasi-conservative G/P
// terms), re-organised around cells instead of faces:
//
//  * primitives (rho,u,v,w,p,G,P) are recovered once per cell, not once per
//    stencil point per face as in source/GPU/GPUkernels.cu:209-244;
//  * each cell is reconstructed once per direction, giving both its
//    plus-state at face j-1/2 and its minus-state at face j+1/2.  The
//    smoothness indicators of _weno_minus(j-2..j+2) and _weno_pluss(j-2..j+2)
//    (source/GPU/GPUonly.cuh:99-159) are the same three expressions in the
//    same operand order (minus is0 == plus is2, is1 == is1, is2 == is0), so
//    they and the binary64 reciprocals 1/(10 is^2) are computed once;
//  * the HLLC helpers share the side ratios chi = (s-vn)/(s-s*).
// These are exact re-uses of identical expressions: with MPCF_FAST == 0 and
// -ffp-contract=off every value is bit-identical to the reference formula.
//
// MPCF_FAST == 1 keeps the same structure but lets the compiler contract
// a*b+c into FMA (built with -ffp-contract=fast); results then differ from the
// reference in the last bits and are checked at a stated tolerance.
#pragma once

#include <hip/hip_runtime.h>
#include "mpcf_geom.h"

#ifndef MPCF_FAST
#define MPCF_FAST 0
#endif

namespace mpcf {

typedef float Real;

struct Prim { Real r, u, v, w, p, G, P; };

// primitive recovery (GPUkernels.cu:209-244: u=ru/r, ..., p=(E-0.5 r|u|^2-P)/G)
__device__ __forceinline__ Prim to_prim(Real r, Real ru, Real rv, Real rw, Real E, Real G, Real P)
{
    Prim q;
    q.r = r;
    q.u = ru / r;
    q.v = rv / r;
    q.w = rw / r;
    q.p = (E - 0.5f*r*(q.u*q.u + q.v*q.v + q.w*q.w) - P) / G;
    q.G = G;
    q.P = P;
    return q;
}

// WENO5 reconstruction of one cell from its window (a,b,c,d,e) = (j-2..j+2).
// m: minus-state at face j+1/2   == _weno_minus_clipped(a,b,c,d,e)
// p: plus-state  at face j-1/2   == _weno_pluss_clipped(a,b,c,d,e)
// (GPUonly.cuh:99-194; WENOEPS = 1e-6, Makefile.config:15)
struct Rec { Real m, p; };

__device__ __forceinline__ Rec weno_cell(const Real a, const Real b, const Real c, const Real d, const Real e)
{
    const Real k4_3 = (Real)(4./3.), k19_3 = (Real)(19./3.), k11_3 = (Real)(11./3.);
    const Real k25_3 = (Real)(25./3.), k31_3 = (Real)(31./3.), k10_3 = (Real)(10./3.);
    const Real k13_3 = (Real)(13./3.), k5_3 = (Real)(5./3.);
    // the three smoothness indicators (minus is0/is1/is2 == plus is2/is1/is0)
    const Real isA = a*(a*k4_3 - b*k19_3 + c*k11_3) + b*(b*k25_3 - c*k31_3) + c*c*k10_3;
    const Real isB = b*(b*k4_3 - c*k13_3 + d*k5_3) + c*(c*k13_3 - d*k13_3) + d*d*k4_3;
    const Real isC = c*(c*k10_3 - d*k31_3 + e*k11_3) + d*(d*k25_3 - e*k19_3) + e*e*k4_3;
    const Real eps = (Real)1e-6;
    const Real iA = isA + eps, iB = isB + eps, iC = isC + eps;

    Real mA0, mA1, mA2, pA0, pA1, pA2;
#if MPCF_FAST
    const Real rA = 1.0f / (10.0f*iA*iA);
    const Real rB = 1.0f / (10.0f*iB*iB);
    const Real rC = 1.0f / (10.0f*iC*iC);
    mA0 = rA;        mA1 = 6.0f*rB; mA2 = 3.0f*rC;
    pA0 = rC;        pA1 = mA1;     pA2 = 3.0f*rA;
#else
    // binary64: the reference writes the literal 10.0 (GPUonly.cuh:111-113,150-152)
    const double rA = 1.0 / (10.0*(double)iA*(double)iA);
    const double rB = 1.0 / (10.0*(double)iB*(double)iB);
    const double rC = 1.0 / (10.0*(double)iC*(double)iC);
    mA0 = (Real)rA;  mA1 = (Real)(6.0*rB); mA2 = (Real)(3.0*rC);
    pA0 = (Real)rC;  pA1 = mA1;            pA2 = (Real)(3.0*rA);
#endif
    const Real k1_3 = (Real)(1./3.), k7_6 = (Real)(7./6.), k11_6 = (Real)(11./6.);
    const Real k1_6 = (Real)(1./6.), k5_6 = (Real)(5./6.);

    Rec out;
    {   // _weno_minus(a,b,c,d,e)
        const Real sum = mA0 + mA1 + mA2;
        const Real inv = (Real)1 / sum;
        const Real w0 = mA0 * inv, w1 = mA1 * inv;
        const Real w2 = 1 - w0 - w1;
        out.m = w0*(k1_3*a - k7_6*b + k11_6*c) + w1*(-k1_6*b + k5_6*c + k1_3*d) + w2*(k1_3*c + k5_6*d - k1_6*e);
    }
    {   // _weno_pluss(a,b,c,d,e)  (its b..f are this cell's a..e)
        const Real sum = pA0 + pA1 + pA2;
        const Real inv = (Real)1 / sum;
        const Real w0 = pA0 * inv, w1 = pA1 * inv;
        const Real w2 = 1 - w0 - w1;
        out.p = w0*(k1_3*e - k7_6*d + k11_6*c) + w1*(-k1_6*d + k5_6*c + k1_3*b) + w2*(k1_3*c + k5_6*b - k1_6*a);
    }
    // clipping to the three central cells (identical bounds for both states)
    const Real mn = fminf(fminf(b, c), d);
    const Real mx = fmaxf(fmaxf(b, c), d);
    out.m = fminf(fmaxf(out.m, mn), mx);
    out.p = fminf(fmaxf(out.p, mn), mx);
    return out;
}

// one cell's reconstructed face states in one direction
struct CellRec { Prim m, p; };

__device__ __forceinline__ CellRec reconstruct(const Prim& A, const Prim& B, const Prim& C, const Prim& D, const Prim& E)
{
    CellRec o;
    Rec t;
    t = weno_cell(A.r, B.r, C.r, D.r, E.r); o.m.r = t.m; o.p.r = t.p;
    t = weno_cell(A.u, B.u, C.u, D.u, E.u); o.m.u = t.m; o.p.u = t.p;
    t = weno_cell(A.v, B.v, C.v, D.v, E.v); o.m.v = t.m; o.p.v = t.p;
    t = weno_cell(A.w, B.w, C.w, D.w, E.w); o.m.w = t.m; o.p.w = t.p;
    t = weno_cell(A.p, B.p, C.p, D.p, E.p); o.m.p = t.m; o.p.p = t.p;
    t = weno_cell(A.G, B.G, C.G, D.G, E.G); o.m.G = t.m; o.p.G = t.p;
    t = weno_cell(A.P, B.P, C.P, D.P, E.P); o.m.P = t.m; o.p.P = t.p;
    return o;
}

// fluxes at one face, in field order, plus the HLLC face velocity of the
// quasi-conservative G/P terms
struct FaceFlux { Real f[NVAR]; Real vel; };

// D = 0,1,2: normal direction.  L = minus-state (left cell), Rt = plus-state.
// GPUonly.cuh:802-1042 with argument orders of GPUkernels.cu:259-283 (x),
// 638-662 (y), 934-957 (z).
template <int D>
__device__ __forceinline__ FaceFlux riemann(const Prim& L, const Prim& Rt)
{
    // normal and the two transverse velocities in the order _hllc_e receives them
    const Real vnm  = D == 0 ? L.u  : D == 1 ? L.v  : L.w;
    const Real vnp  = D == 0 ? Rt.u : D == 1 ? Rt.v : Rt.w;
    const Real v1m  = D == 0 ? L.v  : L.u;
    const Real v1p  = D == 0 ? Rt.v : Rt.u;
    const Real v2m  = D == 2 ? L.v  : L.w;
    const Real v2p  = D == 2 ? Rt.v : Rt.w;
    const Real rm = L.r, rp = Rt.r, pm = L.p, pp = Rt.p;
    const Real Gm = L.G, Gp = Rt.G, Pm = L.P, Pp = Rt.P;

    // _char_vel_einfeldt (GPUonly.cuh:823-846)
    const Real Rr   = sqrtf(rp / rm);
    const Real Rinv = 1.0f / (1.0f + Rr);
    const Real cm2 = ((pm + Pm)/Gm + pm) / rm;
    const Real cp2 = ((pp + Pp)/Gp + pp) / rp;
    const Real cm  = sqrtf(cm2);
    const Real cp  = sqrtf(cp2);
    const Real eta_2 = 0.5f*Rr*Rinv*Rinv;
    const Real d2    = (cm2 + Rr*cp2)*Rinv + eta_2*(vnp - vnm)*(vnp - vnm);
    const Real dd    = sqrtf(d2);
    const Real ubar  = (vnm + Rr*vnp)*Rinv;
    const Real sm = fminf(ubar - dd, vnm - cm);
    const Real sp = fmaxf(ubar + dd, vnp + cp);

    // _char_vel_star (GPUonly.cuh:863-865)
    const Real facm = rm * (sm - vnm);
    const Real facp = rp * (sp - vnp);
    const Real ss = (pp - pm + vnm*facm - vnp*facp) / (facm - facp);

    // shared HLLC terms (GPUonly.cuh:883-889 and the identical lines in the
    // other helpers)
    const int sign_star = ss == 0 ? 0 : (ss < 0 ? -1 : 1);
    const Real s_minus = fminf(0.0f, sm);
    const Real s_pluss = fmaxf(0.0f, sp);
    const Real chi_m = (sm - vnm) / (sm - ss);
    const Real chi_p = (sp - vnp) / (sp - ss);
    const Real wl = 0.5f*(1 + sign_star);
    const Real wr = 0.5f*(1 - sign_star);

    FaceFlux o;
    // _hllc_rho for rho, G, P: right weight is binary64 (GPUonly.cuh:900)
    auto hllc_rho = [&](const Real qm, const Real qp) -> Real {
        const Real q_deltam = qm*chi_m - qm;
        const Real q_deltap = qp*chi_p - qp;
        const Real fm = qm*vnm;
        const Real fp = qp*vnp;
#if MPCF_FAST
        return wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
#else
        return (Real)((double)(wl * (fm + s_minus*q_deltam)) + (0.5*(1 - sign_star)) * (double)(fp + s_pluss*q_deltap));
#endif
    };
    // _hllc_vel for a transverse momentum (GPUonly.cuh:926-938)
    auto hllc_vel = [&](const Real vtm, const Real vtp) -> Real {
        const Real qm = rm*vtm;
        const Real qp = rp*vtp;
        const Real q_deltam = qm*chi_m - qm;
        const Real q_deltap = qp*chi_p - qp;
        const Real fm = qm*vnm;
        const Real fp = qp*vnp;
        return wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
    };
    // _hllc_pvel, normal momentum (GPUonly.cuh:967-979)
    Real fpv;
    {
        const Real qm = rm*vnm;
        const Real qp = rp*vnp;
        const Real q_deltam = rm*ss*chi_m - qm;
        const Real q_deltap = rp*ss*chi_p - qp;
        const Real fm = qm*vnm + pm;
        const Real fp = qp*vnp + pp;
        fpv = wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
    }
    // _hllc_e (GPUonly.cuh:1011-1023)
    Real fe;
    {
        const Real qm = Gm*pm + Pm + 0.5f*rm*(vnm*vnm + v1m*v1m + v2m*v2m);
        const Real qp = Gp*pp + Pp + 0.5f*rp*(vnp*vnp + v1p*v1p + v2p*v2p);
        const Real q_deltam = chi_m*(qm + (ss - vnm)*(rm*ss + pm/(sm - vnm))) - qm;
        const Real q_deltap = chi_p*(qp + (ss - vnp)*(rp*ss + pp/(sp - vnp))) - qp;
        const Real fm = vnm*(qm + pm);
        const Real fp = vnp*(qp + pp);
        fe = wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
    }
    o.f[FR] = hllc_rho(rm, rp);
    if (D == 0) { o.f[FU] = fpv;                o.f[FV] = hllc_vel(L.v, Rt.v); o.f[FW] = hllc_vel(L.w, Rt.w); }
    if (D == 1) { o.f[FU] = hllc_vel(L.u, Rt.u); o.f[FV] = fpv;               o.f[FW] = hllc_vel(L.w, Rt.w); }
    if (D == 2) { o.f[FU] = hllc_vel(L.u, Rt.u); o.f[FV] = hllc_vel(L.v, Rt.v); o.f[FW] = fpv; }
    o.f[FE] = fe;
    o.f[FG] = hllc_rho(Gm, Gp);
    o.f[FP] = hllc_rho(Pm, Pp);
    // _extraterm_hllc_vel (GPUonly.cuh:1035-1041)
    {
        const Real chim1 = chi_m - 1.0f;
        const Real chip1 = chi_p - 1.0f;
        o.vel = wl*(vnm + s_minus*chim1) + wr*(vnp + s_pluss*chip1);
    }
    return o;
}

} // namespace mpcf
