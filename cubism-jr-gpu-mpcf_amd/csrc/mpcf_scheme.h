// mpcf_scheme.h -- arithmetic of the stiffened-gas RHS (host + device).
//
// The scheme is the reference's (WENO5-JS reconstruction of primitives with
// clipping, Einfeldt/Batten wave speeds, HLLC fluxes, quasi-conservative G/P
// terms), re-organised around cells instead of faces:
//
//  * primitives (rho,u,v,w,p,G,P) are recovered once per cell and sweep, not
//    once per stencil point per face as in source/GPU/GPUkernels.cu:209-244;
//  * each cell is reconstructed once per direction, giving both its
//    plus-state at face j-1/2 and its minus-state at face j+1/2.  The
//    smoothness indicators of _weno_minus(j-2..j+2) and _weno_pluss(j-2..j+2)
//    (source/GPU/GPUonly.cuh:99-159) are the same three expressions in the
//    same operand order (minus is0 == plus is2, is1 == is1, is2 == is0), so
//    they and the binary64 reciprocals 1/(10 is^2) are computed once;
//  * the HLLC helpers share the side ratios chi = (s-vn)/(s-s*).
// These are exact re-uses of identical expressions: with FAST == false and
// -ffp-contract=off every value is bit-identical to the reference formula.
//
// FAST == true evaluates the WENO weights in binary32, replaces divisions and
// square roots by the hardware reciprocal / square root (1 ulp) and is built
// with FMA contraction; results then differ from the reference in the last
// bits and are checked at a stated tolerance.
//
// Every function is templated on the scalar type T (float on the device) so
// that the host can instantiate it with a counting type and report the exact
// operation count of the scheme as implemented here.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include "mpcf_geom.h"

namespace mpcf {

typedef float Real;

// binary64 companion of a binary32 scalar type
template <typename T> struct Wide { typedef double type; };

MPCF_HD float mmin(float a, float b) { return fminf(a, b); }
MPCF_HD float mmax(float a, float b) { return fmaxf(a, b); }
MPCF_HD float msqrt(float a) { return sqrtf(a); }
MPCF_HD double mmin(double a, double b) { return fmin(a, b); }
MPCF_HD double mmax(double a, double b) { return fmax(a, b); }
MPCF_HD double msqrt(double a) { return sqrt(a); }
MPCF_HD double mrcp(double a) { return 1.0 / a; }
MPCF_HD double msqrt_fast(double a) { return sqrt(a); }
// approximate reciprocal / square root of the FAST build: the hardware's
// v_rcp_f32 / v_sqrt_f32 (1 ulp) instead of the correctly rounded sequences
MPCF_HD float mrcp(float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(a);
#else
    return 1.0f / a;
#endif
}
MPCF_HD float msqrt_fast(float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(a);
#else
    return sqrtf(a);
#endif
}

// x clamped to [lo, hi] for lo <= hi: one v_med3_f32 on the device (FAST)
MPCF_HD float mclamp(float x, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
    return fminf(fmaxf(x, lo), hi);
#endif
}
MPCF_HD double mclamp(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

// a / b: IEEE division (bit-exact build) or a * rcp(b) (FAST)
template <bool FAST, typename T> MPCF_HD T qdiv(const T a, const T b) { return FAST ? a * mrcp(b) : a / b; }
template <bool FAST, typename T> MPCF_HD T qsqrt(const T a) { return FAST ? msqrt_fast(a) : msqrt(a); }

template <typename T>
struct PrimT { T r, u, v, w, p, G, P; };
typedef PrimT<Real> Prim;

// primitive recovery (GPUkernels.cu:209-244: u=ru/r, ..., p=(E-0.5 r|u|^2-P)/G)
//
// P64 (selective binary64, binary32 storage): the pressure is recovered in
// binary64 from the stored binary32 values and rounded once.  In liquid cells
// E and P are both close to gamma*pc/(gamma-1) (4828.79 for water) while
// G*p is a few units, so E - KE - P cancels all but the last few bits of a
// binary32 evaluation; in binary64 that cancellation is exact and p carries
// only the storage error of E.
//
// EPAIR (E stored as a binary32 pair, E = hi + lo; implies P64): Elo is the
// low part, and the recovery uses hi + lo, exact in binary64.
template <typename T, bool FAST = false, bool P64 = false, bool EPAIR = false>
MPCF_HD PrimT<T> to_prim(T r, T ru, T rv, T rw, T E, T G, T P, T Elo = T())
{
    PrimT<T> q;
    q.r = r;
    if constexpr (P64 || EPAIR) {
        const double rd = r;
        const double ird = 1.0 / rd;
        const double ud = ru * ird, vd = rv * ird, wd = rw * ird;
        q.u = T(ud);
        q.v = T(vd);
        q.w = T(wd);
        const double ke = 0.5 * (double(ru)*ud + double(rv)*vd + double(rw)*wd);
        const double Ed = EPAIR ? double(E) + double(Elo) : double(E);
        q.p = T(((Ed - ke) - double(P)) / double(G));
    } else if (FAST) {
        const T ir = mrcp(r);
        q.u = ru * ir;
        q.v = rv * ir;
        q.w = rw * ir;
        q.p = (E - 0.5f*r*(q.u*q.u + q.v*q.v + q.w*q.w) - P) * mrcp(G);
    } else {
        q.u = ru / r;
        q.v = rv / r;
        q.w = rw / r;
        q.p = (E - 0.5f*r*(q.u*q.u + q.v*q.v + q.w*q.w) - P) / G;
    }
    q.G = G;
    q.P = P;
    return q;
}

// WENO5 reconstruction of one cell from its window (a,b,c,d,e) = (j-2..j+2).
// m: minus-state at face j+1/2   == _weno_minus_clipped(a,b,c,d,e)
// p: plus-state  at face j-1/2   == _weno_pluss_clipped(a,b,c,d,e)
// (GPUonly.cuh:99-194; WENOEPS = 1e-6, Makefile.config:15)
template <typename T> struct RecT { T m, p; };

// the FAST reconstruction from the window's centre three values and its four
// first differences dab = b-a, dbc = c-b, dcd = d-c, dde = e-d (weno_cell
// with FAST == true is this on the differences it forms; a caller that has
// the differences already, e.g. from neighbouring lanes, passes them in)
template <typename T>
MPCF_HD RecT<T> weno_fast_diff(const T b, const T c, const T d, const T dab, const T dbc, const T dcd, const T dde)
{
    RecT<T> out;
    const T Db = dbc - dab, Dc = dcd - dbc, Dd = dde - dcd;   // second differences
    const T tA = Db + 2.0f*dbc;                               // a - 4b + 3c
    const T tB = dbc + dcd;                                   // d - b
    const T tC = Dd - 2.0f*dcd;                               // 3c - 4d + e
    // the indicators times 4, 13/3 D^2 + t^2 + 4 eps: each is exactly 4 times
    // 13/12 D^2 + 1/4 t^2 + eps (a power-of-two scale of every operand), with
    // one product fewer; the normalised weights depend only on their ratios
    const T k13_3 = 4.0f*T(13./12.);
    const T eps4 = 4.0f*T(1e-6);
    const T iA = Db*(k13_3*Db) + (tA*tA + eps4);
    const T iB = Dc*(k13_3*Dc) + (tB*tB + eps4);
    const T iC = Dd*(k13_3*Dd) + (tC*tC + eps4);
    const T rA = mrcp(iA*iA), rB = mrcp(iB*iB), rC = mrcp(iC*iC);
    // The minus state is c + sum_k w_k n_k / 6 with candidates 6 (q_k - c)
    // n0 = 2 tA - (c-b), n1 = (c-b) + 2 (d-c), n2 = 4 (d-c) - (e-d) and
    // weights w ~ (rA, 6 rB, 3 rC).  Written about the central candidate,
    // sum_k w_k n_k = n1 + w0 (n0 - n1) + w2 (n2 - n1), where n0 - n1 = 2 X
    // and n2 - n1 = Y are third differences, X = Db - Dc and Y = Dc - Dd:
    //   m = c + n1/6 + (rA X / 3 + rC Y / 2) / (rA + 6 rB + 3 rC).
    // The plus state is the mirror (weights ~ (rC, 6 rB, 3 rA), candidates
    // 2 tC + (d-c), -((d-c) + 2 (c-b)), (b-a) - 4 (c-b), differences -2 Y
    // and -X about the central one):
    //   p = c - ((d-c) + 2 (c-b))/6 - (rC Y / 3 + rA X / 2) / (rC + 6 rB + 3 rA).
    // Equal to the weighted sum of the candidates in exact arithmetic, with
    // four operations fewer per reconstruction.
    const T X = Db - Dc, Y = Dc - Dd;
    const T k1_3 = T(1./3.), k1_6 = T(1./6.);
    const T rB6 = 6.0f*rB;
    const T Sm = (rA + rB6) + 3.0f*rC;
    const T Sp = (rC + rB6) + 3.0f*rA;
    const T AX = rA*X, CY = rC*Y;
    const T qm = AX*k1_3 + 0.5f*CY;
    const T qp = CY*k1_3 + 0.5f*AX;
    const T bm = (tB + dcd)*k1_6 + c;              // c + n1/6
    const T bp = c - (tB + dbc)*k1_6;              // c - ((d-c) + 2 (c-b))/6
    const T om = qm*mrcp(Sm) + bm;
    const T op = bp - qp*mrcp(Sp);
    const T mn = mmin(mmin(b, c), d);
    const T mx = mmax(mmax(b, c), d);
    out.m = mclamp(om, mn, mx);
    out.p = mclamp(op, mn, mx);
    return out;
}

template <typename T, bool FAST>
MPCF_HD RecT<T> weno_cell(const T a, const T b, const T c, const T d, const T e)
{
    typedef typename Wide<T>::type W;
    const T k4_3 = T(4./3.), k19_3 = T(19./3.), k11_3 = T(11./3.);
    const T k25_3 = T(25./3.), k31_3 = T(31./3.), k10_3 = T(10./3.);
    const T k13_3 = T(13./3.), k5_3 = T(5./3.);
    RecT<T> out;
    const T eps = T(1e-6);
    if constexpr (FAST) {
        // The same arithmetic rearranged (equal in exact arithmetic, rounded
        // differently), on the first differences of the window:
        //  * the indicators in Jiang-Shu form 13/12 D^2 + 1/4 t^2 (expanding
        //    gives the reference's coefficients 4/3, -19/3, 11/3, 25/3, -31/3,
        //    10/3 and 4/3, -13/3, 5/3, 13/3, -13/3, 4/3 term by term), eps
        //    folded in, and evaluated times 4 (13/3 D^2 + t^2 + 4 eps);
        //  * the weights only normalised, omega_k = alpha_k / sum(alpha): the
        //    common factors of alpha_k = d_k / (10 is_k^2) dropped, one
        //    reciprocal of the sum per state;
        //  * the candidates as c + (differences)/6, e.g. (2a - 7b + 11c)/6 =
        //    c + (-2 (b-a) + 5 (c-b))/6, so a state is
        //    c + sum(alpha_k n_k) / (6 sum(alpha));
        //  * the clip to the three central cells as one median-of-three.
        return weno_fast_diff<T>(b, c, d, b - a, c - b, d - c, e - d);
    } else {
        // the three smoothness indicators (minus is0/is1/is2 == plus is2/is1/is0)
        const T isA = a*(a*k4_3 - b*k19_3 + c*k11_3) + b*(b*k25_3 - c*k31_3) + c*c*k10_3;
        const T isB = b*(b*k4_3 - c*k13_3 + d*k5_3) + c*(c*k13_3 - d*k13_3) + d*d*k4_3;
        const T isC = c*(c*k10_3 - d*k31_3 + e*k11_3) + d*(d*k25_3 - e*k19_3) + e*e*k4_3;
        const T iA = isA + eps, iB = isB + eps, iC = isC + eps;
        const T k1_3 = T(1./3.), k7_6 = T(7./6.), k11_6 = T(11./6.);
        const T k1_6 = T(1./6.), k5_6 = T(5./6.), mk1_6 = T(-(1./6.));
        // binary64: the reference writes the literal 10.0 (GPUonly.cuh:111-113,150-152)
        const W rA = W(1.0) / (10.0*W(iA)*W(iA));
        const W rB = W(1.0) / (10.0*W(iB)*W(iB));
        const W rC = W(1.0) / (10.0*W(iC)*W(iC));
        const T mA0 = T(rA), mA1 = T(6.0*rB), mA2 = T(3.0*rC);
        const T pA0 = T(rC), pA1 = mA1,       pA2 = T(3.0*rA);
        {   // _weno_minus(a,b,c,d,e)
            const T sum = mA0 + mA1 + mA2;
            const T inv = qdiv<FAST>(T(1.0f), sum);
            const T w0 = mA0 * inv, w1 = mA1 * inv;
            const T w2 = T(1.0f) - w0 - w1;
            out.m = w0*(k1_3*a - k7_6*b + k11_6*c) + w1*(mk1_6*b + k5_6*c + k1_3*d) + w2*(k1_3*c + k5_6*d - k1_6*e);
        }
        {   // _weno_pluss(a,b,c,d,e)  (its b..f are this cell's a..e)
            const T sum = pA0 + pA1 + pA2;
            const T inv = qdiv<FAST>(T(1.0f), sum);
            const T w0 = pA0 * inv, w1 = pA1 * inv;
            const T w2 = T(1.0f) - w0 - w1;
            out.p = w0*(k1_3*e - k7_6*d + k11_6*c) + w1*(mk1_6*d + k5_6*c + k1_3*b) + w2*(k1_3*c + k5_6*b - k1_6*a);
        }
        // clipping to the three central cells (identical bounds for both states)
        const T mn = mmin(mmin(b, c), d);
        const T mx = mmax(mmax(b, c), d);
        out.m = mmin(mmax(out.m, mn), mx);
        out.p = mmin(mmax(out.p, mn), mx);
        return out;
    }
}

// one cell's reconstructed face states in one direction
template <typename T> struct CellRecT { PrimT<T> m, p; };

template <typename T, bool FAST>
MPCF_HD CellRecT<T> reconstruct(const PrimT<T>& A, const PrimT<T>& B, const PrimT<T>& C, const PrimT<T>& D,
                                const PrimT<T>& E)
{
    CellRecT<T> o;
    RecT<T> t;
    t = weno_cell<T, FAST>(A.r, B.r, C.r, D.r, E.r); o.m.r = t.m; o.p.r = t.p;
    t = weno_cell<T, FAST>(A.u, B.u, C.u, D.u, E.u); o.m.u = t.m; o.p.u = t.p;
    t = weno_cell<T, FAST>(A.v, B.v, C.v, D.v, E.v); o.m.v = t.m; o.p.v = t.p;
    t = weno_cell<T, FAST>(A.w, B.w, C.w, D.w, E.w); o.m.w = t.m; o.p.w = t.p;
    t = weno_cell<T, FAST>(A.p, B.p, C.p, D.p, E.p); o.m.p = t.m; o.p.p = t.p;
    t = weno_cell<T, FAST>(A.G, B.G, C.G, D.G, E.G); o.m.G = t.m; o.p.G = t.p;
    t = weno_cell<T, FAST>(A.P, B.P, C.P, D.P, E.P); o.m.P = t.m; o.p.P = t.p;
    return o;
}

// fluxes at one face, in field order, plus the HLLC face velocity of the
// quasi-conservative G/P terms
template <typename T> struct FaceFluxT { T f[NVAR]; T vel; };

// D = 0,1,2: normal direction.  L = minus-state (left cell), Rt = plus-state.
// GPUonly.cuh:802-1042 with argument orders of GPUkernels.cu:259-283 (x),
// 638-662 (y), 934-957 (z).  The states are taken by value: the FAST path
// selects fields of one side, and selects between loads through references
// would keep the callers' register-resident march state in memory.
template <int D, typename T, bool FAST>
MPCF_HD FaceFluxT<T> riemann(const PrimT<T> L, const PrimT<T> Rt)
{
    typedef typename Wide<T>::type W;
    // normal and the two transverse velocities in the order _hllc_e receives them
    const T vnm  = D == 0 ? L.u  : D == 1 ? L.v  : L.w;
    const T vnp  = D == 0 ? Rt.u : D == 1 ? Rt.v : Rt.w;
    const T v1m  = D == 0 ? L.v  : L.u;
    const T v1p  = D == 0 ? Rt.v : Rt.u;
    const T v2m  = D == 2 ? L.v  : L.w;
    const T v2p  = D == 2 ? Rt.v : Rt.w;
    const T rm = L.r, rp = Rt.r, pm = L.p, pp = Rt.p;
    const T Gm = L.G, Gp = Rt.G, Pm = L.P, Pp = Rt.P;

    // _char_vel_einfeldt (GPUonly.cuh:823-846)
    const T Rr   = qsqrt<FAST>(qdiv<FAST>(rp, rm));
    const T Rinv = qdiv<FAST>(T(1.0f), 1.0f + Rr);
    const T cm2 = qdiv<FAST>(qdiv<FAST>(pm + Pm, Gm) + pm, rm);
    const T cp2 = qdiv<FAST>(qdiv<FAST>(pp + Pp, Gp) + pp, rp);
    const T cm  = qsqrt<FAST>(cm2);
    const T cp  = qsqrt<FAST>(cp2);
    const T eta_2 = 0.5f*Rr*Rinv*Rinv;
    const T d2    = (cm2 + Rr*cp2)*Rinv + eta_2*(vnp - vnm)*(vnp - vnm);
    const T dd    = qsqrt<FAST>(d2);
    const T ubar  = (vnm + Rr*vnp)*Rinv;
    const T sm = mmin(ubar - dd, vnm - cm);
    const T sp = mmax(ubar + dd, vnp + cp);

    // _char_vel_star (GPUonly.cuh:863-865)
    const T facm = rm * (sm - vnm);
    const T facp = rp * (sp - vnp);
    const T ss = qdiv<FAST>(pp - pm + vnm*facm - vnp*facp, facm - facp);

    FaceFluxT<T> o;
    if constexpr (FAST) {
        // One side's star state gives every flux: the left for ss > 0, the
        // right for ss < 0.  At ss == 0 the reference averages the two, which
        // are equal there (F*_R - F*_L = ss (U*_R - U*_L)), so the left is
        // taken.  For rho, G, P and the transverse momenta the HLLC flux
        // q vn + s (q chi - q) is q times the face velocity vel.
        const bool lt = !(ss < 0.0f);
        const T vn = lt ? vnm : vnp, r = lt ? rm : rp, p = lt ? pm : pp;
        const T G = lt ? Gm : Gp, P = lt ? Pm : Pp;
        const T v1 = lt ? v1m : v1p, v2 = lt ? v2m : v2p;
        const T sk = lt ? sm : sp;
        const T s = lt ? mmin(T(0.0f), sm) : mmax(T(0.0f), sp);
        const T chi = qdiv<FAST>(sk - vn, sk - ss);
        const T vel = vn + s*(chi - 1.0f);                      // _extraterm_hllc_vel
        const T rv = r*vel;
        const T rvn = r*vn;
        const T fpv = rvn*vn + p + s*(r*ss*chi - rvn);          // _hllc_pvel
        const T E = G*p + P + 0.5f*r*(vn*vn + v1*v1 + v2*v2);   // _hllc_e
        const T fe = vn*(E + p) + s*(chi*(E + (ss - vn)*(r*ss + qdiv<FAST>(p, sk - vn))) - E);
        o.f[FR] = rv;
        if (D == 0) { o.f[FU] = fpv;    o.f[FV] = rv*v1; o.f[FW] = rv*v2; }
        if (D == 1) { o.f[FU] = rv*v1;  o.f[FV] = fpv;   o.f[FW] = rv*v2; }
        if (D == 2) { o.f[FU] = rv*v1;  o.f[FV] = rv*v2; o.f[FW] = fpv; }
        o.f[FE] = fe;
        o.f[FG] = G*vel;
        o.f[FP] = P*vel;
        o.vel = vel;
        return o;
    }

    // shared HLLC terms (GPUonly.cuh:883-889 and the identical lines in the
    // other helpers)
    const int sign_star = ss == 0.0f ? 0 : (ss < 0.0f ? -1 : 1);
    const T s_minus = mmin(T(0.0f), sm);
    const T s_pluss = mmax(T(0.0f), sp);
    const T chi_m = qdiv<FAST>(sm - vnm, sm - ss);
    const T chi_p = qdiv<FAST>(sp - vnp, sp - ss);
    const float wl = 0.5f*(1 + sign_star);
    const float wr = 0.5f*(1 - sign_star);
    const double wr_d = 0.5*(1 - sign_star);

    // _hllc_rho for rho, G, P: right weight is binary64 (GPUonly.cuh:900)
    auto hllc_rho = [&](const T qm, const T qp) -> T {
        const T q_deltam = qm*chi_m - qm;
        const T q_deltap = qp*chi_p - qp;
        const T fm = qm*vnm;
        const T fp = qp*vnp;
        if (FAST) return wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
        return T(W(wl * (fm + s_minus*q_deltam)) + wr_d * W(fp + s_pluss*q_deltap));
    };
    // _hllc_vel for a transverse momentum (GPUonly.cuh:926-938)
    auto hllc_vel = [&](const T vtm, const T vtp) -> T {
        const T qm = rm*vtm;
        const T qp = rp*vtp;
        const T q_deltam = qm*chi_m - qm;
        const T q_deltap = qp*chi_p - qp;
        const T fm = qm*vnm;
        const T fp = qp*vnp;
        return wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
    };
    // _hllc_pvel, normal momentum (GPUonly.cuh:967-979)
    T fpv;
    {
        const T qm = rm*vnm;
        const T qp = rp*vnp;
        const T q_deltam = rm*ss*chi_m - qm;
        const T q_deltap = rp*ss*chi_p - qp;
        const T fm = qm*vnm + pm;
        const T fp = qp*vnp + pp;
        fpv = wl * (fm + s_minus*q_deltam) + wr * (fp + s_pluss*q_deltap);
    }
    // _hllc_e (GPUonly.cuh:1011-1023)
    T fe;
    {
        const T qm = Gm*pm + Pm + 0.5f*rm*(vnm*vnm + v1m*v1m + v2m*v2m);
        const T qp = Gp*pp + Pp + 0.5f*rp*(vnp*vnp + v1p*v1p + v2p*v2p);
        const T q_deltam = chi_m*(qm + (ss - vnm)*(rm*ss + qdiv<FAST>(pm, sm - vnm))) - qm;
        const T q_deltap = chi_p*(qp + (ss - vnp)*(rp*ss + qdiv<FAST>(pp, sp - vnp))) - qp;
        const T fm = vnm*(qm + pm);
        const T fp = vnp*(qp + pp);
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
        const T chim1 = chi_m - 1.0f;
        const T chip1 = chi_p - 1.0f;
        o.vel = wl*(vnm + s_minus*chim1) + wr*(vnp + s_pluss*chip1);
    }
    return o;
}

} // namespace mpcf
