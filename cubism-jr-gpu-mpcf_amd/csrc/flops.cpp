// flops.cpp -- host instantiations of the scheme (mpcf_scheme.h).
//
//  * mpcf_flop_count: runs the scheme's per-cell work with a counting scalar
//    and reports the exact operation count per cell of each direction sweep,
//    per cell-stage and per cell-step; this is the algorithmic FLOP figure the
//    roofline uses.  A FLOP is one binary32 or binary64 add, subtract,
//    multiply, divide, reciprocal or square root; min/max and conversions
//    are reported apart and not counted as FLOPs.
//  * mpcf_host_weno / mpcf_host_face: the same per-cell code compiled for the
//    host in binary32, so the per-cell restructuring can be checked against
//    the face-based oracle without a GPU.
#include <cmath>
#include <cstring>

#include "mpcf_hip.h"
#include "mpcf_scheme.h"

namespace {

struct Tally
{
    double add, mul, div, sqrt, minmax, dadd, dmul, ddiv, cvt;
};
Tally* g_t = nullptr;

struct CD;

struct CF
{
    float v;
    CF() : v(0) {}
    CF(float x) : v(x) {}
    explicit CF(double x) : v((float)x) {}
    explicit CF(const CD& d);
};

struct CD
{
    double v;
    CD() : v(0) {}
    CD(double x) : v(x) {}
    explicit CD(const CF& f) : v(f.v) { if (g_t) g_t->cvt += 1; }
};

inline CF::CF(const CD& d) : v((float)d.v) { if (g_t) g_t->cvt += 1; }

inline CF operator+(CF a, CF b) { if (g_t) g_t->add += 1; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { if (g_t) g_t->add += 1; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { if (g_t) g_t->mul += 1; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { if (g_t) g_t->div += 1; return CF(a.v / b.v); }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline CD operator+(CD a, CD b) { if (g_t) g_t->dadd += 1; return CD(a.v + b.v); }
inline CD operator*(CD a, CD b) { if (g_t) g_t->dmul += 1; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { if (g_t) g_t->ddiv += 1; return CD(a.v / b.v); }
// found by argument-dependent lookup from the scheme templates
inline CF mmin(CF a, CF b) { if (g_t) g_t->minmax += 1; return CF(fminf(a.v, b.v)); }
inline CF mmax(CF a, CF b) { if (g_t) g_t->minmax += 1; return CF(fmaxf(a.v, b.v)); }
inline CF msqrt(CF a) { if (g_t) g_t->sqrt += 1; return CF(sqrtf(a.v)); }
inline CF msqrt_fast(CF a) { if (g_t) g_t->sqrt += 1; return CF(sqrtf(a.v)); }
inline CF mrcp(CF a) { if (g_t) g_t->div += 1; return CF(1.0f / a.v); }
inline CF mclamp(CF x, CF lo, CF hi) { if (g_t) g_t->minmax += 1; return CF(fminf(fmaxf(x.v, lo.v), hi.v)); }

} // namespace

namespace mpcf {
template <> struct Wide<CF> { typedef CD type; };
} // namespace mpcf

namespace {

using mpcf::PrimT;

PrimT<CF> sample_prim(float s)
{
    // any physical state: counts do not depend on values except through the
    // sign of s* (all branches of the HLLC weights are arithmetic here)
    return mpcf::to_prim<CF>(CF(1.0f + s), CF(0.1f), CF(0.2f - s), CF(0.05f), CF(3.0f), CF(2.5f), CF(0.0f));
}

template <int D, bool FAST>
void count_direction(int mode, Tally& t)
{
    // work of one cell in one sweep: primitives of the incoming cell,
    // reconstruction of one cell, one face, and the accumulator update
    Tally local = {};
    g_t = &local;
    PrimT<CF> w[5];
    for (int k = 0; k < 5; ++k) {
        g_t = nullptr;
        w[k] = sample_prim(0.01f * k);
    }
    g_t = &local;
    (void)mpcf::to_prim<CF, FAST>(CF(1.0f), CF(0.1f), CF(0.2f), CF(0.05f), CF(3.0f), CF(2.5f), CF(0.0f));
    const mpcf::CellRecT<CF> r = mpcf::reconstruct<CF, FAST>(w[0], w[1], w[2], w[3], w[4]);
    const mpcf::FaceFluxT<CF> F = mpcf::riemann<D, CF, FAST>(r.m, r.p);
    const mpcf::FaceFluxT<CF> Fl = F;
    // accumulator arithmetic of the sweep kernels (emit in mpcf_kernels.hip)
    CF acc[mpcf::NACC];
    for (int q = 0; q < mpcf::NACC; ++q) acc[q] = CF(0.5f);
    const CF cG = r.p.G + r.m.G, cP = r.p.P + r.m.P;
    if (mode == 0) {
        for (int q = 0; q < 7; ++q) acc[q] = F.f[q] - Fl.f[q];
        acc[9] = F.vel - Fl.vel;
    } else if (mode == 1) {
        for (int q = 0; q < 7; ++q) acc[q] = (acc[q] + F.f[q]) - Fl.f[q];
        acc[7] = acc[7] + cG; acc[8] = acc[8] + cP; acc[9] = acc[9] + (F.vel - Fl.vel);
    } else {
        const CF sG = acc[7] + cG, sP = acc[8] + cP, dU = acc[9] + (F.vel - Fl.vel);
        const CF f6 = CF(1.0f / 6.0f), dtinvh = CF(0.01f), a = CF(-0.5f), b = CF(0.75f);
        for (int q = 0; q < 7; ++q) {
            const CF tq = (acc[q] + F.f[q]) - Fl.f[q];
            CF rhs;
            if (q == 5) rhs = dtinvh * (tq - dU * sG * f6);
            else if (q == 6) rhs = dtinvh * (tq - dU * sP * f6);
            else rhs = dtinvh * tq;
            const CF Rn = a * CF(0.1f) - rhs;
            (void)(b * Rn + CF(1.0f));
        }
    }
    g_t = nullptr;
    t.add += local.add; t.mul += local.mul; t.div += local.div; t.sqrt += local.sqrt;
    t.minmax += local.minmax; t.dadd += local.dadd; t.dmul += local.dmul; t.ddiv += local.ddiv;
    t.cvt += local.cvt;
}

double flops_of(const Tally& t) { return t.add + t.mul + t.div + t.sqrt + t.dadd + t.dmul + t.ddiv; }

} // namespace

extern "C" int mpcf_flop_count(unsigned flags, double out[16])
{
    if (!out) return MPCF_EINVAL;
    const bool fast = (flags & MPCF_FLAG_FAST) != 0;
    Tally tx = {}, ty = {}, tz = {};
    if (fast) { count_direction<0, true>(0, tx); count_direction<1, true>(1, ty); count_direction<2, true>(2, tz); }
    else      { count_direction<0, false>(0, tx); count_direction<1, false>(1, ty); count_direction<2, false>(2, tz); }
    Tally st = {};
    for (const Tally* t : {&tx, &ty, &tz}) {
        st.add += t->add; st.mul += t->mul; st.div += t->div; st.sqrt += t->sqrt; st.minmax += t->minmax;
        st.dadd += t->dadd; st.dmul += t->dmul; st.ddiv += t->ddiv; st.cvt += t->cvt;
    }
    // max speed of sound per cell (GPUkernels.cu:1206-1209): 3 mul + 2 add
    // (|m|^2), 1 div (0.5/r), 1 mul, 2 sub, 1 div (p); 2 add, 2 div, 1 sqrt (c);
    // 1 div, 1 add; 3 max
    const double sos = 3 + 2 + 1 + 1 + 2 + 1 + 2 + 2 + 1 + 1 + 1;
    std::memset(out, 0, 16 * sizeof(double));
    out[0] = flops_of(tx);
    out[1] = flops_of(ty);
    out[2] = flops_of(tz);
    out[3] = flops_of(st);
    out[4] = sos;
    out[5] = 3 * out[3] + sos;
    out[6] = st.add; out[7] = st.mul; out[8] = st.div; out[9] = st.sqrt;
    out[10] = st.dadd; out[11] = st.dmul; out[12] = st.ddiv; out[13] = st.minmax; out[14] = st.cvt;
    return MPCF_OK;
}

extern "C" int mpcf_host_weno(int n, const float* in5, float* minus, float* plus, unsigned flags)
{
    if (n < 0 || (n > 0 && (!in5 || !minus || !plus))) return MPCF_EINVAL;
    const bool fast = (flags & MPCF_FLAG_FAST) != 0;
    for (int i = 0; i < n; ++i) {
        const float* a = in5 + 5 * (size_t)i;
        const mpcf::RecT<float> r = fast ? mpcf::weno_cell<float, true>(a[0], a[1], a[2], a[3], a[4])
                                         : mpcf::weno_cell<float, false>(a[0], a[1], a[2], a[3], a[4]);
        minus[i] = r.m;
        plus[i] = r.p;
    }
    return MPCF_OK;
}

namespace {
template <int D, bool FAST>
void host_face_one(const float* s, float* o)
{
    PrimT<float> c[6];
    for (int k = 0; k < 6; ++k)
        c[k] = mpcf::to_prim<float, FAST>(s[0 * 6 + k], s[1 * 6 + k], s[2 * 6 + k], s[3 * 6 + k], s[4 * 6 + k],
                                    s[5 * 6 + k], s[6 * 6 + k]);
    const mpcf::CellRecT<float> left = mpcf::reconstruct<float, FAST>(c[0], c[1], c[2], c[3], c[4]);
    const mpcf::CellRecT<float> right = mpcf::reconstruct<float, FAST>(c[1], c[2], c[3], c[4], c[5]);
    const mpcf::FaceFluxT<float> F = mpcf::riemann<D, float, FAST>(left.m, right.p);
    for (int q = 0; q < 7; ++q) o[q] = F.f[q];
    o[7] = F.vel; o[8] = left.m.G; o[9] = right.p.G; o[10] = left.m.P; o[11] = right.p.P;
}
} // namespace

extern "C" int mpcf_host_face(int n, int dir, const float* st, float* out, unsigned flags)
{
    if (n < 0 || dir < 0 || dir > 2 || (n > 0 && (!st || !out))) return MPCF_EINVAL;
    const bool fast = (flags & MPCF_FLAG_FAST) != 0;
    for (int i = 0; i < n; ++i) {
        const float* s = st + 42 * (size_t)i;
        float* o = out + 12 * (size_t)i;
        if (fast) {
            if (dir == 0) host_face_one<0, true>(s, o);
            else if (dir == 1) host_face_one<1, true>(s, o);
            else host_face_one<2, true>(s, o);
        } else {
            if (dir == 0) host_face_one<0, false>(s, o);
            else if (dir == 1) host_face_one<1, false>(s, o);
            else host_face_one<2, false>(s, o);
        }
    }
    return MPCF_OK;
}
