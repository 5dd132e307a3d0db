// sim_ic.cpp -- the initial conditions of the reference's Sod, 2D
// shock-bubble and static-index drivers (host C++): Sim_SodMPI::_ic
// (source/Sim/Sim_SodMPI.cpp:19-95), Sim_2DSBIMPI::_ic
// (source/Sim/Sim_2DSBIMPI.cpp:86-233) with SimTools::getPostShockRatio and
// heaviside / heaviside_smooth (source/Types.h:88-109), and
// Sim_StaticIC::_ic (source/Sim/Sim_StaticIC.cpp:76-97).
//
// Each follows the reference's arithmetic in its default build, whose Real is
// binary32 (Makefile.config:14, ap ?= float): Sod computes in binary64 and
// rounds on store; the shock-bubble case rounds where the reference holds a
// Real (post-shock state, bubble distance and indicator, pressure, kinetic
// energy) and forms E = p G + P + ke in binary32 from the stored fields.
// Reference behaviours kept as they are:
//  * Sod: the "u" it stores is the momentum rho u, and E = G p + P + u^2/(2 rho)
//    with that momentum (Sim_SodMPI.cpp:75-85);
//  * 2DSBI: the bubble indicator of a cell is that of the first bubble whose
//    indicator is positive (:188-193), distances are in the (dims[0], dims[1])
//    plane (cylinders along dims[2]), and -reverse point-mirrors the block
//    through its centre with the momenta negated, leaving the middle slab of
//    an odd extent along dims[0] unmirrored (:210-231).
// Cell centres are at h (i + 1/2) (NodeBlock.h:90-96).  Built with
// -ffp-contract=off so that no product is fused into a sum.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mpcf_hip.h"

namespace mpcf {
namespace detail {
int fail(int code, const std::string& msg);   // mpcf_capi.cpp: sets mpcf_last_error
}
}

namespace {

int fail(const std::string& m) { return mpcf::detail::fail(MPCF_EINVAL, m); }

bool fields_ok(float* const f[7])
{
    if (!f) return false;
    for (int q = 0; q < 7; ++q)
        if (!f[q]) return false;
    return true;
}

// SimTools::heaviside_smooth (Types.h:95-99) with EPSILON = eps: phi, the
// angle and the result are Real (binary32 here), so the cosine is the
// binary32 one
inline float heaviside_smooth(const float phi, const double eps)
{
    const float alpha = (float)(M_PI * std::min(1., std::max(0., 0.5 * (phi + eps) / eps)));
    return (float)(0.5 + 0.5 * std::cos(alpha));
}

// SimTools::getPostShockRatio (Types.h:101-109) for Real = T: the Mach-number
// square and the speeds of sound are formed in T, the ratios in binary64
template <typename T>
void post_shock(const T pre[3], const T mach, const T gamma, const T pc, T post[3])
{
    const T m2 = std::pow(mach, (T)2);
    const double gm1 = gamma - 1., gp1 = gamma + 1.;
    const double Mpost = std::sqrt((m2 * gm1 + 2.) / (2. * gamma * m2 - gm1));
    post[0] = (T)(gp1 * m2 / (gm1 * m2 + 2.) * pre[0]);
    post[2] = (T)(1. / gp1 * (2. * gamma * m2 - gm1) * pre[2]);
    const double c_pre = mach * std::sqrt(gamma * (pc + pre[2]) / pre[0]);
    const double c_post = Mpost * std::sqrt(gamma * (pc + post[2]) / post[0]);
    post[1] = (T)(c_pre - c_post);
}

}  // namespace

extern "C" void mpcf_post_shock_ratio(const double pre[3], double mach, double gamma, double pc, int real32,
                                      double post[3])
{
    if (!pre || !post) return;
    if (real32) {
        const float pf[3] = {(float)pre[0], (float)pre[1], (float)pre[2]};
        float qf[3];
        post_shock<float>(pf, (float)mach, (float)gamma, (float)pc, qf);
        for (int i = 0; i < 3; ++i) post[i] = qf[i];
    } else {
        post_shock<double>(pre, mach, gamma, pc, post);
    }
}

extern "C" void mpcf_sod_defaults(mpcf_sod_params* p)
{
    if (!p) return;
    // Sim_SodMPI.cpp:27-40
    p->x0 = 0.5; p->rho1 = 1.0; p->rho2 = 0.125; p->u1 = 0.0; p->u2 = 0.0; p->p1 = 1.0; p->p2 = 0.1;
    p->g1 = 1.4; p->g2 = 1.4; p->pc1 = 0.0; p->pc2 = 0.0;
    p->reverse = 0; p->dir = 0;
}

extern "C" int mpcf_ic_sod(int nx, int ny, int nz, double h, const mpcf_sod_params* p, float* const fields[7])
{
    if (!p || nx < 1 || ny < 1 || nz < 1 || !(h > 0) || !fields_ok(fields) || p->dir < 0 || p->dir > 2)
        return fail("mpcf_ic_sod: bad argument");
    // -config x|y|z (Sim_SodMPI.cpp:42-56): dims[0] is the tube's axis
    const int d0 = p->dir;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double pos[3] = {h * (ix + 0.5), h * (iy + 0.5), h * (iz + 0.5)};
                const bool x = p->reverse ? pos[d0] > p->x0 : pos[d0] < p->x0;
                const double r = x ? p->rho1 : p->rho2;
                const double pr = x ? p->p1 : p->p2;
                const double mom = x ? p->rho1 * p->u1 : p->rho2 * p->u2;
                const double G = x ? 1. / (p->g1 - 1.) : 1. / (p->g2 - 1.);
                const double P = x ? p->pc1 * p->g1 / (p->g1 - 1.) : p->pc2 * p->g2 / (p->g2 - 1.);
                const size_t i = ((size_t)iz * ny + iy) * nx + ix;
                fields[0][i] = (float)r;
                for (int d = 0; d < 3; ++d) fields[1 + d][i] = d == d0 ? (float)mom : 0.0f;
                fields[4][i] = (float)(G * pr + P + 0.5 * mom * mom / r);
                fields[5][i] = (float)G;
                fields[6][i] = (float)P;
            }
    return MPCF_OK;
}

extern "C" void mpcf_sbi_defaults(mpcf_sbi_params* p)
{
    if (!p) return;
    // Sim_2DSBIMPI.cpp:95-108
    p->x0 = 0.1; p->mach = 1.22; p->rho = 1.0; p->rhobubble = 0.138; p->u = 0.0; p->p = 1.0; p->pbubble = 1.0;
    p->g1 = 1.4; p->g2 = 1.67; p->pc1 = 0.0; p->pc2 = 0.0;
    p->mollfactor = 1; p->reverse = 0;
    p->dims[0] = 0; p->dims[1] = 1; p->dims[2] = 2;
}

extern "C" int mpcf_sbi_config(const char* conf, int dims[3])
{
    // -config (Sim_2DSBIMPI.cpp:20-63): dims = {principal, minor, dummy}
    static const struct { const char* name; int d[3]; } table[] = {
        {"xy", {0, 1, 2}}, {"xz", {0, 2, 1}}, {"yx", {1, 0, 2}}, {"yz", {1, 2, 0}}, {"zx", {2, 0, 1}}, {"zy", {2, 1, 0}}};
    if (!conf || !dims) return fail("mpcf_sbi_config: NULL argument");
    for (const auto& t : table)
        if (std::strcmp(conf, t.name) == 0) {
            std::memcpy(dims, t.d, sizeof(t.d));
            return MPCF_OK;
        }
    return fail(std::string("mpcf_sbi_config: -config ") + conf + " is not possible");
}

extern "C" int mpcf_read_bubbles(const char* path, double* bubbles, int cap, int* n)
{
    // xybubbles.dat & co. (Sim_2DSBIMPI.cpp:126-155): the count, then x y z r per bubble
    if (!path || !n || (cap > 0 && !bubbles)) return fail("mpcf_read_bubbles: bad argument");
    std::ifstream f(path);
    if (!f.good()) return fail(std::string("mpcf_read_bubbles: cannot read ") + path);
    long nb;
    if (!(f >> nb) || nb < 0) return fail(std::string("mpcf_read_bubbles: no bubble count in ") + path);
    for (long i = 0; i < nb; ++i) {
        double b[4];
        if (!(f >> b[0] >> b[1] >> b[2] >> b[3])) return fail(std::string("mpcf_read_bubbles: short file ") + path);
        if (i < cap)
            for (int j = 0; j < 4; ++j) bubbles[4 * i + j] = b[j];
    }
    *n = (int)nb;
    return MPCF_OK;
}

extern "C" int mpcf_ic_sbi(int nx, int ny, int nz, double h, const mpcf_sbi_params* p, const double* bubbles, int nb,
                           float* const fields[7])
{
    if (!p || nx < 1 || ny < 1 || nz < 1 || !(h > 0) || nb < 0 || (nb > 0 && !bubbles) || !fields_ok(fields))
        return fail("mpcf_ic_sbi: bad argument");
    const int* dims = p->dims;
    {
        int seen = 0;
        for (int i = 0; i < 3; ++i) {
            if (dims[i] < 0 || dims[i] > 2) return fail("mpcf_ic_sbi: bad dims");
            seen |= 1 << dims[i];
        }
        if (seen != 7) return fail("mpcf_ic_sbi: dims must be a permutation of 0, 1, 2");
    }
    const double eps = (double)p->mollfactor * std::sqrt(3.) * h;   // SimTools::EPSILON (:106)
    const float pre[3] = {(float)p->rho, (float)p->u, (float)p->p};
    float post[3];
    post_shock<float>(pre, (float)p->mach, (float)p->g1, (float)p->pc1, post);
    const double rho1 = post[0], u1 = post[1], p1 = post[2];
    const double G1 = p->g1 - 1, G2 = p->g2 - 1, F1 = p->g1 * p->pc1, F2 = p->g2 * p->pc2;
    std::vector<float> bub((size_t)4 * nb);                        // bubbles held as Real
    for (size_t i = 0; i < bub.size(); ++i) bub[i] = (float)bubbles[i];
    const int ext[3] = {nx, ny, nz};
    auto at = [&](const int c[3]) { return ((size_t)c[2] * ny + c[1]) * nx + c[0]; };
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double pos[3] = {h * (ix + 0.5), h * (iy + 0.5), h * (iz + 0.5)};
                const float shock = (float)(pos[dims[0]] - p->x0) > 0 ? 0.0f : 1.0f;   // SimTools::heaviside
                float bubble = 0;
                for (int i = 0; i < nb; ++i) {
                    const double a = pos[dims[0]] - bub[4 * i + dims[0]];
                    const double b = pos[dims[1]] - bub[4 * i + dims[1]];
                    const float r = (float)std::sqrt(a * a + b * b);
                    bubble = heaviside_smooth(r - bub[4 * i + 3], eps);
                    if (bubble > 0) break;
                }
                const float nbub = 1 - bubble, nshock = 1 - shock;
                const int c[3] = {ix, iy, iz};
                const size_t e = at(c);
                const float R = (float)(shock * rho1 + nshock * (p->rhobubble * bubble + p->rho * nbub));
                const float M = (float)((shock * u1 + nshock * p->u) * R);
                const float G = (float)(1. / G1 * nbub + 1. / G2 * bubble);
                const float P = (float)(F1 / G1 * nbub + F2 / G2 * bubble);
                const float pressure = (float)(shock * p1 + nshock * (p->pbubble * bubble + p->p * nbub));
                const float ke = (float)(0.5 * ((double)M * M) / R);
                fields[0][e] = R;
                for (int d = 0; d < 3; ++d) fields[1 + d][e] = d == dims[0] ? M : 0.0f;
                fields[4][e] = pressure * G + P + ke;
                fields[5][e] = G;
                fields[6][e] = P;
            }
    if (p->reverse) {
        // point mirror through the centre, momenta negated, over the first
        // half along dims[0] (the middle slab of an odd extent stays)
        int idx[3], rev[3];
        for (idx[dims[0]] = 0; idx[dims[0]] < ext[dims[0]] / 2; ++idx[dims[0]])
            for (idx[dims[2]] = 0; idx[dims[2]] < ext[dims[2]]; ++idx[dims[2]])
                for (idx[dims[1]] = 0; idx[dims[1]] < ext[dims[1]]; ++idx[dims[1]]) {
                    for (int d = 0; d < 3; ++d) rev[d] = ext[d] - 1 - idx[d];
                    const size_t a = at(idx), b = at(rev);
                    for (int q = 0; q < 7; ++q) {
                        const float s = (q >= 1 && q <= 3) ? -1.0f : 1.0f;
                        const float va = fields[q][a], vb = fields[q][b];
                        fields[q][a] = s * vb;
                        fields[q][b] = s * va;
                    }
                }
    }
    return MPCF_OK;
}

extern "C" int mpcf_ic_static(int nx, int ny, int nz, const int dims[3], float* const fields[7])
{
    if (nx < 1 || ny < 1 || nz < 1 || !dims || !fields_ok(fields)) return fail("mpcf_ic_static: bad argument");
    int seen = 0;
    for (int i = 0; i < 3; ++i) {
        if (dims[i] < 0 || dims[i] > 2) return fail("mpcf_ic_static: bad dims");
        seen |= 1 << dims[i];
    }
    if (seen != 7) return fail("mpcf_ic_static: dims must be a permutation of 0, 1, 2");
    const int ext[3] = {nx, ny, nz};
    int idx[3];
    unsigned cnt = 0;   // uint_t in the reference, converted to Real on store
    for (idx[dims[2]] = 0; idx[dims[2]] < ext[dims[2]]; ++idx[dims[2]])
        for (idx[dims[1]] = 0; idx[dims[1]] < ext[dims[1]]; ++idx[dims[1]])
            for (idx[dims[0]] = 0; idx[dims[0]] < ext[dims[0]]; ++idx[dims[0]]) {
                const size_t e = ((size_t)idx[2] * ny + idx[1]) * nx + idx[0];
                for (int q = 0; q < 7; ++q) fields[q][e] = (float)cnt;
                ++cnt;
            }
    return MPCF_OK;
}

extern "C" int mpcf_ic_steady(int nx, int ny, int nz, const double state[7], float* const fields[7])
{
    // Sim_SteadyStateMPI::_ic (Sim_SteadyStateMPI.cpp:95-134): one uniform
    // state, state = rho, u, v, w, p, gamma, pc
    if (nx < 1 || ny < 1 || nz < 1 || !state || !fields_ok(fields)) return fail("mpcf_ic_steady: bad argument");
    const double r = state[0], u = state[1], v = state[2], w = state[3], p = state[4], g = state[5], pc = state[6];
    const double G = 1.0 / (g - 1.0);
    const double P = g * pc * G;
    const double e = G * p + P + 0.5 * r * (u * u + v * v + w * w);
    const float q[7] = {(float)r, (float)(r * u), (float)(r * v), (float)(r * w), (float)e, (float)G, (float)P};
    const size_t n = (size_t)nx * ny * nz;
    for (int k = 0; k < 7; ++k) std::fill(fields[k], fields[k] + n, q[k]);
    return MPCF_OK;
}
