// sic.cpp -- the shock-induced cloud collapse initial condition and its input
// files (host C++): the counterpart of Sim_SICCloudMPI::_ic
// (source/Sim/Sim_SICCloudMPI.cpp:43-128), _initialize_cloud (:134-160),
// shape::make_many (Sim_SICCloudMPI.h:148-181) and _ic_quad (:306-357).
//
//  * Post-shock liquid state from the pre-shock state and the pressure ratio
//    psi across a planar shock in a stiffened gas (Johnsen's relations,
//    Sim_SICCloudMPI.cpp:87-104):
//        c0   = sqrt(g (p0 + pc) / rho0)
//        M    = sqrt((g + 1) / (2 g) (psi - 1) p0 / (p0 + pc) + 1)
//        u1   = u0 + c0 (psi - 1) p0 / (g (p0 + pc) M)
//        rho1 = rho0 (t1 t2 + 1) / (t1 + t2),  t1 = (g+1)/(g-1), t2 = (p1+pc)/(p0+pc)
//    with the shock moving at u0 + c0 M along its normal.
//  * Shock indicator: 1 behind the plane through S with normal n
//    (n . (x - S) <= 0), 0 ahead (SimTools::heaviside, Types.h:90-93).
//  * Bubble indicator from the signed distance d to the nearest sphere:
//    0.5 + 0.5 cos(pi clamp((d + eps) / (2 eps), 0, 1)) (Sim_SICCloudMPI.h:
//    263-272).  The reference's cloud case never sets SimTools::EPSILON, so eps
//    is 0 there and the indicator is sharp: 1 where d <= 0, 0 where d > 0 (the
//    limits the formula takes, d = 0 included); eps = 0 reproduces that.
//  * State: set_IC (:247-276) -- rho and the x-velocity mixed by the shock and
//    bubble indicators (the y/z momenta are zero whatever the shock normal),
//    G and P by the bubble indicator with the liquid's (g1, pc1) and the gas's
//    (g2, pc2), E = p G + P + rho u^2 / 2.
//  * Each cell is the 27-point Simpson average (weights 1/6, 2/3, 1/6 per axis)
//    of the conserved state sampled at the cell centre and +-h/2 (:306-331),
//    cell centres at h (i + 1/2).
// Arithmetic is binary64 throughout; the fields are rounded to binary32 once.
// Shapes are culled per z-slab and row before the per-sample loop (the
// reference tests every shape of a block at every sample).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "mpcf_hip.h"

namespace {

struct Shape { double c[3], r; };

struct Conserved { double r, ru, rv, rw, E, G, P; };

inline Conserved axpy(const double a, const Conserved& x, const Conserved& y)
{
    return {a * x.r + y.r, a * x.ru + y.ru, a * x.rv + y.rv, a * x.rw + y.rw, a * x.E + y.E, a * x.G + y.G,
            a * x.P + y.P};
}

struct Post { double rho1, u1, p1, mach, c0; };

Post post_shock(const mpcf_sic_params& p)
{
    const double g = p.g1, pc = p.pc1, p0 = p.p0, rho0 = p.rho0, u0 = p.u0, psi = p.pressure_ratio;
    const double p1 = p0 * psi;
    const double t1 = (g + 1.0) / (g - 1.0);
    const double t2 = (p1 + pc) / (p0 + pc);
    Post s;
    s.c0 = std::sqrt(g * (p0 + pc) / rho0);
    s.mach = std::sqrt((g + 1.0) / (2.0 * g) * (psi - 1.0) * p0 / (p0 + pc) + 1.0);
    s.u1 = u0 + s.c0 * (psi - 1.0) * p0 / (g * (p0 + pc) * s.mach);
    s.rho1 = rho0 * (t1 * t2 + 1.0) / (t1 + t2);
    s.p1 = p1;
    return s;
}

inline double bubble_indicator(const double d, const double eps)
{
    if (eps <= 0.0) return d > 0.0 ? 0.0 : 1.0;
    const double a = M_PI * std::min(1.0, std::max(0.0, (d + eps) / (2.0 * eps)));
    return 0.5 + 0.5 * std::cos(a);
}

struct Setup
{
    mpcf_sic_params p;
    Post s;
    double n[3];   // unit shock normal
};

// set_IC at one sample point
inline Conserved sample(const Setup& S, const double x[3], const std::vector<const Shape*>& cand)
{
    const mpcf_sic_params& p = S.p;
    const double dn = S.n[0] * (x[0] - p.shock_s[0]) + S.n[1] * (x[1] - p.shock_s[1]) + S.n[2] * (x[2] - p.shock_s[2]);
    const double shock = dn > 0.0 ? 0.0 : 1.0;
    double bubble = 0.0;
    if (!cand.empty()) {
        double d = HUGE_VAL;
        for (const Shape* sh : cand) {
            const double dx = x[0] - sh->c[0], dy = x[1] - sh->c[1], dz = x[2] - sh->c[2];
            d = std::min(d, std::sqrt(dx * dx + dy * dy + dz * dz) - sh->r);
        }
        bubble = bubble_indicator(d, p.eps);
    }
    const double G1 = p.g1 - 1.0, G2 = p.g2 - 1.0;
    const double F1 = p.g1 * p.pc1, F2 = p.g2 * p.pc2;
    const double rho = shock * S.s.rho1 + (1 - shock) * (p.rhoB * bubble + p.rho0 * (1 - bubble));
    const double u = shock * S.s.u1 + (1 - shock) * (p.uB * bubble + p.u0 * (1 - bubble));
    const double pres = shock * S.s.p1 + (1 - shock) * (p.pB * bubble + p.p0 * (1 - bubble));
    Conserved c;
    c.r = rho;
    // set_IC puts the whole velocity into x whatever the shock normal
    // (Sim_SICCloudMPI.cpp:261-263)
    c.ru = rho * u;
    c.rv = 0.0;
    c.rw = 0.0;
    c.G = 1.0 / G1 * (1 - bubble) + 1.0 / G2 * bubble;
    c.P = F1 / G1 * (1 - bubble) + F2 / G2 * bubble;
    c.E = pres * c.G + c.P + 0.5 * rho * u * u;
    return c;
}

}  // namespace

namespace mpcf {
namespace detail {
int fail(int code, const std::string& msg);   // mpcf_capi.cpp: sets mpcf_last_error
}
}

namespace {
int fail(const std::string& msg) { return mpcf::detail::fail(MPCF_EINVAL, msg); }
}  // namespace

extern "C" void mpcf_sic_defaults(mpcf_sic_params* p)
{
    if (!p) return;
    // Sim_SICCloudMPI.cpp:53-77
    p->rho0 = 1000; p->u0 = 0; p->p0 = 1;
    p->rhoB = 1; p->uB = 0; p->pB = 0.0234;
    p->pressure_ratio = 40;
    p->shock_n[0] = 1; p->shock_n[1] = 0; p->shock_n[2] = 0;
    p->shock_s[0] = 0.05; p->shock_s[1] = 0; p->shock_s[2] = 0;
    p->g1 = 6.59; p->g2 = 1.4; p->pc1 = 4096.0; p->pc2 = 1.0;
    p->eps = 0.0;
}

extern "C" int mpcf_sic_post_shock(const mpcf_sic_params* p, double out[5])
{
    if (!p || !out) return fail("mpcf_sic_post_shock: NULL argument");
    if (!(p->g1 > 1.0) || !(p->rho0 > 0) || !(p->p0 + p->pc1 > 0) || !(p->pressure_ratio >= 1.0))
        return fail("mpcf_sic_post_shock: bad parameters");
    const Post s = post_shock(*p);
    out[0] = s.rho1; out[1] = s.u1; out[2] = s.p1; out[3] = s.mach; out[4] = s.c0;
    return MPCF_OK;
}

extern "C" int mpcf_read_cloud_config(const char* path, double out[11])
{
    // cloud_config.dat (Sim_SICCloudMPI.cpp:136-152): n_shapes n_small,
    // min_rad max_rad, seed_s[3], seed_e[3], optionally n_sensors
    if (!path || !out) return fail("mpcf_read_cloud_config: NULL argument");
    std::ifstream f(path);
    if (!f.good()) return fail(std::string("mpcf_read_cloud_config: cannot read ") + path);
    double v[11] = {0};
    for (int i = 0; i < 10; ++i)
        if (!(f >> v[i])) return fail(std::string("mpcf_read_cloud_config: short file ") + path);
    if (!(f >> v[10])) v[10] = 0;
    std::memcpy(out, v, sizeof(v));
    return MPCF_OK;
}

extern "C" int mpcf_read_cloud(const char* path, double* shapes, int cap, int* n)
{
    // cloud.dat (Sim_SICCloudMPI.h:148-181): one "idx x y z r" per line
    if (!path || !n || (cap > 0 && !shapes)) return fail("mpcf_read_cloud: bad argument");
    std::ifstream f(path);
    if (!f.good()) return fail(std::string("mpcf_read_cloud: cannot read ") + path);
    std::string line;
    int k = 0;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        long idx;
        double c[4];
        if (!(in >> idx >> c[0] >> c[1] >> c[2] >> c[3])) continue;   // blank or partial line
        if (k < cap)
            for (int j = 0; j < 4; ++j) shapes[4 * k + j] = c[j];
        ++k;
    }
    *n = k;
    return MPCF_OK;
}

extern "C" int mpcf_ic_sic_cloud(int nx, int ny, int nz, double h, const mpcf_sic_params* prm, const double* shapes,
                                 int nshapes, float* const fields[7])
{
    if (!prm || nx < 1 || ny < 1 || nz < 1 || !(h > 0) || nshapes < 0 || (nshapes > 0 && !shapes) || !fields)
        return fail("mpcf_ic_sic_cloud: bad argument");
    for (int q = 0; q < 7; ++q)
        if (!fields[q]) return fail("mpcf_ic_sic_cloud: NULL field");
    Setup S;
    S.p = *prm;
    const double mag = std::sqrt(prm->shock_n[0] * prm->shock_n[0] + prm->shock_n[1] * prm->shock_n[1] +
                                 prm->shock_n[2] * prm->shock_n[2]);
    if (!(mag > 0) || !(prm->g1 > 1) || !(prm->g2 > 1) || !(prm->pressure_ratio >= 1.0))
        return fail("mpcf_ic_sic_cloud: bad parameters");
    for (int i = 0; i < 3; ++i) S.n[i] = prm->shock_n[i] / mag;
    S.s = post_shock(*prm);
    std::vector<Shape> sh(nshapes);
    for (int i = 0; i < nshapes; ++i) {
        for (int j = 0; j < 3; ++j) sh[i].c[j] = shapes[4 * i + j];
        sh[i].r = shapes[4 * i + 3];
    }
    // a shape can affect a cell's samples if its sphere, widened by the
    // mollifier and the half-cell sample offset, reaches the cell
    const double reach = std::max(prm->eps, 0.0) + h;
    const size_t plane = (size_t)nx * ny;
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min<unsigned>(nt ? nt : 4, (unsigned)std::min(nz, 64)));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            std::vector<const Shape*> zc, yc, xc;
            for (int iz = (int)t; iz < nz; iz += (int)nt) {
                const double zc0 = h * (iz + 0.5);
                zc.clear();
                for (const Shape& s : sh)
                    if (std::fabs(zc0 - s.c[2]) <= s.r + reach) zc.push_back(&s);
                for (int iy = 0; iy < ny; ++iy) {
                    const double yc0 = h * (iy + 0.5);
                    yc.clear();
                    for (const Shape* s : zc)
                        if (std::fabs(yc0 - s->c[1]) <= s->r + reach) yc.push_back(s);
                    for (int ix = 0; ix < nx; ++ix) {
                        const double xc0 = h * (ix + 0.5);
                        xc.clear();
                        for (const Shape* s : yc)
                            if (std::fabs(xc0 - s->c[0]) <= s->r + reach) xc.push_back(s);
                        // 27-point Simpson average over the cell (Sim_SICCloudMPI.cpp:306-331)
                        static const double w[3] = {1.0 / 6, 2.0 / 3, 1.0 / 6};
                        Conserved zint[3][3];
                        for (int jy = 0; jy < 3; ++jy)
                            for (int jx = 0; jx < 3; ++jx) {
                                Conserved acc = {0, 0, 0, 0, 0, 0, 0};
                                for (int jz = 0; jz < 3; ++jz) {
                                    const double x[3] = {xc0 + (jx - 1) * 0.5 * h, yc0 + (jy - 1) * 0.5 * h,
                                                         zc0 + (jz - 1) * 0.5 * h};
                                    acc = axpy(w[jz], sample(S, x, xc), acc);
                                }
                                zint[jy][jx] = acc;
                            }
                        Conserved cell = {0, 0, 0, 0, 0, 0, 0};
                        for (int jx = 0; jx < 3; ++jx) {
                            Conserved yint = {0, 0, 0, 0, 0, 0, 0};
                            for (int jy = 0; jy < 3; ++jy) yint = axpy(w[jy], zint[jy][jx], yint);
                            cell = axpy(w[jx], yint, cell);
                        }
                        const size_t i = (size_t)iz * plane + (size_t)iy * nx + ix;
                        fields[0][i] = (float)cell.r;
                        fields[1][i] = (float)cell.ru;
                        fields[2][i] = (float)cell.rv;
                        fields[3][i] = (float)cell.rw;
                        fields[4][i] = (float)cell.E;
                        fields[5][i] = (float)cell.G;
                        fields[6][i] = (float)cell.P;
                    }
                }
            }
        });
    }
    for (auto& x : th) x.join();
    return MPCF_OK;
}
