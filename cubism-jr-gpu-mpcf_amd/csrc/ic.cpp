// ic.cpp -- initial conditions of the benchmark cases (host C++).
//
// These are the counterparts of the reference's Sim_*::_ic routines
// (source/Sim/Sim_SodMPI.cpp:19-95, Sim_2DSBIMPI.cpp:86-233,
// Sim_SICCloudMPI.cpp:43-357): a two-material stiffened-gas field built from a
// shock indicator and a bubble indicator exactly as set_IC blends them
// (Sim_SICCloudMPI.cpp:249-276): rho, u and p are mixed by the indicators,
// G = 1/(gamma-1) and P = gamma pc/(gamma-1) by the bubble indicator, and
// E = p G + P + rho |u|^2 / 2.  The indicators are tanh-smoothed over a stated
// number of cells (the reference uses a cosine mollifier, Types.h:95-99) and
// the noise is counter-based (splitmix64 of seed and cell index), so a field
// is reproducible independent of thread count.  Cell centres are at
// h (i + 1/2) with h = 1/max(NX,NY,NZ) (NodeBlock.h:71,90-96).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mpcf_hip.h"

namespace {

struct Mat { double gamma, pc; };
const Mat AIR = {1.4, 0.0};
const Mat WATER = {6.59, 4096.0};

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform in [-1, 1) from (seed, stream, index)
inline double unoise(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t r = splitmix64(seed ^ splitmix64(stream * 0x632BE59BD9B4E019ull + i));
    return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

struct Rng
{
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    double uniform() { s = splitmix64(s); return (double)(s >> 11) / 9007199254740992.0; }
    double normal()
    {
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct State { double rho, u, v, w, p, G, P; };

inline double Gof(const Mat& m) { return 1.0 / (m.gamma - 1.0); }
inline double Pof(const Mat& m) { return m.gamma * m.pc / (m.gamma - 1.0); }

// set_IC of Sim_SICCloudMPI.cpp:249-276 with a liquid, a gas and a post-shock state
inline State mix(double shock, double bubble, const double post[4], const double pre[4], const double gas[4])
{
    State s;
    s.rho = shock * post[0] + (1 - shock) * (gas[0] * bubble + pre[0] * (1 - bubble));
    s.u = shock * post[1] + (1 - shock) * (gas[1] * bubble + pre[1] * (1 - bubble));
    s.v = 0; s.w = 0;
    s.p = shock * post[3] + (1 - shock) * (gas[3] * bubble + pre[3] * (1 - bubble));
    s.G = Gof(WATER) * (1 - bubble) + Gof(AIR) * bubble;
    s.P = Pof(WATER) * (1 - bubble) + Pof(AIR) * bubble;
    return s;
}

inline void store(float* const f[7], size_t i, const State& s)
{
    f[0][i] = (float)s.rho;
    f[1][i] = (float)(s.rho * s.u);
    f[2][i] = (float)(s.rho * s.v);
    f[3][i] = (float)(s.rho * s.w);
    f[4][i] = (float)(s.G * s.p + s.P + 0.5 * s.rho * (s.u * s.u + s.v * s.v + s.w * s.w));
    f[5][i] = (float)s.G;
    f[6][i] = (float)s.P;
}

template <typename F>
void parallel_z(int nz, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    nt = std::min<unsigned>(nt, 32u);
    nt = std::min<unsigned>(nt, (unsigned)nz);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=]() { for (int z = (int)t; z < nz; z += (int)nt) fn(z); });
    for (auto& x : th) x.join();
}

// smoothed indicator: 1 where d < 0, 0 where d > 0, tanh over `w` cells of size h
inline double inside(double d, double w, double h) { return 0.5 * (1.0 - std::tanh(d / (w * h))); }

bool fields_ok(float* const f[7])
{
    if (!f) return false;
    for (int q = 0; q < 7; ++q)
        if (!f[q]) return false;
    return true;
}

} // namespace

extern "C" int mpcf_ic_shock_tube(int nx, int ny, int nz, int axis, double noise, unsigned long long seed,
                                  float* const fields[7])
{
    if (nx < 1 || ny < 1 || nz < 1 || axis < 0 || axis > 2 || !fields_ok(fields)) return MPCF_EINVAL;
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    const double pre[4] = {1000.0, 0.0, 0.0, 0.1};   // water (right)
    const double gas[4] = {1.0, 0.0, 0.0, 1.0};      // air (left)
    parallel_z(nz, [&](int z) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double c[3] = {h * (x + 0.5), h * (y + 0.5), h * (z + 0.5)};
                const double air = inside(c[axis] - 0.5, 2.0, h);
                State s = mix(0.0, air, pre, pre, gas);
                const size_t i = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                s.rho *= 1.0 + noise * unoise(seed, 1, i);
                store(fields, i, s);
            }
    });
    return MPCF_OK;
}

extern "C" int mpcf_ic_shock_bubble(int nx, int ny, int nz, unsigned long long seed, float* const fields[7])
{
    if (nx < 1 || ny < 1 || nz < 1 || !fields_ok(fields)) return MPCF_EINVAL;
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    const double post[4] = {1000.0, 10.0, 0.0, 1000.0};
    const double pre[4] = {1000.0, 0.0, 0.0, 1.0};
    const double gas[4] = {1.0, 0.0, 0.0, 1.0};
    parallel_z(nz, [&](int z) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double cx = h * (x + 0.5), cy = h * (y + 0.5), cz = h * (z + 0.5);
                const double r = std::sqrt((cx - 0.5) * (cx - 0.5) + (cy - 0.5) * (cy - 0.5) + (cz - 0.5) * (cz - 0.5));
                const double bubble = inside(r - 0.1, 1.5, h);
                const double shock = inside(cx - 0.3, 1.5, h);
                State s = mix(shock, bubble, post, pre, gas);
                const size_t i = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                s.p *= 1.0 + 1e-4 * unoise(seed, 2, i);
                store(fields, i, s);
            }
    });
    return MPCF_OK;
}

// Bubble centres uniform in a ball of radius 0.3 about the domain centre,
// radii lognormal(mu = ln 0.02, sigma = 0.3), overlaps rejected.  bubbles
// (optional) receives n x 4 doubles (x, y, z, r).
extern "C" int mpcf_ic_bubble_cloud(int nx, int ny, int nz, int nbubbles, unsigned long long seed,
                                    double* bubbles, float* const fields[7])
{
    if (nx < 1 || ny < 1 || nz < 1 || nbubbles < 0 || nbubbles > 100000 || !fields_ok(fields)) return MPCF_EINVAL;
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    Rng rng(seed);
    std::vector<double> b;   // x, y, z, r
    int tries = 0;
    while ((int)(b.size() / 4) < nbubbles && tries < 1000000) {
        ++tries;
        const double r = std::exp(std::log(0.02) + 0.3 * rng.normal());
        double px, py, pz;
        do {
            px = 2 * rng.uniform() - 1; py = 2 * rng.uniform() - 1; pz = 2 * rng.uniform() - 1;
        } while (px * px + py * py + pz * pz > 1.0);
        px = 0.5 + 0.3 * px; py = 0.5 + 0.3 * py; pz = 0.5 + 0.3 * pz;
        bool ok = true;
        for (size_t k = 0; k < b.size(); k += 4) {
            const double dx = px - b[k], dy = py - b[k + 1], dz = pz - b[k + 2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) < r + b[k + 3]) { ok = false; break; }
        }
        if (ok) { b.push_back(px); b.push_back(py); b.push_back(pz); b.push_back(r); }
    }
    if ((int)(b.size() / 4) < nbubbles) return MPCF_EINVAL;
    if (bubbles) std::memcpy(bubbles, b.data(), b.size() * sizeof(double));
    const double pre[4] = {1000.0, 0.0, 0.0, 100.0};
    const double gas[4] = {1.0, 0.0, 0.0, 1.0};
    const int nb = (int)(b.size() / 4);
    parallel_z(nz, [&](int z) {
        const double cz = h * (z + 0.5);
        for (int y = 0; y < ny; ++y) {
            const double cy = h * (y + 0.5);
            for (int x = 0; x < nx; ++x) {
                const double cx = h * (x + 0.5);
                double dmin = 1e30;
                for (int k = 0; k < nb; ++k) {
                    const double dx = cx - b[4 * k], dy = cy - b[4 * k + 1], dz = cz - b[4 * k + 2];
                    const double d = std::sqrt(dx * dx + dy * dy + dz * dz) - b[4 * k + 3];
                    dmin = std::min(dmin, d);
                }
                const double bubble = inside(dmin, 1.5, h);
                const State s = mix(0.0, bubble, pre, pre, gas);
                store(fields, (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z), s);
            }
        }
    });
    return MPCF_OK;
}

// One air bubble (p = 1, rho = 1) of radius `radius` centred at (cx, cy, cz)
// in water at p = pratio (rho = 1000), at rest.
extern "C" int mpcf_ic_bubble_collapse(int nx, int ny, int nz, double radius, double pratio, double cx0,
                                       double cy0, double cz0, float* const fields[7])
{
    if (nx < 1 || ny < 1 || nz < 1 || !(radius > 0) || !(pratio > 0) || !fields_ok(fields)) return MPCF_EINVAL;
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    const double pre[4] = {1000.0, 0.0, 0.0, pratio};
    const double gas[4] = {1.0, 0.0, 0.0, 1.0};
    parallel_z(nz, [&](int z) {
        const double cz = h * (z + 0.5);
        for (int y = 0; y < ny; ++y) {
            const double cy = h * (y + 0.5);
            for (int x = 0; x < nx; ++x) {
                const double cx = h * (x + 0.5);
                const double r = std::sqrt((cx - cx0) * (cx - cx0) + (cy - cy0) * (cy - cy0) + (cz - cz0) * (cz - cz0));
                const State s = mix(0.0, inside(r - radius, 1.5, h), pre, pre, gas);
                store(fields, (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z), s);
            }
        }
    });
    return MPCF_OK;
}
