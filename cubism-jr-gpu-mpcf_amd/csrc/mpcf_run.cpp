// mpcf_run.cpp -- command-line driver of the benchmark cases (C++ host over
// mpcf.hpp), the counterpart of the reference's main + Sim_* run loops
// (source/main.cpp:24-79, source/Sim/Sim_SteadyStateMPI.cpp:182-225):
//
//   mpcf_run -sim CASE -n 64 [-nx -ny -nz] -steps 10
//            [-cfl 0.3] [-tend T] [-mode parity|fast] [-bc absorbing|reflecting]
//            [-seed 0] [-nbubbles 50] [-dump out.bin] [-verb 1]
//            [-fpath dir] [-dumpinterval T | -dumpperiod K] [-saveperiod K]
//            [-restart 0|1]
//   CASE: tube | bubble | cloud | collapse (the benchmark inputs), or the
//   reference's drivers under their own names or short ones:
//     SteadyStateMPI | steady   uniform state: -rho -u -v -w -p -g -pc
//     SodMPI | sod               -x0 -rho1 -rho2 -u1 -u2 -p1 -p2 -g1 -g2 -pc1
//                                -pc2 -reverse -config x|y|z
//     2DSBIMPI | 2dsbi           -x0 -mach -rho -rhobubble -u -p -pbubble -g1
//                                -g2 -pc1 -pc2 -mollfactor -reverse
//                                -config xy|xz|yx|yz|zx|zy, bubbles from
//                                <config>bubbles.dat in the working directory;
//                                reflecting on the faces of the minor axis
//     SICCloudMPI | sic          cloud_config.dat and cloud.dat from the working
//                                directory; -rho0 -u0 -p0 -rhoB -uB -pB
//                                -pressureratio -shockNx -shockNy -shockNz
//                                -shockSx -shockSy -shockSz -g1 -g2 -pc1 -pc2
//     StaticIC | static          the index probe (-config as 2dsbi), dumped to
//                                -fpath as <fname>_0000 without stepping
//   Boundaries: -bc sets all six faces; without it each case takes the
//   reference's (absorbing, 2dsbi as above).
//   -dumpinterval T: the reference's loop -- dt is cut to land on every
//            multiple of T, a tensorial-streamer dump data_NNNN follows each,
//            and one more is written before the first step and after the
//            last (Sim_SteadyStateMPI.cpp:74-75,197-221); -dumpperiod K: a dump
//            every K steps instead; -saveperiod K (or -saveinterval K): a
//            checkpoint every K steps and at the end; -restart 1: continue
//            from the checkpoint in -fpath (Sim_SteadyStateMPI.cpp:138-179).
//   -nsteps K: the reference's step limit (0 = none, -tend bounds the run).
//
// Arguments follow the reference's "-key value" convention
// (source/IO/ArgumentParser.h:68-130); an unknown key or a malformed value is
// an error (exit 2).  -dump writes the final state as
//   "MPCF0001" | int32 nx, ny, nz, steps | float64 t | 7 x nx*ny*nz float32 (ID3)
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include <algorithm>

#include "mpcf.hpp"

namespace {

struct Args
{
    std::map<std::string, std::string> kv;
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    std::string str(const std::string& k, const std::string& d) const { return has(k) ? kv.at(k) : d; }
    long num(const std::string& k, long d) const
    {
        if (!has(k)) return d;
        char* end = nullptr;
        const long v = std::strtol(kv.at(k).c_str(), &end, 10);
        if (!end || *end) { std::fprintf(stderr, "bad integer for -%s: %s\n", k.c_str(), kv.at(k).c_str()); std::exit(2); }
        return v;
    }
    double real(const std::string& k, double d) const
    {
        if (!has(k)) return d;
        char* end = nullptr;
        const double v = std::strtod(kv.at(k).c_str(), &end);
        if (!end || *end) { std::fprintf(stderr, "bad number for -%s: %s\n", k.c_str(), kv.at(k).c_str()); std::exit(2); }
        return v;
    }
    // ArgumentParser::asBool (ArgumentParser.h:51-58): "0" and "false" are false
    bool flag(const std::string& k, bool d) const
    {
        if (!has(k) || kv.at(k).empty()) return d;
        return !(kv.at(k) == "0" || kv.at(k) == "false");
    }
};

Args parse(int argc, char** argv)
{
    static const char* known[] = {"sim", "n", "nx", "ny", "nz", "steps", "nsteps", "cfl", "tend", "mode", "bc",
                                  "seed", "nbubbles", "dump", "verb", "fpath", "dumpperiod", "dumpinterval",
                                  "saveperiod", "saveinterval", "restart", "rho0", "u0", "p0", "rhoB", "uB", "pB",
                                  "pressureratio", "shockNx", "shockNy", "shockNz", "shockSx", "shockSy", "shockSz",
                                  "g1", "g2", "pc1", "pc2", "x0", "rho1", "rho2", "u1", "u2", "p1", "p2", "reverse",
                                  "config", "mach", "rho", "rhobubble", "u", "v", "w", "p", "pbubble", "g", "pc",
                                  "mollfactor", "fname", nullptr};
    Args a;
    for (int i = 1; i < argc; i += 2) {
        if (argv[i][0] != '-' || i + 1 >= argc) { std::fprintf(stderr, "expected -key value at '%s'\n", argv[i]); std::exit(2); }
        const std::string k = argv[i] + 1;
        bool ok = false;
        for (int j = 0; known[j]; ++j) ok = ok || k == known[j];
        if (!ok) { std::fprintf(stderr, "unknown argument -%s\n", k.c_str()); std::exit(2); }
        a.kv[k] = argv[i + 1];
    }
    return a;
}

// Sim_SICCloudMPI::_ic (Sim_SICCloudMPI.cpp:43-128): parameters from the
// command line, the cloud from cloud_config.dat / cloud.dat
int sic_ic(const Args& a, int nx, int ny, int nz, float* const* f, int verb)
{
    mpcf_sic_params p;
    mpcf_sic_defaults(&p);
    p.rho0 = a.real("rho0", p.rho0); p.u0 = a.real("u0", p.u0); p.p0 = a.real("p0", p.p0);
    p.rhoB = a.real("rhoB", p.rhoB); p.uB = a.real("uB", p.uB); p.pB = a.real("pB", p.pB);
    p.pressure_ratio = a.real("pressureratio", p.pressure_ratio);
    p.shock_n[0] = a.real("shockNx", p.shock_n[0]); p.shock_n[1] = a.real("shockNy", p.shock_n[1]);
    p.shock_n[2] = a.real("shockNz", p.shock_n[2]);
    p.shock_s[0] = a.real("shockSx", p.shock_s[0]); p.shock_s[1] = a.real("shockSy", p.shock_s[1]);
    p.shock_s[2] = a.real("shockSz", p.shock_s[2]);
    p.g1 = a.real("g1", p.g1); p.g2 = a.real("g2", p.g2); p.pc1 = a.real("pc1", p.pc1); p.pc2 = a.real("pc2", p.pc2);
    double cfg[11];
    int rc = mpcf_read_cloud_config("cloud_config.dat", cfg);
    if (rc != MPCF_OK) return rc;
    int n = 0;
    rc = mpcf_read_cloud("cloud.dat", nullptr, 0, &n);
    if (rc != MPCF_OK) return rc;
    std::vector<double> shapes(4 * (size_t)n);
    rc = mpcf_read_cloud("cloud.dat", shapes.data(), n, &n);
    if (rc != MPCF_OK) return rc;
    if (n != (int)cfg[0])
        std::fprintf(stderr, "cloud.dat holds %d shapes, cloud_config.dat announces %d\n", n, (int)cfg[0]);
    double post[5];
    rc = mpcf_sic_post_shock(&p, post);
    if (rc != MPCF_OK) return rc;
    if (verb)
        std::printf("shock: p-ratio %g, Mach %.6f, speed %.6f; post-shock rho %.6f u %.6f p %.6f; %d bubbles\n",
                    p.pressure_ratio, post[3], p.u0 + post[4] * post[3], post[0], post[1], post[2], n);
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    return mpcf_ic_sic_cloud(nx, ny, nz, h, &p, shapes.data(), n, f);
}

int sod_ic(const Args& a, int nx, int ny, int nz, float* const* f)
{
    mpcf_sod_params p;
    mpcf_sod_defaults(&p);
    p.x0 = a.real("x0", p.x0);
    p.rho1 = a.real("rho1", p.rho1); p.u1 = a.real("u1", p.u1); p.p1 = a.real("p1", p.p1);
    p.rho2 = a.real("rho2", p.rho2); p.u2 = a.real("u2", p.u2); p.p2 = a.real("p2", p.p2);
    p.g1 = a.real("g1", p.g1); p.g2 = a.real("g2", p.g2); p.pc1 = a.real("pc1", p.pc1); p.pc2 = a.real("pc2", p.pc2);
    p.reverse = a.flag("reverse", false);
    const std::string conf = a.str("config", "x");
    if (conf != "x" && conf != "y" && conf != "z") { std::fprintf(stderr, "-config x|y|z\n"); std::exit(2); }
    p.dir = conf[0] - 'x';
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    return mpcf_ic_sod(nx, ny, nz, h, &p, f);
}

int sbi_ic(const Args& a, int nx, int ny, int nz, float* const* f, int dims[3], int verb)
{
    mpcf_sbi_params p;
    mpcf_sbi_defaults(&p);
    const std::string conf = a.str("config", "");
    int rc = mpcf_sbi_config(conf.c_str(), p.dims);   // mandatory, as in the reference
    if (rc != MPCF_OK) return rc;
    std::memcpy(dims, p.dims, sizeof(p.dims));
    p.x0 = a.real("x0", p.x0); p.mach = a.real("mach", p.mach);
    p.rho = a.real("rho", p.rho); p.u = a.real("u", p.u); p.p = a.real("p", p.p);
    p.rhobubble = a.real("rhobubble", p.rhobubble); p.pbubble = a.real("pbubble", p.pbubble);
    p.g1 = a.real("g1", p.g1); p.g2 = a.real("g2", p.g2); p.pc1 = a.real("pc1", p.pc1); p.pc2 = a.real("pc2", p.pc2);
    p.mollfactor = (int)a.num("mollfactor", p.mollfactor);
    p.reverse = a.flag("reverse", false);
    const std::string file = std::string(1, "xyz"[p.dims[0]]) + "xyz"[p.dims[1]] + "bubbles.dat";
    int n = 0;
    rc = mpcf_read_bubbles(file.c_str(), nullptr, 0, &n);
    if (rc != MPCF_OK) return rc;
    std::vector<double> b(4 * (size_t)n);
    rc = mpcf_read_bubbles(file.c_str(), b.data(), n, &n);
    if (rc != MPCF_OK) return rc;
    if (verb) std::printf("2D shock-bubble, -config %s: %d bubbles from %s\n", conf.c_str(), n, file.c_str());
    const double h = 1.0 / std::max(nx, std::max(ny, nz));
    return mpcf_ic_sbi(nx, ny, nz, h, &p, b.data(), n, f);
}

int steady_ic(const Args& a, int nx, int ny, int nz, float* const* f)
{
    const double st[7] = {a.real("rho", 1.0), a.real("u", 0.0), a.real("v", 0.0), a.real("w", 0.0),
                          a.real("p", 1.0),   a.real("g", 1.4), a.real("pc", 0.0)};
    return mpcf_ic_steady(nx, ny, nz, st, f);
}

} // namespace

int main(int argc, char** argv)
{
    const Args a = parse(argc, argv);
    const int n = (int)a.num("n", 64);
    const int nx = (int)a.num("nx", n), ny = (int)a.num("ny", n), nz = (int)a.num("nz", n);
    // -steps K counts the steps of this run; -nsteps K is the reference's
    // limit on the absolute step number (0 = none)
    const long nsteps_abs = a.num("nsteps", 0);
    const long steps = a.has("nsteps") && !a.has("steps") ? LONG_MAX : a.num("steps", 10);
    const double cfl = a.real("cfl", 0.3);
    const double tend = a.real("tend", 1e30);
    std::string sim = a.str("sim", "tube");
    const std::string mode = a.str("mode", "parity"), bc = a.str("bc", "absorbing");
    const unsigned long long seed = (unsigned long long)a.num("seed", 0);
    const int verb = (int)a.num("verb", 1);
    if (mode != "parity" && mode != "fast") { std::fprintf(stderr, "-mode parity|fast\n"); return 2; }
    if (bc != "absorbing" && bc != "reflecting") { std::fprintf(stderr, "-bc absorbing|reflecting\n"); return 2; }
    if (steps == LONG_MAX && nsteps_abs <= 0 && !a.has("tend")) {
        std::fprintf(stderr, "-nsteps 0 needs -tend\n");
        return 2;
    }
    static const char* alias[][2] = {{"SteadyStateMPI", "steady"}, {"SodMPI", "sod"}, {"2DSBIMPI", "2dsbi"},
                                     {"SICCloudMPI", "sic"},        {"StaticIC", "static"}};
    for (const auto& al : alias)
        if (sim == al[0]) sim = al[1];

    try {
        mpcf::HostBlock blk(nx, ny, nz);
        float* const* f = blk.pdata().data();
        int kinds[6] = {MPCF_BC_ABSORBING, MPCF_BC_ABSORBING, MPCF_BC_ABSORBING,
                        MPCF_BC_ABSORBING, MPCF_BC_ABSORBING, MPCF_BC_ABSORBING};
        int rc = MPCF_EINVAL;
        if (sim == "tube") rc = mpcf_ic_shock_tube(nx, ny, nz, 0, 1e-3, seed, f);
        else if (sim == "bubble") rc = mpcf_ic_shock_bubble(nx, ny, nz, seed, f);
        else if (sim == "cloud") rc = mpcf_ic_bubble_cloud(nx, ny, nz, (int)a.num("nbubbles", 50), seed, nullptr, f);
        else if (sim == "collapse") rc = mpcf_ic_bubble_collapse(nx, ny, nz, 0.2, 100.0, 0.5, 0.5, 0.5, f);
        else if (sim == "sic") rc = sic_ic(a, nx, ny, nz, f, verb);
        else if (sim == "sod") rc = sod_ic(a, nx, ny, nz, f);
        else if (sim == "steady") rc = steady_ic(a, nx, ny, nz, f);
        else if (sim == "2dsbi") {
            int dims[3] = {0, 1, 2};
            rc = sbi_ic(a, nx, ny, nz, f, dims, verb);
            // GPUlab2DSBI_{x,y,z}reflect: reflecting across the minor axis
            if (rc == MPCF_OK) kinds[2 * dims[1]] = kinds[2 * dims[1] + 1] = MPCF_BC_REFLECTING;
        } else if (sim == "static") {
            int dims[3] = {0, 1, 2};
            rc = mpcf_sbi_config(a.str("config", "").c_str(), dims);
            if (rc == MPCF_OK) rc = mpcf_ic_static(nx, ny, nz, dims, f);
        } else {
            std::fprintf(stderr, "-sim tube|bubble|cloud|collapse|sic|sod|2dsbi|steady|static "
                                 "(or SICCloudMPI|SodMPI|2DSBIMPI|SteadyStateMPI|StaticIC)\n");
            return 2;
        }
        mpcf::check(rc);
        blk.clear_tmp();
        if (a.has("bc")) {
            const int k = bc == "absorbing" ? MPCF_BC_ABSORBING : MPCF_BC_REFLECTING;
            for (int i = 0; i < 6; ++i) kinds[i] = k;
        }

        mpcf::DeviceLab lab(blk, mode == "fast" ? MPCF_FLAG_FAST : 0u);
        lab.set_bc(kinds);
        const std::string fpath = a.str("fpath", ".");
        int fcount = 0;
        auto dump = [&](const char* base, int step) {
            char name[64];
            std::snprintf(name, sizeof(name), "%s_%04d", base, fcount++);
            lab.dump(MPCF_STREAM_TENSOR, fpath, name, step, blk.h_gridpoint());
        };
        if (sim == "static") {   // Sim_StaticIC::_setup: the IC, one dump, no steps
            dump(a.str("fname", "staticIC").c_str(), 0);
            lab.synchronize();
            return 0;
        }
        mpcf::LSRK3_Integrator stepper(lab, cfl);

        const double dumpinterval = a.real("dumpinterval", 0);
        const int dumpperiod = (int)a.num("dumpperiod", 0);
        const int saveperiod = (int)(a.has("saveinterval") ? a.num("saveinterval", 0) : a.num("saveperiod", 0));
        double t = 0;
        long done = 0;
        int step0 = 0;
        const bool restarted = a.flag("restart", false);
        if (restarted) {
            lab.restart(fpath, t, step0, fcount);
            if (verb) std::printf("restart from %s: t = %.9e, step %d\n", fpath.c_str(), t, step0);
            if (dumpinterval > 0) dump("restart_ic", step0);
        } else if (dumpinterval > 0) {
            dump("data", 0);
        }
        // Sim_SteadyStateMPI.cpp:53,64: the next dump time, after the restart dump
        double tnextdump = dumpinterval > 0 ? (restarted ? fcount * dumpinterval : dumpinterval) : 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (done < steps && t < tend) {
            double dt_max = tend - t;
            if (dumpinterval > 0 && tnextdump - t < dt_max) dt_max = tnextdump - t;
            const double dt = stepper(dt_max);
            t += dt;
            ++done;
            const int step = step0 + (int)done;
            if (verb > 1) std::printf("step %d  t = %.9e  dt = %.9e  sos = %.6e\n", step, t, dt, stepper.sos());
            if (dumpinterval > 0 && (float)t == (float)tnextdump) {
                tnextdump += dumpinterval;
                dump("data", step);
            }
            if (dumpperiod > 0 && step % dumpperiod == 0) dump("data", step);
            if (saveperiod > 0 && step % saveperiod == 0) lab.save(fpath, t, step, fcount, blk.h_gridpoint());
            if (nsteps_abs > 0 && step == nsteps_abs) break;
        }
        if (dumpinterval > 0) dump("data", step0 + (int)done);
        if (saveperiod > 0) lab.save(fpath, t, step0 + (int)done, fcount, blk.h_gridpoint());
        lab.synchronize();
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (verb) std::printf("%s %dx%dx%d %s: %ld steps, t = %.9e, %.3f s, %.3e cell-updates/s\n", sim.c_str(), nx, ny,
                              nz, mode.c_str(), done, t, el, (double)blk.cells() * 3 * done / el);
        if (a.has("dump")) {
            lab.download(blk);
            FILE* fp = std::fopen(a.str("dump", "").c_str(), "wb");
            if (!fp) { std::fprintf(stderr, "cannot open dump file\n"); return 1; }
            const int hdr[4] = {nx, ny, nz, (int)done};
            bool ok = std::fwrite("MPCF0001", 1, 8, fp) == 8 && std::fwrite(hdr, sizeof(int), 4, fp) == 4 &&
                      std::fwrite(&t, sizeof(double), 1, fp) == 1;
            for (int q = 0; q < 7 && ok; ++q) ok = std::fwrite(blk.pdata()[q], sizeof(float), blk.cells(), fp) == blk.cells();
            std::fclose(fp);
            if (!ok) { std::fprintf(stderr, "dump write failed\n"); return 1; }
        }
    } catch (const mpcf::Error& e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
        return 1;
    }
    return 0;
}
