// mpcf_run.cpp -- command-line driver of the benchmark cases (C++ host over
// mpcf.hpp), the counterpart of the reference's main + Sim_* run loops
// (source/main.cpp:24-79, source/Sim/Sim_SteadyStateMPI.cpp:182-225):
//
//   mpcf_run -sim tube|bubble|cloud|collapse|sic -n 64 [-nx -ny -nz] -steps 10
//            [-cfl 0.3] [-tend T] [-mode parity|fast] [-bc absorbing|reflecting]
//            [-seed 0] [-nbubbles 50] [-dump out.bin] [-verb 1]
//            [-fpath dir] [-dumpperiod K] [-saveperiod K] [-restart 0|1]
//   -sim sic: the shock-induced cloud collapse of Sim_SICCloudMPI, reading
//            cloud_config.dat and cloud.dat from the working directory and the
//            reference's options -rho0 -u0 -p0 -rhoB -uB -pB -pressureratio
//            -shockNx -shockNy -shockNz -shockSx -shockSy -shockSz -g1 -g2
//            -pc1 -pc2 (Sim_SICCloudMPI.cpp:53-77)
//   -dumpperiod K: a tensorial-streamer dump data_NNNN every K steps into
//            -fpath; -saveperiod K: a checkpoint every K steps (and at the end);
//            -restart 1: continue from the checkpoint in -fpath
//            (Sim_SteadyStateMPI.cpp:138-179)
//
// Arguments follow the reference's "-key value" convention
// (source/IO/ArgumentParser.h:68-130); an unknown key or a malformed value is
// an error (exit 2).  -dump writes the final state as
//   "MPCF0001" | int32 nx, ny, nz, steps | float64 t | 7 x nx*ny*nz float32 (ID3)
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
};

Args parse(int argc, char** argv)
{
    static const char* known[] = {"sim", "n", "nx", "ny", "nz", "steps", "cfl", "tend", "mode", "bc", "seed",
                                  "nbubbles", "dump", "verb", "fpath", "dumpperiod", "saveperiod", "restart",
                                  "rho0", "u0", "p0", "rhoB", "uB", "pB", "pressureratio", "shockNx", "shockNy",
                                  "shockNz", "shockSx", "shockSy", "shockSz", "g1", "g2", "pc1", "pc2", nullptr};
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

} // namespace

int main(int argc, char** argv)
{
    const Args a = parse(argc, argv);
    const int n = (int)a.num("n", 64);
    const int nx = (int)a.num("nx", n), ny = (int)a.num("ny", n), nz = (int)a.num("nz", n);
    const int steps = (int)a.num("steps", 10);
    const double cfl = a.real("cfl", 0.3);
    const double tend = a.real("tend", 1e30);
    const std::string sim = a.str("sim", "tube"), mode = a.str("mode", "parity"), bc = a.str("bc", "absorbing");
    const unsigned long long seed = (unsigned long long)a.num("seed", 0);
    const int verb = (int)a.num("verb", 1);
    if (mode != "parity" && mode != "fast") { std::fprintf(stderr, "-mode parity|fast\n"); return 2; }
    if (bc != "absorbing" && bc != "reflecting") { std::fprintf(stderr, "-bc absorbing|reflecting\n"); return 2; }

    try {
        mpcf::HostBlock blk(nx, ny, nz);
        float* const* f = blk.pdata().data();
        int rc = MPCF_EINVAL;
        if (sim == "tube") rc = mpcf_ic_shock_tube(nx, ny, nz, 0, 1e-3, seed, f);
        else if (sim == "bubble") rc = mpcf_ic_shock_bubble(nx, ny, nz, seed, f);
        else if (sim == "cloud") rc = mpcf_ic_bubble_cloud(nx, ny, nz, (int)a.num("nbubbles", 50), seed, nullptr, f);
        else if (sim == "collapse") rc = mpcf_ic_bubble_collapse(nx, ny, nz, 0.2, 100.0, 0.5, 0.5, 0.5, f);
        else if (sim == "sic") rc = sic_ic(a, nx, ny, nz, f, verb);
        else { std::fprintf(stderr, "-sim tube|bubble|cloud|collapse|sic\n"); return 2; }
        mpcf::check(rc);
        blk.clear_tmp();

        mpcf::DeviceLab lab(blk, mode == "fast" ? MPCF_FLAG_FAST : 0u);
        const int k = bc == "absorbing" ? MPCF_BC_ABSORBING : MPCF_BC_REFLECTING;
        const int kinds[6] = {k, k, k, k, k, k};
        lab.set_bc(kinds);
        mpcf::LSRK3_Integrator stepper(lab, cfl);

        const std::string fpath = a.str("fpath", ".");
        const int dumpperiod = (int)a.num("dumpperiod", 0), saveperiod = (int)a.num("saveperiod", 0);
        double t = 0;
        int done = 0, step0 = 0, fcount = 0;
        if (a.num("restart", 0)) {
            lab.restart(fpath, t, step0, fcount);
            if (verb) std::printf("restart from %s: t = %.9e, step %d\n", fpath.c_str(), t, step0);
        }
        const auto t0 = std::chrono::steady_clock::now();
        while (done < steps && t < tend) {
            const double dt = stepper(tend - t);
            t += dt;
            ++done;
            const int step = step0 + done;
            if (verb > 1) std::printf("step %d  t = %.9e  dt = %.9e  sos = %.6e\n", step, t, dt, stepper.sos());
            if (dumpperiod > 0 && step % dumpperiod == 0) {
                char name[32];
                std::snprintf(name, sizeof(name), "data_%04d", fcount++);
                lab.dump(MPCF_STREAM_TENSOR, fpath, name, step, blk.h_gridpoint());
            }
            if (saveperiod > 0 && step % saveperiod == 0) lab.save(fpath, t, step, fcount, blk.h_gridpoint());
        }
        if (saveperiod > 0) lab.save(fpath, t, step0 + done, fcount, blk.h_gridpoint());
        lab.synchronize();
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (verb) std::printf("%s %dx%dx%d %s: %d steps, t = %.9e, %.3f s, %.3e cell-updates/s\n", sim.c_str(), nx, ny,
                              nz, mode.c_str(), done, t, el, (double)blk.cells() * 3 * done / el);
        if (a.has("dump")) {
            lab.download(blk);
            FILE* fp = std::fopen(a.str("dump", "").c_str(), "wb");
            if (!fp) { std::fprintf(stderr, "cannot open dump file\n"); return 1; }
            const int hdr[4] = {nx, ny, nz, done};
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
