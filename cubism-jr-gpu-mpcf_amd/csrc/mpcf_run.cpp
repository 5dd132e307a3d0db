// mpcf_run.cpp -- command-line driver of the benchmark cases (C++ host over
// mpcf.hpp), the counterpart of the reference's main + Sim_* run loops
// (source/main.cpp:24-79, source/Sim/Sim_SteadyStateMPI.cpp:182-225):
//
//   mpcf_run -sim tube|bubble|cloud|collapse -n 64 [-nx -ny -nz] -steps 10
//            [-cfl 0.3] [-tend T] [-mode parity|fast] [-bc absorbing|reflecting]
//            [-seed 0] [-nbubbles 50] [-dump out.bin] [-verb 1]
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
                                  "nbubbles", "dump", "verb", nullptr};
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
        else { std::fprintf(stderr, "-sim tube|bubble|cloud|collapse\n"); return 2; }
        mpcf::check(rc);
        blk.clear_tmp();

        mpcf::DeviceLab lab(blk, mode == "fast" ? MPCF_FLAG_FAST : 0u);
        const int k = bc == "absorbing" ? MPCF_BC_ABSORBING : MPCF_BC_REFLECTING;
        const int kinds[6] = {k, k, k, k, k, k};
        lab.set_bc(kinds);
        mpcf::LSRK3_Integrator stepper(lab, cfl);

        double t = 0;
        int done = 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (done < steps && t < tend) {
            const double dt = stepper(tend - t);
            t += dt;
            ++done;
            if (verb > 1) std::printf("step %d  t = %.9e  dt = %.9e  sos = %.6e\n", done, t, dt, stepper.sos());
        }
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
