// mpcf.hpp -- C++ host interface over the C-ABI (include/mpcf_hip.h), shaped
// like the reference's kernel facades, lab and integrator so that code
// written against them ports by renaming:
//
//   reference (source/...)                         here
//   NodeBlock (NodeBlock.h:22-119)                 mpcf::HostBlock
//   GPUlab (GPUlab.h:33-265)                       mpcf::DeviceLab
//     load_ghosts / max_sos / process_all          same names and arguments
//     virtual _apply_bc(t) per case                set_bc(kinds[6]) (on-device fill)
//   Convection_CUDA(a, dtinvh).compute(...)        mpcf::Convection_HIP(lab, a, dtinvh).compute()
//   Update_CUDA(b).compute(...)                    mpcf::Update_HIP(lab, b).compute()
//   MaxSpeedOfSound_CUDA().compute(...)            mpcf::MaxSpeedOfSound_HIP(lab).compute() -> float
//   LSRK3_IntegratorMPI(grid, GPU, CFL, parser)    mpcf::LSRK3_Integrator(lab, CFL)
//     double operator()(double dt_max)             same
//
// Differences by design: the facades take the lab explicitly (the reference
// keeps device state in file-scope globals, GPU.h:39-85); there is no nslices
// (the block is resident in HBM); failures throw mpcf::Error carrying the
// C-ABI status and message instead of being ignored.
#pragma once

#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpcf_hip.h"

namespace mpcf {

class Error : public std::runtime_error
{
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

inline void check(int rc)
{
    if (rc != MPCF_OK) throw Error(rc, std::string("mpcf: ") + mpcf_last_error());
}

// Seven SoA fields (u) and the RK register (tmp), NX*NY*NZ floats each,
// ID3 order (source/Types.h:31), spacing h = extent / max(NX,NY,NZ)
// (NodeBlock.h:71).
class HostBlock
{
public:
    enum PRIM { R = 0, U, V, W, E, G, P, NPRIMITIVES = 7 };
    static const int NVAR = NPRIMITIVES;

    HostBlock(int nx, int ny, int nz, double extent = 1.0)
        : nx_(nx), ny_(ny), nz_(nz), store_(2 * NVAR * (size_t)nx * ny * nz, 0.0f), data_(NVAR), tmp_(NVAR)
    {
        const size_t n = (size_t)nx * ny * nz;
        for (int q = 0; q < NVAR; ++q) {
            data_[q] = &store_[q * n];
            tmp_[q] = &store_[(NVAR + q) * n];
        }
        int m = nx > ny ? nx : ny;
        m = m > nz ? m : nz;
        h_ = extent / m;
    }

    int sizeX() const { return nx_; }
    int sizeY() const { return ny_; }
    int sizeZ() const { return nz_; }
    double h_gridpoint() const { return h_; }
    size_t cells() const { return (size_t)nx_ * ny_ * nz_; }

    std::vector<float*>& pdata() { return data_; }
    std::vector<float*>& ptmp() { return tmp_; }
    const std::vector<float*>& pdata() const { return data_; }
    const std::vector<float*>& ptmp() const { return tmp_; }

    float& operator()(int ix, int iy, int iz, PRIM p)
    {
        return data_[p][(size_t)ix + (size_t)nx_ * ((size_t)iy + (size_t)ny_ * iz)];
    }

    void clear_tmp() { std::memset(tmp_[0], 0, NVAR * cells() * sizeof(float)); }

private:
    int nx_, ny_, nz_;
    double h_;
    std::vector<float> store_;
    std::vector<float*> data_, tmp_;
};

// The resident block and its device step state (GPUlab counterpart).
class DeviceLab
{
public:
    DeviceLab(int nx, int ny, int nz, double h, unsigned flags = 0) { check(mpcf_create(&ctx_, nx, ny, nz, h, flags)); }
    explicit DeviceLab(const HostBlock& b, unsigned flags = 0)
    {
        check(mpcf_create(&ctx_, b.sizeX(), b.sizeY(), b.sizeZ(), b.h_gridpoint(), flags));
        upload(b);
    }
    ~DeviceLab() { if (ctx_) mpcf_destroy(ctx_); }
    DeviceLab(const DeviceLab&) = delete;
    DeviceLab& operator=(const DeviceLab&) = delete;

    mpcf_ctx* ctx() { return ctx_; }

    // boundary kind per face (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi), the choice a
    // GPUlab<Case>::_apply_bc override makes (e.g. Sim_SodMPI.h:30-39)
    void set_bc(const int kinds[6]) { check(mpcf_set_bc(ctx_, kinds)); }

    void upload(const HostBlock& b)
    {
        check(mpcf_upload(ctx_, MPCF_DATA, b.pdata().data()));
        check(mpcf_upload(ctx_, MPCF_TMP, b.ptmp().data()));
    }
    void download(HostBlock& b)
    {
        check(mpcf_download(ctx_, MPCF_DATA, b.pdata().data()));
        check(mpcf_download(ctx_, MPCF_TMP, b.ptmp().data()));
    }

    // GPUlab::load_ghosts (GPUlab.cpp:622-644); t is unused by the supported BCs
    void load_ghosts(const double t = 0) { (void)t; check(mpcf_load_ghosts(ctx_)); }

    // GPUlab::max_sos (GPUlab.cpp:647-700): returns seconds, sos by reference
    double max_sos(float& sos)
    {
        const auto t0 = std::chrono::steady_clock::now();
        check(mpcf_max_sos(ctx_, &sos));
        return seconds_since(t0);
    }

    // GPUlab::process_all (GPUlab.cpp:703-793): one stage; blocks until done
    // like the reference (which copies the stage's results back) and returns
    // seconds
    double process_all(const float a, const float b, const float dtinvh)
    {
        const auto t0 = std::chrono::steady_clock::now();
        check(mpcf_process_all(ctx_, a, b, dtinvh));
        check(mpcf_synchronize(ctx_));
        return seconds_since(t0);
    }

    void synchronize() { check(mpcf_synchronize(ctx_)); }

    // the device side of GPUlab's halo exchange (GPUlab.cpp:76-112,636-641):
    // a face's three boundary cells out to a device buffer, a neighbour's
    // into the ghost band (faces set to MPCF_BC_HALO); asynchronous
    long long halo_count(const int face)
    {
        long long n = 0;
        check(mpcf_halo_count(ctx_, face, &n));
        return n;
    }
    void halo_export(const int face, void* dev_dst) { check(mpcf_halo_export(ctx_, face, dev_dst)); }
    void halo_import(const int face, const void* dev_src) { check(mpcf_halo_import(ctx_, face, dev_src)); }

    // Sim_SteadyStateMPI::_dump / _save / _restart (Sim_SteadyStateMPI.cpp:138-179):
    // raw [Z][Y][X][C] + XDMF dumps of a streamer, and save.info + save.data
    void dump(int kind, const std::string& dir, const std::string& name, int step, double h)
    {
        check(mpcf_dump(ctx_, kind, dir.c_str(), name.c_str(), step, h));
    }
    void save(const std::string& dir, double t, int step, int fcount, double h)
    {
        check(mpcf_save(ctx_, dir.c_str(), t, step, fcount, h));
    }
    void restart(const std::string& dir, double& t, int& step, int& fcount)
    {
        check(mpcf_restart(ctx_, dir.c_str(), &t, &step, &fcount));
    }

private:
    static double seconds_since(std::chrono::steady_clock::time_point t0)
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    mpcf_ctx* ctx_ = nullptr;
};

// Convection_CUDA (Convection_CUDA.h:16-24): R <- a R - dtinvh L(u)
class Convection_HIP
{
public:
    const float a, dtinvh;
    Convection_HIP(DeviceLab& lab, const float a_, const float dtinvh_) : a(a_), dtinvh(dtinvh_), lab_(lab) {}
    void compute() { check(mpcf_convection(lab_.ctx(), a, dtinvh)); }
private:
    DeviceLab& lab_;
};

// Update_CUDA (Update_CUDA.h:14-26): u <- b R + u
class Update_HIP
{
public:
    Update_HIP(DeviceLab& lab, const float b = 1) : m_b(b), lab_(lab) {}
    void compute() { check(mpcf_update(lab_.ctx(), m_b)); }
protected:
    float m_b;
private:
    DeviceLab& lab_;
};

// MaxSpeedOfSound_CUDA (MaxSpeedOfSound_CUDA.h:13-17), returning the value
// like MaxSpeedOfSound_CPP::compute (MaxSpeedOfSound.h:15-19)
class MaxSpeedOfSound_HIP
{
public:
    explicit MaxSpeedOfSound_HIP(DeviceLab& lab) : lab_(lab) {}
    float compute()
    {
        float s = 0;
        check(mpcf_max_sos(lab_.ctx(), &s));
        return s;
    }
private:
    DeviceLab& lab_;
};

// LSRK3_IntegratorMPI (LSRK3_IntegratorMPI.h:21-37): one call = one step,
// dt = min(CFL h / sos, dt_max) computed on the device; returns dt
class LSRK3_Integrator
{
public:
    LSRK3_Integrator(DeviceLab& lab, const double CFL_) : CFL(CFL_), lab_(lab) {}
    double operator()(const double dt_max)
    {
        double dt = 0;
        check(mpcf_step(lab_.ctx(), CFL, dt_max, &dt));
        return dt;
    }
    // last step's max speed of sound
    float sos()
    {
        double dt;
        float s;
        check(mpcf_last_dt(lab_.ctx(), &dt, &s));
        return s;
    }
    const double CFL;
private:
    DeviceLab& lab_;
};

} // namespace mpcf
