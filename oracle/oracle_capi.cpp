/*
 * oracle_capi.cpp -- extern "C" face of the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Loaded by tests/ (ctypes), __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg only.  Arrays cross this boundary as one contiguous block
 * of 7 fields, field-major, each NX*NY*NZ in the reference's ID3 order
 * (source/Types.h:31).
 */
#include "mpcf_oracle.hpp"
#include <omp.h>

using namespace oracle;

namespace {

template <typename Real>
void set_fields(Block<Real>* b, int which, const Real* src)
{
    const size_t n = (size_t)b->NX*b->NY*b->NZ;
    for (int q = 0; q < NVAR; ++q) {
        std::vector<Real>& dst = which ? b->tmp[q] : b->data[q];
        std::memcpy(dst.data(), src + q*n, n*sizeof(Real));
    }
}

template <typename Real>
void get_fields(const Block<Real>* b, int which, Real* dst)
{
    const size_t n = (size_t)b->NX*b->NY*b->NZ;
    for (int q = 0; q < NVAR; ++q) {
        const std::vector<Real>& src = which ? b->tmp[q] : b->data[q];
        std::memcpy(dst + q*n, src.data(), n*sizeof(Real));
    }
}

// halo of one face (face = 2*dir + side), 7 fields, reference layout
template <typename Real>
void get_halo(const Block<Real>* b, int face, Real* dst)
{
    for (int q = 0; q < NVAR; ++q) {
        const std::vector<Real>* h = nullptr;
        switch (face) {
            case 0: h = &b->hxl[q]; break; case 1: h = &b->hxr[q]; break;
            case 2: h = &b->hyl[q]; break; case 3: h = &b->hyr[q]; break;
            case 4: h = &b->hzl[q]; break; default: h = &b->hzr[q]; break;
        }
        std::memcpy(dst + q*h->size(), h->data(), h->size()*sizeof(Real));
    }
}

template <typename Real>
void set_halo(Block<Real>* b, int face, const Real* src)
{
    for (int q = 0; q < NVAR; ++q) {
        std::vector<Real>* h = nullptr;
        switch (face) {
            case 0: h = &b->hxl[q]; break; case 1: h = &b->hxr[q]; break;
            case 2: h = &b->hyl[q]; break; case 3: h = &b->hyr[q]; break;
            case 4: h = &b->hzl[q]; break; default: h = &b->hzr[q]; break;
        }
        std::memcpy(h->data(), src + q*h->size(), h->size()*sizeof(Real));
    }
}

// face arrays of one direction: 12 arrays (7 fluxes, vel, Gm, Gp, Pm, Pp),
// each (N_d+1) along d, ID3 order with that extent
template <typename Real>
void faces(Block<Real>* b, int dir, Real* dst)
{
    FaceArrays<Real> fa;
    compute_faces(*b, fa);
    const size_t nf = fa.vel[dir].size();
    for (int q = 0; q < NVAR; ++q) std::memcpy(dst + q*nf, fa.flux[dir][q].data(), nf*sizeof(Real));
    std::memcpy(dst + 7*nf,  fa.vel[dir].data(), nf*sizeof(Real));
    std::memcpy(dst + 8*nf,  fa.Gm[dir].data(),  nf*sizeof(Real));
    std::memcpy(dst + 9*nf,  fa.Gp[dir].data(),  nf*sizeof(Real));
    std::memcpy(dst + 10*nf, fa.Pm[dir].data(),  nf*sizeof(Real));
    std::memcpy(dst + 11*nf, fa.Pp[dir].data(),  nf*sizeof(Real));
}

// n independent faces; st is n x 7 x 6 (conserved quantities at the six
// stencil points i-3..i+2); out is n x 12
template <typename Real>
void face_batch(int n, int dir, const Real* st, Real* out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Real s[NVAR][6];
        for (int q = 0; q < NVAR; ++q)
            for (int k = 0; k < 6; ++k) s[q][k] = st[(size_t)i*42 + q*6 + k];
        const FaceResult<Real> r = face_flux<Real>(dir, s);
        Real* o = out + (size_t)i*12;
        for (int q = 0; q < NVAR; ++q) o[q] = r.f[q];
        o[7] = r.vel; o[8] = r.Gm; o[9] = r.Gp; o[10] = r.Pm; o[11] = r.Pp;
    }
}

} // namespace

#define ORC_DEFINE(SUF, T)                                                                   \
extern "C" void* orc_create_##SUF(int nx, int ny, int nz) { return new Block<T>(nx, ny, nz); } \
extern "C" void  orc_destroy_##SUF(void* b) { delete (Block<T>*)b; }                           \
extern "C" void  orc_set_##SUF(void* b, int which, const T* src) { set_fields((Block<T>*)b, which, src); } \
extern "C" void  orc_get_##SUF(void* b, int which, T* dst) { get_fields((const Block<T>*)b, which, dst); } \
extern "C" void  orc_apply_bc_##SUF(void* b, const int* bc) { apply_bc(*(Block<T>*)b, bc); }  \
extern "C" void  orc_get_halo_##SUF(void* b, int face, T* dst) { get_halo((const Block<T>*)b, face, dst); } \
extern "C" void  orc_set_halo_##SUF(void* b, int face, const T* src) { set_halo((Block<T>*)b, face, src); } \
extern "C" void  orc_rhs_##SUF(void* b, T a, T dtinvh) { rhs(*(Block<T>*)b, a, dtinvh); }      \
extern "C" void  orc_update_##SUF(void* b, T bb) { update(*(Block<T>*)b, bb); }                \
extern "C" T     orc_max_sos_##SUF(void* b, int cpu_form) { return max_sos(*(const Block<T>*)b, cpu_form); } \
extern "C" double orc_step_##SUF(void* b, double h, double cfl, double dt_max, const int* bc, int cpu_sos, float* sos) \
    { return lsrk3_step(*(Block<T>*)b, h, cfl, dt_max, bc, cpu_sos, sos); }                   \
extern "C" void  orc_faces_##SUF(void* b, int dir, T* dst) { faces((Block<T>*)b, dir, dst); } \
extern "C" void  orc_face_batch_##SUF(int n, int dir, const T* st, T* out) { face_batch<T>(n, dir, st, out); } \
extern "C" void  orc_weno_##SUF(int n, const T* in5, T* minus, T* plus)                       \
{                                                                                           \
    for (int i = 0; i < n; ++i) {                                                           \
        const T* a = in5 + (size_t)5*i;                                                     \
        minus[i] = weno_minus_clipped<T>(a[0], a[1], a[2], a[3], a[4]);                     \
        plus[i]  = weno_pluss_clipped<T>(a[0], a[1], a[2], a[3], a[4]);                     \
    }                                                                                       \
}                                                                                           \
extern "C" void  orc_weno_raw_##SUF(int n, const T* in5, T* minus, T* plus)                   \
{                                                                                           \
    for (int i = 0; i < n; ++i) {                                                           \
        const T* a = in5 + (size_t)5*i;                                                     \
        minus[i] = weno_minus<T>(a[0], a[1], a[2], a[3], a[4]);                             \
        plus[i]  = weno_pluss<T>(a[0], a[1], a[2], a[3], a[4]);                             \
    }                                                                                       \
}

ORC_DEFINE(f, float)
ORC_DEFINE(d, double)

extern "C" void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
extern "C" int  orc_get_threads(void) { return omp_get_max_threads(); }
