/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE.  Compiled (by oracle/Makefile,
 * target `ref`) together with the reference's own, unmodified sources
 *   source/MaxSpeedOfSound.cpp   (MaxSpeedOfSound_CPP::compute)
 *   source/NodeBlock.cpp, source/Types.cpp
 * and the reference's header-only BoundaryConditions<TGrid>
 * (source/BoundaryConditions.h), instantiated with the reference's own
 * NodeBlock as TGrid.  The block size is fixed at compile time by the
 * reference's _BLOCKSIZE{X,Y,Z}_ knobs (Makefile.config:252), so this
 * harness only serves blocks of that size.  Output: oracle/_ref/libref.so.
 *
 * These are the only parts of the path that build from the reference's
 * sources with the tools in this image (the CUDA kernels need nvcc and the
 * removed texture-reference API; the integrator and GPUlab need MPI).
 */
#include "Types.h"
#include "MaxSpeedOfSound.h"
#include "BoundaryConditions.h"

extern "C" void ref_sizes(int* n)
{
    n[0] = NodeBlock::sizeX; n[1] = NodeBlock::sizeY; n[2] = NodeBlock::sizeZ;
}

// fields: 7 contiguous NX*NY*NZ arrays
extern "C" float ref_max_sos_cpp(float* fields)
{
    const size_t n = (size_t)NodeBlock::sizeX*NodeBlock::sizeY*NodeBlock::sizeZ;
    RealPtrVec_t src(7);
    for (int q = 0; q < 7; ++q) src[q] = fields + q*n;
    MaxSpeedOfSound_CPP k;
    return k.compute(src);
}

// kind: 0 absorbing, 1 reflecting; face = 2*dir + side; halo: 7 contiguous
// arrays of the face's ghost-buffer size (GPUlab.cpp:33-35)
template <int dir, int side>
static void apply_one(int kind, RealPtrVec_t& data, RealPtrVec_t& halo)
{
    BoundaryConditions<NodeBlock> bc(data);
    if (dir == 0) {
        if (kind == 0) bc.template applyBC_absorbing<dir, side, ghostmap::X>(halo);
        else           bc.template applyBC_reflecting<dir, side, ghostmap::X>(halo);
    } else if (dir == 1) {
        if (kind == 0) bc.template applyBC_absorbing<dir, side, ghostmap::Y>(halo);
        else           bc.template applyBC_reflecting<dir, side, ghostmap::Y>(halo);
    } else {
        if (kind == 0) bc.template applyBC_absorbing<dir, side, ghostmap::Z>(halo);
        else           bc.template applyBC_reflecting<dir, side, ghostmap::Z>(halo);
    }
}

extern "C" void ref_apply_bc(int face, int kind, float* fields, float* halo_out)
{
    const int NX = NodeBlock::sizeX, NY = NodeBlock::sizeY, NZ = NodeBlock::sizeZ;
    const size_t n = (size_t)NX*NY*NZ;
    const int dir = face / 2;
    const size_t nh = dir == 0 ? (size_t)3*NY*NZ : dir == 1 ? (size_t)NX*3*NZ : (size_t)NX*NY*3;
    RealPtrVec_t data(7), halo(7);
    for (int q = 0; q < 7; ++q) { data[q] = fields + q*n; halo[q] = halo_out + q*nh; }
    switch (face) {
        case 0: apply_one<0,0>(kind, data, halo); break;
        case 1: apply_one<0,1>(kind, data, halo); break;
        case 2: apply_one<1,0>(kind, data, halo); break;
        case 3: apply_one<1,1>(kind, data, halo); break;
        case 4: apply_one<2,0>(kind, data, halo); break;
        default: apply_one<2,1>(kind, data, halo); break;
    }
}
