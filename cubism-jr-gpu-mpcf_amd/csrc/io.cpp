// io.cpp -- dumps and restart files (host C++ over the C-ABI), the
// counterpart of DumpHDF5_MPI / ReadHDF5_MPI (source/IO/HDF5Dumper_MPI.h:30-186)
// and Sim_SteadyStateMPI::_dump/_save/_restart
// (source/Sim/Sim_SteadyStateMPI.cpp:138-179).
//
// A dump is the streamer's dense [Z][Y][X][C] binary32 array (the reference's
// HDF5 dataset "data", HDF5Dumper_MPI.h:118-160) written as raw little-endian
// <name>.raw, with the reference's XDMF wrapper <name>.xmf
// (HDF5Dumper_MPI.h:162-186: 3DCORECTMesh, ORIGIN_DXDYDZ geometry, one
// attribute "data") pointing at it as a Binary DataItem instead of an HDF
// one (HDF5 is not used here).  A checkpoint is the save streamer's dump named
// "save.data" plus save.info holding t (15 significant digits), the step and
// the dump counter, one per line.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "mpcf_hip.h"

namespace mpcf {
namespace detail {
int fail(int code, const std::string& msg);   // mpcf_capi.cpp: sets mpcf_last_error
}
}

namespace {

int fail(const std::string& msg) { return mpcf::detail::fail(MPCF_EINVAL, msg); }

std::string join(const char* dir, const std::string& name)
{
    std::string d = dir && *dir ? dir : ".";
    if (d.back() != '/') d += '/';
    return d + name;
}

const char* attribute_name(const int kind)
{
    // Streamer::getAttributeName (Types.h:214,235,282)
    return kind == MPCF_STREAM_TENSOR ? "Tensor" : kind == MPCF_STREAM_SAVE ? "Save Data" : "Scalar";
}

}  // namespace

extern "C" int mpcf_write_xdmf(const char* xmf_path, const char* raw_name, int nx, int ny, int nz, int channels,
                               double h, int step, const char* attribute)
{
    if (!xmf_path || !raw_name || !attribute || nx < 1 || ny < 1 || nz < 1 || channels < 1 || !(h > 0))
        return fail("mpcf_write_xdmf: bad argument");
    FILE* f = std::fopen(xmf_path, "w");
    if (!f) return fail(std::string("mpcf_write_xdmf: cannot open ") + xmf_path);
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n");
    std::fprintf(f, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
    std::fprintf(f, "<Xdmf Version=\"2.0\">\n");
    std::fprintf(f, " <Domain>\n");
    std::fprintf(f, "   <Grid GridType=\"Uniform\">\n");
    std::fprintf(f, "     <Time Value=\"%05d\"/>\n", step);
    std::fprintf(f, "     <Topology TopologyType=\"3DCORECTMesh\" Dimensions=\"%d %d %d\"/>\n", nz, ny, nx);
    std::fprintf(f, "     <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n");
    std::fprintf(f, "       <DataItem Name=\"Origin\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n");
    std::fprintf(f, "        %e %e %e\n", 0., 0., 0.);
    std::fprintf(f, "       </DataItem>\n");
    std::fprintf(f, "       <DataItem Name=\"Spacing\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n");
    std::fprintf(f, "        %e %e %e\n", h, h, h);
    std::fprintf(f, "       </DataItem>\n");
    std::fprintf(f, "     </Geometry>\n");
    std::fprintf(f, "     <Attribute Name=\"data\" AttributeType=\"%s\" Center=\"Node\">\n", attribute);
    std::fprintf(f, "       <DataItem Dimensions=\"%d %d %d %d\" NumberType=\"Float\" Precision=\"4\" Format=\"Binary\" "
                    "Endian=\"Little\">\n", nz, ny, nx, channels);
    std::fprintf(f, "        %s\n", raw_name);
    std::fprintf(f, "       </DataItem>\n");
    std::fprintf(f, "     </Attribute>\n");
    std::fprintf(f, "   </Grid>\n");
    std::fprintf(f, " </Domain>\n");
    std::fprintf(f, "</Xdmf>\n");
    const bool ok = std::fclose(f) == 0;
    return ok ? MPCF_OK : fail(std::string("mpcf_write_xdmf: write failed: ") + xmf_path);
}

extern "C" int mpcf_write_save_info(const char* path, double t, int step, int fcount)
{
    if (!path) return fail("mpcf_write_save_info: NULL path");
    std::ofstream f(path);
    if (!f.good()) return fail(std::string("mpcf_write_save_info: cannot open ") + path);
    f << std::setprecision(15) << t << "\n" << step << "\n" << fcount << "\n";
    f.close();
    return f.good() ? MPCF_OK : fail(std::string("mpcf_write_save_info: write failed: ") + path);
}

extern "C" int mpcf_read_save_info(const char* path, double* t, int* step, int* fcount)
{
    if (!path || !t || !step || !fcount) return fail("mpcf_read_save_info: NULL argument");
    std::ifstream f(path);
    if (!f.good()) return fail(std::string("mpcf_read_save_info: cannot read ") + path);
    if (!(f >> *t >> *step >> *fcount)) return fail(std::string("mpcf_read_save_info: malformed ") + path);
    return MPCF_OK;
}

extern "C" int mpcf_dump(mpcf_ctx* ctx, int kind, const char* dir, const char* name, int step, double h)
{
    if (!ctx || !name) return fail("mpcf_dump: NULL argument");
    const int nc = mpcf_stream_channels(kind);
    if (!nc) return fail("mpcf_dump: unknown streamer kind");
    long long q[8];
    int rc = mpcf_query(ctx, q);
    if (rc != MPCF_OK) return rc;
    const int nx = (int)q[0], ny = (int)q[1], nz = (int)q[2];
    std::vector<float> buf((size_t)nx * ny * nz * nc);
    rc = mpcf_stream_out(ctx, kind, buf.data());
    if (rc != MPCF_OK) return rc;
    const std::string raw = std::string(name) + ".raw";
    FILE* f = std::fopen(join(dir, raw).c_str(), "wb");
    if (!f) return fail("mpcf_dump: cannot open " + join(dir, raw) + ": " + std::strerror(errno));
    const bool ok = std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) return fail("mpcf_dump: write failed: " + join(dir, raw));
    return mpcf_write_xdmf(join(dir, std::string(name) + ".xmf").c_str(), raw.c_str(), nx, ny, nz, nc, h, step,
                           attribute_name(kind));
}

extern "C" int mpcf_save(mpcf_ctx* ctx, const char* dir, double t, int step, int fcount, double h)
{
    const int rc = mpcf_write_save_info(join(dir, "save.info").c_str(), t, step, fcount);
    if (rc != MPCF_OK) return rc;
    return mpcf_dump(ctx, MPCF_STREAM_SAVE, dir, "save.data", step, h);
}

extern "C" int mpcf_restart(mpcf_ctx* ctx, const char* dir, double* t, int* step, int* fcount)
{
    if (!ctx || !t || !step || !fcount) return fail("mpcf_restart: NULL argument");
    int rc = mpcf_read_save_info(join(dir, "save.info").c_str(), t, step, fcount);
    if (rc != MPCF_OK) return rc;
    long long q[8];
    rc = mpcf_query(ctx, q);
    if (rc != MPCF_OK) return rc;
    const size_t n = (size_t)q[0] * q[1] * q[2] * 7;
    std::vector<float> buf(n);
    const std::string raw = join(dir, "save.data.raw");
    FILE* f = std::fopen(raw.c_str(), "rb");
    if (!f) return fail("mpcf_restart: cannot open " + raw);
    const size_t got = std::fread(buf.data(), sizeof(float), n, f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != n || extra) return fail("mpcf_restart: " + raw + " does not match the block's size");
    return mpcf_stream_in(ctx, buf.data());
}
