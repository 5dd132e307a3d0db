/*
 * mpcf_hip.h -- C-ABI of the MI355X RHS / LSRK3 path for the 7-field
 * stiffened-gas two-phase Euler system (rho, rho u, rho v, rho w, E,
 * G = 1/(gamma-1), P = gamma pc/(gamma-1)).
 *
 * One context owns one node block resident in HBM: the solution u, the
 * low-storage RK register R (the reference's "tmp"), sweep accumulators and a
 * device step state (max speed of sound, dt, dt/h).  Every call returns
 * MPCF_OK or an error code; mpcf_last_error() gives the message of the last
 * failure on the calling thread.  Calls on one context must come from one
 * host thread at a time; contexts are independent.  Work is queued on the
 * context's HIP stream; calls that hand results to the host synchronize it.
 *
 * Host arrays use the reference's storage: seven separate binary32 arrays
 * (NodeBlock::pdata()/ptmp(), source/NodeBlock.h:47-48,110-113), each
 * NX*NY*NZ values indexed ID3(ix,iy,iz,NX,NY) = ix + NX*(iy + NY*iz)
 * (source/Types.h:31), in the order R, U, V, W, E, G, P (NodeBlock.h:26-29).
 */
#ifndef MPCF_HIP_H
#define MPCF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpcf_ctx mpcf_ctx;

/* status codes */
enum {
    MPCF_OK = 0,
    MPCF_EINVAL = 1,   /* bad argument (null pointer, size, enum value)        */
    MPCF_EHIP = 2,     /* a HIP runtime call failed                             */
    MPCF_ENOMEM = 3,   /* device allocation failed                              */
    MPCF_ENUMERIC = 4, /* non-finite or non-positive dt / speed of sound, or a
                          cell that fails the reference's state assertions:
                          NaN / inf, rho <= 0, E <= 0, G <= 0, c <= 0
                          (GPUkernels.cu:1171-1173,1218, MaxSpeedOfSound.cpp:31-48);
                          the kernels flag such cells and finish normally      */
    MPCF_EBUSY = 5     /* a wait timed out with work still outstanding          */
};

/* boundary kinds per face (face = 2*dir + side; dir 0,1,2 = x,y,z; side 0 = low) */
enum {
    MPCF_BC_ABSORBING = 0,   /* BoundaryConditions::applyBC_absorbing  (BoundaryConditions.h:60-82)  */
    MPCF_BC_REFLECTING = 1,  /* BoundaryConditions::applyBC_reflecting (BoundaryConditions.h:85-109) */
    MPCF_BC_HALO = 2         /* a face shared with a neighbouring block: its three ghost layers hold
                                the neighbour's cells, written by mpcf_halo_import before each
                                stage (the MPI halo faces of GPUlab::load_ghosts, GPUlab.cpp:622-644) */
};

/* context flags */
enum {
    MPCF_FLAG_FAST = 1u,     /* FMA-contracted kernels (tolerance parity) instead of the bit-exact build */
    MPCF_FLAG_SOS_CPU = 2u,  /* speed of sound with p recovered in binary64 as MaxSpeedOfSound_CPP
                                (source/MaxSpeedOfSound.cpp:43) instead of the GPU form (GPUkernels.cu:1206) */
    MPCF_FLAG_SPLIT_XY = 4u, /* run x and y as two sweeps ...                                          */
    MPCF_FLAG_FUSED_XY = 8u, /* ... or as one fused x+y sweep (same results either way).  Default: fused,
                                except split with MPCF_FLAG_FP64.                                      */
    MPCF_FLAG_P64 = 16u,     /* selective binary64 on binary32 storage: the FAST arithmetic, with the
                                pressure recovery p = (E - KE - P)/G and the speed-of-sound / dt
                                reduction in binary64 (implies MPCF_FLAG_FAST)                      */
    MPCF_FLAG_NO_PIPELINE = 128u, /* start with mpcf_set_pipeline(ctx, 0): every stage of mpcf_step as
                                whole-block sweeps                                                */
    MPCF_FLAG_E_PAIR = 256u, /* MPCF_FLAG_P64 with the total energy stored as a binary32 pair, E = hi +
                                lo, exact in binary64 (the other fields stay binary32): the pressure
                                recovery, the register update of E and the speed of sound use hi + lo,
                                so p no longer carries the binary32 storage error of E (half an ulp of
                                E ~ gamma pc/(gamma-1) in liquid cells, divided by G).  Implies P64;
                                excludes FP64 and SPLIT_XY.  Transfers, steps, mpcf_compare and
                                mpcf_integrals see E = hi + lo; the ghost band, halo, streamer, dump,
                                checkpoint and wavelet entry points refuse such a context.
                                No reference counterpart (the reference stores E in binary32,
                                NodeBlock.h:47-48).                                                */
    MPCF_FLAG_FP64 = 32u     /* the whole scheme in binary64 (storage, arithmetic, reductions): the
                                high-precision run the binary32 builds are measured against.  Host
                                arrays stay binary32 for mpcf_upload/mpcf_download (rounded on the way
                                out); mpcf_upload_f64/mpcf_download_f64 move binary64 arrays.       */
};

/* which host array set */
enum { MPCF_DATA = 0, MPCF_TMP = 1 };

/* Message of the last failed call on this thread ("" if none). */
const char* mpcf_last_error(void);

/* Library version string. */
const char* mpcf_version(void);

/* Create a context for an NX x NY x NZ block with spacing h on the current
 * HIP device, with all boundaries absorbing.  Replaces the GPUlab constructor
 * and GPU::alloc (source/GPUlab.cpp:25-71, source/GPU/GPUhousehold.cu:81-184);
 * there is no nslices: the whole block stays in HBM.  NX, NY, NZ >= 5
 * (GPUkernels.cu:18-24 asks for >= 5 in x and y). */
int mpcf_create(mpcf_ctx** out, int nx, int ny, int nz, double h, unsigned flags);

/* Release a context (GPUlab::~GPUlab -> GPU::dealloc, GPUlab.h:251). */
int mpcf_destroy(mpcf_ctx* ctx);

/* How mpcf_step runs its stages (same results either way).  With slabs >= 2
 * the block is cut into that many z-slabs and the z sweep of a slab runs on
 * a second stream of the context while the x and y sweeps (fused or split) of
 * the next slab run on the context's stream; the context's stream is ordered
 * after the step's last sweep.  0: whole-block sweeps.  -1 (default; 0 under
 * MPCF_FLAG_NO_PIPELINE): in the FAST builds, for blocks of 2^24 cells or
 * more, NZ/64 slabs for NZ < 320, NZ/100 for NZ < 512 and min(8, NZ/64)
 * above (the fastest counts measured on cubes), whole-block when that is
 * below 2 or the block is smaller; in the bit-exact and binary64 builds and
 * with the split sweeps, whole-block (their sweeps are all VALU-bound).  slabs <= min(64, NZ).  No counterpart in the
 * reference, whose stages run chunk by chunk on the host (GPUlab.cpp:703-793). */
int mpcf_set_pipeline(mpcf_ctx* ctx, int slabs);

/* Use the given hipStream_t (NULL = create and own one).  Default: own stream. */
int mpcf_set_stream(mpcf_ctx* ctx, void* hip_stream);
void* mpcf_get_stream(mpcf_ctx* ctx);

/* Boundary kind per face, bc[6] in face order x-lo, x-hi, y-lo, y-hi, z-lo, z-hi
 * (the per-face choice the reference makes in each GPUlab<Case>::_apply_bc,
 * e.g. source/Sim/Sim_SodMPI.h:30-39). */
int mpcf_set_bc(mpcf_ctx* ctx, const int bc[6]);

/* Copy seven host arrays into the block (which = MPCF_DATA: u, MPCF_TMP: R),
 * or back out.  Synchronous.  (The reference streams these over PCIe every
 * stage, GPU::h2d_3DArray / h2d_tmp / d2h_*, GPU.h:52-57.) */
int mpcf_upload(mpcf_ctx* ctx, int which, const float* const fields[7]);
int mpcf_download(mpcf_ctx* ctx, int which, float* const fields[7]);

/* The same transfers with binary64 host arrays (exact for an MPCF_FLAG_FP64
 * context; rounded to nearest on upload / exact on download otherwise).
 * No reference counterpart: the reference stores binary32 only (NodeBlock.h:26). */
int mpcf_upload_f64(mpcf_ctx* ctx, int which, const double* const fields[7]);
int mpcf_download_f64(mpcf_ctx* ctx, int which, double* const fields[7]);

/* Fill the three-cell ghost band of u from the boundary conditions
 * (GPUlab::load_ghosts -> _apply_bc, source/GPUlab.cpp:622-644); halo faces
 * keep what mpcf_halo_import wrote. */
int mpcf_load_ghosts(mpcf_ctx* ctx);

/* Face halos of a block that is one part of a decomposed domain: the device
 * side of GPUlab's MPI halo exchange (GPUlab::_copysend_halos,
 * GPUlab.cpp:76-112, and the receives of GPUlab::load_ghosts, :636-641).
 * face = 2*dir + side.  A halo buffer holds mpcf_halo_count() elements of the
 * context's storage type (binary32; binary64 for MPCF_FLAG_FP64): the seven
 * fields one after the other, each in the reference's ghost map order
 * (Types.h:31-69) -- x faces [NZ][NY][3], y faces [NZ][3][NX], z faces
 * [3][NY][NX], layer l = the l-th cell in increasing index.
 *   mpcf_halo_export: the three boundary cells of the current u (0..2 on the
 *     low face, n-3..n-1 on the high face) into dst, the neighbour's import;
 *   mpcf_halo_import: src into the ghost band of the current u (-3..-1 low,
 *     n..n+2 high), read by the next stage on faces set to MPCF_BC_HALO.
 * dst/src are device pointers on the context's device; both calls are
 * asynchronous on the context's stream (order them against another
 * context's stream before that context reads the buffer).  A stage needs
 * fresh halos, so a block with halo faces steps with mpcf_max_sos (the
 * maximum over all blocks), dt = min(cfl h / sos, dt_max) on the host, and per
 * stage: export, exchange, import, mpcf_process_all(a, b, (float)(dt/h));
 * mpcf_step refuses halo faces. */
int mpcf_halo_count(mpcf_ctx* ctx, int face, long long* count);
int mpcf_halo_export(mpcf_ctx* ctx, int face, void* dst);
int mpcf_halo_import(mpcf_ctx* ctx, int face, const void* src);

/* The ghost band of `face` (the three ghost layers mpcf_load_ghosts or
 * mpcf_halo_import wrote) into the device buffer dst, in the halo buffer order
 * above: the values GPUlab::load_ghosts leaves in its halo buffers
 * (GPUlab.cpp:622-644, Types.h:31-69).  Asynchronous on the context's stream. */
int mpcf_ghosts_read(mpcf_ctx* ctx, int face, void* dst);

/* Max over the block of c + max(|rho u|,|rho v|,|rho w|)/rho
 * (GPUlab::max_sos / MaxSpeedOfSound_CUDA::compute, GPUlab.cpp:647-700,
 * MaxSpeedOfSound_CUDA.cpp:13-18).  Synchronous.  MPCF_ENUMERIC (sos still
 * written) when a cell is non-finite or non-physical. */
int mpcf_max_sos(mpcf_ctx* ctx, float* sos);

/* R <- a R - dtinvh (div F + quasi-conservative G/P terms), from the current u
 * with its ghosts (Convection_CUDA(a, dtinvh).compute(), Convection_CUDA.cpp:15-23).
 * Asynchronous. */
int mpcf_convection(mpcf_ctx* ctx, float a, float dtinvh);

/* u <- b R + u (Update_CUDA(b).compute(), Update_CUDA.cpp:12-17).  Asynchronous. */
int mpcf_update(mpcf_ctx* ctx, float b);

/* One full stage: ghosts must be loaded; R <- a R - dtinvh L(u), then
 * u <- u + b R, all reads of u at the old time (GPUlab::process_all,
 * GPUlab.cpp:703-793).  Same results as mpcf_convection + mpcf_update; the
 * updated u goes to a second resident buffer.  Asynchronous. */
int mpcf_process_all(mpcf_ctx* ctx, float a, float b, float dtinvh);

/* One LSRK3 step (LSRK3_IntegratorMPI::operator(), LSRK3_IntegratorMPI.cpp:27-62):
 * max speed of sound, dt = min(cfl h / sos, dt_max) in binary64, then three
 * stages (a,b) = (0,1/4), (-17/32,8/9), (-32/27,3/4) with ghosts refilled
 * before each.  dt never leaves the device between the reduction and the
 * stages.  Writes dt (synchronizes) unless dt is NULL (asynchronous). */
int mpcf_step(mpcf_ctx* ctx, double cfl, double dt_max, double* dt);

/* nsteps LSRK3 steps, queued asynchronously: the time loop of the reference's
 * drivers (Sim_*::run calling LSRK3_IntegratorMPI::operator() once per step,
 * Sim_SteadyStateMPI.cpp:182-222) without a host round trip per step.
 * Results equal nsteps calls of mpcf_step bit for bit.  In the FAST builds
 * consecutive pipelined steps overlap: the first stage's x+y sweeps of a step
 * need no dt, so they start on each z-slab as soon as the previous step's last
 * z sweep of that slab is done, while that step's remaining slabs and the dt
 * reduction finish.  When the call returns, the context's stream is ordered
 * after every step (as after mpcf_step).  dt / sos of the last step:
 * mpcf_last_dt. */
int mpcf_steps(mpcf_ctx* ctx, int nsteps, double cfl, double dt_max);

/* dt and max speed of sound of the last mpcf_step (synchronizes).
 * MPCF_ENUMERIC (dt and sos still written) when the state the step started
 * from, or the state it wrote, has a non-finite or non-physical cell; with a
 * dt pointer, mpcf_step returns the same. */
int mpcf_last_dt(mpcf_ctx* ctx, double* dt, float* sos);

/* Max speed of sound of the last mpcf_step before rounding to binary32: the
 * binary64 reduction of MPCF_FLAG_P64 / MPCF_FLAG_FP64 contexts (the binary32
 * value widened otherwise).  Synchronizes. */
int mpcf_last_sos64(mpcf_ctx* ctx, double* sos);

/* Allocation-free speed of sound and dt.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * The reduction of mpcf_max_sos, and the dt that mpcf_step takes from it
 * (dt = min(cfl h / sos, dt_max) in binary64), over a state the caller keeps
 * in its own device memory in the host layout: field q of cell (ix,iy,iz) at
 * element q*field_stride + ix + nx*(iy + ny*iz) of u (binary32, or binary64
 * with MPCF_FLAG_FP64).  flags: MPCF_FLAG_FAST, MPCF_FLAG_SOS_CPU,
 * MPCF_FLAG_P64, MPCF_FLAG_FP64, with the meaning they have for a context;
 * the results equal mpcf_max_sos / mpcf_last_sos64 and mpcf_step's dt for
 * the same state and flags, bit for bit.  Outputs are caller device memory
 * (any may be NULL): sos_out[0] the maximum (binary64; the binary32 maximum
 * widened in the binary32 builds), dt_out[0], bad_out[0] nonzero when a cell
 * fails the state assertions of MPCF_ENUMERIC.  Asynchronous on hip_stream
 * (NULL: the null stream) of the current device.
 * It calls no hipMalloc or hipFree: the reduction's working state (256 bytes)
 * is a static device array of the library, one per GPU, located once per
 * device on the first call there.  Each call clears it with a kernel (no
 * memset node), so the call can be graph-captured.  Calls that overlap on one
 * GPU race on that array: order them.  Limits: nx*ny*nz <= 2^31 - 1 cells,
 * field_stride >= nx*ny*nz; larger states are refused with MPCF_EINVAL. */
int mpcf_max_sos_noalloc(const void* u, long long field_stride, int nx, int ny, int nz, double h, double cfl,
                         double dt_max, unsigned flags, double* sos_out, double* dt_out, int* bad_out,
                         void* hip_stream);

/* Field-wise difference of two blocks of equal extent on one device, in
 * binary64 (any storage precision on either side): out[4q + k] for field q
 * is k=0: sum |a-b|, 1: max |a-b|, 2: sum |b|, 3: max |b| over the interior
 * cells.  Reproducible (fixed reduction order).  Synchronizes both.  No
 * reference counterpart: it measures a binary32 run against a binary64 one. */
int mpcf_compare(mpcf_ctx* a, mpcf_ctx* b, int which, double out[28]);

/* Conserved integrals: out[q] = h^3 * sum over interior cells of field q
 * (mass, momenta, total energy, and the G / P sums), accumulated in binary64
 * in a fixed order.  Synchronizes.  The reference computes no integrals. */
int mpcf_integrals(mpcf_ctx* ctx, int which, double out[7]);

/* Block until the context's queued work is done. */
int mpcf_synchronize(mpcf_ctx* ctx);

/* Sizes: out[0..2] = NX, NY, NZ; out[3..5] = padded extents; out[6] = field
 * stride in elements (low 31 bits); out[7] = device bytes held by the context. */
int mpcf_query(mpcf_ctx* ctx, long long out[8]);

/* Per-kernel timing.  While enabled, every kernel the context launches is
 * bracketed by HIP events on the context stream (the reference's
 * tCUDA_START/tCUDA_STOP timers, source/GPU/GPU.h:25-35).  mpcf_timing_read
 * synchronizes and returns, per kernel class, the summed milliseconds and the
 * number of launches since the last read:
 *   0 ghost fill, 1 max speed of sound, 2 dt, 3 x sweep (or the fused x+y
 *   sweep), 4 y sweep (split variant only), 5 z sweep (+ register update),
 *   6 update, 7 other. */
int mpcf_timing_enable(mpcf_ctx* ctx, int enable);
int mpcf_timing_read(mpcf_ctx* ctx, double ms[8], long long count[8]);

/* Initial conditions (host arrays, seven fields as above; h = 1/max(NX,NY,NZ),
 * cell centres h (i + 1/2)).  Material states blended like set_IC of
 * source/Sim/Sim_SICCloudMPI.cpp:249-276 with tanh-smoothed indicators;
 * air gamma 1.4, pc 0; water gamma 6.59, pc 4096.
 *   shock_tube:      air (rho 1, p 1) | water (rho 1000, p 0.1) across 0.5 on
 *                    `axis`, tanh over 2 cells, rho noise +-`noise` relative
 *                    (the Sod-type tube of source/Sim/Sim_SodMPI.cpp:19-95).
 *   shock_bubble:    air bubble r 0.1 at the centre in water (p 1), post-shock
 *                    water (p 1e3, u 10) for x < 0.3, tanh over 1.5 cells,
 *                    p noise +-1e-4 (source/Sim/Sim_2DSBIMPI.cpp:86-233, in 3D).
 *   bubble_cloud:    `nbubbles` air bubbles (p 1), radii lognormal(ln 0.02, 0.3),
 *                    centres uniform in a ball of radius 0.3 about the centre,
 *                    overlaps rejected, in water at p 100; bubbles (optional)
 *                    receives n x 4 doubles x,y,z,r (Sim_SICCloudMPI.cpp:43-357).
 *   bubble_collapse: one air bubble (p 1) of the given radius and centre in
 *                    water at p = pratio. */
int mpcf_ic_shock_tube(int nx, int ny, int nz, int axis, double noise, unsigned long long seed,
                       float* const fields[7]);
int mpcf_ic_shock_bubble(int nx, int ny, int nz, unsigned long long seed, float* const fields[7]);
int mpcf_ic_bubble_cloud(int nx, int ny, int nz, int nbubbles, unsigned long long seed, double* bubbles,
                         float* const fields[7]);
int mpcf_ic_bubble_collapse(int nx, int ny, int nz, double radius, double pratio, double cx, double cy,
                            double cz, float* const fields[7]);

/* ---------------------------------------------------------------------------
 * Batched ensembles.  No counterpart in the reference (which advances one
 * node block per process); this is the batching part of the MI355X build.
 * A batch holds independent cases of uneven sizes (each a full resident block
 * with its own dt).  mpcf_batch_step advances every live case by one LSRK3
 * step with one launch per kernel kind per stage (the cases' blocks share
 * each grid), entirely on the device; with mpcf_batch_use_graph the step is
 * captured once as a hipGraph and replayed (re-captured when the case set,
 * the march lengths, cfl or dt_max change).  Cases may be added and removed
 * between steps (continuous batching).  Each case gets the same arithmetic
 * as a context of its own, so its results are identical alone and in a batch.
 * --------------------------------------------------------------------------- */
typedef struct mpcf_batch mpcf_batch;

/* HBM bytes one case of this size holds (admission by footprint). */
int mpcf_case_bytes(int nx, int ny, int nz, long long* bytes);

int mpcf_batch_create(mpcf_batch** out, unsigned flags);   /* MPCF_FLAG_FAST / _SOS_CPU */
int mpcf_batch_destroy(mpcf_batch* batch);
/* Add a case (u uploaded, R zeroed; bc may be NULL = all absorbing); returns its id. */
int mpcf_batch_add(mpcf_batch* batch, int nx, int ny, int nz, double h, const int bc[6],
                   const float* const fields[7], int* case_id);
int mpcf_batch_remove(mpcf_batch* batch, int case_id);
int mpcf_batch_download(mpcf_batch* batch, int case_id, int which, float* const fields[7]);
/* One LSRK3 step of every live case (asynchronous). */
int mpcf_batch_step(mpcf_batch* batch, double cfl, double dt_max);
int mpcf_batch_use_graph(mpcf_batch* batch, int enable);
/* dt and max speed of sound of a case's last step (synchronizes). */
int mpcf_batch_last_dt(mpcf_batch* batch, int case_id, double* dt, float* sos);
int mpcf_batch_synchronize(mpcf_batch* batch);
void* mpcf_batch_get_stream(mpcf_batch* batch);
/* Force the march lengths of one case (0 = tuned / default). */
int mpcf_batch_set_march(mpcf_batch* batch, int case_id, int len_xy, int len_z);

/* How mpcf_batch_step launches a step (same results either way):
 *   MPCF_BATCH_MODE_PERSISTENT: the speed of sound / dt of every case, then
 *     per RK stage two persistent launches, one over the x+y sweep items of
 *     all cases and one over their z sweep items; each workgroup pulls
 *     (case, tile) items from a device queue -- one workgroup's share of a
 *     case's x+y sweep (threads/64 consecutive 58-cell x-tiles over one
 *     y-chunk) or z sweep (`threads` columns over one z-chunk) -- largest case
 *     first, until the queue is empty.  threads: workgroup size, 64, 128 or
 *     256 (0: 256).
 *   MPCF_BATCH_MODE_GROUPS (default): the cases in four groups of about equal
 *     cells, each on its own stream, with one flat launch per kernel kind per
 *     stage. */
enum { MPCF_BATCH_MODE_GROUPS = 0, MPCF_BATCH_MODE_PERSISTENT = 1 };
int mpcf_batch_set_mode(mpcf_batch* batch, int mode, int threads);

/* Autotune the batch on its live cases: the launch mode and the persistent
 * launches' workgroup size, then for each case extent (largest first) the
 * x+y sweep's march length (by NY) and the z sweep's (by NZ) among 8, 16,
 * 32, 64 and the whole extent, each by coordinate descent.  Every candidate
 * is timed over `steps` steps after a warm-up step, from a saved state that
 * is restored afterwards (the cases do not advance).  The choice is kept per
 * extent for the cases added later.  best_len (may be NULL): the x+y march
 * length chosen for the largest extent. */
int mpcf_batch_autotune(mpcf_batch* batch, int steps, double cfl, int* best_len);

/* The tuning in force: out[0] mode, [1] persistent workgroup size, [2] x+y
 * and [3] z march length of cases of extent `extent` (0: no tuned value, the
 * size-based default). */
int mpcf_batch_tuning(mpcf_batch* batch, int extent, int out[4]);

/* out[0] live cases, [1] interior cells, [2] device bytes, [3] graph captures,
 * [4] steps since the last autotune, [5] tuned extents. */
int mpcf_batch_query(mpcf_batch* batch, long long out[6]);

/* ------------------------------------------------------------------------
 * Dumps and checkpoints (source/IO/HDF5Dumper_MPI.h:30-186, Types.h:188-280,
 * Sim_SteadyStateMPI.cpp:138-179).  The streamers turn the resident state
 * into the reference's dense channel-interleaved [Z][Y][X][C] binary32 array;
 * dumps are written as raw data with the reference's XDMF wrapper (a Binary
 * DataItem in place of the HDF5 dataset).
 * ------------------------------------------------------------------------ */
enum {
    MPCF_STREAM_TENSOR = 0,  /* myTensorialStreamer: rho, u, v, w, p, G, P, 0, 0 (9 channels) */
    MPCF_STREAM_SAVE = 1,    /* mySaveStreamer: the 7 conserved fields                       */
    MPCF_STREAM_SCALAR = 2   /* myScalarStreamer: rho                                        */
};

/* Channels of a streamer kind (0 for an unknown kind). */
int mpcf_stream_channels(int kind);

/* The current solution through a streamer into out[NX*NY*NZ*C] (synchronizes). */
int mpcf_stream_out(mpcf_ctx* ctx, int kind, float* out);

/* A save-streamer array (7 channels) into the current solution (synchronizes). */
int mpcf_stream_in(mpcf_ctx* ctx, const float* in);

/* The reference's XDMF wrapper for a [Z][Y][X][C] binary32 raw file. */
int mpcf_write_xdmf(const char* xmf_path, const char* raw_name, int nx, int ny, int nz, int channels, double h,
                    int step, const char* attribute);

/* save.info: t (15 significant digits), step and dump counter, one per line. */
int mpcf_write_save_info(const char* path, double t, int step, int fcount);
int mpcf_read_save_info(const char* path, double* t, int* step, int* fcount);

/* <dir>/<name>.raw + <dir>/<name>.xmf of the current solution (Sim_SteadyStateMPI::_dump). */
int mpcf_dump(mpcf_ctx* ctx, int kind, const char* dir, const char* name, int step, double h);

/* <dir>/save.info + <dir>/save.data.{raw,xmf} (Sim_SteadyStateMPI::_save). */
int mpcf_save(mpcf_ctx* ctx, const char* dir, double t, int step, int fcount, double h);

/* Read <dir>/save.info and save.data.raw into the context (Sim_SteadyStateMPI::_restart). */
int mpcf_restart(mpcf_ctx* ctx, const char* dir, double* t, int* step, int* fcount);

/* ------------------------------------------------------------------------
 * Shock-induced cloud collapse input (Sim_SICCloudMPI, source/Sim/
 * Sim_SICCloudMPI.cpp:43-357): parameters with the reference's defaults and
 * command-line meaning, the post-shock state, the cloud files, and the
 * 27-point Simpson initial condition.  Host-side, binary64 arithmetic.
 * ------------------------------------------------------------------------ */
typedef struct {
    double rho0, u0, p0;        /* pre-shock liquid (-rho0 -u0 -p0)                        */
    double rhoB, uB, pB;        /* gas in the bubbles (-rhoB -uB -pB)                      */
    double pressure_ratio;      /* p1/p0 across the shock (-pressureratio)                 */
    double shock_n[3];          /* shock normal, normalized by the callee (-shockNx/Ny/Nz) */
    double shock_s[3];          /* a point on the shock plane (-shockSx/Sy/Sz)             */
    double g1, g2, pc1, pc2;    /* liquid / gas stiffened-gas gamma and pc (-g1 -g2 -pc1 -pc2) */
    double eps;                 /* bubble mollifier half-width (SimTools::EPSILON; the
                                   reference's cloud case leaves it 0: sharp interfaces)   */
} mpcf_sic_params;

/* Fill p with the reference's defaults (Sim_SICCloudMPI.cpp:53-77). */
void mpcf_sic_defaults(mpcf_sic_params* p);

/* Post-shock liquid state (Sim_SICCloudMPI.cpp:87-104): out = rho1, u1, p1,
 * shock Mach number, pre-shock speed of sound c0. */
int mpcf_sic_post_shock(const mpcf_sic_params* p, double out[5]);

/* cloud_config.dat (Sim_SICCloudMPI.cpp:134-160): out = n_shapes, n_small,
 * min_rad, max_rad, seed_s[3], seed_e[3], n_sensors (0 when absent). */
int mpcf_read_cloud_config(const char* path, double out[11]);

/* cloud.dat (Sim_SICCloudMPI.h:148-181), lines "idx x y z r": writes up to
 * `cap` spheres as (x, y, z, r) into shapes[4*cap], the count read into *n. */
int mpcf_read_cloud(const char* path, double* shapes, int cap, int* n);

/* The initial condition of _ic_quad (Sim_SICCloudMPI.cpp:306-357) on an
 * NX x NY x NZ block with spacing h (cell centres at h(i+1/2)): every cell is
 * the 27-point Simpson average of set_IC over the cell.  shapes: nshapes
 * spheres (x, y, z, r).  fields: seven NX*NY*NZ binary32 arrays. */
int mpcf_ic_sic_cloud(int nx, int ny, int nz, double h, const mpcf_sic_params* p, const double* shapes, int nshapes,
                      float* const fields[7]);

/* ------------------------------------------------------------------------
 * The other drivers' initial conditions (host-side, binary64 arithmetic,
 * binary32 stores; cell centres at h(i+1/2); fields: seven NX*NY*NZ
 * binary32 arrays, x fastest).
 * ------------------------------------------------------------------------ */

/* SimTools::getPostShockRatio (source/Types.h:101-109): the state behind a
 * shock of Mach number `mach` running into pre = (rho, u, p) of a stiffened
 * gas (gamma, pc); post = (rho, u, p).  The density and pressure ratios are
 * the ideal-gas ones (pc enters only the speeds of sound).  real32 != 0 rounds
 * like the reference's default build (Real = binary32: inputs, Mach-number
 * square, speeds of sound and results in binary32), as mpcf_ic_sbi does. */
void mpcf_post_shock_ratio(const double pre[3], double mach, double gamma, double pc, int real32, double post[3]);

/* Sim_SodMPI (source/Sim/Sim_SodMPI.cpp:19-95): two states split at x0 along
 * one axis.  Like the reference, u1/u2 are velocities, the stored momentum is
 * rho*u along that axis and E = G p + P + (rho u)^2 / (2 rho). */
typedef struct {
    double x0;                  /* interface position (-x0)                               */
    double rho1, u1, p1;        /* state at pos < x0 (-rho1 -u1 -p1)                       */
    double rho2, u2, p2;        /* state at pos > x0 (-rho2 -u2 -p2)                       */
    double g1, g2, pc1, pc2;    /* stiffened-gas gamma and pc of the two states            */
    int reverse;                /* -reverse: state 1 at pos > x0                           */
    int dir;                    /* tube axis, 0/1/2 for -config x/y/z                      */
} mpcf_sod_params;

/* Fill p with the reference's defaults (Sim_SodMPI.cpp:27-40). */
void mpcf_sod_defaults(mpcf_sod_params* p);

int mpcf_ic_sod(int nx, int ny, int nz, double h, const mpcf_sod_params* p, float* const fields[7]);

/* Sim_2DSBIMPI (source/Sim/Sim_2DSBIMPI.cpp:86-233): a planar shock at x0
 * along dims[0] (post-shock state from mpcf_post_shock_ratio with g1, pc1)
 * running into gas 1 holding cylindrical bubbles of gas 2 (axes along
 * dims[2]), mollified over mollfactor*sqrt(3)*h. */
typedef struct {
    double x0, mach;            /* shock position and Mach number (-x0 -mach)             */
    double rho, u, p;           /* pre-shock gas (-rho -u -p)                              */
    double rhobubble, pbubble;  /* bubble gas (-rhobubble -pbubble)                        */
    double g1, g2, pc1, pc2;    /* stiffened-gas gamma and pc of the gas and the bubbles   */
    int mollfactor;             /* -mollfactor                                             */
    int reverse;                /* -reverse: point-mirror the block, momenta negated       */
    int dims[3];                /* principal, minor, dummy axis (mpcf_sbi_config)          */
} mpcf_sbi_params;

/* Fill p with the reference's defaults (Sim_2DSBIMPI.cpp:95-108), -config xy. */
void mpcf_sbi_defaults(mpcf_sbi_params* p);

/* dims for -config xy|xz|yx|yz|zx|zy (Sim_2DSBIMPI.cpp:20-63); MPCF_EINVAL
 * for anything else. */
int mpcf_sbi_config(const char* conf, int dims[3]);

/* <config>bubbles.dat (Sim_2DSBIMPI.cpp:126-155): the count, then "x y z r"
 * per bubble; writes up to `cap` bubbles into bubbles[4*cap], the count read
 * into *n. */
int mpcf_read_bubbles(const char* path, double* bubbles, int cap, int* n);

int mpcf_ic_sbi(int nx, int ny, int nz, double h, const mpcf_sbi_params* p, const double* bubbles, int nb,
                float* const fields[7]);

/* Sim_SteadyStateMPI (source/Sim/Sim_SteadyStateMPI.cpp:95-134): one
 * uniform state, state = rho, u, v, w, p, gamma, pc (reference defaults
 * 1, 0, 0, 0, 1, 1.4, 0). */
int mpcf_ic_steady(int nx, int ny, int nz, const double state[7], float* const fields[7]);

/* Sim_StaticIC (source/Sim/Sim_StaticIC.cpp:76-97): every field of a cell
 * holds the cell's running index in the dims[0]-fastest order (a layout
 * probe for dumps), rounded to binary32. */
int mpcf_ic_static(int nx, int ny, int nz, const int dims[3], float* const fields[7]);

/* ------------------------------------------------------------------------
 * Wavelet-compressed dumps (SerializerIO_WaveletCompression_MPI_Simple,
 * source/IO/SerializerIO_WaveletCompression_MPI_Simple.h:26-808, over
 * source/WaveletCompression/): 4th-order interpolating wavelets on the whole
 * block, survivors |c| > threshold as a bitset plus binary32 (or the
 * reference's truncating binary16) values, zlib-encoded, in the reference's
 * file layout.  The block must be a cube of power-of-two edge 8..1024.
 * ------------------------------------------------------------------------ */

/* Write channel 0..6 of StreamerGridPointIterative (rho, u, v, w, p, G, P;
 * Types.h:157-186) of the current state to "<path>.StreamerGridPointIterative.
 * channel<k>" (SerializerIO_...::Write<channel>, :785-794).  stats (may be
 * NULL): stream bytes before zlib, file bytes, and the reference's
 * "compr. rate" (binary32 bytes of the channel / ocean bytes). */
int mpcf_wavelet_dump(mpcf_ctx* ctx, int channel, const char* path, float threshold, int halffloat,
                      double stats[3]);

/* Read such a file back (decompress and inverse transform): *n = the block
 * edge; cube (n^3 binary32, x fastest) is filled when cap >= n^3, else only
 * *n is set. */
int mpcf_wavelet_read(const char* file, float* cube, long long cap, int* n);

/* The compressor on a host cube of edge n (WaveletCompressorGeneric::compress
 * / decompress, WaveletCompressor.cpp:117-180): the stream handed to the
 * encoder, bitset bytes then survivors; *nbytes = its size, copied to out up
 * to cap bytes.  Decode reconstructs the cube from such a stream. */
int mpcf_wavelet_encode(const float* cube, int n, float threshold, int halffloat, unsigned char* out,
                        long long cap, long long* nbytes);
int mpcf_wavelet_decode(const unsigned char* in, long long nbytes, int n, int halffloat, float* cube);

/* Operation count of the scheme as implemented here, from a host run of the
 * same per-cell code with a counting scalar (binary32/64 add, sub, mul, div,
 * sqrt = 1 FLOP each; min/max and conversions listed apart):
 *   out[0..2] FLOPs per cell of the x, y, z sweep; out[3] per cell-stage;
 *   out[4] max speed of sound per cell; out[5] per cell-step (3 stages + SOS);
 *   out[6..14] per cell-stage: fp32 add/sub, mul, div, sqrt, fp64 add, mul,
 *   div, min/max, conversions.  flags: MPCF_FLAG_FAST selects that build. */
int mpcf_flop_count(unsigned flags, double out[16]);

/* The per-cell scheme compiled for the host in binary32 (same arguments as
 * the device hooks below); lets the restructured scheme be checked against a
 * CPU reference without a GPU. */
int mpcf_host_weno(int n, const float* in5, float* minus, float* plus, unsigned flags);
int mpcf_host_face(int n, int dir, const float* st, float* out, unsigned flags);

/* Device-function test hooks (the counterpart of GPU::TestKernel,
 * source/GPU/GPUkernels.cu:1354-1382).  Host arrays in and out; synchronous.
 *   mpcf_test_weno: in5 = n x 5 cell values (j-2..j+2); minus[i], plus[i] =
 *     _weno_minus_clipped / _weno_pluss_clipped of that window
 *     (GPUonly.cuh:177-194).
 *   mpcf_test_face: st = n x 7 x 6 conserved values at the stencil points
 *     i-3..i+2 of face i-1/2 in direction dir; out = n x 12: the seven fluxes,
 *     the HLLC face velocity, Gm, Gp, Pm, Pp (GPUkernels.cu:203-287). */
int mpcf_test_weno(int n, const float* in5, float* minus, float* plus, unsigned flags);
int mpcf_test_face(int n, int dir, const float* st, float* out, unsigned flags);

/* ------------------------------------------------------------------------
 * Serving: a stream of independent requests advanced with continuous
 * batching by an engine thread (cubism-jr-gpu-mpcf_amd/csrc/mpcf_serve.cpp).
 * No reference counterpart -- the reference runs one case per process
 * (source/main.cpp) -- this is the request loop of the serving workload.
 *
 * A request is one block (binary32 host fields in, results out), its
 * boundary kinds and a number of LSRK3 steps at the server's CFL.  Requests
 * are admitted in arrival order while their footprint (mpcf_case_bytes) fits
 * the HBM budget, which the server allocates once; each engine step advances
 * an active subset of the admitted requests with one launch per kernel kind
 * per stage (replayed as a hipGraph while the subset is unchanged).  Each
 * request's result equals mpcf_step run alone on it for the same number of
 * steps, bit for bit, under every policy.
 * ------------------------------------------------------------------------ */
typedef struct mpcf_server mpcf_server;

/* active-subset policies */
enum {
    MPCF_SERVE_ALL = 0,        /* every admitted request steps every time                    */
    MPCF_SERVE_FIFO_FILL = 1,  /* oldest admitted first, added while the cells fit fill_cells */
    MPCF_SERVE_SRPT_FILL = 2,  /* least remaining work (cells x steps) first, same fill       */
    MPCF_SERVE_HYBRID = 3,     /* the oldest request, then least remaining work, same fill    */
    MPCF_SERVE_FIFO_SLACK = 4  /* oldest first; budget max(fill_cells, 1.25 x the oldest's
                                  cells): the oldest steps at close to its solo speed.  With
                                  two live requests or fewer, the oldest steps alone unless
                                  the other's remaining work is at most a tenth of its own */
};

/* server flag: launch each step eagerly instead of replaying a hipGraph.
 * Without it, the batch steps are replayed from a graph; a request of 2^18
 * cells or more that steps alone runs the whole-block kernels of mpcf_step,
 * enqueued eagerly either way. */
enum { MPCF_FLAG_SERVE_EAGER = 64u };

/* Create a server on the current device: `hbm_bytes` of device memory for
 * request state, `max_live` admitted requests at most, `fill_cells` the
 * active-subset cell budget of the *_FILL policies (<= 0: unbounded), `cfl`
 * for every request.  flags: MPCF_FLAG_FAST, MPCF_FLAG_SOS_CPU,
 * MPCF_FLAG_P64, MPCF_FLAG_SERVE_EAGER. */
int mpcf_server_create(mpcf_server** out, unsigned flags, long long hbm_bytes, int policy,
                       long long fill_cells, int max_live, double cfl);

/* Queue a request (thread-safe, returns at once).  `in` and `out` (seven
 * NX*NY*NZ binary32 fields each) must stay valid until the request is
 * reported done.  bc may be NULL (all absorbing). */
int mpcf_server_submit(mpcf_server* s, int nx, int ny, int nz, double h, const int bc[6], int steps,
                       const float* const in[7], float* const out[7], long long* id);

/* Up to `cap` completed requests since the last poll: ids and completion
 * times on the mpcf_server_clock() clock (the result is in `out` by then). */
int mpcf_server_poll(mpcf_server* s, long long* ids, double* t_done, int cap, int* n);

/* Block until every submitted request is done (timeout_s < 0: no limit;
 * MPCF_EBUSY on timeout). */
int mpcf_server_wait(mpcf_server* s, double timeout_s);

/* out: submitted, completed, engine steps, graph captures, active-set
 * changes, most requests stepped together, peak bytes in use, cells of the
 * current active set. */
int mpcf_server_stats(mpcf_server* s, long long out[8]);

/* Seconds on the steady clock the server timestamps completions with. */
double mpcf_server_clock(void);

/* Finish the outstanding requests, then free everything. */
int mpcf_server_destroy(mpcf_server* s);

#ifdef __cplusplus
}
#endif
#endif /* MPCF_HIP_H */
