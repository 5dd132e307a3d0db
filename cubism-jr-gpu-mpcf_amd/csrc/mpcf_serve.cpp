// mpcf_serve.cpp -- serving engine for streams of independent cases
// (mpcf_server_*, include/mpcf_hip.h).
//
// Requests (a block, its boundary kinds and a step count) are submitted from
// any host thread and advanced by an engine thread with continuous batching:
//
//   submit -> waiting -> [staged in] -> live (on the device) -> finishing -> done
//
// * Admission: in arrival order, while the request's footprint fits the HBM
//   budget.  Device memory comes from one allocation managed first-fit on the
//   host (no hipMalloc / hipFree per request, so admission never synchronizes
//   the device), and is released in stream order as soon as a finished
//   request's download has been enqueued.
// * Host copies go through pinned staging buffers and a copier thread: the
//   caller's input is copied into pinned memory while the GPU keeps running
//   the live cases, then moved with an asynchronous copy; results come back
//   the same way.  The engine thread itself only enqueues GPU work.
// * Every engine step advances an active subset of the live cases with one
//   launch per kernel kind per stage (the batch kernels of mpcf_kernels.hip),
//   captured as a hipGraph while the subset is unchanged.  Which cases are
//   active is the policy: all live cases; or, filled up to a cell budget, the
//   oldest first, those with the least remaining work first, or the oldest
//   one and then the least remaining work; or the oldest first with the
//   budget raised to 1.25x the oldest request's cells.  A request of 2^18
//   cells or more that steps alone runs the whole-block kernels of
//   mpcf_step instead, enqueued eagerly.
// * Each case's current-buffer index is mirrored on the host (a step flips it
//   once), so membership changes need no read-back.  At most `depth` steps are
//   in flight; the engine waits on the oldest beyond that.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mpcf_internal.h"

namespace {

double clock_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

// first-fit allocator over one device allocation
struct Arena
{
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   // offset -> length, coalesced

    bool alloc(size_t n, size_t* off)
    {
        n = round256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t len = it->second;
            free_.erase(it);
            if (len > n) free_[*off + n] = len - n;
            used += n;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = round256(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Staging
{
    float* p = nullptr;
    size_t bytes = 0;
    bool busy = false;
};

enum State { S_WAITING, S_STAGING_IN, S_LIVE, S_FINISHING, S_STAGING_OUT };

struct Req
{
    long long id = 0;
    int nx = 0, ny = 0, nz = 0;
    double h = 0;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    int steps = 0;
    const float* in[7] = {};
    float* out[7] = {};
    double t_submit = 0;
    // engine state
    State state = S_WAITING;
    size_t off = 0, bytes = 0;
    int stage_buf = -1;
    std::atomic<bool> copied{false};   // set by the copier when its job is done
    hipEvent_t ev = nullptr;           // device-side milestone (upload done / download done)
    mpcf_ctx* c = nullptr;
    int left = 0;
    long long seq = 0;                 // admission order
    // a step has been enqueued: its last z sweep folded the maximum speed of
    // sound of the state it wrote into the step state's next_sos_bits
    bool stepped = false;
    // the step state's device flag `folded` is set (the batched speed-of-sound
    // launch then takes dt from the folded maximum instead of reading u)
    bool dev_folded = false;
    long long cells() const { return (long long)nx * ny * nz; }
};

struct CopyJob
{
    Req* r;
    bool to_pinned;   // user input -> pinned (stage in) or pinned -> user output (stage out)
    float* pinned;
};

}  // namespace

struct mpcf_server
{
    unsigned flags = 0;
    const Launchers* L = nullptr;
    int device = 0;
    int policy = 0;
    long long fill_cells = 0;
    int max_live = 0;
    int depth = 4;
    bool use_graph = true;
    double cfl = 0.3;
    hipStream_t stream = nullptr;
    Arena arena;

    std::mutex mu;
    std::condition_variable cv_engine, cv_copier, cv_done;
    bool stop = false;
    std::deque<Req*> submitted;                        // guarded by mu
    std::deque<CopyJob> copy_jobs;                     // guarded by mu
    std::deque<std::pair<long long, double>> done;     // guarded by mu
    long long n_submitted = 0, n_completed = 0;        // guarded by mu
    int error = MPCF_OK;                               // guarded by mu
    std::string error_msg;

    // engine-owned
    std::vector<Req*> reqs;                            // every request not yet completed
    std::vector<Staging> staging;
    std::deque<hipEvent_t> inflight;
    std::vector<hipEvent_t> ev_free;
    long long next_seq = 0;
    mpcf::CaseDesc* d_cases = nullptr;
    int* d_cur = nullptr;
    int* d_off = nullptr;
    // pinned host tables: a ring of slots, each reusable once its upload is done
    enum { NSLOT = 4 };
    mpcf::CaseDesc* h_cases[NSLOT] = {};
    int* h_cur[NSLOT] = {};
    int* h_off[NSLOT] = {};
    hipEvent_t slot_ev[NSLOT] = {};
    int slot = 0;
    std::vector<Req*> active;
    std::vector<long long> active_seq;
    // the live cases in kServerGroups groups of about equal cells, each stepped
    // on its own stream (groups 1.. on side streams forked from and joined
    // back to `stream`), so that one group's x+y sweep runs beside another's
    // z sweep
    mpcf::BatchDev dev[kServerGroups]{};
    int tot[kServerGroups][3] = {};   // per group: blocks of the x+y sweep, z sweep, speed of sound
    hipStream_t side[kServerGroups] = {};
    hipEvent_t fork = nullptr, join[kServerGroups] = {};
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> retired;
    // statistics (engine-owned, read under mu via snapshot)
    std::atomic<long long> st_steps{0}, st_captures{0}, st_rebuilds{0}, st_peak_live{0};
    std::atomic<long long> st_active_cells{0};

    std::thread engine, copier;
};

namespace {

void set_error(mpcf_server* S, int code, const std::string& msg)
{
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->error == MPCF_OK) {
        S->error = code;
        S->error_msg = msg;
    }
    S->stop = true;
    S->cv_done.notify_all();
    S->cv_copier.notify_all();
}

hipEvent_t take_event(mpcf_server* S)
{
    if (!S->ev_free.empty()) {
        hipEvent_t e = S->ev_free.back();
        S->ev_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

int staging_get(mpcf_server* S, size_t bytes)
{
    int best = -1;
    for (size_t i = 0; i < S->staging.size(); ++i)
        if (!S->staging[i].busy && S->staging[i].bytes >= bytes &&
            (best < 0 || S->staging[i].bytes < S->staging[best].bytes))
            best = (int)i;
    if (best >= 0) {
        S->staging[best].busy = true;
        return best;
    }
    Staging st;
    if (hipHostMalloc((void**)&st.p, bytes, hipHostMallocDefault) != hipSuccess) return -1;
    st.bytes = bytes;
    st.busy = true;
    S->staging.push_back(st);
    return (int)S->staging.size() - 1;
}

// a context over arena memory, on the server's stream (no own stream, no
// pinned mirror: the engine never reads a case's step state back)
mpcf_ctx* make_case(mpcf_server* S, Req* r)
{
    mpcf_ctx* c = new mpcf_ctx();
    c->g = mpcf::make_geom(r->nx, r->ny, r->nz);
    c->h = r->h;
    c->flags = S->flags;
    for (int i = 0; i < 6; ++i) c->bc[i] = r->bc[i];
    c->device = S->device;
    c->esize = sizeof(float);
    c->base = S->arena.base + r->off;
    const size_t fb = (size_t)c->g.fstride * c->esize;
    c->u[0] = c->base;
    c->u[1] = c->base + 7 * fb;
    c->R = c->base + 14 * fb;
    c->acc = c->base + 21 * fb;
    c->st = (StepState*)(c->base + (7 + 7 + 7 + mpcf::NACC) * fb);
    c->cur = 0;
    c->st_host = nullptr;
    c->stream = S->stream;
    c->own_stream = false;
    c->bytes = r->bytes;
    c->L = S->L;
    c->timing = false;
    return c;
}

size_t case_bytes(int nx, int ny, int nz)
{
    const Geom g = mpcf::make_geom(nx, ny, nz);
    return (7 + 7 + 7 + mpcf::NACC) * (size_t)g.fstride * sizeof(float) + 256;
}

void copier_main(mpcf_server* S)
{
    for (;;) {
        CopyJob j;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_copier.wait(lk, [&] { return S->stop || !S->copy_jobs.empty(); });
            if (S->copy_jobs.empty()) return;   // stop requested and nothing left
            j = S->copy_jobs.front();
            S->copy_jobs.pop_front();
        }
        // split over the library's pool of copy threads (the copy and the
        // first touch of the caller's pages are the cost of a large request)
        const size_t n = (size_t)j.r->cells();
        for (int q = 0; q < 7; ++q) {
            if (j.to_pinned) mpcf::detail::parallel_copy(j.pinned + q * n, j.r->in[q], n * sizeof(float));
            else             mpcf::detail::parallel_copy(j.r->out[q], j.pinned + q * n, n * sizeof(float));
        }
        j.r->copied.store(true, std::memory_order_release);
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv_engine.notify_all();
    }
}

void post_copy(mpcf_server* S, Req* r, bool to_pinned)
{
    r->copied.store(false, std::memory_order_relaxed);
    std::lock_guard<std::mutex> lk(S->mu);
    S->copy_jobs.push_back({r, to_pinned, S->staging[r->stage_buf].p});
    S->cv_copier.notify_one();
}

bool event_done(hipEvent_t e) { return hipEventQuery(e) == hipSuccess; }

// advance requests whose host copy or device milestone has completed;
// returns false on a HIP error
bool progress(mpcf_server* S, bool* changed)
{
    std::vector<Req*> keep;
    keep.reserve(S->reqs.size());
    for (Req* r : S->reqs) {
        bool finished = false;
        if (r->state == S_STAGING_IN && r->copied.load(std::memory_order_acquire)) {
            // pinned input ready: upload, pack into u[0], clear R
            mpcf_ctx* c = make_case(S, r);
            r->c = c;
            const size_t n = (size_t)r->cells();
            if (hipMemcpyAsync(c->acc, S->staging[r->stage_buf].p, 7 * n * sizeof(float), hipMemcpyHostToDevice,
                               S->stream) != hipSuccess) return false;
            S->L->pack(c->g, c->u[0], c->acc, 1, 7, 0, S->stream);
            // R starts at zero; the step state's reduction targets and counters too
            if (hipMemsetAsync(c->R, 0, 7 * (size_t)c->g.fstride * sizeof(float), S->stream) != hipSuccess) return false;
            if (hipMemsetAsync(c->st, 0, sizeof(StepState), S->stream) != hipSuccess) return false;
            r->ev = take_event(S);
            if (!r->ev || hipEventRecord(r->ev, S->stream) != hipSuccess) return false;
            r->state = S_LIVE;
            r->left = r->steps;
            r->seq = S->next_seq++;
            *changed = true;
        } else if (r->state == S_LIVE && r->ev && event_done(r->ev)) {
            // the upload has landed: the staging buffer is free again
            S->ev_free.push_back(r->ev);
            r->ev = nullptr;
            S->staging[r->stage_buf].busy = false;
            r->stage_buf = -1;
        } else if (r->state == S_FINISHING && event_done(r->ev)) {
            S->ev_free.push_back(r->ev);
            r->ev = nullptr;
            r->state = S_STAGING_OUT;
            post_copy(S, r, false);
        } else if (r->state == S_STAGING_OUT && r->copied.load(std::memory_order_acquire)) {
            S->staging[r->stage_buf].busy = false;
            const double t = clock_s();
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->done.emplace_back(r->id, t);
                ++S->n_completed;
                S->cv_done.notify_all();
            }
            delete r->c;
            delete r;
            finished = true;
        }
        if (!finished) keep.push_back(r);
    }
    S->reqs.swap(keep);
    return true;
}

// admit waiting requests in arrival order while memory and slots allow
bool admit(mpcf_server* S)
{
    int live = 0;
    for (Req* r : S->reqs)
        if (r->state == S_STAGING_IN || r->state == S_LIVE) ++live;
    for (Req* r : S->reqs) {
        if (r->state != S_WAITING) continue;
        if (live >= S->max_live) break;
        size_t off;
        if (!S->arena.alloc(r->bytes, &off)) break;   // arrival order: later ones wait too
        const int sb = staging_get(S, 7 * (size_t)r->cells() * sizeof(float));
        if (sb < 0) {
            S->arena.release(off, r->bytes);
            return false;
        }
        r->off = off;
        r->stage_buf = sb;
        r->state = S_STAGING_IN;
        post_copy(S, r, true);
        ++live;
    }
    return true;
}

void choose_active(mpcf_server* S, std::vector<Req*>& act)
{
    act.clear();
    std::vector<Req*> live;
    for (Req* r : S->reqs)
        if (r->state == S_LIVE && r->left > 0) live.push_back(r);
    if (S->policy == MPCF_SERVE_ALL || live.empty()) {
        act = live;
    } else {
        auto by_age = [](const Req* a, const Req* b) { return a->seq < b->seq; };
        auto by_work = [](const Req* a, const Req* b) { return a->cells() * a->left < b->cells() * b->left; };
        if (S->policy == MPCF_SERVE_FIFO_FILL || S->policy == MPCF_SERVE_FIFO_SLACK) {
            std::stable_sort(live.begin(), live.end(), by_age);
        } else {
            std::stable_sort(live.begin(), live.end(), by_work);
            if (S->policy == MPCF_SERVE_HYBRID) {
                // the oldest request always steps, so no request starves
                auto oldest = std::min_element(live.begin(), live.end(), by_age);
                std::rotate(live.begin(), oldest, oldest + 1);
            }
        }
        long long budget = S->fill_cells;
        if (S->policy == MPCF_SERVE_FIFO_SLACK) {
            std::stable_sort(live.begin(), live.end(), by_age);
            if (live.size() <= 2) {
                // light load: the oldest request steps as if alone, except for
                // a companion whose remaining work is at most a tenth of its
                // own (it finishes quickly and delays the oldest by at most
                // that much), which would otherwise wait for the whole of it
                auto work = [](const Req* r) { return r->cells() * (long long)r->left; };
                act.push_back(live[0]);
                if (live.size() == 2 && 10 * work(live[1]) <= work(live[0])) act.push_back(live[1]);
                std::sort(act.begin(), act.end(), [](const Req* a, const Req* b) { return a->seq < b->seq; });
                return;
            }
            // the oldest request steps at close to its own speed: companions
            // may add at most a quarter of its cells (or fill the floor budget
            // when it is small)
            budget = std::max(budget, live.front()->cells() + live.front()->cells() / 4);
        }
        long long cells = 0;
        for (Req* r : live) {
            if (!act.empty() && cells + r->cells() > budget) continue;
            act.push_back(r);
            cells += r->cells();
        }
    }
    // a fixed order inside the launch: admission order
    std::sort(act.begin(), act.end(), [](const Req* a, const Req* b) { return a->seq < b->seq; });
}

void retire_graph(mpcf_server* S)
{
    if (!S->exec) return;
    hipEvent_t e = take_event(S);
    if (e) hipEventRecord(e, S->stream);
    S->retired.emplace_back(S->exec, e);
    S->exec = nullptr;
}

void reap_graphs(mpcf_server* S)
{
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> keep;
    for (auto& g : S->retired) {
        if (!g.second || event_done(g.second)) {
            hipGraphExecDestroy(g.first);
            if (g.second) S->ev_free.push_back(g.second);
        } else {
            keep.push_back(g);
        }
    }
    S->retired.swap(keep);
}

// write the device tables of the active set (stream-ordered after the steps
// already enqueued, which read the previous contents)
bool rebuild(mpcf_server* S)
{
    // next slot of the pinned ring; it may still be the source of an
    // earlier table copy (four rebuilds ago)
    const int sl = S->slot;
    S->slot = (S->slot + 1) % mpcf_server::NSLOT;
    if (hipEventSynchronize(S->slot_ev[sl]) != hipSuccess) return false;
    mpcf::CaseDesc* hc = S->h_cases[sl];
    int* hcur = S->h_cur[sl];
    const int n = (int)S->active.size();
    const int cap = S->max_live;
    const int tab = 3 * (cap + 1);
    int* hoff = S->h_off[sl];
    long long cells = 0, cols_xy = 0, cols_z = 0;
    std::vector<long long> cc(n);
    for (int i = 0; i < n; ++i) {
        const mpcf::Geom& g = S->active[i]->c->g;
        cols_xy += (long long)(g.nx + 6) * g.nz;
        cols_z += (long long)g.nx * g.ny;
        cc[i] = g.nx * (long long)g.ny * g.nz;
        cells += cc[i];
    }
    int first[kServerGroups + 1];
    split_groups(cc.data(), n, kServerGroups, first);
    for (int grp = 0; grp < kServerGroups; ++grp)
        hoff[grp * tab] = hoff[grp * tab + (cap + 1)] = hoff[grp * tab + 2 * (cap + 1)] = 0;
    for (int i = 0; i < n; ++i) {
        Req* r = S->active[i];
        mpcf_ctx* c = r->c;
        if (r->stepped && !r->dev_folded) {
            // its steps so far ran alone on the whole-block kernels, which
            // fold the maximum but leave the batch's flag to the host
            if (hipMemsetD32Async((hipDeviceptr_t)&c->st->folded, 1, 1, S->stream) != hipSuccess) return false;
            r->dev_folded = true;
        }
        int grp = 0;
        while (i >= first[grp + 1]) ++grp;
        const int j = i - first[grp];                                      // index within the group
        int* oxy = hoff + grp * tab;
        int* oz = oxy + (cap + 1);
        int* os = oxy + 2 * (cap + 1);
        int blocks[3];
        // a lone case marches as a whole-block single-case launch would
        // (shorter chunks, more waves to fill the chip); a mix shares the
        // chip and takes the longer, cheaper chunks of adaptive_march
        const int lxy = n == 1 ? mpcf::parity::whole_chunk(c->g.ny, (long long)c->g.nx * c->g.nz)
                               : mpcf::detail::adaptive_march(c->g.ny, cols_xy);
        const int lz = n == 1 ? mpcf::parity::whole_chunk(c->g.nz, (long long)c->g.nx * c->g.ny)
                              : mpcf::detail::adaptive_march(c->g.nz, cols_z);
        mpcf::detail::case_desc(c, c->cur, lxy, lz, hc[i], blocks);
        oxy[j + 1] = oxy[j] + blocks[0];
        oz[j + 1] = oz[j] + blocks[1];
        os[j + 1] = os[j] + blocks[2];
        hcur[i] = c->cur;
    }
    if (n > 0) {
        if (hipMemcpyAsync(S->d_cases, hc, n * sizeof(mpcf::CaseDesc), hipMemcpyHostToDevice, S->stream) != hipSuccess ||
            hipMemcpyAsync(S->d_cur, hcur, n * sizeof(int), hipMemcpyHostToDevice, S->stream) != hipSuccess ||
            hipMemcpyAsync(S->d_off, hoff, kServerGroups * tab * sizeof(int), hipMemcpyHostToDevice, S->stream) != hipSuccess)
            return false;
        if (hipEventRecord(S->slot_ev[sl], S->stream) != hipSuccess) return false;
    }
    for (int grp = 0; grp < kServerGroups; ++grp) {
        const int count = first[grp + 1] - first[grp];
        const int* ho = hoff + grp * tab;
        S->tot[grp][0] = ho[count]; S->tot[grp][1] = ho[(cap + 1) + count]; S->tot[grp][2] = ho[2 * (cap + 1) + count];
        mpcf::BatchDev& d = S->dev[grp];
        d.cases = S->d_cases + first[grp]; d.cur = S->d_cur + first[grp]; d.n = count;
        d.off_xy = S->d_off + grp * tab; d.off_z = d.off_xy + (cap + 1); d.off_sos = d.off_xy + 2 * (cap + 1);
    }
    S->st_active_cells = cells;
    retire_graph(S);
    ++S->st_rebuilds;
    return true;
}

bool enqueue_step_kernels(mpcf_server* S)
{
    // the sweeps apply each case's boundary conditions as they load: seven
    // launches per step and group; groups 1.. run on side streams
    const int cpu_form = (S->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0;
    bool forked = false;
    for (int grp = kServerGroups - 1; grp >= 0; --grp) {
        const mpcf::BatchDev& d = S->dev[grp];
        if (d.n == 0) continue;
        hipStream_t st = S->stream;
        if (grp > 0) {
            if (!forked && hipEventRecord(S->fork, S->stream) != hipSuccess) return false;
            forked = true;
            st = S->side[grp];
            if (hipStreamWaitEvent(st, S->fork, 0) != hipSuccess) return false;
        }
        S->L->b_sos(d, S->tot[grp][2], cpu_form, S->cfl, 1e30, st);
        for (int s = 0; s < 3; ++s) {
            S->L->b_xy(d, S->tot[grp][0], kA[s], kB[s], st);
            // a request's R is never handed back, and the next step's first
            // stage multiplies it by a = 0: the last stage does not store it
            S->L->b_z(d, S->tot[grp][1], kA[s], kB[s], s == 2, cpu_form, s != 2, st);
        }
        if (grp > 0 && hipEventRecord(S->join[grp], st) != hipSuccess) return false;
    }
    for (int grp = 1; grp < kServerGroups; ++grp)
        if (S->dev[grp].n > 0 && hipStreamWaitEvent(S->stream, S->join[grp], 0) != hipSuccess) return false;
    return true;
}

// A request stepped alone that is large enough to fill the GPU by itself runs
// the kernels a whole-block mpcf_step launches (the batched launches cost it
// 6-12 % at 128^3-192^3: the per-workgroup case lookup and the tile order of
// a mixed batch).  Its buffers alternate between steps, so these launches are
// enqueued eagerly, not replayed from a graph; below this size the launch
// overhead outweighs the difference and the batch graph steps it.
constexpr long long kSoloCells = 1LL << 18;

bool solo_step(const mpcf_server* S)
{
    return S->active.size() == 1 && S->active[0]->cells() >= kSoloCells;
}

// one step of a lone request on the whole-block kernels: dt (from the maximum
// its previous step folded, or a full reduction before its first step), then
// per stage the x+y sweep(s) and the z sweep, the last stage without the R
// store and folding the next step's maximum
bool enqueue_solo_step(mpcf_server* S, Req* r)
{
    mpcf_ctx* c = r->c;
    const int cpu_form = (S->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0;
    if (r->stepped) S->L->sos_folded(c->st, S->cfl, c->h, 1e30, S->stream);
    else S->L->sos(c->g, c->u[c->cur], c->st, cpu_form, 1, S->cfl, c->h, 1e30, S->stream);
    int cur = c->cur;
    for (int s = 0; s < 3; ++s) {
        mpcf::SweepParams p;
        p.u = c->u[cur]; p.R = c->R; p.acc = c->acc; p.uo = c->u[1 - cur];
        p.a = kA[s]; p.b = kB[s]; p.dtinvh = 0.0; p.st = c->st;
        p.sos_next = s == 2 ? c->st : nullptr;
        p.sos_cpu = cpu_form;
        p.bcmask = mpcf::detail::bc_mask(c->bc);
        S->L->sweep(c->g, 3, p, 1, S->stream);                // the fused x+y sweep
        S->L->sweep(c->g, 2, p, s == 2 ? 2 : 1, S->stream);
        cur = 1 - cur;
    }
    if (hipGetLastError() != hipSuccess) return false;
    ++S->st_steps;
    return true;
}

bool enqueue_step(mpcf_server* S)
{
    if (S->use_graph) {
        if (!S->exec) {
            if (hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
            const bool ok = enqueue_step_kernels(S);
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(S->stream, &g) != hipSuccess) return false;
            if (!ok) { if (g) hipGraphDestroy(g); return false; }
            const hipError_t e = hipGraphInstantiate(&S->exec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return false;
            ++S->st_captures;
        }
        if (hipGraphLaunch(S->exec, S->stream) != hipSuccess) return false;
    } else {
        if (!enqueue_step_kernels(S)) return false;
        if (hipGetLastError() != hipSuccess) return false;
    }
    ++S->st_steps;
    return true;
}

// a finished case: unpack into its accumulator area, copy to pinned memory,
// release its device memory (later uploads are ordered after this download)
bool finish(mpcf_server* S, Req* r)
{
    if (r->ev) {
        // the upload's milestone was not observed yet; it precedes this
        // case's steps, so it is complete or about to be
        if (hipEventSynchronize(r->ev) != hipSuccess) return false;
        S->ev_free.push_back(r->ev);
        r->ev = nullptr;
        S->staging[r->stage_buf].busy = false;
        r->stage_buf = -1;
    }
    mpcf_ctx* c = r->c;
    const size_t n = (size_t)r->cells();
    const int sb = staging_get(S, 7 * n * sizeof(float));
    if (sb < 0) return false;
    r->stage_buf = sb;
    S->L->pack(c->g, c->u[c->cur], c->acc, 0, 7, 0, S->stream);
    if (hipMemcpyAsync(S->staging[sb].p, c->acc, 7 * n * sizeof(float), hipMemcpyDeviceToHost, S->stream) != hipSuccess)
        return false;
    r->ev = take_event(S);
    if (!r->ev || hipEventRecord(r->ev, S->stream) != hipSuccess) return false;
    S->arena.release(r->off, r->bytes);
    r->state = S_FINISHING;
    return true;
}

void engine_main(mpcf_server* S)
{
    if (hipSetDevice(S->device) != hipSuccess) {
        set_error(S, MPCF_EHIP, "server: hipSetDevice failed");
        return;
    }
    std::vector<Req*> act;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            while (!S->submitted.empty()) {
                S->reqs.push_back(S->submitted.front());
                S->submitted.pop_front();
            }
            if (S->stop && S->reqs.empty()) break;
        }
        bool changed = false;
        if (!admit(S)) { set_error(S, MPCF_ENOMEM, "server: no pinned staging memory"); break; }
        if (!progress(S, &changed)) { set_error(S, MPCF_EHIP, "server: HIP error while moving a case"); break; }
        reap_graphs(S);
        choose_active(S, act);
        std::vector<long long> seq(act.size());
        for (size_t i = 0; i < act.size(); ++i) seq[i] = act[i]->seq;
        if (act.empty()) {
            S->active.clear();
            S->active_seq.clear();
            // nothing to step: wait for a submission, a copy, or a device milestone
            bool gpu_pending = false;
            for (Req* r : S->reqs)
                if (r->ev) { gpu_pending = true; break; }
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->reqs.empty()) break;
            S->cv_engine.wait_for(lk, std::chrono::microseconds(gpu_pending ? 50 : 2000));
            continue;
        }
        if (seq != S->active_seq) {
            S->active = act;
            S->active_seq = seq;
            if (!rebuild(S)) { set_error(S, MPCF_EHIP, "server: table upload failed"); break; }
        }
        const bool solo = solo_step(S);
        if (!(solo ? enqueue_solo_step(S, S->active[0]) : enqueue_step(S))) {
            set_error(S, MPCF_EHIP, "server: step launch failed");
            break;
        }
        S->st_peak_live = std::max<long long>(S->st_peak_live, (long long)S->active.size());
        for (Req* r : S->active) {
            r->stepped = true;
            if (!solo) r->dev_folded = true;   // the batched z sweep sets the flag
            r->c->cur ^= 1;
            if (--r->left == 0 && !finish(S, r)) { set_error(S, MPCF_EHIP, "server: download failed"); return; }
        }
        // throttle the queue depth
        hipEvent_t e = take_event(S);
        if (!e || hipEventRecord(e, S->stream) != hipSuccess) { set_error(S, MPCF_EHIP, "server: event"); break; }
        S->inflight.push_back(e);
        while ((int)S->inflight.size() > S->depth) {
            hipEvent_t o = S->inflight.front();
            S->inflight.pop_front();
            if (hipEventSynchronize(o) != hipSuccess) { set_error(S, MPCF_EHIP, "server: step failed"); return; }
            S->ev_free.push_back(o);
        }
        while (!S->inflight.empty() && event_done(S->inflight.front())) {
            S->ev_free.push_back(S->inflight.front());
            S->inflight.pop_front();
        }
    }
}

}  // namespace

extern "C" double mpcf_server_clock(void) { return clock_s(); }

extern "C" int mpcf_server_create(mpcf_server** out, unsigned flags, long long hbm_bytes, int policy,
                                  long long fill_cells, int max_live, double cfl)
{
    if (!out) return fail(MPCF_EINVAL, "mpcf_server_create: out is NULL");
    *out = nullptr;
    if (flags & ~(unsigned)(MPCF_FLAG_FAST | MPCF_FLAG_SOS_CPU | MPCF_FLAG_P64 | MPCF_FLAG_SERVE_EAGER))
        return fail(MPCF_EINVAL, "mpcf_server_create: unsupported flag (binary32 storage only)");
    if (policy < MPCF_SERVE_ALL || policy > MPCF_SERVE_FIFO_SLACK) return fail(MPCF_EINVAL, "mpcf_server_create: bad policy");
    if (hbm_bytes <= 0 || max_live < 1 || !(cfl > 0)) return fail(MPCF_EINVAL, "mpcf_server_create: bad budget / slots / cfl");
    if (flags & MPCF_FLAG_P64) flags |= MPCF_FLAG_FAST;
    mpcf_server* S = new mpcf_server();
    S->flags = flags & ~(unsigned)MPCF_FLAG_SERVE_EAGER;
    S->use_graph = !(flags & MPCF_FLAG_SERVE_EAGER);
    S->L = pick_launchers(S->flags);
    S->policy = policy;
    S->fill_cells = fill_cells > 0 ? fill_cells : (1LL << 62);
    S->max_live = max_live;
    S->cfl = cfl;
    auto bail = [&](int code, const std::string& m) {
        if (S->arena.base) hipFree(S->arena.base);
        if (S->d_cases) hipFree(S->d_cases);
        if (S->d_cur) hipFree(S->d_cur);
        if (S->d_off) hipFree(S->d_off);
        for (int k = 0; k < mpcf_server::NSLOT; ++k) {
            if (S->h_cases[k]) hipHostFree(S->h_cases[k]);
            if (S->h_cur[k]) hipHostFree(S->h_cur[k]);
            if (S->h_off[k]) hipHostFree(S->h_off[k]);
            if (S->slot_ev[k]) hipEventDestroy(S->slot_ev[k]);
        }
        if (S->stream) hipStreamDestroy(S->stream);
        for (int grp = 1; grp < kServerGroups; ++grp) {
            if (S->side[grp]) hipStreamDestroy(S->side[grp]);
            if (S->join[grp]) hipEventDestroy(S->join[grp]);
        }
        if (S->fork) hipEventDestroy(S->fork);
        delete S;
        return fail(code, m);
    };
    if (hipGetDevice(&S->device) != hipSuccess) return bail(MPCF_EHIP, "mpcf_server_create: hipGetDevice");
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&S->fork, hipEventDisableTiming) != hipSuccess)
        return bail(MPCF_EHIP, "mpcf_server_create: streams");
    for (int grp = 1; grp < kServerGroups; ++grp)
        if (hipStreamCreateWithFlags(&S->side[grp], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&S->join[grp], hipEventDisableTiming) != hipSuccess)
            return bail(MPCF_EHIP, "mpcf_server_create: streams");
    S->arena.size = round256((size_t)hbm_bytes);
    if (hipMalloc((void**)&S->arena.base, S->arena.size) != hipSuccess)
        return bail(MPCF_ENOMEM, "mpcf_server_create: cannot allocate the HBM budget");
    S->arena.free_[0] = S->arena.size;
    const int cap = max_live;
    if (hipMalloc((void**)&S->d_cases, cap * sizeof(mpcf::CaseDesc)) != hipSuccess ||
        hipMalloc((void**)&S->d_cur, cap * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&S->d_off, kServerGroups * 3 * (cap + 1) * sizeof(int)) != hipSuccess)
        return bail(MPCF_ENOMEM, "mpcf_server_create: tables");
    for (int k = 0; k < mpcf_server::NSLOT; ++k)
        if (hipHostMalloc((void**)&S->h_cases[k], cap * sizeof(mpcf::CaseDesc), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&S->h_cur[k], cap * sizeof(int), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&S->h_off[k], kServerGroups * 3 * (cap + 1) * sizeof(int), hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&S->slot_ev[k], hipEventDisableTiming) != hipSuccess)
            return bail(MPCF_ENOMEM, "mpcf_server_create: tables");
    S->copier = std::thread(copier_main, S);
    S->engine = std::thread(engine_main, S);
    *out = S;
    return MPCF_OK;
}

extern "C" int mpcf_server_submit(mpcf_server* S, int nx, int ny, int nz, double h, const int bc[6], int steps,
                                  const float* const in[7], float* const out[7], long long* id)
{
    if (!S || !in || !out || !id) return fail(MPCF_EINVAL, "mpcf_server_submit: NULL argument");
    if (nx < 5 || ny < 5 || nz < 5 || steps < 1 || !(h > 0)) return fail(MPCF_EINVAL, "mpcf_server_submit: bad extents / steps / h");
    for (int q = 0; q < 7; ++q)
        if (!in[q] || !out[q]) return fail(MPCF_EINVAL, "mpcf_server_submit: NULL field pointer");
    if (bc)
        for (int i = 0; i < 6; ++i)
            if (bc[i] != MPCF_BC_ABSORBING && bc[i] != MPCF_BC_REFLECTING)
                return fail(MPCF_EINVAL, "mpcf_server_submit: unknown boundary kind");
    const size_t bytes = case_bytes(nx, ny, nz);
    if (bytes > S->arena.size) return fail(MPCF_ENOMEM, "mpcf_server_submit: the case does not fit the HBM budget");
    const Geom g = mpcf::make_geom(nx, ny, nz);
    if ((int64_t)g.px * g.py * g.pz >= (int64_t(1) << 31)) return fail(MPCF_EINVAL, "mpcf_server_submit: block too large");
    Req* r = new Req();
    r->nx = nx; r->ny = ny; r->nz = nz; r->h = h; r->steps = steps;
    for (int i = 0; i < 6; ++i) r->bc[i] = bc ? bc[i] : MPCF_BC_ABSORBING;
    for (int q = 0; q < 7; ++q) { r->in[q] = in[q]; r->out[q] = out[q]; }
    r->bytes = bytes;
    r->t_submit = clock_s();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->error != MPCF_OK) {
        delete r;
        return fail(S->error, S->error_msg);
    }
    r->id = ++S->n_submitted;
    *id = r->id;
    S->submitted.push_back(r);
    S->cv_engine.notify_all();
    return MPCF_OK;
}

extern "C" int mpcf_server_poll(mpcf_server* S, long long* ids, double* t_done, int cap, int* n)
{
    if (!S || !n || (cap > 0 && (!ids || !t_done))) return fail(MPCF_EINVAL, "mpcf_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    int k = 0;
    while (k < cap && !S->done.empty()) {
        ids[k] = S->done.front().first;
        t_done[k] = S->done.front().second;
        S->done.pop_front();
        ++k;
    }
    *n = k;
    if (S->error != MPCF_OK) return fail(S->error, S->error_msg);
    return MPCF_OK;
}

extern "C" int mpcf_server_wait(mpcf_server* S, double timeout_s)
{
    if (!S) return fail(MPCF_EINVAL, "mpcf_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto pred = [&] { return S->error != MPCF_OK || S->n_completed == S->n_submitted; };
    if (timeout_s < 0) S->cv_done.wait(lk, pred);
    else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (S->error != MPCF_OK) return fail(S->error, S->error_msg);
    return S->n_completed == S->n_submitted ? MPCF_OK : fail(MPCF_EBUSY, "mpcf_server_wait: timed out");
}

extern "C" int mpcf_server_stats(mpcf_server* S, long long out[8])
{
    if (!S || !out) return fail(MPCF_EINVAL, "mpcf_server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = S->n_submitted;
    out[1] = S->n_completed;
    out[2] = S->st_steps;
    out[3] = S->st_captures;
    out[4] = S->st_rebuilds;
    out[5] = S->st_peak_live;
    out[6] = (long long)S->arena.peak;
    out[7] = S->st_active_cells;
    return MPCF_OK;
}

extern "C" int mpcf_server_destroy(mpcf_server* S)
{
    if (!S) return MPCF_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_engine.notify_all();
    }
    if (S->engine.joinable()) S->engine.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv_copier.notify_all();
    }
    if (S->copier.joinable()) S->copier.join();
    hipStreamSynchronize(S->stream);
    for (Req* r : S->reqs) {
        if (r->ev) hipEventDestroy(r->ev);
        delete r->c;
        delete r;
    }
    for (Req* r : S->submitted) delete r;
    if (S->exec) hipGraphExecDestroy(S->exec);
    for (auto& g : S->retired) {
        hipGraphExecDestroy(g.first);
        if (g.second) hipEventDestroy(g.second);
    }
    for (hipEvent_t e : S->inflight) hipEventDestroy(e);
    for (hipEvent_t e : S->ev_free) hipEventDestroy(e);
    for (auto& st : S->staging) hipHostFree(st.p);
    hipFree(S->arena.base);
    hipFree(S->d_cases);
    hipFree(S->d_cur);
    hipFree(S->d_off);
    for (int k = 0; k < mpcf_server::NSLOT; ++k) {
        hipHostFree(S->h_cases[k]);
        hipHostFree(S->h_cur[k]);
        hipHostFree(S->h_off[k]);
        hipEventDestroy(S->slot_ev[k]);
    }
    for (int grp = 1; grp < kServerGroups; ++grp) {
        hipStreamSynchronize(S->side[grp]);
        hipStreamDestroy(S->side[grp]);
        hipEventDestroy(S->join[grp]);
    }
    hipEventDestroy(S->fork);
    hipStreamDestroy(S->stream);
    delete S;
    return MPCF_OK;
}
