// engine.cpp -- the on-device request engine and fit sessions.
//
// Request engine (continuous batching).  A request is one I(q) curve: a PDB
// leaf (atoms and hydration boxes) or a resident grid, optional symmetry
// placements, a q vector and an orientation average, optionally scored by
// chi^2 against data.  Submission copies the inputs and returns at once.
// Requests are admitted in arrival order whenever a device slot is free and
// their device memory fits the HBM budget; each slot is a worker thread with
// its own HIP stream, so the kernels of admitted requests overlap on the GPU
// while later arrivals wait.  Device buffers are stream-ordered pool
// allocations, so one request's frees never stall the others.  A request's
// result does not depend on what else runs: every kernel reads only that
// request's buffers and its own seeded random streams.
// Grouped builds (optional).  Small mixed PDB builds (up to batch_terms
// terms) are handed to a batcher thread with its own stream: whatever builds
// are ready when the previous grouped launch has been issued go out together
// as one builder launch over up to 16 grids of different structures, grid
// sizes and slice counts (each job waits on its request's upload event, each
// request's stream waits on the launch's event), so small requests fill the
// device together instead of launching one small grid each.
//
// Fixed-count requests (Monte Carlo or VEGAS with eps <= 0) are prepared
// first -- resident structure, grid, form factors, slice partials,
// placements, q and scratch allocated and uploaded on the slot's stream --
// and then their whole kernel chain (builder, slice sums, origin, splines,
// orientation average, intensity, chi^2, download into pinned memory) is
// launched with a single synchronisation at the end.  Captured requests
// (optional): the chain is captured into one HIP graph and launched.  Each
// slot keeps instantiated graphs keyed by the chain's shape (method,
// iterations, precision, phase form, builder workgroup, chi^2); a new
// request updates the kept graph of its key in place (hipGraphExecUpdate)
// and instantiates only when the update is refused.  The kernels and their
// arguments are those of the eager launches, so both give the same bits.
//
// Fit session.  Resident subunit grids plus fixed q, data and sigma; each
// evaluation re-docks the subunits with new placements (symmetry parameters of
// a fitting iteration), averages by fixed-count Monte Carlo and returns I(q)
// and chi^2 computed on the device.  The evaluation is one captured HIP graph
// (placement upload, MC kernels, intensity, chi^2, result download), replayed
// while the number of rotations stays the same and re-captured when it changes.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "device.hpp"

namespace dplus {

void launch_mc(bool fp64, bool single, const void *const *grids, const RotDev *rots, int nrot, const double4 *trans,
               const ZCopyDev *zc, int nz, const GroupDev *groups, const int *active, int nactive, int mstride, const double *qs,
               double step, uint64_t seed, long long dir0, int nblk, int chunk, double *partial, double *out,
               hipStream_t s);
void launch_mc_finish(const GroupDev *groups, int ngroups, int mstride, const double *out, double total, double *I,
                      hipStream_t s);
void launch_chi2_param(const double *I, const double *y, const double *sig, int n, const double *scale_constant,
                       double *chi2, hipStream_t s);

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// 128-bit content hash (multiply-rotate mixing over 8-byte words): two seeds
// times four interleaved lanes, eight independent chains per pass over the
// data, folded at the end of each array
uint64_t mix64(uint64_t h, uint64_t v) {
    h ^= v * 0x9E3779B97F4A7C15ULL;
    h = (h << 29 | h >> 35) * 0xBF58476D1CE4E5B9ULL;
    return h ^ (h >> 31);
}

void hash_bytes(uint64_t key[2], const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h[2][4];
    for (int k = 0; k < 2; k++)
        for (int l = 0; l < 4; l++) h[k][l] = mix64(key[k], (uint64_t)l + 1);
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        uint64_t v[4];
        memcpy(v, b + i, 32);
        for (int l = 0; l < 4; l++) {
            h[0][l] = mix64(h[0][l], v[l]);
            h[1][l] = mix64(h[1][l], v[l]);
        }
    }
    uint64_t t[4] = {0, 0, 0, 0};
    for (size_t k = 0; i + k < n; k++) t[k / 8] |= (uint64_t)b[i + k] << (8 * (k % 8));
    for (int k = 0; k < 2; k++) {
        uint64_t a = key[k];
        for (int l = 0; l < 4; l++) a = mix64(a, mix64(h[k][l], t[l]));
        key[k] = mix64(a, (uint64_t)n);
    }
}

struct Request {
    long long id = 0;
    bool has_pdb = false;
    uint64_t key[2] = {0, 0};  // content hash of the PDB description (structure cache)
    std::vector<double> xyz, coef9, rad, box_com, excl_com;
    std::vector<int> ion, box_dims, excl_dims;
    dplus_pdb_desc desc{};
    Grid *grid = nullptr;
    int grid_size = 0, precision = DPLUS_PRECISION_MIXED;
    double qmax = 0.0;
    std::vector<double> rot, trans;
    std::vector<int> ntrans;
    std::vector<double> q, y, sigma;
    double fit_scale = 1.0, fit_constant = 0.0;
    bool want_chi = false;
    dplus_oa_params oa{};
    size_t need = 0;
    // results
    std::vector<double> I;
    double chi2 = 0.0;
    bool chi_on_device = false;
    bool graph = false;  // ran as a captured graph
    dplus_oa_stats st{};
    int code = 0;
    std::string message;
    int slot = -1;
    bool done = false;
    Clock::time_point t_submit, t_start, t_end;
};

// device bytes a request holds while it runs: its grid and structure, the
// orientation-average scratch, and a fixed margin for small tables
size_t request_bytes(const Request &r) {
    size_t b = 16u << 20;
    int gs = r.grid_size;
    double step;
    bool fp64;
    if (r.has_pdb) {
        if (gs % 2) gs++;
        const long long P = total_points(gs);
        fp64 = r.precision == DPLUS_PRECISION_FP64;
        b += (size_t)P * (fp64 ? 32 : 16);
        // FP64 partial sums of the paired builder's atom slices (as many as a
        // tuned launch shape may take; FP64 grids too, unless built plainly)
        if (!fp64 || !fp64_plain_builder()) {
            const long long units = pair_units(half_base(gs / 2 + 4));
            b += (size_t)kMaxSlices * 12 * units * sizeof(double);
        }
        b += (size_t)(r.desc.natoms + r.desc.nbox + r.desc.nexcl) * 48;
        step = r.qmax / (double)(gs / 2);
    } else {
        fp64 = r.grid->fp64();
        step = r.grid->step;
    }
    (void)fp64;
    const std::vector<GroupDev> groups = make_groups(r.q.data(), (int)r.q.size(), step);
    size_t ng = groups.size(), mstride = 1;
    for (auto &g : groups) mstride = std::max<size_t>(mstride, (size_t)g.m);
    const dplus_oa_params &p = r.oa;
    if (p.method == DPLUS_OA_MC) {
        const long long chunk = p.chunk > 0 ? p.chunk : 8192;
        const long long dirs = p.eps <= 0.0 ? ((std::max<long long>(1, p.max_dirs) + chunk - 1) / chunk) * chunk : chunk;
        b += ng * (size_t)(dirs / 256) * mstride * 8 + ng * mstride * 8;
    } else if (p.method == DPLUS_OA_VEGAS) {
        b += ng * mstride * 1024 * 8 + ng * 1024 * 8 + (size_t)std::max(1, p.vegas_iterations) * r.q.size() * 8;
    } else {
        b += gk_device_bytes(r.q.size());
    }
    b += (r.rot.size() / 4 + 1) * sizeof(RotDev) + (r.trans.size() / 3 + 1) * (sizeof(double4) + sizeof(ZCopyDev));
    return b;
}

}  // namespace

struct Engine {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending, admitted;
    std::map<long long, std::unique_ptr<Request>> all;
    long long next_id = 1, submitted = 0, completed = 0, max_concurrent = 0;
    size_t budget = 0, reserved = 0, peak_reserved = 0;
    // structure cache: device-resident atoms/boxes of recent PDB descriptions,
    // shared by requests with the same content; at most cache_cap bytes, which
    // admission keeps free of request reservations
    struct Cached {
        uint64_t key[2];
        std::shared_ptr<Structure> s;
        size_t bytes;
        unsigned long long last;
    };
    std::mutex cache_mu;
    std::vector<Cached> cache;
    size_t cache_cap = 0, cache_bytes = 0;
    // the slots' arenas: a fixed share of the budget (like the cache, kept out of
    // request admission), split evenly over the slots; a request whose small
    // buffers exceed its slot's share allocates them per request instead
    size_t arena_cap = 0, slot_arena_max = 0;
    unsigned long long cache_tick = 0;
    long long cache_hits = 0, cache_misses = 0;
    int busy = 0;
    bool stopping = false;
    std::vector<hipStream_t> streams;
    std::vector<std::thread> workers;
    // grouped builds: tickets of ready builds, the batcher's thread and stream
    struct BuildTicket {
        BuildJob job;
        bool reduce = false;
        hipEvent_t ready = nullptr, done = nullptr;
        bool launched = false;
        int code = 0;
        std::string message;
    };
    bool batch_builds = false, autotune = false, graphs = false;
    // captured requests: per slot, pinned result staging and the instantiated
    // graphs kept by chain key (least recently used dropped past kKeptGraphs)
    static constexpr int kKeptGraphs = 16;
    static constexpr int kMaxSlots = 32;  // worker threads, each with its own stream
    struct KeptGraph {
        std::string key;
        hipGraphExec_t exec = nullptr;
        unsigned long long last = 0;
    };
    // each slot also keeps a device arena and its pinned staging for the small
    // buffers of its requests (leaf tables, q, groups, partial sums, bins, chi^2
    // inputs), grown to the largest request it has run: no allocation per
    // request, and their uploads go up in one or two copies
    struct Slot {
        double *h_out = nullptr;
        size_t cap = 0;
        std::vector<KeptGraph> kept;
        unsigned long long tick = 0;
        Arena arena;
        DevBuf arena_dev;
        char *arena_pin = nullptr;
        size_t pin_cap = 0;
    };
    std::deque<Slot> slots;  // never relocated (slots hold device buffers)
    std::atomic<long long> graph_launches{0}, graph_instantiations{0}, graph_updates{0};
    // host time of captured requests (ns): waiting for the capture lock and
    // capturing; updating or instantiating the executable graph
    std::atomic<long long> graph_capture_ns{0}, graph_exec_ns{0};
    double batch_terms = 1e10;
    std::mutex bmu;
    std::condition_variable bcv;
    std::deque<BuildTicket *> bq;
    bool bstop = false;
    hipStream_t bstream = nullptr;
    std::thread bthread;
    long long batch_launches = 0, batch_jobs = 0;
    std::mutex tune_mu;  // one autotuning sweep at a time

    Engine(long long budget_bytes, int nstreams, bool batch, long long bterms, bool tune, bool graph) {
        Device &d = Device::get();
        d.require();
        budget = budget_bytes > 0 ? (size_t)budget_bytes : (size_t)(0.85 * (double)d.mem);
        cache_cap = std::min<size_t>(budget / 16, (size_t)8 << 30);
        arena_cap = std::min<size_t>(budget / 8, (size_t)1 << 30);
        batch_builds = batch;
        autotune = tune;
        graphs = graph;
        if (bterms > 0) batch_terms = (double)bterms;
        nstreams = std::max(1, std::min(kMaxSlots, nstreams));
        for (int i = 0; i < nstreams; i++) {
            hipStream_t s;
            check_hip(hipStreamCreate(&s), "hipStreamCreate");
            streams.push_back(s);
        }
        if (batch_builds) {
            check_hip(hipStreamCreate(&bstream), "hipStreamCreate");
            bthread = std::thread([this] { batcher(); });
        }
        slots.resize(nstreams);
        slot_arena_max = arena_cap / (size_t)nstreams;
        for (int i = 0; i < nstreams; i++) workers.emplace_back([this, i] { work(i); });
    }

    ~Engine() {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_done.wait(lk, [&] { return pending.empty() && admitted.empty() && busy == 0; });
            stopping = true;
        }
        cv_work.notify_all();
        for (auto &t : workers) t.join();
        if (bthread.joinable()) {
            {
                std::lock_guard<std::mutex> lk(bmu);
                bstop = true;
            }
            bcv.notify_all();
            bthread.join();
        }
        cache.clear();  // frees are ordered on the worker streams: release before destroying them
        for (size_t i = 0; i < slots.size(); i++) {
            (void)hipStreamSynchronize(streams[i]);
            for (auto &k : slots[i].kept) (void)hipGraphExecDestroy(k.exec);
            std::shared_lock<CaptureMutex> lk(capture_mutex());
            if (slots[i].h_out) (void)hipHostFree(slots[i].h_out);
            if (slots[i].arena_pin) (void)hipHostFree(slots[i].arena_pin);
            lk.unlock();
            slots[i].arena_dev.release();  // ordered on the slot's stream, before it is destroyed
        }
        for (auto s : streams) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (bstream) {
            (void)hipStreamSynchronize(bstream);
            (void)hipStreamDestroy(bstream);
        }
    }

    // The batcher: takes every ready build (up to 16), launches them as one
    // grouped builder launch per phase form on its stream after each job's
    // upload event, and records each job's completion event.  It keeps at
    // most two grouped launches in flight, so builds that become ready while
    // the device works on one launch gather into the next.
    void batcher() {
        hipEvent_t prev = nullptr;
        for (;;) {
            std::vector<BuildTicket *> take;
            {
                std::unique_lock<std::mutex> lk(bmu);
                bcv.wait(lk, [&] { return bstop || !bq.empty(); });
                if (bq.empty()) break;
                while (!bq.empty() && (int)take.size() < kMaxBuildJobs) {
                    take.push_back(bq.front());
                    bq.pop_front();
                }
            }
            hipEvent_t pace = nullptr;
            for (int red = 0; red < 2; red++) {
                std::vector<BuildJob> jobs;
                std::vector<BuildTicket *> ts;
                for (BuildTicket *t : take)
                    if ((int)t->reduce == red) {
                        jobs.push_back(t->job);
                        ts.push_back(t);
                    }
                if (jobs.empty()) continue;
                try {
                    for (BuildTicket *t : ts) check_hip(hipStreamWaitEvent(bstream, t->ready, 0), "hipStreamWaitEvent");
                    launch_build_batch(red != 0, 256, jobs.data(), (int)jobs.size(), bstream);
                    check_hip(hipGetLastError(), "grouped build launch");
                    for (BuildTicket *t : ts) check_hip(hipEventRecord(t->done, bstream), "hipEventRecord");
                } catch (const Error &e) {
                    // nothing of this launch may still read the jobs' buffers when they hear of it
                    (void)hipStreamSynchronize(bstream);
                    for (BuildTicket *t : ts) {
                        t->code = e.code;
                        t->message = e.what();
                    }
                }
            }
            if (hipEventCreateWithFlags(&pace, hipEventDisableTiming) == hipSuccess &&
                hipEventRecord(pace, bstream) != hipSuccess) {
                (void)hipEventDestroy(pace);
                pace = nullptr;
            }
            {
                std::lock_guard<std::mutex> lk(bmu);
                for (BuildTicket *t : take) t->launched = true;
                batch_launches++;
                batch_jobs += (long long)take.size();
            }
            bcv.notify_all();
            if (prev) {
                (void)hipEventSynchronize(prev);
                (void)hipEventDestroy(prev);
            }
            prev = pace;
        }
        if (prev) {
            (void)hipEventSynchronize(prev);
            (void)hipEventDestroy(prev);
        }
    }

    // mixed build of a request's grid: autotuned on first sight of its class
    // when asked, grouped with other small builds when batching, else alone.
    // Without autotuning the engine takes the occupancy model's shapes: the
    // shape table's entries are timed on an otherwise idle device, and on the
    // concurrent request stream they were slower than the model (a 512-request
    // stream 619-637 against 660-672 curves/s).
    void build(Structure &S, Grid &g, hipStream_t s) {
        BuildShape tuned;
        const bool mixed = !g.fp64() && getenv("DPLUS_PHASE64") == nullptr;
        if (mixed && autotune && !shape_lookup(g.N, S.ntiles, &tuned)) {
            std::lock_guard<std::mutex> lk(tune_mu);
            if (!shape_lookup(g.N, S.ntiles, &tuned)) {
                S.autotune(g, s, nullptr, nullptr);
                return;
            }
        }
        const double terms = (double)half_base(g.N + 1) * (double)S.nsources;
        if (!mixed || !batch_builds || terms > batch_terms) {
            S.build(g, false, s, autotune);
            return;
        }
        Structure::Pending P;
        BuildTicket t;
        t.job = S.prepare(g, S.shape_for(g, autotune).slices, P, s);
        t.reduce = P.reduce;
        check_hip(hipEventCreateWithFlags(&t.ready, hipEventDisableTiming), "hipEventCreate");
        if (hipEventCreateWithFlags(&t.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipEventDestroy(t.ready);
            throw Error(kErrGeneral, "hipEventCreate failed");
        }
        struct Events {
            BuildTicket &t;
            hipStream_t s;
            ~Events() {
                (void)hipStreamSynchronize(s);
                (void)hipEventDestroy(t.ready);
                (void)hipEventDestroy(t.done);
            }
        } ev{t, s};
        check_hip(hipEventRecord(t.ready, s), "hipEventRecord");
        {
            std::unique_lock<std::mutex> lk(bmu);
            bq.push_back(&t);
            bcv.notify_all();
            bcv.wait(lk, [&] { return t.launched; });
        }
        if (t.code) throw Error(t.code, t.message);
        check_hip(hipStreamWaitEvent(s, t.done, 0), "hipStreamWaitEvent");
        S.finish(g, P, false, s);
    }

    // admit queued requests in arrival order while a slot is free and the HBM budget holds
    void admit_locked() {
        while (!pending.empty()) {
            Request *r = pending.front();
            if ((int)admitted.size() + busy >= (int)streams.size()) break;
            if (reserved + r->need + cache_cap + arena_cap > budget) break;
            pending.pop_front();
            reserved += r->need;
            peak_reserved = std::max(peak_reserved, reserved);
            admitted.push_back(r);
            cv_work.notify_one();
        }
    }

    void work(int w) {
        hipStream_t s = streams[w];
        try {
            Device::get().require();
        } catch (...) {
        }
        set_alloc_stream(s);
        for (;;) {
            Request *r = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !admitted.empty(); });
                if (admitted.empty()) return;
                r = admitted.front();
                admitted.pop_front();
                busy++;
                max_concurrent = std::max<long long>(max_concurrent, busy);
            }
            r->slot = w;
            r->t_start = Clock::now();
            // a fixed-count request runs its whole chain from prepared buffers (as one
            // captured graph when asked); the rest -- Gauss-Kronrod, convergence-tested
            // averages, builds for the grouped batcher -- take the general path
            const bool fixed = r->oa.eps <= 0.0 && (r->oa.method == DPLUS_OA_MC || r->oa.method == DPLUS_OA_VEGAS);
            const bool batched = batch_builds && r->has_pdb && r->precision != DPLUS_PRECISION_FP64;
            if (fixed && !batched)
                run_fixed(*r, s, slots[w], graphs);
            else
                run(*r, s);
            r->t_end = Clock::now();
            {
                std::lock_guard<std::mutex> lk(mu);
                busy--;
                reserved -= r->need;
                r->done = true;
                completed++;
                admit_locked();
            }
            cv_done.notify_all();
        }
    }

    // content key of a request's PDB description: every array and scalar
    static void structure_key(Request &r) {
        const dplus_pdb_desc &d = r.desc;
        uint64_t key[2] = {0x13198A2E03707344ULL, 0x243F6A8885A308D3ULL};
        const double sc[7] = {d.sol_ed, d.c1, d.scale, d.vox_step, d.box_contrast, (double)d.flags, d.excl_contrast};
        const long long n[4] = {d.natoms, d.nion, d.nbox, d.nexcl};
        hash_bytes(key, n, sizeof n);
        hash_bytes(key, sc, sizeof sc);
        hash_bytes(key, r.xyz.data(), r.xyz.size() * sizeof(double));
        hash_bytes(key, r.ion.data(), r.ion.size() * sizeof(int));
        hash_bytes(key, r.coef9.data(), r.coef9.size() * sizeof(double));
        hash_bytes(key, r.rad.data(), r.rad.size() * sizeof(double));
        hash_bytes(key, r.box_dims.data(), r.box_dims.size() * sizeof(int));
        hash_bytes(key, r.box_com.data(), r.box_com.size() * sizeof(double));
        hash_bytes(key, r.excl_dims.data(), r.excl_dims.size() * sizeof(int));
        hash_bytes(key, r.excl_com.data(), r.excl_com.size() * sizeof(double));
        r.key[0] = key[0];
        r.key[1] = key[1];
    }

    // the resident structure of a request's PDB description: from the cache, or
    // built on the request's stream and cached (least recently used entries that
    // no running request holds are dropped past the cap)
    std::shared_ptr<Structure> structure_for(Request &r, hipStream_t s) {
        structure_key(r);
        auto same = [&](const Cached &c) { return c.key[0] == r.key[0] && c.key[1] == r.key[1]; };
        {
            std::lock_guard<std::mutex> lk(cache_mu);
            for (auto &c : cache)
                if (same(c)) {
                    c.last = ++cache_tick;
                    cache_hits++;
                    return c.s;
                }
        }
        std::shared_ptr<Structure> S(new Structure(&r.desc, s));
        const size_t bytes = S->blob.n + 4096;
        std::lock_guard<std::mutex> lk(cache_mu);
        cache_misses++;
        for (auto &c : cache)
            if (same(c)) return c.s;  // another worker built it meanwhile
        if (bytes > cache_cap) return S;
        cache.push_back({{r.key[0], r.key[1]}, S, bytes, ++cache_tick});
        cache_bytes += bytes;
        while (cache_bytes > cache_cap) {
            int victim = -1;
            for (size_t i = 0; i < cache.size(); i++)
                if (cache[i].s.use_count() == 1 && (victim < 0 || cache[i].last < cache[victim].last)) victim = (int)i;
            if (victim < 0) break;
            cache_bytes -= cache[victim].bytes;
            cache.erase(cache.begin() + victim);
        }
        return S;
    }

    void run(Request &r, hipStream_t s) {
        try {
            std::shared_ptr<Structure> S;
            std::unique_ptr<Grid> G;
            Grid *g = r.grid;
            if (r.has_pdb) {
                S = structure_for(r, s);
                G.reset(new Grid(r.grid_size, r.qmax, r.precision));
                build(*S, *G, s);
                g = G.get();
            }
            const double single[4] = {0, 0, 0, 1}, origin[3] = {0, 0, 0};
            const int one = 1;
            dplus_leaf leaf{};
            leaf.grid = (dplus_grid *)g;
            if (r.rot.empty()) {
                leaf.nrot = 1;
                leaf.rot = single;
                leaf.ntrans = &one;
                leaf.trans = origin;
            } else {
                leaf.nrot = (int)(r.rot.size() / 4);
                leaf.rot = r.rot.data();
                leaf.ntrans = r.ntrans.data();
                leaf.trans = r.trans.data();
            }
            DockSet D(1, &leaf, s);
            r.I.assign(r.q.size(), 0.0);
            Chi2Request chi;
            if (r.want_chi) {
                chi.y = r.y.data();
                chi.sigma = r.sigma.data();
                chi.scale = r.fit_scale;
                chi.constant = r.fit_constant;
            }
            orientation_average(D, (int)r.q.size(), r.q.data(), &r.oa, r.I.data(), &r.st, s,
                                r.want_chi ? &chi : nullptr);
            check_hip(hipStreamSynchronize(s), "request");
            r.chi2 = chi.chi2;
            r.chi_on_device = chi.on_device;
            r.code = 0;
        } catch (const Error &e) {
            r.code = e.code;
            r.message = e.what();
        } catch (const std::exception &e) {
            r.code = kErrGeneral;
            r.message = e.what();
        }
    }

    // results of a fixed-count request from the slot's staging
    static void finish_fixed(Request &r, const FixedAverage &F, const Slot &sl, bool graph) {
        const int nq = (int)r.q.size();
        r.I.assign(sl.h_out, sl.h_out + nq);
        r.chi2 = r.want_chi ? sl.h_out[nq] : 0.0;
        r.chi_on_device = r.want_chi;
        r.graph = graph;
        if (r.oa.method == DPLUS_OA_MC)
            r.st.dirs_used = F.total;
        else
            r.st.iterations_used = F.iters;
        r.code = 0;
    }

    // the slot's arena sized for dev_bytes of buffers and host_bytes of staged
    // uploads, emptied for the next request (its previous request has finished)
    void slot_arena(Slot &sl, size_t dev_bytes, size_t host_bytes) const {
        if (dev_bytes > sl.arena_dev.n)
            sl.arena_dev.ensure(std::min(slot_arena_max,
                                         std::max(dev_bytes, std::max<size_t>(2 * sl.arena_dev.n, (size_t)4 << 20))));
        if (host_bytes > sl.pin_cap) {
            const size_t cap = std::max(host_bytes, std::max<size_t>(2 * sl.pin_cap, (size_t)1 << 20));
            std::shared_lock<CaptureMutex> lk(capture_mutex());
            if (sl.arena_pin) (void)hipHostFree(sl.arena_pin);
            sl.arena_pin = nullptr;
            sl.pin_cap = 0;
            check_hip(hipHostMalloc((void **)&sl.arena_pin, cap, hipHostMallocDefault), "hipHostMalloc");
            sl.pin_cap = cap;
        }
        sl.arena.base = (char *)sl.arena_dev.p;
        sl.arena.size = sl.arena_dev.n;
        sl.arena.staging = sl.arena_pin;
        sl.arena.staging_size = sl.pin_cap;
        sl.arena.reset();
    }

    // pinned staging of a slot's results (nq intensities, then chi^2)
    static void slot_staging(Slot &sl, size_t n) {
        if (n <= sl.cap) return;
        std::shared_lock<CaptureMutex> lk(capture_mutex());
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        sl.h_out = nullptr;
        sl.cap = 0;
        const size_t cap = std::max<size_t>(n, 2048);
        check_hip(hipHostMalloc((void **)&sl.h_out, cap * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        sl.cap = cap;
    }

    // the slot's executable graph for `graph`: the kept graph of its key updated
    // in place, else a new instantiation that replaces it (or the least recently
    // used of the kept graphs)
    hipGraphExec_t slot_exec(Slot &sl, const std::string &key, hipGraph_t graph) {
        sl.tick++;
        auto it = std::find_if(sl.kept.begin(), sl.kept.end(), [&](const KeptGraph &k) { return k.key == key; });
        if (it != sl.kept.end()) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(it->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
                it->last = sl.tick;
                graph_updates++;
                return it->exec;
            }
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(it->exec);
            sl.kept.erase(it);
        }
        KeptGraph k;
        k.key = key;
        k.last = sl.tick;
        check_hip(hipGraphInstantiate(&k.exec, graph, nullptr, nullptr, 0), "graph instantiate");
        graph_instantiations++;
        if ((int)sl.kept.size() >= kKeptGraphs) {
            auto old = std::min_element(sl.kept.begin(), sl.kept.end(),
                                        [](const KeptGraph &a, const KeptGraph &b) { return a.last < b.last; });
            (void)hipGraphExecDestroy(old->exec);
            sl.kept.erase(old);
        }
        sl.kept.push_back(k);
        return k.exec;
    }

    // a fixed-count request: everything allocated and uploaded first, then the
    // kernel chain launched with one synchronisation at the end -- eagerly, or
    // captured into a graph that is instantiated (or an instantiation of the
    // slot updated) and launched; the same kernels and arguments either way
    void run_fixed(Request &r, hipStream_t s, Slot &sl, bool graph_mode) {
        try {
            std::shared_ptr<Structure> S;
            std::unique_ptr<Grid> G;
            Grid *g = r.grid;
            Structure::Pending P;
            BuildJob J{};
            BuildShape sh;
            bool build_in_chain = false;
            if (r.has_pdb) {
                S = structure_for(r, s);
                G.reset(new Grid(r.grid_size, r.qmax, r.precision, false));
                g = G.get();
                // the paired builder (mixed, and FP64 unless the plain form is selected)
                // joins the chain; the plain FP64 builder, the FP64-phase diagnostic and
                // an autotuning sweep build the grid before it
                const bool paired = g->fp64() ? !fp64_plain_builder() : getenv("DPLUS_PHASE64") == nullptr;
                BuildShape tuned;
                if (!paired) {
                    S->build(*g, false, s, autotune);
                } else if (!g->fp64() && autotune && !shape_lookup(g->N, S->ntiles, &tuned)) {
                    std::lock_guard<std::mutex> lk(tune_mu);
                    if (!shape_lookup(g->N, S->ntiles, &tuned))
                        S->autotune(*g, s, nullptr, nullptr);  // times the shapes and builds g with the best
                    else
                        build_in_chain = true;
                } else {
                    build_in_chain = true;
                }
                if (build_in_chain) {
                    sh = S->shape_for(*g, autotune);
                    J = S->prepare(*g, sh.slices, P, s);
                    g->prepare_splines(s);
                }
            }
            const double single[4] = {0, 0, 0, 1}, origin[3] = {0, 0, 0};
            const int one = 1;
            dplus_leaf leaf{};
            leaf.grid = (dplus_grid *)g;
            if (r.rot.empty()) {
                leaf.nrot = 1;
                leaf.rot = single;
                leaf.ntrans = &one;
                leaf.trans = origin;
            } else {
                leaf.nrot = (int)(r.rot.size() / 4);
                leaf.rot = r.rot.data();
                leaf.ntrans = r.ntrans.data();
                leaf.trans = r.trans.data();
            }
            Chi2Request chi;
            if (r.want_chi) {
                chi.y = r.y.data();
                chi.sigma = r.sigma.data();
                chi.scale = r.fit_scale;
                chi.constant = r.fit_constant;
            }
            const int nq = (int)r.q.size();
            // leaf tables, q, groups, scratch and chi^2 inputs from the slot's arena,
            // their uploads staged and sent together
            long long ntr = 0;
            for (int k = 0; k < leaf.nrot; k++) ntr += leaf.ntrans[k];
            size_t up = 0;
            const size_t need =
                oa_arena_bytes(1, leaf.nrot, ntr, nq, r.q.data(), g->step, &r.oa, r.want_chi, &up);
            // (rounded up as slot_arena grows it; past the slot's share: allocated per request)
            const bool in_arena = std::max(need, (size_t)4 << 20) <= slot_arena_max;
            if (in_arena) slot_arena(sl, need, up);
            std::unique_ptr<DockSet> Dp;
            FixedAverage F;
            {
                struct Install {
                    explicit Install(Arena *a) { set_arena(a); }
                    ~Install() { set_arena(nullptr); }
                } install(in_arena ? &sl.arena : nullptr);
                Dp.reset(new DockSet(1, &leaf, s));
                F.prepare(*Dp, nq, r.q.data(), &r.oa, r.want_chi ? &chi : nullptr, s);
            }
            if (in_arena) sl.arena.flush(s);
            const DockSet &D = *Dp;
            slot_staging(sl, (size_t)nq + 1);
            // the chain's shape: kernels (and so the graph's nodes) follow from these
            std::string key = std::to_string(r.oa.method) + "/" + std::to_string(F.iters) + "/" +
                              std::to_string(F.ppb) + (D.fp64 ? "d" : "f") + (D.single ? "s" : "r") +
                              (r.want_chi ? "c" : "-");
            // (grid sizes, launch sizes and buffers may differ within a key: an update
            // takes them; rows beyond the one-wave splines add a kernel)
            if (build_in_chain)
                key += "/b" + std::to_string(sh.wg) + (P.reduce ? "r" : "u") + (g->fp64() ? "d" : "f") +
                       (g->N > kWaveShellMax ? "L" : "S");
            if (!graph_mode) {
                if (build_in_chain) {
                    {
                        KTimer kt("build", s);
                        launch_build_batch(P.reduce, sh.wg, &J, 1, s);
                        check_hip(hipGetLastError(), "build_atoms launch");
                    }
                    S->enqueue_finish(*g, P, s, true);
                }
                F.enqueue(D, sl.h_out, sl.h_out + nq, s, true);
                check_hip(hipStreamSynchronize(s), "request");
                timing_collect();
                finish_fixed(r, F, sl, false);
                return;
            }
            hipGraph_t graph = nullptr;
            const Clock::time_point t0 = Clock::now();
            {
                // no allocation anywhere in the process while this stream captures
                std::unique_lock<CaptureMutex> lk(capture_mutex());
                check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
                // a failure inside the capture still ends it, so the stream stays usable
                struct EndOnThrow {
                    hipStream_t s;
                    bool armed = true;
                    ~EndOnThrow() {
                        if (!armed) return;
                        hipGraph_t gg = nullptr;
                        if (hipStreamEndCapture(s, &gg) == hipSuccess && gg) (void)hipGraphDestroy(gg);
                    }
                } end_on_throw{s};
                if (build_in_chain) {
                    launch_build_batch(P.reduce, sh.wg, &J, 1, s);
                    check_hip(hipGetLastError(), "build_atoms launch");
                    S->enqueue_finish(*g, P, s, false);
                }
                F.enqueue(D, sl.h_out, sl.h_out + nq, s, false);
                end_on_throw.armed = false;
                check_hip(hipStreamEndCapture(s, &graph), "end capture");
            }
            struct DropGraph {
                hipGraph_t g;
                ~DropGraph() {
                    if (g) (void)hipGraphDestroy(g);
                }
            } drop{graph};
            const Clock::time_point t1 = Clock::now();
            hipGraphExec_t exec = slot_exec(sl, key, graph);
            const Clock::time_point t2 = Clock::now();
            graph_capture_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
            graph_exec_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t2 - t1).count();
            check_hip(hipGraphLaunch(exec, s), "graph launch");
            check_hip(hipStreamSynchronize(s), "request");
            graph_launches++;
            finish_fixed(r, F, sl, true);
        } catch (const Error &e) {
            r.code = e.code;
            r.message = e.what();
        } catch (const std::exception &e) {
            r.code = kErrGeneral;
            r.message = e.what();
        }
    }
};

// ------------------------------------------------------------------ fit session
struct Session {
    hipStream_t s = nullptr;
    std::vector<Grid *> leaves;
    bool fp64 = false;
    double step = 0.0;
    int nq = 0, ng = 0, mstride = 1, nblk = 0;
    long long total = 0;
    int chunk = 0;  // MC stratification period (directions)
    unsigned long long seed = 0;
    std::vector<GroupDev> groups;
    DevBuf bgrids, bq, bg, bact, bpart, bout, bI, by, bsig, bchi, bparam, brot, btr, bz;
    size_t cap_rot = 0, cap_tr = 0;
    RotDev *h_rot = nullptr;
    double4 *h_tr = nullptr;
    ZCopyDev *h_z = nullptr;  // cap_tr records: the copies when all turn about z
    double *h_param = nullptr, *h_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nrot = -1, graph_nz = -1;
    long long evaluations = 0, captures = 0;

    ~Session() {
        if (s) (void)hipStreamSynchronize(s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : {(void *)h_rot, (void *)h_tr, (void *)h_z, (void *)h_param, (void *)h_out})
            if (p) host_free(p);
        bgrids.release(), bq.release(), bg.release(), bact.release(), bpart.release(), bout.release();
        bI.release(), by.release(), bsig.release(), bchi.release(), bparam.release(), brot.release(), btr.release();
        bz.release();
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }

    static void host_alloc(void **p, size_t bytes) {
        std::shared_lock<CaptureMutex> lk(capture_mutex());
        check_hip(hipHostMalloc(p, bytes, hipHostMallocDefault), "hipHostMalloc");
    }
    static void host_free(void *p) {
        std::shared_lock<CaptureMutex> lk(capture_mutex());
        (void)hipHostFree(p);
    }

    void grow(size_t nrot, size_t ntr) {
        if (nrot > cap_rot) {
            if (h_rot) host_free(h_rot);
            cap_rot = std::max<size_t>(nrot, 2 * cap_rot);
            host_alloc((void **)&h_rot, cap_rot * sizeof(RotDev));
            brot.ensure(cap_rot * sizeof(RotDev));
            graph_nrot = -1;
        }
        if (ntr > cap_tr) {
            if (h_tr) host_free(h_tr);
            if (h_z) host_free(h_z);
            cap_tr = std::max<size_t>(ntr, 2 * cap_tr);
            host_alloc((void **)&h_tr, cap_tr * sizeof(double4));
            host_alloc((void **)&h_z, cap_tr * sizeof(ZCopyDev));
            btr.ensure(cap_tr * sizeof(double4));
            bz.ensure(cap_tr * sizeof(ZCopyDev));
            graph_nrot = -1;
        }
    }

    void capture(int nrot, int nz) {
        if (exec) (void)hipGraphExecDestroy(exec), exec = nullptr;
        if (graph) (void)hipGraphDestroy(graph), graph = nullptr;
        // no allocation anywhere in the process while this stream captures
        std::unique_lock<CaptureMutex> lk(capture_mutex());
        check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        // a failure inside the capture still ends it, so the stream stays usable
        struct EndOnThrow {
            hipStream_t s;
            bool armed = true;
            ~EndOnThrow() {
                if (!armed) return;
                hipGraph_t g = nullptr;
                if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            }
        } end_on_throw{s};
        check_hip(hipMemcpyAsync(brot.p, h_rot, cap_rot * sizeof(RotDev), hipMemcpyHostToDevice, s), "H2D");
        check_hip(hipMemcpyAsync(btr.p, h_tr, cap_tr * sizeof(double4), hipMemcpyHostToDevice, s), "H2D");
        if (nz > 0)
            check_hip(hipMemcpyAsync(bz.p, h_z, cap_tr * sizeof(ZCopyDev), hipMemcpyHostToDevice, s), "H2D");
        check_hip(hipMemcpyAsync(bparam.p, h_param, 2 * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
        launch_mc(fp64, false, (const void *const *)bgrids.p, (const RotDev *)brot.p, nrot, (const double4 *)btr.p,
                  (const ZCopyDev *)bz.p, nz, (const GroupDev *)bg.p, (const int *)bact.p, ng, mstride, (const double *)bq.p, step, seed, 0, nblk,
                  chunk, (double *)bpart.p, (double *)bout.p, s);
        launch_mc_finish((const GroupDev *)bg.p, ng, mstride, (const double *)bout.p, (double)total, (double *)bI.p, s);
        launch_chi2_param((const double *)bI.p, (const double *)by.p, (const double *)bsig.p, nq,
                          (const double *)bparam.p, (double *)bchi.p, s);
        check_hip(hipMemcpyAsync(h_out, bI.p, nq * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        check_hip(hipMemcpyAsync(h_out + nq, bchi.p, sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        hipGraph_t gph = nullptr;
        end_on_throw.armed = false;
        const hipError_t e = hipStreamEndCapture(s, &gph);
        check_hip(e, "end capture");
        graph = gph;
        check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        graph_nrot = nrot;
        graph_nz = nz;
        captures++;
    }
};

}  // namespace dplus

using namespace dplus;

#define ENGINE_GUARD(...)                                   \
    try {                                                   \
        __VA_ARGS__;                                        \
        return kOk;                                         \
    } catch (const Error &e) {                              \
        return set_error(e.code, e.what());                 \
    } catch (const std::exception &e) {                     \
        return set_error(kErrGeneral, e.what());            \
    } catch (...) {                                         \
        return set_error(kErrGeneral, "unknown exception"); \
    }

extern "C" {

int dplus_engine_create(const dplus_engine_config *cfg, dplus_engine **out) {
    ENGINE_GUARD({
        if (!out) throw Error(kErrInvalidArgs, "null output");
        const long long budget = cfg ? cfg->hbm_budget_bytes : 0;
        const int streams = cfg && cfg->streams > 0 ? cfg->streams : 16;
        *out = (dplus_engine *)new Engine(budget, streams, cfg && cfg->batch_builds != 0, cfg ? cfg->batch_terms : 0,
                                          cfg && cfg->autotune != 0, cfg && cfg->graphs != 0);
    })
}

void dplus_engine_destroy(dplus_engine *e) { delete (Engine *)e; }

int dplus_engine_submit(dplus_engine *eh, const dplus_request *rq, long long *ticket) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e || !rq || !ticket) throw Error(kErrInvalidArgs, "null argument");
        if (rq->nq <= 0 || !rq->q) throw Error(kErrInvalidArgs, "a request needs q values");
        std::unique_ptr<Request> r(new Request());
        if (rq->pdb) {
            const dplus_pdb_desc &d = *rq->pdb;
            if (d.natoms < 0 || d.nion <= 0 || d.nbox < 0 || d.nexcl < 0)
                throw Error(kErrInvalidArgs, "bad PDB description");
            r->has_pdb = true;
            r->xyz.assign(d.xyz, d.xyz + 3 * (size_t)d.natoms);
            r->ion.assign(d.ion, d.ion + d.natoms);
            r->coef9.assign(d.coef9, d.coef9 + 9 * (size_t)d.nion);
            r->rad.assign(d.rad, d.rad + d.nion);
            if (d.nbox > 0) {
                r->box_dims.assign(d.box_dims, d.box_dims + 3 * (size_t)d.nbox);
                r->box_com.assign(d.box_com, d.box_com + 3 * (size_t)d.nbox);
            }
            if (d.nexcl > 0) {
                r->excl_dims.assign(d.excl_dims, d.excl_dims + 3 * (size_t)d.nexcl);
                r->excl_com.assign(d.excl_com, d.excl_com + 3 * (size_t)d.nexcl);
            }
            r->desc = d;
            r->desc.xyz = r->xyz.data();
            r->desc.ion = r->ion.data();
            r->desc.coef9 = r->coef9.data();
            r->desc.rad = r->rad.data();
            r->desc.box_dims = d.nbox > 0 ? r->box_dims.data() : nullptr;
            r->desc.box_com = d.nbox > 0 ? r->box_com.data() : nullptr;
            r->desc.excl_dims = d.nexcl > 0 ? r->excl_dims.data() : nullptr;
            r->desc.excl_com = d.nexcl > 0 ? r->excl_com.data() : nullptr;
            // the structure cache's key is hashed by the worker that runs the request
            r->grid_size = rq->grid_size;
            r->qmax = rq->qmax;
            r->precision = rq->precision;
            if (r->grid_size < 2 || !(r->qmax > 0)) throw Error(kErrInvalidArgs, "bad grid size or qmax");
        } else if (rq->grid) {
            r->grid = (Grid *)rq->grid;
        } else {
            throw Error(kErrInvalidArgs, "a request needs a PDB description or a resident grid");
        }
        if (rq->nrot > 0) {
            if (!rq->rot || !rq->ntrans || !rq->trans) throw Error(kErrInvalidArgs, "placements need rot/ntrans/trans");
            r->rot.assign(rq->rot, rq->rot + 4 * (size_t)rq->nrot);
            r->ntrans.assign(rq->ntrans, rq->ntrans + rq->nrot);
            size_t nt = 0;
            for (int i = 0; i < rq->nrot; i++) {
                if (rq->ntrans[i] < 0) throw Error(kErrInvalidArgs, "negative translation count");
                nt += (size_t)rq->ntrans[i];
            }
            r->trans.assign(rq->trans, rq->trans + 3 * nt);
        }
        r->q.assign(rq->q, rq->q + rq->nq);
        for (int i = 1; i < rq->nq; i++)
            if (r->q[i] < r->q[i - 1]) throw Error(kErrInvalidArgs, "q must be ascending");
        if (rq->data) {
            if (!rq->sigma) throw Error(kErrInvalidArgs, "chi^2 needs sigma with the data");
            r->want_chi = true;
            r->y.assign(rq->data, rq->data + rq->nq);
            r->sigma.assign(rq->sigma, rq->sigma + rq->nq);
            r->fit_scale = rq->fit_scale;
            r->fit_constant = rq->fit_constant;
        }
        r->oa = rq->oa;
        r->need = request_bytes(*r);
        if (r->need + e->cache_cap + e->arena_cap > e->budget)
            throw Error(kErrOutOfMemory, "the request needs " + std::to_string(r->need) +
                                             " device bytes, more than the engine's budget of " +
                                             std::to_string(e->budget));
        r->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(e->mu);
        r->id = e->next_id++;
        *ticket = r->id;
        e->pending.push_back(r.get());
        e->all[r->id] = std::move(r);
        e->submitted++;
        e->admit_locked();
    })
}

int dplus_engine_poll(dplus_engine *eh, long long ticket, int *done) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e || !done) throw Error(kErrInvalidArgs, "null argument");
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->all.find(ticket);
        if (it == e->all.end()) throw Error(kErrJobNotFound, "unknown ticket " + std::to_string(ticket));
        *done = it->second->done ? 1 : 0;
    })
}

int dplus_engine_wait(dplus_engine *eh, long long ticket, double *intensity, dplus_request_result *res) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e) throw Error(kErrInvalidArgs, "null engine");
        std::unique_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(e->mu);
            auto it = e->all.find(ticket);
            if (it == e->all.end()) throw Error(kErrJobNotFound, "unknown ticket " + std::to_string(ticket));
            Request *rp = it->second.get();
            e->cv_done.wait(lk, [&] { return rp->done; });
            r = std::move(it->second);
            e->all.erase(it);
        }
        if (res) {
            memset(res, 0, sizeof *res);
            res->queue_ms = ms_between(r->t_submit, r->t_start);
            res->run_ms = ms_between(r->t_start, r->t_end);
            res->latency_ms = ms_between(r->t_submit, r->t_end);
            res->chi2 = r->chi2;
            res->chi2_on_device = r->chi_on_device ? 1 : 0;
            res->reserved_bytes = (long long)r->need;
            res->slot = r->slot;
            res->oa = r->st;
            res->graph = r->graph ? 1 : 0;
        }
        if (r->code) throw Error(r->code, r->message);
        if (intensity) std::copy(r->I.begin(), r->I.end(), intensity);
    })
}

int dplus_engine_stats(dplus_engine *eh, long long *submitted, long long *completed, long long *peak_reserved,
                       long long *budget, long long *max_concurrent) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e) throw Error(kErrInvalidArgs, "null engine");
        std::lock_guard<std::mutex> lk(e->mu);
        if (submitted) *submitted = e->submitted;
        if (completed) *completed = e->completed;
        if (peak_reserved) *peak_reserved = (long long)e->peak_reserved;
        if (budget) *budget = (long long)e->budget;
        if (max_concurrent) *max_concurrent = e->max_concurrent;
    })
}

int dplus_engine_cache_stats(dplus_engine *eh, long long *hits, long long *misses, long long *bytes,
                             long long *capacity) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e) throw Error(kErrInvalidArgs, "null engine");
        std::lock_guard<std::mutex> lk(e->cache_mu);
        if (hits) *hits = e->cache_hits;
        if (misses) *misses = e->cache_misses;
        if (bytes) *bytes = (long long)e->cache_bytes;
        if (capacity) *capacity = (long long)e->cache_cap;
    })
}

int dplus_engine_batch_stats(dplus_engine *eh, long long *launches, long long *builds) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e) throw Error(kErrInvalidArgs, "null engine");
        std::lock_guard<std::mutex> lk(e->bmu);
        if (launches) *launches = e->batch_launches;
        if (builds) *builds = e->batch_jobs;
    })
}

int dplus_engine_graph_stats(dplus_engine *eh, long long *launches, long long *instantiations, long long *updates,
                             double *capture_ms, double *exec_ms) {
    ENGINE_GUARD({
        Engine *e = (Engine *)eh;
        if (!e) throw Error(kErrInvalidArgs, "null engine");
        if (launches) *launches = e->graph_launches.load();
        if (instantiations) *instantiations = e->graph_instantiations.load();
        if (updates) *updates = e->graph_updates.load();
        if (capture_ms) *capture_ms = 1e-6 * (double)e->graph_capture_ns.load();
        if (exec_ms) *exec_ms = 1e-6 * (double)e->graph_exec_ns.load();
    })
}

int dplus_session_create(int nleaf, dplus_grid *const *leaves, int nq, const double *q, const double *data,
                         const double *sigma, unsigned long long seed, long long directions, dplus_session **out) {
    ENGINE_GUARD({
        if (nleaf <= 0 || !leaves || nq <= 0 || !q || !data || !sigma || !out)
            throw Error(kErrInvalidArgs, "bad session arguments");
        Device::get().require();
        std::unique_ptr<Session> S(new Session());
        check_hip(hipStreamCreate(&S->s), "hipStreamCreate");
        set_alloc_stream(S->s);
        struct Restore {
            ~Restore() { set_alloc_stream(nullptr); }
        } restore;
        std::vector<const void *> gp;
        for (int l = 0; l < nleaf; l++) {
            Grid *g = (Grid *)leaves[l];
            if (!g) throw Error(kErrInvalidArgs, "null leaf grid");
            if (l == 0) {
                S->fp64 = g->fp64();
                S->step = g->step;
            } else if (g->fp64() != S->fp64 || g->step != S->step || g->N != ((Grid *)leaves[0])->N) {
                throw Error(kErrInvalidArgs, "session leaves must share grid size, qmax and precision");
            }
            S->leaves.push_back(g);
            gp.push_back(g->dev.p);
        }
        for (int i = 1; i < nq; i++)
            if (q[i] < q[i - 1]) throw Error(kErrInvalidArgs, "q must be ascending");
        // the orientation average's range checks, before anything is allocated or
        // captured: a q past shell grid_size/2 + 1 would make the four-plane lookup
        // read shells beyond the grid
        if (!(q[0] >= 0)) throw Error(kErrInvalidArgs, "q must be non-negative");
        {
            const Grid *g0 = S->leaves[0];
            if (!(q[nq - 1] <= (g0->grid_size / 2 + 1) * g0->step)) throw Error(kErrInvalidArgs, "q beyond the grid");
        }
        S->nq = nq;
        S->seed = seed;
        S->groups = make_groups(q, nq, S->step);
        S->ng = (int)S->groups.size();
        for (auto &g : S->groups) S->mstride = std::max(S->mstride, g.m);
        S->total = ((std::max<long long>(256, directions) + 255) / 256) * 256;
        S->nblk = (int)(S->total / 256);
        // the orientation average's default chunk of 8192 directions when it
        // divides the count (so a session equals that eager average), else one
        // chunk of all directions
        if (S->total % 8192 != 0 && S->total > (1LL << 30))
            throw Error(kErrInvalidArgs, "a session's direction count above 2^30 must be a multiple of 8192");
        S->chunk = (int)(S->total % 8192 == 0 ? 8192 : S->total);
        std::vector<int> act(S->ng);
        for (int i = 0; i < S->ng; i++) act[i] = i;
        hipStream_t s = S->s;
        S->bgrids.upload(gp.data(), gp.size() * sizeof(void *), s);
        S->bq.upload(q, nq * sizeof(double), s);
        S->bg.upload(S->groups.data(), S->groups.size() * sizeof(GroupDev), s);
        S->bact.upload(act.data(), act.size() * sizeof(int), s);
        S->bpart.ensure((size_t)S->ng * S->nblk * S->mstride * sizeof(double));
        S->bout.ensure((size_t)S->ng * S->mstride * sizeof(double));
        S->bI.ensure((size_t)nq * sizeof(double));
        S->by.upload(data, nq * sizeof(double), s);
        S->bsig.upload(sigma, nq * sizeof(double), s);
        S->bchi.ensure(sizeof(double));
        S->bparam.ensure(2 * sizeof(double));
        S->host_alloc((void **)&S->h_param, 2 * sizeof(double));
        S->host_alloc((void **)&S->h_out, (nq + 1) * sizeof(double));
        check_hip(hipStreamSynchronize(s), "session setup");
        *out = (dplus_session *)S.release();
    })
}

int dplus_session_evaluate(dplus_session *sh, const dplus_leaf *placements, int nleaf, double fit_scale,
                           double fit_constant, double *intensity, double *chi2) {
    ENGINE_GUARD({
        Session *S = (Session *)sh;
        if (!S || !placements || nleaf != (int)S->leaves.size()) throw Error(kErrInvalidArgs, "bad evaluation arguments");
        Device::get().require();
        set_alloc_stream(S->s);
        struct Restore {
            ~Restore() { set_alloc_stream(nullptr); }
        } restore;
        size_t nrot = 0, ntr = 0;
        for (int l = 0; l < nleaf; l++) {
            if (placements[l].grid && (Grid *)placements[l].grid != S->leaves[l])
                throw Error(kErrInvalidArgs, "placement leaf differs from the session's leaf");
            nrot += (size_t)placements[l].nrot;
            for (int r = 0; r < placements[l].nrot; r++) {
                if (placements[l].ntrans[r] < 0) throw Error(kErrInvalidArgs, "negative translation count");
                ntr += (size_t)placements[l].ntrans[r];
            }
        }
        if (nrot == 0) throw Error(kErrInvalidArgs, "no placements");
        S->grow(nrot, std::max<size_t>(1, ntr));
        size_t ri = 0, ti = 0;
        for (int l = 0; l < nleaf; l++) {
            const dplus_leaf &L = placements[l];
            size_t t0 = 0;
            for (int r = 0; r < L.nrot; r++) {
                RotDev &rd = S->h_rot[ri++];
                euler_rt(L.rot[4 * r], L.rot[4 * r + 1], L.rot[4 * r + 2], rd.m);
                rd.scale = L.rot[4 * r + 3];
                rd.leaf = l;
                rd.t0 = (int)ti;
                rd.nt = L.ntrans[r];
                rd.pad = 0;
                for (int t = 0; t < L.ntrans[r]; t++, t0++)
                    S->h_tr[ti++] = make_double4(L.trans[3 * t0], L.trans[3 * t0 + 1], L.trans[3 * t0 + 2], 0.0);
            }
        }
        S->h_param[0] = fit_scale;
        S->h_param[1] = fit_constant;
        // copies that all turn about z run the copy-major MC kernel (a graph of its own)
        std::vector<ZCopyDev> zc;
        int nz = 0;
        if (z_copies(S->h_rot, (int)nrot, S->h_tr, zc) && (int)zc.size() >= kZCopyMin) {
            nz = (int)zc.size();
            std::copy(zc.begin(), zc.end(), S->h_z);
        }
        if (S->graph_nrot != (int)nrot || S->graph_nz != nz) S->capture((int)nrot, nz);
        check_hip(hipGraphLaunch(S->exec, S->s), "graph launch");
        check_hip(hipStreamSynchronize(S->s), "session evaluation");
        if (intensity) std::copy(S->h_out, S->h_out + S->nq, intensity);
        if (chi2) *chi2 = S->h_out[S->nq];
        S->evaluations++;
    })
}

int dplus_session_info(dplus_session *sh, long long *evaluations, long long *graph_captures) {
    ENGINE_GUARD({
        Session *S = (Session *)sh;
        if (!S) throw Error(kErrInvalidArgs, "null session");
        if (evaluations) *evaluations = S->evaluations;
        if (graph_captures) *graph_captures = S->captures;
    })
}

void dplus_session_destroy(dplus_session *s) {
    Session *S = (Session *)s;
    if (!S) return;
    set_alloc_stream(S->s);
    delete S;
    set_alloc_stream(nullptr);
}

}  // extern "C"
