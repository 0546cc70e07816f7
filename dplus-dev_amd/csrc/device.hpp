// device.hpp -- host-side objects of libdplus_mi355x shared by the C ABI,
// the orientation-average drivers and the backend: device handle, device
// buffers, grids and docked-leaf sets, plus the records the kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <thread>
#include <string>
#include <vector>

#include <condition_variable>
#include <mutex>
#include <shared_mutex>

#include "../../include/dplus_mi355x.h"
#include "common.hpp"

namespace dplus {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

int set_error(int code, const std::string &msg);
std::string json_escape(const std::string &s);
void check_hip(hipError_t e, const char *what);

struct Device {
    bool ok = false;
    int id = 0;
    std::string name;
    int cus = 0;
    size_t mem = 0;
    hipStream_t stream = nullptr;
    static Device &get();
    void require();

   private:
    Device();
};

// Stream that device buffers of the calling thread are allocated and released
// on (stream-ordered allocation from the device memory pool): the library's
// stream unless the thread has set its own, as each worker of the request
// engine does.  Frees never synchronise the device, so work of concurrent
// requests on other streams keeps running.
hipStream_t alloc_stream();
void set_alloc_stream(hipStream_t s);
// Graph capture against allocation: a thread capturing a stream holds this
// exclusively from begin to end of the capture; stream-ordered and pinned
// allocations and frees take it shared, so none runs during another thread's
// capture (HIP rejects them there).  A waiting capture goes first: once a
// thread asks for it exclusively, new shared holders wait, so a capture is
// not held back by a steady stream of allocations on other threads.
class CaptureMutex {
   public:
    void lock() {
        std::unique_lock<std::mutex> lk(m_);
        writers_waiting_++;
        cv_.wait(lk, [&] { return !writer_ && readers_ == 0; });
        writers_waiting_--;
        writer_ = true;
    }
    void unlock() {
        {
            std::lock_guard<std::mutex> lk(m_);
            writer_ = false;
        }
        cv_.notify_all();
    }
    void lock_shared() {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return !writer_ && writers_waiting_ == 0; });
        readers_++;
    }
    void unlock_shared() {
        bool last;
        {
            std::lock_guard<std::mutex> lk(m_);
            last = --readers_ == 0;
        }
        if (last) cv_.notify_all();
    }

   private:
    std::mutex m_;
    std::condition_variable cv_;
    int readers_ = 0, writers_waiting_ = 0;
    bool writer_ = false;
};
CaptureMutex &capture_mutex();

// owning device allocation that only grows (or, while the calling thread has
// a scratch arena installed, a region of that arena)
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;  // stream the allocation is ordered on
    bool borrowed = false;    // p lies in a scratch arena: nothing to free
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void release();
    void ensure(size_t bytes);
    void upload(const void *src, size_t bytes, hipStream_t s);
};

// Scratch arena (the allocation-free entry point, the request engine's
// slots): while a thread has one installed, DevBuf::ensure carves its buffers
// from [base, base + size) in call order (256-byte aligned) instead of
// allocating, failing with ERROR_INVALID_ARGS when the arena is exhausted, and
// release() frees nothing.  With a pinned staging buffer, DevBuf::upload
// copies into staging instead of issuing a copy, and flush() sends the staged
// bytes with one copy per run of uploads that are contiguous on the device
// (nothing reads them on the device before flush).
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    char *staging = nullptr;  // pinned host memory of staging_size bytes, or none
    size_t staging_size = 0, staged = 0;
    struct Run {
        size_t dev, host, bytes;
    };
    std::vector<Run> runs;
    void reset() {
        used = staged = 0;
        runs.clear();
    }
    // stage bytes for device offset dev (false: no room, copy directly)
    bool stage(size_t dev, const void *src, size_t bytes);
    void flush(hipStream_t s);
};
void set_arena(Arena *a);  // nullptr: allocate again
// device bytes of a Monte Carlo or VEGAS average's buffers (docked leaves,
// q, groups, partial sums, bins, chi^2 inputs) carved from an arena, each
// 256-byte aligned; host bytes they upload
size_t oa_arena_bytes(int nleaf, long long nrot, long long ntr, int nq, const double *q, double step,
                      const dplus_oa_params *p, bool chi = false, size_t *upload_bytes = nullptr);
// the allocation-free orientation average's device array (scratch.hip): its
// size, the address of this device's copy (looked up on first use, with the
// device current), and the lock a call holds while it uses it
size_t scratch_capacity();
void *scratch_address(int device);
std::mutex &scratch_lock(int device);
// zero n doubles on the stream with a fill kernel
void launch_fill_zero(double *p, size_t n, hipStream_t s);
// device buffers allocated so far by DevBuf (pool allocations, not arena regions)
long long device_allocations();

// Rotation record of a docked leaf: m = R^T row-major (R = Rx(a) Ry(b) Rz(c) for
// Euler angles, Amplitude.h:557-573, or a composition of such), the amplitude
// scale, translations [t0, t0+nt) of the translation table, leaf = grid index.
struct RotDev {
    double m[9];
    double scale;
    int leaf, t0, nt, pad;
};

// R^T of R = Rx(a) Ry(b) Rz(c), row-major
void euler_rt(double a, double b, double c, double m[9]);

// One placed copy of a leaf turned about z alone (R = Rz(alpha)): R^T q has the
// polar angle of q and the azimuth phi - alpha, so the copy's lookup is the
// leaf's grid at (theta, phi - alpha).  One record per translation; alpha in
// [0, 2 pi).  aux of the first record: the largest shell whose row segments
// fit the copy-major kernel's collapse window for every chunk of 64 records
// (the shell groups up to it take the collapsed lookup).
struct ZCopyDev {
    double alpha;
    double t[3];
    double scale;
    int leaf, aux;
};

// the copies of rotation records that all turn about z alone, one per
// translation, ordered by (leaf, alpha) so that neighbouring records look up
// neighbouring nodes of the same theta rows, with out[0].aux set; false (out
// untouched) when a record has another rotation
bool z_copies(const RotDev *rots, int nrot, const double4 *tr, std::vector<ZCopyDev> &out);

// the copy-major MC kernel takes a docking whose copies all turn about z when
// it has at least this many of them (a wave's lanes are its copies)
constexpr int kZCopyMin = 32;

// one shell group: q-values q0..q0+m-1 lie in (layer*step, (layer+1)*step];
// uniform = 1 when they are equally spaced (every q_j within 1e-12 of
// q_0 + j (q_{m-1} - q_0) / (m - 1), relative to the largest), so the mixed
// kernels step a translation's phase from one q to the next
struct GroupDev {
    int layer, q0, m, uniform;
};

// the 15 Kronrod abscissae on [-1, 1] with the Kronrod and Gauss weights
// (in device memory for the integrand kernel)
struct GkNodes {
    double x[15];
    double kw[15];
    double gw[7];
};

struct LeafGeo {
    int N;
    double step;
};

// Gauss-Kronrod phi recursion on the device: a node of the phi trees of one
// batch of theta tasks (roots first, 15 per theta task) and an open phi interval
struct GkNode {
    double value;
    int parent;  // -1 for a root
    int child0, child1;
    int pending;  // children not yet resolved
};
struct GkInterval {
    double pmin, pmax;
    int node, maxd, mind, pad;
};

struct Grid {
    int grid_size = 0, N = 0, precision = DPLUS_PRECISION_MIXED;
    double qmax = 0, step = 0;
    long long P = 0;
    DevBuf dev;  // P nodes of (Re, Im, dRe, dIm) in float4 or double4
    std::vector<double> host_data, host_interp;
    // zero = false leaves the nodes unset: for a grid a build fills at once (a
    // build writes every node and every spline coefficient)
    Grid(int gs, double qmax, int precision, bool zero = true);
    bool fp64() const { return precision == DPLUS_PRECISION_FP64; }
    size_t node_bytes() const { return fp64() ? sizeof(double4) : sizeof(float4); }
    void ensure_host();
    // spline coefficients from the node values (timed: HIP-event timing, never
    // while capturing); prepare_splines makes the multiplier tables resident
    // and keeps their address, so a captured splines() takes no lock and
    // uploads nothing
    void splines(hipStream_t s, bool timed = true);
    void prepare_splines(hipStream_t s);
    const void *spline_tables = nullptr;
    void download(hipStream_t s);
    void upload(hipStream_t s);
};

// one placed copy of a leaf: amplitude scale * A_grid(R^T q) e^{i q.t};
// rt = R^T row-major
struct Placement {
    Grid *grid = nullptr;
    double rt[9];
    double t[3];
    double scale = 1.0;
};

// leaves of one orientation average, flattened for the kernels
struct DockSet {
    int grid_size = 0, nrot = 0;
    double qmax = 0, step = 0;
    bool fp64 = false;
    // one leaf placed once, unrotated, at the origin: the orientation-average
    // kernels look up the drawn angles directly
    bool single = false;
    // nz > 0: every copy turns about z alone; zcb holds the nz ZCopyDev records
    // (z_copies) for the copy-major MC kernel
    int nz = 0;
    DevBuf grids, geos, rotb, trb, zcb;
    DockSet(int nleaf, const dplus_leaf *leaves, hipStream_t s);
    // placements sharing grid, rotation and scale become one rotation record
    // with several translations
    DockSet(const std::vector<Placement> &pl, hipStream_t s);

   private:
    void finish(std::vector<Grid *> &gl, std::vector<RotDev> &rots, std::vector<double4> &tr, hipStream_t s);
};

// run fn(i) for i in [0, n) on up to 16 host threads (dynamic chunks)
template <typename F>
void parallel_for(long long n, F fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 2 || nt == 1) {
        for (long long i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<long long> next(0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&]() {
            for (long long i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto &x : th) x.join();
}

// HIP-event timing of one kernel family on a stream (active when enabled)
struct KTimer {
    const char *name;
    hipStream_t s;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KTimer(const char *n, hipStream_t st);
    ~KTimer();
};
void timing_collect();
void timing_set(bool on);
void timing_clear();
void timing_read(const char *name, long long *n, double *ms);
void microbench_valu(hipStream_t s, double *fma32, double *sin32, double *cos32, double *fma64);
void copy_probe(hipStream_t s, long long bytes, double *ms);

// atoms and solvent boxes of one PDB leaf, sorted, tiled and resident in HBM
// One grid's share of a paired-builder launch (read by the kernels from the
// launch arguments): its resident atoms/tiles, per-shell form factors, slice
// partials and output grid, with the tile range of each atom slice.
constexpr int kMaxSlices = 8;
constexpr int kMaxBuildJobs = 16;
struct BuildJob {
    const float4 *atoms;
    const double4 *atoms64;  // FP64 jobs: the sources in double
    const int4 *tiles;
    const double *ff;
    const int4 *bdims;
    const double *bcon;  // contrast of each box shape
    double *partial;
    float4 *grid;
    double4 *grid64;  // FP64 jobs: the grid's nodes
    int fp64;         // 1: FP64 phases, sincos and sums into grid64
    long long nunits;
    double vox, step, scale;
    int nion, nslices;
    int xblocks;    // workgroups per slice (set by the launcher)
    int wg_begin;   // first builder workgroup of the job (set by the launcher)
    int fin_begin;  // first finishing workgroup of the job (set by the launcher)
    int slice_begin[kMaxSlices + 1];
};
struct BuildBatch {
    int njobs;
    BuildJob job[kMaxBuildJobs];
};
// launch shape of the paired builder: workgroup size (64, 128, 256 or 512
// units) and atom slices (1..kMaxSlices)
struct BuildShape {
    int wg = 256;
    int slices = 1;
};

struct Structure {
    dplus_pdb_desc desc{};
    std::vector<double> coef9, rad;
    std::vector<int> ion_counts;
    double box_origin = 0.0;  // sum over box sets of contrast * total volume (the boxes' A(0))
    double rmax = 0.0;  // largest |position| of an atom or box centre (nm)
    bool atoms_on = false, boxes_on = false;
    int ntiles = 0;
    std::vector<int> tile_len;  // terms of each tile
    long long nsources = 0;
    // device copies, immutable after construction: one allocation holding the
    // sources in float and double, the tile table, box shapes and contrasts
    DevBuf blob;
    const float4 *d32 = nullptr;
    const double4 *d64 = nullptr;
    const int4 *dtiles = nullptr, *ddims = nullptr;
    const double *dcon = nullptr;
    Structure(const dplus_pdb_desc *d, hipStream_t s);
    // use_table: the launch shape from the autotuner's table where it has the grid's
    // class (shapes timed on a device to themselves), else the occupancy model's
    void build(Grid &g, bool host_copy, hipStream_t s, bool use_table = true);
    // The mixed build in three steps, for callers that launch the builder of
    // several grids together: prepare uploads this grid's form factors and
    // sizes the slice partials (buffers owned by `bufs`, used on stream s) and
    // returns the builder's job; after the builder ran, finish sets the origin,
    // takes the splines and drains s.
    struct Pending {
        DevBuf ff, part;
        double origin = 0.0;
        bool reduce = false;
    };
    BuildJob prepare(Grid &g, int slices, Pending &bufs, hipStream_t s) const;
    void finish(Grid &g, const Pending &bufs, bool host_copy, hipStream_t s) const;
    // finish's kernels alone (origin, splines): no upload, allocation or
    // synchronisation once g.prepare_splines ran, so they can be captured
    void enqueue_finish(Grid &g, const Pending &bufs, hipStream_t s, bool timed) const;
    // launch shape of the paired builder for grid g: the tuned one when the
    // shape table holds it, else the occupancy model's
    BuildShape shape_for(const Grid &g, bool use_table = true) const;
    // times every launch shape on grid g, records the fastest in the shape
    // table and returns it (ms: its time and the occupancy model's)
    BuildShape autotune(Grid &g, hipStream_t s, double *best_ms, double *model_ms);
};

std::vector<GroupDev> make_groups(const double *q, int nq, double step);
// paired builder: work units of a grid with nhalf canonical nodes, the number
// of atom slices the occupancy model splits a launch of workgroup size wg over
// nunits units into, and one launch over several grids' jobs (njobs <=
// kMaxBuildJobs, the same phase form `reduce` and workgroup size)
long long pair_units(long long nhalf);
int pair_slices(long long nunits, int ntiles, int wg, bool fp64 = false);
// one launch of up to kMaxBuildJobs jobs, all mixed or all FP64 (jobs[0].fp64)
void launch_build_batch(bool reduce, int wg, BuildJob *jobs, int njobs, hipStream_t s);
// FP64 grids: the paired builder (default) or the plain one node at a time
bool fp64_plain_builder();
void set_fp64_plain_builder(bool plain);
// shape table of the autotuner, keyed by (grid shells, tile-count class)
bool shape_lookup(int nshells, int ntiles, BuildShape *out);
void shape_record(int nshells, int ntiles, BuildShape shape);
void pdb_build(Grid &g, const dplus_pdb_desc *d, bool host_copy, hipStream_t s);
// optional goodness of fit of an average: chi2 = sum ((scale I + constant - y) / sigma)^2
// over q with sigma > 0; computed on the device from the device-resident
// intensities when the average has a fixed count (eps <= 0, MC or VEGAS),
// else from the host intensities
struct Chi2Request {
    const double *y = nullptr, *sigma = nullptr;  // host arrays, nq each
    double scale = 1.0, constant = 0.0;
    double chi2 = 0.0;                            // result
    bool on_device = false;                       // result: where it was computed
};
void orientation_average(const DockSet &D, int nq, const double *q, const dplus_oa_params *p, double *out,
                         dplus_oa_stats *st, hipStream_t s, Chi2Request *chi = nullptr);
// the q checks of an orientation average: ascending, non-negative, at most
// grid_size/2 + 1 shells
void check_q_range(const DockSet &D, int nq, const double *q);
// A fixed-count average (Monte Carlo or VEGAS with eps <= 0) in two steps:
// prepare allocates and uploads everything it needs on stream s; enqueue then
// issues its kernels and the download of I(q) (and chi^2) with no allocation
// and no host synchronisation, so a caller can capture enqueue into a HIP
// graph.  orientation_average runs every fixed-count average through it, so a
// captured average gives the bits of an eager one.
struct FixedAverage {
    int nq = 0, ng = 0, mstride = 1, method = 0, iters = 0, ppb = 0, chunk = 0, nblk = 0;
    long long total = 0;
    uint64_t seed = 0;
    bool chi = false;
    double chi_scale = 1.0, chi_constant = 0.0;
    DevBuf bq, bg, bact, bpart, bout, bI, bb, bsum, bvar, bh, by, bsig, bchi;
    void prepare(const DockSet &D, int nq, const double *q, const dplus_oa_params *p, const Chi2Request *c,
                 hipStream_t s);
    // I: nq doubles, chi2: one double (with a chi^2 request), written when the
    // stream reaches the downloads (pinned host memory for a captured graph);
    // timed: HIP-event kernel timing (never while capturing)
    void enqueue(const DockSet &D, double *I, double *chi2, hipStream_t s, bool timed) const {
        enqueue_kernels(D, s, timed);
        download(I, chi2, s);
    }
    // the kernels alone, and the downloads alone
    void enqueue_kernels(const DockSet &D, hipStream_t s, bool timed) const;
    void download(double *I, double *chi2, hipStream_t s) const;
    // device intensities the kernels leave
    const double *result() const {
        return method == DPLUS_OA_MC ? (const double *)bI.p : (const double *)bh.p + (size_t)(iters - 1) * nq;
    }
};
void gk_average(const DockSet &D, int nq, const double *q, double eps, long long max_depth, double *out,
                dplus_oa_stats *st, hipStream_t s);
// azimuthal averages of |A|^2 at npts (q, theta) points over nphi stratified phi
// (the 2-D generate path); q outside the grid (beyond grid_size/2 + 1 shells) is rejected
void azimuthal_average(const DockSet &D, int npts, const double *q, const double *theta, int nphi, uint64_t seed,
                       double *out, hipStream_t s);
// Gauss-Kronrod levels are evaluated on the device in batches of at most this
// many phi intervals (32 B task + 16 B sums each)
constexpr size_t kGkBatch = (size_t)1 << 20;
// device phi recursion: at most kGkBatch open intervals per level and
// kGkNodes tree nodes per theta level (beyond them the average runs its
// recursion on the host, with the same results)
constexpr size_t kGkNodes = (size_t)1 << 23;
// upper bound of the device memory of one Gauss-Kronrod average over nq q values:
// two interval levels (task + record), the Kronrod/Gauss sums, the node arena,
// the phi integrals of a theta level and the counters
inline size_t gk_device_bytes(size_t nq) {
    return kGkBatch * (2 * (32 + sizeof(GkInterval)) + 16 + 8) + kGkNodes * sizeof(GkNode) + nq * 32 + 8192;
}
// space-filling lattice of a docked grid: basis vectors v[3k..3k+2] (nm) and
// repetitions n[k]; on = 0 for none
struct LatticeDev {
    double v[9];
    double n[3];
    int on;
};
void dock_grid(const DockSet &D, Grid &out, double scale, hipStream_t s, const LatticeDev *lattice = nullptr);
// basis of a space-filling symmetry from distances (nm), angles (degrees) and repetitions
LatticeDev space_filling_lattice(const double *distances, const double *angles_deg, const double *reps);
void synth_walk(unsigned long long seed, int n, double step, double min_sep, double ax, double ay, double az,
                double *xyz);

}  // namespace dplus
