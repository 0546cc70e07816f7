// gk.cpp -- adaptive Gauss-Kronrod orientation average on the GPU.
//
// DomainModel::GaussKron2DSphereRecurs / GaussLegTheta / GaussLegPhi
// (Backend/Backend/Amplitude.cpp:1416-1467) nest an adaptive G7/K15 rule in
// phi inside one in theta over the Friedel half sphere theta in [0, pi/2]
// and divide by 2 pi.  The recursion is replayed level by level here: every
// open phi interval of every q and every theta node is evaluated in one batch
// of integrand calls on the device (the host sends one task per interval, the
// device forms its 15 points and their Kronrod and Gauss sums, in the
// reference's arithmetic); accept/split decisions and the sums left + right in
// the reference's order are taken on the host.
// The host work of a level is split over threads at theta-task boundaries (a
// theta task's open phi intervals are contiguous and no other task touches its
// trees), and the threads' new intervals are concatenated in order, so the
// result does not depend on the thread count.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <vector>

#include "device.hpp"

namespace dplus {

void launch_docked_cart(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                        const double4 *trans, int n, const double *qxyz, double *out, hipStream_t s);
void launch_docked_gk(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                      const double4 *trans, long long ntasks, const double4 *tasks, const GkNodes *nodes,
                      double *out, hipStream_t s);
void launch_gk_decide(long long n, const GkInterval *cur, const double *kvgv, double eps, GkNode *nodes, int *cnt,
                      int cap_int, int cap_nodes, const double4 *task, double4 *next_task, GkInterval *next,
                      hipStream_t s);
void launch_gk_tree_sums(GkNode *nodes, int lo, int hi, int nroots, double *phival, hipStream_t s);

// Gauss-Kronrod 15/7 tables (Common/mathfuncs.h:100-125)
static const double kX[15] = {
    -0.9914553711208126392068547, -0.9491079123427585245261897, -0.8648644233597690727897128,
    -0.7415311855993944398638648, -0.5860872354676911302941448, -0.4058451513773971669066064,
    -0.2077849550078984676006894, 0.0,
    0.2077849550078984676006894,  0.4058451513773971669066064,  0.5860872354676911302941448,
    0.7415311855993944398638648,  0.8648644233597690727897128,  0.9491079123427585245261897,
    0.9914553711208126392068547};
static const double kW[15] = {
    0.0229353220105292249637320, 0.0630920926299785532907007, 0.1047900103222501838398763,
    0.1406532597155259187451896, 0.1690047266392679028265834, 0.1903505780647854099132564,
    0.2044329400752988924141620, 0.2094821410847278280129992, 0.2044329400752988924141620,
    0.1903505780647854099132564, 0.1690047266392679028265834, 0.1406532597155259187451896,
    0.1047900103222501838398763, 0.0630920926299785532907007, 0.0229353220105292249637320};
static const double gW[7] = {0.1294849661688696932706114, 0.2797053914892766679014678,
                             0.3818300505051189449503698, 0.4179591836734693877551020,
                             0.3818300505051189449503698, 0.2797053914892766679014678,
                             0.1294849661688696932706114};

namespace {
// summation tree: a node's value is its own estimate (leaf) or left + right
struct Node {
    double value = 0.0;
    int parent = -1, child[2] = {-1, -1}, pending = 0;
};

struct Tree {
    std::vector<Node> n;
    int add(int parent) {
        n.push_back(Node());
        n.back().parent = parent;
        return (int)n.size() - 1;
    }
    // resolve node v with value x; returns the root index if the root resolved
    int resolve(int v, double x) {
        n[v].value = x;
        while (true) {
            const int p = n[v].parent;
            if (p < 0) return v;
            if (--n[p].pending > 0) return -1;
            n[p].value = n[n[p].child[0]].value + n[n[p].child[1]].value;
            v = p;
        }
    }
};

struct ThetaTask {
    int q;
    double tmin, tmax;
    long long maxd, mind;
    int node;          // node in the theta tree of q
    double phival[15]; // phi integrals of the 15 theta nodes
    double qs[15], qc[15];  // q sin(theta_k), q cos(theta_k) of the 15 theta nodes
    int left = 15;
};

struct PhiTask {
    int theta_task, k;  // theta task and its theta node
    double pmin, pmax;
    long long maxd, mind;
    int node;  // node in the phi tree
    int tree;  // node pool of the phi trees (one per theta task, holding its 15 trees)
};

// growable array of phi tasks without value-initialisation (millions per level)
struct TaskBuf {
    std::unique_ptr<PhiTask[]> p;
    size_t n = 0, cap = 0;
    void reserve(size_t c) {
        if (c <= cap) return;
        c = std::max(c, 2 * cap);
        std::unique_ptr<PhiTask[]> q(new PhiTask[c]);
        if (n) std::copy(p.get(), p.get() + n, q.get());
        p.swap(q);
        cap = c;
    }
    void push_back(const PhiTask &t) {
        reserve(n + 1);
        p[n++] = t;
    }
    PhiTask &operator[](size_t i) { return p[i]; }
    const PhiTask &operator[](size_t i) const { return p[i]; }
    bool empty() const { return n == 0; }
    size_t size() const { return n; }
};

// pinned host staging buffer (grows; DMA without a pageable bounce)
template <typename V>
struct Pinned {
    V *p = nullptr;
    size_t cap = 0;
    V *get(size_t n) {
        if (n > cap) {
            std::shared_lock<CaptureMutex> lk(capture_mutex());
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            check_hip(hipHostMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(V), hipHostMallocDefault),
                      "hipHostMalloc");
            cap = n;
        }
        return p;
    }
    ~Pinned() {
        if (p) {
            std::shared_lock<CaptureMutex> lk(capture_mutex());
            (void)hipHostFree(p);
        }
    }
};

// worker threads that run fn(0..n-1) with the caller taking part
class Pool {
   public:
    explicit Pool(int n) {
        for (int i = 1; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return (int)th_.size() + 1; }
    void run(int n, const std::function<void(int)> &fn) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            njob_ = n;
            next_ = 0;
            remaining_ = n;
        }
        cv_.notify_all();
        std::unique_lock<std::mutex> lk(mu_);
        while (next_ < njob_) {
            const int k = next_++;
            lk.unlock();
            fn(k);
            lk.lock();
            --remaining_;
        }
        done_.wait(lk, [&] { return remaining_ == 0; });
        job_ = nullptr;
    }

   private:
    void loop() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || (job_ && next_ < njob_); });
            if (stop_) return;
            const int k = next_++;
            const std::function<void(int)> *fn = job_;
            lk.unlock();
            (*fn)(k);
            lk.lock();
            if (--remaining_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int njob_ = 0, next_ = 0, remaining_ = 0;
    bool stop_ = false;
};
}  // namespace

// q sin(theta_k), q cos(theta_k) of the 15 Kronrod nodes of a theta interval
static void theta_nodes(ThetaTask &t, double qv) {
    const double ba2 = (t.tmax - t.tmin) * 0.5, apb2 = (t.tmax + t.tmin) * 0.5;
    for (int k = 0; k < 15; k++) {
        const double th = apb2 + ba2 * kX[k];
        t.qs[k] = qv * std::sin(th);
        t.qc[k] = qv * std::cos(th);
    }
}

// Kronrod and Gauss sums of a theta interval over its 15 phi integrals (GaussLegTheta, Amplitude.cpp:1424-1432)
static void theta_sums(const ThetaTask &T, double &kv, double &gv) {
    const double ba2 = (T.tmax - T.tmin) * 0.5, apb2 = (T.tmax + T.tmin) * 0.5;
    kv = 0.0;
    gv = 0.0;
    for (int i = 0; i < 15; i++) {
        const double sT = std::sin(apb2 + ba2 * kX[i]);
        kv += T.phival[i] * kW[i] * ba2 * sT;
        if (i % 2 == 1) gv += T.phival[i] * gW[i / 2] * ba2 * sT;
    }
}

// |A(0)|^2 of the q = 0 entries (Amplitude.cpp:1479-1484); returns the evaluations
static long long gk_zero_q(const DockSet &D, int nq, const double *q, double *out, hipStream_t s) {
    std::vector<int> zero_q;
    for (int i = 0; i < nq; i++)
        if (q[i] == 0.0) zero_q.push_back(i);
    if (zero_q.empty()) return 0;
    std::vector<double> pts(3 * zero_q.size(), 0.0), vals(zero_q.size());
    DevBuf bq, bo;
    bq.upload(pts.data(), pts.size() * sizeof(double), s);
    bo.ensure(vals.size() * sizeof(double));
    KTimer kt("gk", s);
    launch_docked_cart(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p, (const RotDev *)D.rotb.p,
                       D.nrot, (const double4 *)D.trb.p, (int)vals.size(), (const double *)bq.p, (double *)bo.p, s);
    check_hip(hipGetLastError(), "gk launch");
    check_hip(hipMemcpyAsync(vals.data(), bo.p, vals.size() * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    check_hip(hipStreamSynchronize(s), "gk batch");
    for (size_t k = 0; k < zero_q.size(); k++) out[zero_q[k]] = vals[k];
    return (long long)vals.size();
}

// The recursion with every phi interval on the device: the host keeps the theta
// intervals (their nodes' sines and cosines, decisions and tree); each theta
// level's 15 phi trees per interval are grown level by level on the device
// (docked_gk_batch evaluates, gk_decide accepts and splits; gk_tree_sums adds
// each tree's values level by level from the leaves), and their integrals come
// back once per theta level.  Every sum is the host replay's:
// the same phi intervals, Kronrod/Gauss sums, decisions and left + right tree
// sums, so results and evaluation counts are equal.  Returns false when a level
// outgrows the device arenas (the caller then runs the host replay).
static bool gk_average_device(const DockSet &D, int nq, const double *q, double eps, long long max_depth,
                              double *out, long long &evals, hipStream_t s) {
    const int cap_int = (int)kGkBatch, cap_nodes = (int)kGkNodes;
    std::vector<Tree> thetaTrees(nq);
    std::vector<ThetaTask> level;
    for (int i = 0; i < nq; i++) {
        if (q[i] == 0.0) continue;
        ThetaTask t;
        t.q = i;
        t.tmin = 0.0;
        t.tmax = kPi / 2.0;
        t.maxd = max_depth - 1;
        t.mind = -1;
        t.node = thetaTrees[i].add(-1);
        theta_nodes(t, q[i]);
        level.push_back(t);
    }
    if (level.empty()) return true;
    GkNodes nodes;
    for (int i = 0; i < 15; i++) {
        nodes.x[i] = kX[i];
        nodes.kw[i] = kW[i];
    }
    for (int i = 0; i < 7; i++) nodes.gw[i] = gW[i];
    DevBuf bnodes, btask[2], biv[2], bkv, barena, bphi, bcnt;
    bnodes.upload(&nodes, sizeof nodes, s);
    for (int b = 0; b < 2; b++) {
        btask[b].ensure((size_t)cap_int * sizeof(double4));
        biv[b].ensure((size_t)cap_int * sizeof(GkInterval));
    }
    bkv.ensure((size_t)cap_int * 2 * sizeof(double));
    barena.ensure((size_t)cap_nodes * sizeof(GkNode));
    bphi.ensure((size_t)cap_int * sizeof(double));
    bcnt.ensure(4 * sizeof(int));
    // pinned staging, kept by the thread across averages (an engine worker runs many)
    static thread_local Pinned<int> hcnt;
    static thread_local Pinned<double> hphi;
    static thread_local Pinned<double4> hroot_t;
    static thread_local Pinned<GkInterval> hroot_i;
    static thread_local Pinned<GkNode> hroot_n;
    int *cnt = hcnt.get(4);
    while (!level.empty()) {
        const size_t T = level.size(), nroots = 15 * T;
        if (nroots > (size_t)cap_int) return false;
        double4 *rt = hroot_t.get(nroots);
        GkInterval *ri = hroot_i.get(nroots);
        GkNode *rn = hroot_n.get(nroots);
        for (size_t t = 0; t < T; t++)
            for (int k = 0; k < 15; k++) {
                const size_t r = 15 * t + k;
                const double pmin = 0.0, pmax = k2Pi;
                rt[r] = make_double4(level[t].qs[k], level[t].qc[k], (pmax + pmin) * 0.5, (pmax - pmin) * 0.5);
                ri[r] = GkInterval{pmin, pmax, (int)r, (int)level[t].maxd, (int)level[t].mind, 0};
                rn[r] = GkNode{0.0, -1, -1, -1, 0};
            }
        cnt[0] = 0;
        cnt[1] = (int)nroots;
        cnt[2] = 0;
        int cur = 0;
        check_hip(hipMemcpyAsync(btask[0].p, rt, nroots * sizeof(double4), hipMemcpyHostToDevice, s), "H2D");
        check_hip(hipMemcpyAsync(biv[0].p, ri, nroots * sizeof(GkInterval), hipMemcpyHostToDevice, s), "H2D");
        check_hip(hipMemcpyAsync(barena.p, rn, nroots * sizeof(GkNode), hipMemcpyHostToDevice, s), "H2D");
        check_hip(hipMemcpyAsync(bcnt.p, cnt, 3 * sizeof(int), hipMemcpyHostToDevice, s), "H2D");
        long long n = (long long)nroots;
        std::vector<int> bounds(1, 0);  // node ranges of the levels: roots first
        bounds.push_back((int)nroots);
        while (n > 0) {
            {
                KTimer kt("gk", s);
                launch_docked_gk(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p,
                                 (const RotDev *)D.rotb.p, D.nrot, (const double4 *)D.trb.p, n,
                                 (const double4 *)btask[cur].p, (const GkNodes *)bnodes.p, (double *)bkv.p, s);
                launch_gk_decide(n, (const GkInterval *)biv[cur].p, (const double *)bkv.p, eps, (GkNode *)barena.p,
                                 (int *)bcnt.p, cap_int, cap_nodes, (const double4 *)btask[cur].p,
                                 (double4 *)btask[1 - cur].p, (GkInterval *)biv[1 - cur].p, s);
                check_hip(hipGetLastError(), "gk launch");
            }
            check_hip(hipMemcpyAsync(cnt, bcnt.p, 3 * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
            check_hip(hipStreamSynchronize(s), "gk level");
            evals += 15 * n;
            if (cnt[2]) return false;
            n = cnt[0];
            bounds.push_back(cnt[1]);
            cnt[0] = 0;
            check_hip(hipMemcpyAsync(bcnt.p, cnt, sizeof(int), hipMemcpyHostToDevice, s), "H2D");
            cur = 1 - cur;
        }
        // tree sums from the deepest level to the roots
        for (size_t b = bounds.size() - 1; b-- > 0;)
            launch_gk_tree_sums((GkNode *)barena.p, bounds[b], bounds[b + 1], (int)nroots, (double *)bphi.p, s);
        check_hip(hipGetLastError(), "gk tree sums");
        double *phi = hphi.get(nroots);
        check_hip(hipMemcpyAsync(phi, bphi.p, nroots * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        check_hip(hipStreamSynchronize(s), "gk phi integrals");
        // theta decisions of this level, in order
        std::vector<ThetaTask> next;
        for (size_t t = 0; t < T; t++) {
            ThetaTask &Tk = level[t];
            for (int k = 0; k < 15; k++) Tk.phival[k] = phi[15 * t + k];
            double kv, gv;
            theta_sums(Tk, kv, gv);
            Tree &tr = thetaTrees[Tk.q];
            if (Tk.mind < 0 && (Tk.maxd < 0 || std::fabs(1.0 - (gv / kv)) <= eps || gv == kv)) {
                const int r = tr.resolve(Tk.node, kv);
                if (r >= 0) out[Tk.q] = tr.n[r].value / (2.0 * kPi);
            } else {
                const double mid = (Tk.tmax + Tk.tmin) / 2.0;
                const int l = tr.add(Tk.node), rgt = tr.add(Tk.node);
                tr.n[Tk.node].child[0] = l;
                tr.n[Tk.node].child[1] = rgt;
                tr.n[Tk.node].pending = 2;
                for (int h = 0; h < 2; h++) {
                    ThetaTask c;
                    c.q = Tk.q;
                    c.tmin = h == 0 ? Tk.tmin : mid;
                    c.tmax = h == 0 ? mid : Tk.tmax;
                    c.maxd = Tk.maxd - 1;
                    c.mind = Tk.mind - 1;
                    c.node = h == 0 ? l : rgt;
                    theta_nodes(c, q[Tk.q]);
                    next.push_back(c);
                }
            }
        }
        level.swap(next);
    }
    return true;
}

// the host replay (below) serves when the device arenas would overflow, when a
// depth exceeds the device records' int range, or on request (DPLUS_GK_HOST)
static void gk_average_host(const DockSet &D, int nq, const double *q, double eps, long long max_depth,
                            double *out, dplus_oa_stats *st, hipStream_t s);

void gk_average(const DockSet &D, int nq, const double *q, double eps, long long max_depth, double *out,
                dplus_oa_stats *st, hipStream_t s) {
    if (getenv("DPLUS_GK_HOST") == nullptr && max_depth <= (long long)INT_MAX) {
        long long evals = gk_zero_q(D, nq, q, out, s);
        if (gk_average_device(D, nq, q, eps, max_depth, out, evals, s)) {
            if (st) st->evaluations = evals;
            return;
        }
    }
    gk_average_host(D, nq, q, eps, max_depth, out, st, s);
}

static void gk_average_host(const DockSet &D, int nq, const double *q, double eps, long long max_depth,
                            double *out, dplus_oa_stats *st, hipStream_t s) {
    std::vector<Tree> thetaTrees(nq);
    std::vector<ThetaTask> tt;
    std::vector<Tree> phiTrees;
    TaskBuf open, next;
    long long evals = 0;
    std::vector<int> zero_q;

    auto open_theta = [&](int qi, double a, double b, long long maxd, long long mind, int node) {
        ThetaTask t;
        t.q = qi;
        t.tmin = a;
        t.tmax = b;
        t.maxd = maxd;
        t.mind = mind;
        t.node = node;
        tt.push_back(t);  // its q sin / q cos of the theta nodes: fill_nodes
        const int id = (int)tt.size() - 1;
        phiTrees.push_back(Tree());
        Tree &pt = phiTrees.back();
        pt.n.reserve(64);
        for (int k = 0; k < 15; k++) {
            const int root = pt.add(-1);
            open.push_back({id, k, 0.0, k2Pi, maxd, mind, root, (int)phiTrees.size() - 1});
        }
    };

    Pool pool(std::max(1, std::min(16, (int)std::thread::hardware_concurrency())));
    // q sin(theta_k), q cos(theta_k) of the 15 nodes of theta tasks [from, end)
    auto fill_nodes = [&](size_t from) {
        const size_t n = tt.size() - from;
        const int parts = (int)std::max<size_t>(1, std::min<size_t>((size_t)pool.size(), n / 256));
        pool.run(parts, [&](int part) {
            for (size_t id = from + n * part / parts; id < from + n * (part + 1) / parts; id++) {
                ThetaTask &t = tt[id];
                const double ba2 = (t.tmax - t.tmin) * 0.5, apb2 = (t.tmax + t.tmin) * 0.5;
                for (int k = 0; k < 15; k++) {
                    const double th = apb2 + ba2 * kX[k];
                    t.qs[k] = q[t.q] * std::sin(th);
                    t.qc[k] = q[t.q] * std::cos(th);
                }
            }
        });
    };
    for (int i = 0; i < nq; i++) {
        if (q[i] == 0.0) {
            zero_q.push_back(i);
            continue;
        }
        const int root = thetaTrees[i].add(-1);
        open_theta(i, 0.0, kPi / 2.0, max_depth - 1, -1, root);
    }
    fill_nodes(0);
    // q = 0: |A(0)|^2 (Amplitude.cpp:1479-1484)
    if (!zero_q.empty()) {
        std::vector<double> pts(3 * zero_q.size(), 0.0), vals(zero_q.size());
        DevBuf bq, bo;
        bq.upload(pts.data(), pts.size() * sizeof(double), s);
        bo.ensure(vals.size() * sizeof(double));
        KTimer kt("gk", s);
        launch_docked_cart(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p,
                           (const RotDev *)D.rotb.p, D.nrot, (const double4 *)D.trb.p, (int)vals.size(),
                           (const double *)bq.p, (double *)bo.p, s);
        check_hip(hipGetLastError(), "gk launch");
        check_hip(hipMemcpyAsync(vals.data(), bo.p, vals.size() * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        check_hip(hipStreamSynchronize(s), "gk batch");
        evals += (long long)vals.size();
        for (size_t k = 0; k < zero_q.size(); k++) out[zero_q[k]] = vals[k];
    }

    GkNodes nodes;
    for (int i = 0; i < 15; i++) {
        nodes.x[i] = kX[i];
        nodes.kw[i] = kW[i];
    }
    for (int i = 0; i < 7; i++) nodes.gw[i] = gW[i];
    DevBuf bnodes, btask, bo;  // nodes in device memory: the kernel indexes them per lane
    bnodes.upload(&nodes, sizeof nodes, s);
    Pinned<double4> htask;
    Pinned<double> hval;
    std::vector<unsigned char> split;
    // [begin, end) ranges of `open` for the host threads, cut at theta-task boundaries
    auto ranges = [&]() {
        const size_t n = open.size();
        const int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)pool.size(), n / 8192));
        std::vector<size_t> cut(1, 0);
        for (int k = 1; k < nt; k++) {
            size_t b = std::max(cut.back(), n * (size_t)k / (size_t)nt);
            while (b > 0 && b < n && open[b].theta_task == open[b - 1].theta_task) b++;
            cut.push_back(b);
        }
        cut.push_back(n);
        return cut;
    };
    while (!open.empty()) {
        const std::vector<size_t> cut = ranges();
        const int nt = (int)cut.size() - 1;
        const size_t n = open.size();
        // integrand at the 15 Kronrod nodes of every open phi interval: the
        // interval's task (q sin theta, q cos theta, centre, half width)
        double4 *tasks = htask.get(n);
        pool.run(nt, [&](int part) {
            for (size_t t = cut[part]; t < cut[part + 1]; t++) {
                const PhiTask &P = open[t];
                const ThetaTask &T = tt[P.theta_task];
                tasks[t] = make_double4(T.qs[P.k], T.qc[P.k], (P.pmax + P.pmin) * 0.5, (P.pmax - P.pmin) * 0.5);
            }
        });
        // the device returns each interval's Kronrod and Gauss sums
        double *vals = hval.get(2 * n);
        {
            // device batches of at most kGkBatch intervals, so a level's device
            // memory is bounded (gk_device_bytes) however many intervals it holds
            const size_t nb = std::min<size_t>(n, kGkBatch);
            btask.ensure(nb * sizeof(double4));
            bo.ensure(2 * nb * sizeof(double));
            KTimer kt("gk", s);
            for (size_t b0 = 0; b0 < n; b0 += nb) {
                const size_t m = std::min(nb, n - b0);
                check_hip(hipMemcpyAsync(btask.p, tasks + b0, m * sizeof(double4), hipMemcpyHostToDevice, s), "H2D");
                launch_docked_gk(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p,
                                 (const RotDev *)D.rotb.p, D.nrot, (const double4 *)D.trb.p, (long long)m,
                                 (const double4 *)btask.p, (const GkNodes *)bnodes.p, (double *)bo.p, s);
                check_hip(hipGetLastError(), "gk launch");
                check_hip(hipMemcpyAsync(vals + 2 * b0, bo.p, 2 * m * sizeof(double), hipMemcpyDeviceToHost, s),
                          "D2H");
            }
            check_hip(hipStreamSynchronize(s), "gk batch");
            evals += 15 * (long long)n;
        }
        // pass 1: accept or split each interval, and the split count of every range
        split.resize(n);
        std::vector<size_t> nsplit(nt + 1, 0), nthetas(nt, 0);
        pool.run(nt, [&](int part) {
            size_t c = 0;
            for (size_t t = cut[part]; t < cut[part + 1]; t++) {
                if (t == cut[part] || open[t].theta_task != open[t - 1].theta_task) nthetas[part]++;
                const PhiTask &P = open[t];
                const double kv = vals[2 * t], gv = vals[2 * t + 1];
                const bool accept = P.mind < 0 && (P.maxd < 0 || std::fabs(1.0 - (gv / kv)) <= eps || gv == kv);
                split[t] = accept ? 0 : 1;
                c += accept ? 0 : 1;
            }
            nsplit[part + 1] = c;
        });
        for (int k = 0; k < nt; k++) nsplit[k + 1] += nsplit[k];
        // pass 2: resolve accepted intervals up their trees; write the halves of
        // split ones at their place in the next level (range offsets keep the order)
        // room for the new intervals of theta tasks that split after this level (30 each)
        // so that appending them does not move the level's intervals
        size_t nthe = 0;
        for (int k = 0; k < nt; k++) nthe += nthetas[k];
        next.n = 0;
        next.reserve(2 * nsplit[nt] + 30 * nthe);
        next.n = 2 * nsplit[nt];
        std::vector<std::vector<int>> fin_part(nt);
        pool.run(nt, [&](int part) {
            size_t w = 2 * nsplit[part];
            std::vector<int> &finished_theta = fin_part[part];
            for (size_t t = cut[part]; t < cut[part + 1]; t++) {
                const PhiTask &P = open[t];
                Tree &pt = phiTrees[P.tree];
                if (!split[t]) {
                    const int r = pt.resolve(P.node, vals[2 * t]);
                    if (r >= 0) {
                        ThetaTask &T = tt[P.theta_task];
                        T.phival[P.k] = pt.n[r].value;
                        if (--T.left == 0) finished_theta.push_back(P.theta_task);
                    }
                } else {
                    const double apb2 = (P.pmax + P.pmin) * 0.5;
                    const int l = pt.add(P.node), rgt = pt.add(P.node);
                    pt.n[P.node].child[0] = l;
                    pt.n[P.node].child[1] = rgt;
                    pt.n[P.node].pending = 2;
                    next[w++] = {P.theta_task, P.k, P.pmin, apb2, P.maxd - 1, P.mind - 1, l, P.tree};
                    next[w++] = {P.theta_task, P.k, apb2, P.pmax, P.maxd - 1, P.mind - 1, rgt, P.tree};
                }
            }
        });
        std::swap(open, next);
        // theta intervals whose 15 phi integrals are complete, in interval order:
        // their Kronrod and Gauss sums in parallel, then the decisions in order
        std::vector<int> fin;
        for (int part = 0; part < nt; part++) fin.insert(fin.end(), fin_part[part].begin(), fin_part[part].end());
        std::vector<double> fkv(fin.size()), fgv(fin.size());
        {
            const int parts = (int)std::max<size_t>(1, std::min<size_t>((size_t)pool.size(), fin.size() / 64));
            pool.run(parts, [&](int part) {
                for (size_t f = fin.size() * part / parts; f < fin.size() * (part + 1) / parts; f++) {
                    const ThetaTask &T = tt[fin[f]];
                    const double ba2 = (T.tmax - T.tmin) * 0.5, apb2 = (T.tmax + T.tmin) * 0.5;
                    double kv = 0.0, gv = 0.0;
                    for (int i = 0; i < 15; i++) {
                        const double sT = std::sin(apb2 + ba2 * kX[i]);
                        kv += T.phival[i] * kW[i] * ba2 * sT;
                        if (i % 2 == 1) gv += T.phival[i] * gW[i / 2] * ba2 * sT;
                    }
                    fkv[f] = kv;
                    fgv[f] = gv;
                }
            });
        }
        const size_t tt_before = tt.size();
        for (size_t f = 0; f < fin.size(); f++) {
            {
                const ThetaTask T = tt[fin[f]];
                const double kv = fkv[f], gv = fgv[f];
                Tree &tr = thetaTrees[T.q];
                if (T.mind < 0 && (T.maxd < 0 || std::fabs(1.0 - (gv / kv)) <= eps || gv == kv)) {
                    const int r = tr.resolve(T.node, kv);
                    if (r >= 0) out[T.q] = tr.n[r].value / (2.0 * kPi);
                } else {
                    const double mid = (T.tmax + T.tmin) / 2.0;
                    const int l = tr.add(T.node), rgt = tr.add(T.node);
                    tr.n[T.node].child[0] = l;
                    tr.n[T.node].child[1] = rgt;
                    tr.n[T.node].pending = 2;
                    open_theta(T.q, T.tmin, mid, T.maxd - 1, T.mind - 1, l);
                    open_theta(T.q, mid, T.tmax, T.maxd - 1, T.mind - 1, rgt);
                }
            }
        }
        if (tt.size() > tt_before) fill_nodes(tt_before);
    }
    if (st) st->evaluations = evals;
}

}  // namespace dplus
