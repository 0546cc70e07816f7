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
#include <cmath>
#include <thread>
#include <vector>

#include "device.hpp"

namespace dplus {

void launch_docked_cart(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                        const double4 *trans, int n, const double *qxyz, double *out, hipStream_t s);
void launch_docked_gk(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                      const double4 *trans, long long ntasks, const double4 *tasks, const GkNodes *nodes,
                      double *out, hipStream_t s);

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
    int left = 15;
};

struct PhiTask {
    int theta_task, k;  // theta task and its theta node
    double pmin, pmax;
    long long maxd, mind;
    int node;  // node in the phi tree
    int tree;  // which phi tree (theta_task * 15 + k)
};
}  // namespace

void gk_average(const DockSet &D, int nq, const double *q, double eps, long long max_depth, double *out,
                dplus_oa_stats *st, hipStream_t s) {
    std::vector<Tree> thetaTrees(nq);
    std::vector<ThetaTask> tt;
    std::vector<Tree> phiTrees;
    std::vector<PhiTask> open;
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
        tt.push_back(t);
        const int id = (int)tt.size() - 1;
        for (int k = 0; k < 15; k++) {
            phiTrees.push_back(Tree());
            Tree &pt = phiTrees.back();
            const int root = pt.add(-1);
            open.push_back({id, k, 0.0, k2Pi, maxd, mind, root, (int)phiTrees.size() - 1});
        }
    };

    for (int i = 0; i < nq; i++) {
        if (q[i] == 0.0) {
            zero_q.push_back(i);
            continue;
        }
        const int root = thetaTrees[i].add(-1);
        open_theta(i, 0.0, kPi / 2.0, max_depth - 1, -1, root);
    }
    // q = 0: |A(0)|^2 (Amplitude.cpp:1479-1484)
    std::vector<double> pts;
    std::vector<double> vals;
    DevBuf bq, bo;
    auto evaluate = [&](size_t npts) {
        bq.upload(pts.data(), npts * 3 * sizeof(double), s);
        bo.ensure(npts * sizeof(double));
        KTimer kt("gk", s);
        launch_docked_cart(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p,
                           (const RotDev *)D.rotb.p, D.nrot, (const double4 *)D.trb.p, (int)npts,
                           (const double *)bq.p, (double *)bo.p, s);
        check_hip(hipGetLastError(), "gk launch");
        vals.resize(npts);
        check_hip(hipMemcpyAsync(vals.data(), bo.p, npts * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
        check_hip(hipStreamSynchronize(s), "gk batch");
        evals += (long long)npts;
    };
    if (!zero_q.empty()) {
        pts.assign(3 * zero_q.size(), 0.0);
        evaluate(zero_q.size());
        for (size_t k = 0; k < zero_q.size(); k++) out[zero_q[k]] = vals[k];
    }

    GkNodes nodes;
    for (int i = 0; i < 15; i++) {
        nodes.x[i] = kX[i];
        nodes.kw[i] = kW[i];
    }
    for (int i = 0; i < 7; i++) nodes.gw[i] = gW[i];
    DevBuf bnodes;  // in device memory: the kernel indexes it per lane
    bnodes.upload(&nodes, sizeof nodes, s);
    std::vector<double4> tasks;
    DevBuf btask;
    const int hw = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    // [begin, end) ranges of `open` for the host threads, cut at theta-task boundaries
    auto ranges = [&]() {
        const size_t n = open.size();
        const int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)hw, n / 8192));
        std::vector<size_t> cut(1, 0);
        for (int k = 1; k < nt; k++) {
            size_t b = std::max(cut.back(), n * (size_t)k / (size_t)nt);
            while (b > 0 && b < n && open[b].theta_task == open[b - 1].theta_task) b++;
            cut.push_back(b);
        }
        cut.push_back(n);
        return cut;
    };
    auto parallel = [&](const std::vector<size_t> &cut, const auto &fn) {
        const int nt = (int)cut.size() - 1;
        if (nt == 1) {
            fn(0, cut[0], cut[1]);
            return;
        }
        std::vector<std::thread> th;
        for (int k = 1; k < nt; k++) th.emplace_back([&, k] { fn(k, cut[k], cut[k + 1]); });
        fn(0, cut[0], cut[1]);
        for (auto &t : th) t.join();
    };
    while (!open.empty()) {
        const std::vector<size_t> cut = ranges();
        // integrand at the 15 Kronrod nodes of every open phi interval: the
        // interval's task (q sin theta, q cos theta, centre, half width)
        tasks.resize(open.size());
        parallel(cut, [&](int, size_t b, size_t e) {
            for (size_t t = b; t < e; t++) {
                const PhiTask &P = open[t];
                const ThetaTask &T = tt[P.theta_task];
                const double ba2 = (T.tmax - T.tmin) * 0.5, apb2 = (T.tmax + T.tmin) * 0.5;
                const double th = apb2 + ba2 * kX[P.k];
                const double qq = q[T.q];
                tasks[t] = make_double4(qq * std::sin(th), qq * std::cos(th), (P.pmax + P.pmin) * 0.5,
                                        (P.pmax - P.pmin) * 0.5);
            }
        });
        {
            // the device returns each interval's Kronrod and Gauss sums
            const size_t npts = open.size() * 15, nout = open.size() * 2;
            btask.upload(tasks.data(), tasks.size() * sizeof(double4), s);
            bo.ensure(nout * sizeof(double));
            KTimer kt("gk", s);
            launch_docked_gk(D.fp64, (const void *const *)D.grids.p, (const LeafGeo *)D.geos.p,
                             (const RotDev *)D.rotb.p, D.nrot, (const double4 *)D.trb.p, (long long)open.size(),
                             (const double4 *)btask.p, (const GkNodes *)bnodes.p, (double *)bo.p, s);
            check_hip(hipGetLastError(), "gk launch");
            vals.resize(nout);
            check_hip(hipMemcpyAsync(vals.data(), bo.p, nout * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
            check_hip(hipStreamSynchronize(s), "gk batch");
            evals += (long long)npts;
        }
        const int nt = (int)cut.size() - 1;
        std::vector<std::vector<PhiTask>> next_part(nt);
        std::vector<std::vector<int>> fin_part(nt);
        parallel(cut, [&](int part, size_t b, size_t e) {
            std::vector<PhiTask> &next = next_part[part];
            std::vector<int> &finished_theta = fin_part[part];
            for (size_t t = b; t < e; t++) {
                const PhiTask &P = open[t];
                const double apb2 = (P.pmax + P.pmin) * 0.5;
                const double kv = vals[2 * t], gv = vals[2 * t + 1];
                Tree &pt = phiTrees[P.tree];
                if (P.mind < 0 && (P.maxd < 0 || std::fabs(1.0 - (gv / kv)) <= eps || gv == kv)) {
                    const int r = pt.resolve(P.node, kv);
                    if (r >= 0) {
                        ThetaTask &T = tt[P.theta_task];
                        T.phival[P.k] = pt.n[r].value;
                        if (--T.left == 0) finished_theta.push_back(P.theta_task);
                    }
                } else {
                    const int l = pt.add(P.node), rgt = pt.add(P.node);
                    pt.n[P.node].child[0] = l;
                    pt.n[P.node].child[1] = rgt;
                    pt.n[P.node].pending = 2;
                    next.push_back({P.theta_task, P.k, P.pmin, apb2, P.maxd - 1, P.mind - 1, l, P.tree});
                    next.push_back({P.theta_task, P.k, apb2, P.pmax, P.maxd - 1, P.mind - 1, rgt, P.tree});
                }
            }
        });
        std::vector<PhiTask> next;
        std::vector<int> finished_theta;
        for (int k = 0; k < nt; k++) {
            next.insert(next.end(), next_part[k].begin(), next_part[k].end());
            finished_theta.insert(finished_theta.end(), fin_part[k].begin(), fin_part[k].end());
        }
        open.swap(next);
        // theta intervals whose 15 phi integrals are complete
        for (int id : finished_theta) {
            const ThetaTask T = tt[id];
            const double ba2 = (T.tmax - T.tmin) * 0.5, apb2 = (T.tmax + T.tmin) * 0.5;
            double kv = 0.0, gv = 0.0;
            for (int i = 0; i < 15; i++) {
                const double sT = std::sin(apb2 + ba2 * kX[i]);
                kv += T.phival[i] * kW[i] * ba2 * sT;
                if (i % 2 == 1) gv += T.phival[i] * gW[i / 2] * ba2 * sT;
            }
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
    if (st) st->evaluations = evals;
}

}  // namespace dplus
