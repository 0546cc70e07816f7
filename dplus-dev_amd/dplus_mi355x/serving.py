"""Serving on one GPU over the C ABI: the request engine (continuous batching
of I(q) requests under an HBM budget) and fit sessions (resident subunit grids
re-docked per fitting evaluation, replayed as HIP graphs, chi^2 on the device)."""
import ctypes

import numpy as np

from . import _lib as L
from .grid import _pdb_desc


class Engine:
    """dplus_engine_*: submit() returns a ticket at once; wait(ticket) gives
    (I(q), result dict).  batch_builds groups small mixed builds (at most
    batch_terms canonical-node x source terms; 0: 1e10) into shared builder
    launches; autotune tunes each new (grid, tile class) build shape on first
    sight; graphs runs each fixed-count request (MC or VEGAS with eps <= 0) as
    one captured HIP graph of its whole kernel chain."""

    def __init__(self, hbm_budget_bytes=0, streams=16, batch_builds=False, batch_terms=0, autotune=False,
                 graphs=True):
        h = ctypes.c_void_p()
        cfg = L.EngineConfig(int(hbm_budget_bytes), int(streams), 1 if batch_builds else 0, int(batch_terms),
                             1 if autotune else 0, 1 if graphs else 0)
        L.check(L.lib().dplus_engine_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._nq = {}

    def submit(self, q, pdb=None, grid=None, grid_size=0, qmax=0.0, precision=L.PRECISION_MIXED, placements=None,
               method=L.OA_MC, seed=1, max_dirs=1 << 20, chunk=8192, eps=0.0, vegas_iterations=10,
               vegas_points_per_bin=4, gk_max_depth=1 << 40, data=None, sigma=None, fit_scale=1.0,
               fit_constant=0.0):
        """pdb: dict of Structure/build_pdb arguments (xyz, ion, coef9, rad, sol_ed, c1, flags, scale,
        box_dims, box_com, vox_step, box_contrast, excl_dims, excl_com, excl_contrast); or grid: a resident Grid.  placements:
        (rot [n,4], ntrans [n], trans [m,3]) of the leaf."""
        keep = []
        q = L.f64(q)
        keep.append(q)
        r = L.Request()
        if pdb is not None:
            p = dict(sol_ed=0.0, c1=1.0, flags=L.CALC_ATOMIC_FORMFACTORS, scale=1.0, box_dims=None, box_com=None,
                     vox_step=0.0, box_contrast=0.0, excl_dims=None, excl_com=None, excl_contrast=0.0)
            p.update(pdb)
            desc = _pdb_desc(p["xyz"], p["ion"], p["coef9"], p["rad"], p["sol_ed"], p["c1"], p["flags"], p["scale"],
                             p["box_dims"], p["box_com"], p["vox_step"], p["box_contrast"], keep, p["excl_dims"],
                             p["excl_com"], p["excl_contrast"])
            keep.append(desc)
            r.pdb = ctypes.pointer(desc)
            r.grid_size, r.qmax, r.precision = int(grid_size), float(qmax), int(precision)
        elif grid is not None:
            r.grid = grid.handle
        if placements is not None:
            rot, nt, tr = L.f64(placements[0]).reshape(-1, 4), L.i32(placements[1]), L.f64(placements[2]).reshape(-1, 3)
            keep += [rot, nt, tr]
            r.nrot, r.rot, r.ntrans, r.trans = len(rot), L.dp(rot), L.ip(nt), L.dp(tr)
        r.nq, r.q = len(q), L.dp(q)
        r.oa = L.OAParams(method, seed, max_dirs, chunk, eps, vegas_iterations, vegas_points_per_bin, gk_max_depth)
        if data is not None:
            y, s = L.f64(data), L.f64(sigma)
            keep += [y, s]
            r.data, r.sigma = L.dp(y), L.dp(s)
            r.fit_scale, r.fit_constant = float(fit_scale), float(fit_constant)
        t = ctypes.c_longlong()
        L.check(L.lib().dplus_engine_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        self._nq[t.value] = len(q)
        return t.value

    def done(self, ticket):
        d = ctypes.c_int()
        L.check(L.lib().dplus_engine_poll(self._h, int(ticket), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        out = np.zeros(self._nq.pop(ticket))
        res = L.RequestResult()
        L.check(L.lib().dplus_engine_wait(self._h, int(ticket), L.dp(out), ctypes.byref(res)))
        info = {k: getattr(res, k) for k, _ in L.RequestResult._fields_ if k != "oa"}
        info["dirs_used"], info["iterations_used"] = res.oa.dirs_used, res.oa.iterations_used
        return out, info

    def stats(self):
        v = [ctypes.c_longlong() for _ in range(5)]
        L.check(L.lib().dplus_engine_stats(self._h, *[ctypes.byref(x) for x in v]))
        out = dict(zip(("submitted", "completed", "peak_reserved", "budget", "max_concurrent"), (x.value for x in v)))
        c = [ctypes.c_longlong() for _ in range(4)]
        L.check(L.lib().dplus_engine_cache_stats(self._h, *[ctypes.byref(x) for x in c]))
        out.update(zip(("cache_hits", "cache_misses", "cache_bytes", "cache_capacity"), (x.value for x in c)))
        b = [ctypes.c_longlong() for _ in range(2)]
        L.check(L.lib().dplus_engine_batch_stats(self._h, *[ctypes.byref(x) for x in b]))
        out.update(zip(("batch_launches", "batched_builds"), (x.value for x in b)))
        g = [ctypes.c_longlong() for _ in range(3)] + [ctypes.c_double() for _ in range(2)]
        L.check(L.lib().dplus_engine_graph_stats(self._h, *[ctypes.byref(x) for x in g]))
        out.update(zip(("graph_launches", "graph_instantiations", "graph_updates", "graph_capture_ms",
                        "graph_exec_ms"), (x.value for x in g)))
        return out

    def close(self):
        if self._h:
            L.lib().dplus_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FitSession:
    """dplus_session_*: resident leaf grids, fixed q/data/sigma, fixed-count MC."""

    def __init__(self, leaves, q, data, sigma, seed=1, directions=1 << 16):
        self._leaves = list(leaves)
        arr = (ctypes.c_void_p * len(leaves))(*[g.handle.value for g in leaves])
        self._q, self._y, self._s = L.f64(q), L.f64(data), L.f64(sigma)
        h = ctypes.c_void_p()
        L.check(L.lib().dplus_session_create(len(leaves), arr, len(self._q), L.dp(self._q), L.dp(self._y),
                                             L.dp(self._s), int(seed), int(directions), ctypes.byref(h)))
        self._h = h

    def evaluate(self, placements, fit_scale=1.0, fit_constant=0.0):
        """placements[l] = (rot [n,4], ntrans [n], trans [m,3]) of leaf l; returns (I(q), chi2)."""
        keep = []
        arr = (L.Leaf * len(placements))()
        for i, (rot, nt, tr) in enumerate(placements):
            rot, nt, tr = L.f64(rot).reshape(-1, 4), L.i32(nt), L.f64(tr).reshape(-1, 3)
            keep += [rot, nt, tr]
            arr[i] = L.Leaf(None, len(rot), L.dp(rot), L.ip(nt), L.dp(tr))
        out = np.zeros(len(self._q))
        c2 = ctypes.c_double()
        L.check(L.lib().dplus_session_evaluate(self._h, arr, len(placements), float(fit_scale), float(fit_constant),
                                               L.dp(out), ctypes.byref(c2)))
        return out, c2.value

    def info(self):
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        L.check(L.lib().dplus_session_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return {"evaluations": a.value, "graph_captures": b.value}

    def close(self):
        if self._h:
            L.lib().dplus_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
