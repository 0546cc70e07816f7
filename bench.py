"""Benchmark of the D+ scattering hot path on one MI355X.

Workload (single-GPU configuration of the headline metric): a 50,000-atom
seeded self-avoiding random walk (0.38 nm steps, 0.12 nm minimum separation)
in an ellipsoid with semi-axes 8/6/5 nm, elements C/N/O/S = 63/17/19/1 %,
X-ray form factors with dummy-atom solvent (rho_0 = 333 e/nm^3, radius type 4),
a 0.3 nm hydration layer (probe 0.14 nm, voxel 0.05 nm) at +10 % contrast
(rho_out = 366.3), grid size 250, qmax 7.5 nm^-1, 1,000 q points, orientation
average by VEGAS (10 iterations x 1024 bins x 98 points).

One step = one I(q) curve: amplitude grid of the resident atoms and boxes,
spline coefficients, VEGAS average, I(q) back on the host.  The mixed-precision
path is timed for --steps steps; the FP64 path (FP64 phases, sincos,
accumulation, with the mixed builder's algebra) runs the same requests one at a
time for comparison, and the maximum relative I(q) difference between the two is
reported; the plain FP64 builder (one node at a time) is timed beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dplus-dev_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "I(q) curves/s + p50/p99 latency, mixed vs FP64 GPU at ≤1e-4 rel err"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector/matrix peak
PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(D, S, seed=2024):
    from dplus_mi355x.grid import hydration_boxes
    t0 = time.time()
    xyz, el = S.walk_structure(seed, 50000, (8.0, 6.0, 5.0))
    coef9, rad, ion, order = S.structure_tables(el)
    dims, com = hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.3, vox=0.05)
    log(f"[bench] structure {len(xyz)} atoms, {len(dims)} hydration boxes "
        f"({int(np.prod(dims, axis=1).sum())} voxels), host prep {time.time() - t0:.2f} s")
    return dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, dims=dims, com=com,
                sol=333.0, contrast=366.3 - 333.0, vox=0.05, gs=250, qmax=7.5,
                q=np.linspace(0.0, 7.5, 1000), iters=10, ppb=98)


def resident(D, W):
    return D.Structure(W["xyz"], W["ion"], W["coef9"], W["rad"], sol_ed=W["sol"], c1=1.0,
                       flags=D.CALC_ATOMIC_FORMFACTORS | D.CALC_DUMMY_SOLVENT, scale=1.0,
                       box_dims=W["dims"], box_com=W["com"], vox_step=W["vox"], box_contrast=W["contrast"])


def curve(D, st, g, W, seed):
    st.build(g, host_copy=False)
    I, _ = D.orientation_average(D.Docking.single(g), W["q"], method=D.OA_VEGAS, seed=seed,
                                 vegas_iterations=W["iters"], vegas_points_per_bin=W["ppb"])
    return I


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def O_total_points(gs):
    """Nodes of a grid: 1 + 6 N (N + 1)^2 with N = gs / 2 + 3 (Grid.cpp:184-186)."""
    n = gs // 2 + 3
    return 1 + 6 * n * (n + 1) ** 2


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(W, g64_data, g64_interp, shells=(8, 16, 24)):
    """The FP64 oracle timed on a bounded sample of the same curve: the amplitude
    of every node of whole shells (all atoms and boxes), the splines, and one full
    VEGAS iteration (1024 bins x the workload's points per bin, every shell
    group), scaled to one curve: node time x P / sampled nodes + iteration time x
    iterations."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    P = O.total_points(W["gs"])
    idx = np.concatenate([O.index_from_indices(i, 0, 0) + np.arange(6 * i * (3 * i + 1)) for i in shells])
    qv = np.array([O.qvector_from_index(W["gs"], W["qmax"], int(k)) for k in idx])
    counts = np.bincount(W["ion"], minlength=len(W["coef9"]))
    ions = O.Ions(W["coef9"], W["rad"], counts, W["sol"], 1.0, O.CALC_ATOMIC_FORMFACTORS | O.CALC_DUMMY_SOLVENT)
    t0 = time.time()
    O.amp_atoms(W["xyz"], W["ion"], ions, qv)
    O.amp_boxes(W["dims"], W["com"], W["vox"], W["contrast"], qv)
    t_nodes = time.time() - t0
    t_grid = t_nodes * P / len(idx)
    t0 = time.time()
    co = O.grid_splines(W["gs"], g64_data)
    t_spl = time.time() - t0
    t0 = time.time()
    dock = O.Docked.single(W["gs"], W["qmax"], g64_data, co)
    O.vegas_oa(dock, W["q"], 1, 1, W["ppb"])
    t_iter = time.time() - t0
    total = t_grid + t_spl + t_iter * W["iters"]
    return {"value": 1.0 / total, "unit": "curves/s", "cores": threads, "nproc": os.cpu_count(),
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"FP64 oracle on {threads} threads: every node of shells {list(shells)} ({len(idx)} of {P} "
                      f"nodes) over all {len(W['xyz'])} atoms + {len(W['dims'])} boxes in {t_nodes:.1f} s (scaled "
                      f"x{P / len(idx):.0f}), splines {t_spl:.1f} s, one full VEGAS iteration (1024 x {W['ppb']} "
                      f"points per shell group) in {t_iter:.1f} s (scaled x{W['iters']}); full curve estimate "
                      f"{total:.0f} s"}


def docking_curve(D, S, mb, directions=2_000_000):
    """The docking workload: a 5,000-atom subunit (2.5 nm sphere) on a grid of 200
    (qmax 6) docked as a 1,000-copy helix (radius 12 nm, rise 0.9 nm, twist 27.7
    deg), 800 q on [0, 6], MC over `directions` directions (chunks of 8,192), one
    curve in mixed precision and one in FP64.  The helix's copies all turn about z,
    so the copy-major kernel runs it; the general kernel (DPLUS_ZCOPY=0) is timed
    on the same curve beside it, with the gather ceiling of its lookups."""
    xyz, el = S.walk_structure(3, 5000, (2.5, 2.5, 2.5))
    coef9, rad, ion, _ = S.structure_tables(el)
    rot, nt, tr = S.rows_to_docking(S.helix_rows(1000, 12.0, 0.9, 27.7))
    q = np.linspace(0.0, 6.0, 800)
    st = D.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3)
    out = {}

    def average(name, g, zcopy=True):
        # the docked leaves are flattened (and the kernel chosen) inside the average call
        if not zcopy:
            os.environ["DPLUS_ZCOPY"] = "0"
        try:
            dock = D.Docking([(g, rot, nt, tr)])
            D.timing_reset()
            D.timing_enable(True)
            sync()
            t0 = time.perf_counter()
            I, stt = D.orientation_average(dock, q, seed=5, max_dirs=directions, chunk=8192)
            sync()
            dt = time.perf_counter() - t0
            D.timing_enable(False)
        finally:
            os.environ.pop("DPLUS_ZCOPY", None)
        out[name] = (I, dt, D.timing_get("mc")[1], stt.dirs_used, g)
        log(f"[docking] {name}: {stt.dirs_used} directions in {dt:.2f} s (MC kernels {out[name][2]:.0f} ms)")

    for name, prec in (("mixed", D.PRECISION_MIXED), ("fp64", D.PRECISION_FP64)):
        g = D.Grid(200, 6.0, prec)
        st.build(g)
        average(name, g)
    average("mixed_general", out["mixed"][4], zcopy=False)
    ngroups = len(np.unique(np.maximum(np.ceil(q / out["mixed"][4].step_size) - 1, 0)))
    lookups = float(out["mixed"][3]) * ngroups * len(rot)
    rate = lookups * 512 / (out["mixed"][2] * 1e-3)
    rate_general = lookups * 512 / (out["mixed_general"][2] * 1e-3)
    pms, plk = D.gather_probe(out["mixed"][4], 32768, 0, ngroups, rot)
    ceiling = plk * 512 / (pms * 1e-3)
    err = float(np.max(np.abs(out["mixed"][0] - out["fp64"][0]) / np.abs(out["fp64"][0])))
    err_general = float(np.max(np.abs(out["mixed"][0] - out["mixed_general"][0]) / np.abs(out["mixed_general"][0])))
    log(f"[docking] mixed vs FP64 {err:.2e}, copy-major vs general kernel {err_general:.2e}; node bytes "
        f"{rate / 1e12:.2f} TB/s copy-major, {rate_general / 1e12:.2f} TB/s general (its gather ceiling "
        f"{ceiling / 1e12:.2f} TB/s)")
    # counted: the copy-major launch's VALU wave instructions from rocprofv3 --pmc of this
    # workload at 32,768 directions (profiles/docking_zcopy_valu.json), scaled to the
    # directions run; their issue time at the microbenchmarked rates (FP32-rate, transcendental,
    # FP64) over the live kernel time
    roofline = None
    vjs = os.path.join(ROOT, "profiles", "docking_zcopy_valu.json")
    if os.path.exists(vjs):
        P = json.load(open(vjs))
        wi = P["wave_instructions"]
        k = float(out["mixed"][3]) / P["directions"]
        other = wi["valu"] - wi["valu_trans_f32"] - wi["valu_fp64"]
        issue_s = k * 64.0 * (other / mb["fma32"] + wi["valu_trans_f32"] / mb["sin32"] + wi["valu_fp64"] / mb["fma64"])
        t = out["mixed"][2] * 1e-3
        roofline = {"bound": "valu", "kernel": P["kernel"], "achieved": k * wi["valu"] / t,
                    "peak": k * wi["valu"] / issue_s, "unit": "VALU wave instructions/s (this kernel's mix)",
                    "frac": issue_s / t, "valu_per_64_copies": wi["valu"] / (P["units"] / 64.0)}
        log(f"[docking] copy-major VALU issue {issue_s * 1e3:.0f} ms of {t * 1e3:.0f} ms ({issue_s / t:.2f})")
    return {"workload": "5,000-atom subunit on grid 200 (qmax 6) docked as a 1,000-copy helix (R 12 nm, rise 0.9 nm, "
                        "twist 27.7 deg), 800 q, MC 2e6 directions",
            "value": 1.0 / out["mixed"][1], "unit": "curves/s", "fp64_value": 1.0 / out["fp64"][1],
            "general_kernel_value": 1.0 / out["mixed_general"][1],
            "speedup_vs_fp64": out["fp64"][1] / out["mixed"][1], "max_rel_err_vs_fp64": err,
            "max_rel_err_vs_general_kernel": err_general,
            "directions": int(out["mixed"][3]), "roofline": roofline,
            "gather": {"kernel": "mc_kernel<float, float4, false> (the general kernel, DPLUS_ZCOPY=0)",
                       "node_GBps_general": rate_general / 1e9, "node_GBps_copy_major": rate / 1e9,
                       "gather_ceiling_GBps": ceiling / 1e9, "general_frac": rate_general / ceiling,
                       "lookups": lookups, "bytes_per_lookup": 512,
                       "ceiling": "dplus_gather_probe: the general kernel's lookups' node loads alone, 1,000 rotations"}}


def plain_curve(D, S, steps=5):
    """The 2,000-atom workload: a seeded walk in a 3 nm sphere (C/N/O/S), dummy-atom
    solvent at 333 e/nm^3, grid 100, qmax 5, 500 q, MC over 10^6 directions (123
    chunks of 8,192) with a fixed seed.  A curve is build + splines + MC from the
    resident structure, in mixed precision and in FP64; median of `steps` curves."""
    xyz, el = S.walk_structure(1, 2000, (3.0, 3.0, 3.0))
    coef9, rad, ion, _ = S.structure_tables(el)
    st = D.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=D.CALC_ATOMIC_FORMFACTORS | D.CALC_DUMMY_SOLVENT)
    q = np.linspace(0.0, 5.0, 500)
    out = {}
    for name, prec in (("mixed", D.PRECISION_MIXED), ("fp64", D.PRECISION_FP64)):
        g = D.Grid(100, 5.0, prec)
        ts = []
        for i in range(steps + 1):
            t0 = time.perf_counter()
            st.build(g)
            I, stt = D.orientation_average(D.Docking.single(g), q, seed=1, max_dirs=10 ** 6, chunk=8192)
            if i:
                ts.append(time.perf_counter() - t0)
        out[name] = (I, float(np.median(ts)), stt.dirs_used)
    err = float(np.max(np.abs(out["mixed"][0] - out["fp64"][0]) / np.abs(out["fp64"][0])))
    log(f"[plain] 2,000 atoms, grid 100, 500 q, MC {out['mixed'][2]} directions: mixed {out['mixed'][1] * 1e3:.1f} ms, "
        f"FP64 {out['fp64'][1] * 1e3:.1f} ms per curve; max rel err {err:.2e}")
    return {"workload": "2,000-atom SAW in a 3 nm sphere, solvent 333 e/nm^3, grid 100, qmax 5, 500 q, "
                        "MC 10^6 directions (123 x 8,192), fixed seed",
            "value": 1.0 / out["mixed"][1], "unit": "curves/s", "fp64_value": 1.0 / out["fp64"][1],
            "speedup_vs_fp64": out["fp64"][1] / out["mixed"][1], "max_rel_err_vs_fp64": err,
            "directions": int(out["mixed"][2]), "steps": steps}


def serving_stream(D, S, n_requests=512, rate=2000.0, pool_size=None, seed=4, streams=16, fp64=True,
                   batch_builds=False, graphs=True, compare_modes=True, repeats=5):
    """Batched serving stream: n_requests with Poisson arrivals at `rate`/s; each
    request is its own seeded random-walk structure (or one of `pool_size`; atoms
    log-uniform in 500..200,000 at 50 atoms/nm^3, dummy-atom solvent), qmax
    U(2, 8) nm^-1, grid size round-up-10(2 qmax L / pi + 3) with L the structure's
    diameter, 200..1,000 q points, MC (65,536 directions) or VEGAS (5 x 4,096
    points) 50/50.  The mixed engine (`streams` streams, 200 GB HBM budget)
    serves the stream (at 2,000/s the queue never empties: the value is the
    saturated throughput and the latencies include queueing), then the same
    requests arriving at 70 % of that throughput (latency without a standing
    queue); the FP64 path then runs the same requests one at a time
    (fp64=False skips it).  batch_builds: small mixed builds share grouped
    builder launches.  graphs: each request's kernel chain runs as one captured
    HIP graph; compare_modes also serves the saturated stream the other way
    (eager launches vs graphs) and checks that every curve is bit-identical."""
    from dplus_mi355x.serving import Engine
    t0 = time.time()
    pool, reqs, arrivals = S.serving_workload(n_requests, pool_size, seed, rate)
    nstruct = n_requests if pool_size is None else pool_size
    log(f"[serving] {n_requests} requests over {nstruct} structures prepared in {time.time() - t0:.1f} s")

    def run_mixed(arr, use_graphs=graphs):
        eng = Engine(hbm_budget_bytes=200 * 10 ** 9, streams=streams, batch_builds=batch_builds, graphs=use_graphs)
        tickets = []
        start = time.perf_counter()
        for (q, kw, _), at in zip(reqs, arr):
            lag = start + at - time.perf_counter()
            if lag > 0:
                time.sleep(lag)
            tickets.append(eng.submit(q, precision=D.PRECISION_MIXED, **kw))
        res = [eng.wait(t) for t in tickets]
        span = time.perf_counter() - start
        st = eng.stats()
        eng.close()
        return res, span, np.array([r["latency_ms"] for _, r in res]), st

    def summary(span_, lat_, st_):
        return {"value": n_requests / span_, "p50_ms": float(np.percentile(lat_, 50)),
                "p99_ms": float(np.percentile(lat_, 99)), "graph_launches": st_["graph_launches"],
                "graph_instantiations": st_["graph_instantiations"], "graph_updates": st_["graph_updates"],
                "graph_capture_ms": st_["graph_capture_ms"], "graph_exec_ms": st_["graph_exec_ms"]}

    mode = "graphs" if graphs else "eager"
    order = [mode] + ([("eager" if graphs else "graphs")] if compare_modes else [])
    saturated, loaded_modes = {}, {}
    for m in order:
        # the saturated window is short (~0.4 s): the run of median span out of `repeats`
        runs = sorted((run_mixed(arrivals, m == "graphs") for _ in range(repeats)), key=lambda x: x[1])
        res_m, span_m, lat_m, st_m = runs[len(runs) // 2]
        saturated[m] = summary(span_m, lat_m, st_m)
        saturated[m]["values_all"] = [n_requests / x[1] for x in runs]
        if not all(np.array_equal(a[0], b[0]) for x in runs for a, b in zip(res_m, x[0])):
            saturated["repeat_mismatch"] = True
        qms = np.array([r["queue_ms"] for _, r in res_m])
        rms = np.array([r["run_ms"] for _, r in res_m])
        log(f"[serving] mixed ({m}): {n_requests / span_m:.2f} curves/s over {span_m:.2f} s, "
            f"p50 {np.percentile(lat_m, 50):.1f} ms, p99 {np.percentile(lat_m, 99):.1f} ms "
            f"(queue p50 {np.percentile(qms, 50):.1f}, run p50 {np.percentile(rms, 50):.1f} / p99 "
            f"{np.percentile(rms, 99):.1f} ms), "
            f"max concurrent {st_m['max_concurrent']}, graph launches {st_m['graph_launches']} "
            f"(instantiated {st_m['graph_instantiations']}, updated {st_m['graph_updates']}; host ms capturing "
            f"{st_m['graph_capture_ms']:.1f}, updating/instantiating {st_m['graph_exec_ms']:.1f})")
        if m == mode:
            mixed, span, lat, st = res_m, span_m, lat_m, st_m
        else:
            saturated["bit_identical"] = bool(all(np.array_equal(x[0], y[0]) for x, y in zip(mixed, res_m)))
            log(f"[serving] curves bit-identical across modes: {saturated['bit_identical']}")
    # the same requests arriving at 70 % of the primary mode's saturated throughput:
    # latency without a standing queue, every mode at the same arrival rate
    rate70 = 0.7 * n_requests / span
    for m in order:
        res70, span70, lat70, _ = run_mixed(arrivals * (rate / rate70), m == "graphs")
        loaded_modes[m] = {"rate": rate70, "value": n_requests / span70, "p50_ms": float(np.percentile(lat70, 50)),
                           "p99_ms": float(np.percentile(lat70, 99))}
        rms = np.array([r["run_ms"] for _, r in res70])
        log(f"[serving] mixed ({m}) at {rate70:.0f}/s: {n_requests / span70:.2f} curves/s, "
            f"p50 {loaded_modes[m]['p50_ms']:.1f} ms, p99 {loaded_modes[m]['p99_ms']:.1f} ms "
            f"(run p50 {np.percentile(rms, 50):.1f} / p99 {np.percentile(rms, 99):.1f} ms)")
    loaded = loaded_modes[mode]
    batching = {"batch_builds": bool(batch_builds), "batch_launches": st["batch_launches"],
                "batched_builds": st["batched_builds"], "engine_mode": mode,
                "modes": {"saturated": saturated, "at_70pct_load": loaded_modes}}
    if not fp64:
        return {"value": n_requests / span, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "streams": streams, "at_70pct_load": loaded, **batching}
    eng64 = Engine(hbm_budget_bytes=200 * 10 ** 9, streams=1)
    t64 = time.perf_counter()
    errs, lat64 = [], []
    for (q, kw, _), (Im, _) in zip(reqs, mixed):
        I64, r64 = eng64.wait(eng64.submit(q, precision=D.PRECISION_FP64, **kw))
        lat64.append(r64["latency_ms"])
        errs.append(float(np.max(np.abs(Im - I64) / np.abs(I64))))
    span64 = time.perf_counter() - t64
    eng64.close()
    log(f"[serving] FP64, one request at a time: {n_requests / span64:.2f} curves/s; "
        f"max rel err mixed vs FP64 {max(errs):.2e}")
    return {"workload": f"{n_requests} Poisson requests at {rate:g}/s over {nstruct} seeded structures "
                        "(500..200,000 atoms), qmax U(2,8), 200..1,000 q, MC 65,536 dirs or VEGAS 5x4,096",
            "value": n_requests / span, "unit": "curves/s", "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "fp64_value": n_requests / span64,
            "fp64_p50_ms": float(np.percentile(lat64, 50)), "fp64_p99_ms": float(np.percentile(lat64, 99)),
            "speedup_vs_fp64": (n_requests / span) / (n_requests / span64), "max_rel_err_vs_fp64": max(errs),
            "streams": streams, "max_concurrent": st["max_concurrent"], "peak_reserved_bytes": st["peak_reserved"],
            "hbm_budget_bytes": st["budget"], "atoms_total": int(sum(n for _, _, n in reqs)),
            "at_70pct_load": loaded, **batching}


def fit_sessions(D, S, n_sessions=16, evals=4, directions=8192, nq=600, qmax=5.0, seed=5):
    """Fitting-loop serving: n_sessions concurrent sessions, each with 1..4 cached
    subunit grids of 2,000..20,000 atoms (grid 150..250, FP32) placed by a helix
    (radius 6..12 nm, +-10 %; twist 20..40 deg, +-5 deg) or a 60-fold icosahedral
    shell (radius 10..20 nm, +-10 %), 60..2,000 copies in total; each evaluation
    re-docks, averages `directions` MC directions at nq q and reduces chi^2 on the
    device against Poisson-noised synthetic data.  Sessions run on their own threads."""
    import threading
    from dplus_mi355x.serving import FitSession
    rng = np.random.default_rng(seed)
    q = np.linspace(0.01, qmax, nq)
    t0 = time.time()
    plans = S.session_plans(n_sessions, seed)
    placements = S.session_placements

    def build(p, prec):
        gl = []
        for xyz, ion, coef9, rad in p["leaves"]:
            g = D.Grid(p["gs"], qmax, prec)
            D.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3).build(g)
            gl.append(g)
        return gl

    sessions = []
    for p in plans:
        grids = build(p, D.PRECISION_MIXED)
        prob = FitSession(grids, q, np.ones(nq), np.ones(nq), seed=p["seed"], directions=directions)
        I0, _ = prob.evaluate(placements(p, 1.0, 0.0))
        prob.close()
        k = 1e4 / I0.max()
        y = rng.poisson(I0 * k) / k
        sig = np.sqrt(np.maximum(y * k, 1.0)) / k
        sessions.append((p, grids, FitSession(grids, q, y, sig, seed=p["seed"], directions=directions)))
    log(f"[sessions] {n_sessions} sessions prepared in {time.time() - t0:.1f} s "
        f"(copies {[p['copies'] for p in plans]})")
    lat = []
    lock = threading.Lock()
    last = [time.perf_counter()]

    def run(sess_tuple, idx):
        p, grids, sess = sess_tuple
        r = np.random.default_rng(100 + idx)
        for _ in range(evals):
            pl = placements(p, float(r.uniform(0.9, 1.1)), float(r.uniform(-5.0, 5.0)))
            t = time.perf_counter()
            sess.evaluate(pl, fit_scale=1.0, fit_constant=0.0)
            now = time.perf_counter()
            with lock:
                lat.append((now - t) * 1e3)
                if now - last[0] > 20.0:
                    last[0] = now
                    log(f"[sessions] {len(lat)} of {n_sessions * evals} evaluations")

    threads = [threading.Thread(target=run, args=(st, i)) for i, st in enumerate(sessions)]
    start = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    span = time.perf_counter() - start
    lat = np.array(lat)
    # accuracy against the FP64 path on the first two sessions' data point
    errs = []
    for p, grids, sess in sessions[:2]:
        g64 = build(p, D.PRECISION_FP64)
        pl = placements(p, 1.0, 0.0)
        Im, _ = sess.evaluate(pl)
        s64 = FitSession(g64, q, sess._y, sess._s, seed=p["seed"], directions=directions)
        I64, _ = s64.evaluate(pl)
        s64.close()
        errs.append(float(np.max(np.abs(Im - I64) / np.abs(I64))))
    captures = sum(s.info()["graph_captures"] for _, _, s in sessions)
    for _, _, s in sessions:
        s.close()
    n_eval = n_sessions * evals
    log(f"[sessions] {n_eval / span:.2f} evaluations/s, p50 {np.percentile(lat, 50):.1f} ms, p99 "
        f"{np.percentile(lat, 99):.1f} ms, graph captures {captures}, max rel err vs FP64 {max(errs):.2e}")
    return {"workload": f"{n_sessions} concurrent fit sessions x {evals} evaluations (helix / icosahedral re-docking, "
                        f"{directions} MC directions, {nq} q, chi^2 on device)",
            "value": n_eval / span, "unit": "evaluations/s", "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_rel_err_vs_fp64": max(errs), "graph_captures": captures,
            "copies": [p["copies"] for p in plans]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--fp64-plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream and fit-session measurements")
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--session-evals", type=int, default=300)
    ap.add_argument("--pmc-probe", "--probe", dest="pmc_probe", action="store_true",
                    help="one mixed build of the 50k-atom grid, a 256 MiB device copy (counter calibration), one "
                         "VEGAS curve, one MC average of 262,144 directions and the gather probe on that grid, then "
                         "the 1,000-copy helix docked MC at 32,768 directions and its gather probe, and a "
                         "1,800-copy icosahedral shell's docked MC at 32,768 directions: for rocprofv3")
    ap.add_argument("--no-docking", action="store_true", help="skip the 1,000-copy helix docking curve")
    ap.add_argument("--gk-only", action="store_true",
                    help="the 50k-atom grid built once and averaged by Gauss-Kronrod three times: for rocprofv3")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving stream's mixed legs (saturated and at 70 %% load): for rocprofv3")
    ap.add_argument("--serving-streams", type=int, default=16, help="engine slots (streams) of the serving stream")
    ap.add_argument("--serving-eager", action="store_true",
                    help="the serving stream's primary mode is eager launches (captured graphs beside it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 (the path shards by shell "
                                   "ranges and directions across GPUs, which is not built)"}))
        sys.exit(2)

    import dplus_mi355x as D
    import dplus_mi355x.synth as S

    info = D.device_info()
    log("[bench] device", info)
    if not info.get("ok"):
        raise SystemExit("no usable GPU")
    if args.serving_only:
        print(json.dumps({"serving": serving_stream(D, S, n_requests=args.serving_requests, fp64=False,
                                                    streams=args.serving_streams, graphs=not args.serving_eager)}))
        return
    W = make_workload(D, S)
    st = resident(D, W)
    g32 = D.Grid(W["gs"], W["qmax"], D.PRECISION_MIXED)
    if args.gk_only:
        st.build(g32)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, stg = D.orientation_average(D.Docking.single(g32), W["q"], method=D.OA_GK, eps=1e-4, gk_max_depth=8)
            runs.append(time.perf_counter() - t0)
        print(json.dumps({"gauss_kronrod_s": runs, "evaluations": stg.evaluations}))
        return
    if args.pmc_probe:
        st.build(g32)
        nbytes = 256 * 2 ** 20
        ms = D.copy_probe(nbytes)
        # orientation averages on the built grid: every (direction, shell group) reads
        # 4 planes x 4 theta rows x 2 phi nodes of 16 B (mixed) = 512 B of nodes
        D.timing_reset()
        D.timing_enable(True)
        D.orientation_average(D.Docking.single(g32), W["q"], method=D.OA_VEGAS, seed=7,
                              vegas_iterations=W["iters"], vegas_points_per_bin=W["ppb"])
        D.orientation_average(D.Docking.single(g32), W["q"], method=D.OA_MC, seed=7, max_dirs=1 << 18, chunk=8192)
        D.timing_enable(False)
        # shell groups: q in (l step, (l + 1) step] share layer l, q = 0 joins layer 0
        ngroups = len(np.unique(np.maximum(np.ceil(W["q"] / g32.step_size) - 1, 0)))
        vegas_ms, mc_ms = D.timing_get("vegas")[1], D.timing_get("mc")[1]
        g_ms, g_lk = D.gather_probe(g32, 1 << 18, 0, ngroups)
        # the docking workload's MC at 32,768 directions and its gather probe
        xyz, el = S.walk_structure(3, 5000, (2.5, 2.5, 2.5))
        coef9, rad, ion, _ = S.structure_tables(el)
        rot, nt, tr = S.rows_to_docking(S.helix_rows(1000, 12.0, 0.9, 27.7))
        q3 = np.linspace(0.0, 6.0, 800)
        g3 = D.Grid(200, 6.0, D.PRECISION_MIXED)
        D.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3).build(g3)
        D.timing_reset()
        D.timing_enable(True)
        D.orientation_average(D.Docking([(g3, rot, nt, tr)]), q3, seed=5, max_dirs=32768, chunk=8192)
        D.timing_enable(False)
        ng3 = len(np.unique(np.maximum(np.ceil(q3 / g3.step_size) - 1, 0)))
        docked_ms = D.timing_get("mc")[1]
        d_ms, d_lk = D.gather_probe(g3, 32768, 0, ng3, rot)
        # an icosahedral shell of 1,800 copies (60 rotations x 30 translations: the general
        # kernel with the translations' phases stepped in q) on the same grid and q
        rot_i, nt_i, tr_i = S.rows_to_docking(S.icosahedral_rows(15.0, 1800, seed=3))
        D.timing_reset()
        D.timing_enable(True)
        D.orientation_average(D.Docking([(g3, rot_i, nt_i, tr_i)]), q3, seed=5, max_dirs=32768, chunk=8192)
        D.timing_enable(False)
        ico_ms = D.timing_get("mc")[1]
        print(json.dumps({"probe": "build_atoms_pairs + copy_probe_kernel + vegas_sample + mc_kernel + gather_probe "
                                   "(50k-atom grid 250), mc_zcopy_kernel + gather_probe<true> (1,000-copy helix), "
                                   "mc_kernel docked (1,800-copy icosahedral shell)",
                          "copy_bytes": nbytes, "copy_GBps": 2 * nbytes / (ms * 1e-3) / 1e9,
                          "shell_groups": ngroups,
                          "vegas_ms": vegas_ms, "vegas_units": W["iters"] * 1024 * W["ppb"] * ngroups,
                          "mc_ms": mc_ms, "mc_units": (1 << 18) * ngroups,
                          "gather_probe_ms": g_ms, "gather_probe_units": g_lk,
                          "docked_mc_ms": docked_ms, "docked_mc_units": 32768 * ng3 * len(rot),
                          "docked_gather_probe_ms": d_ms, "docked_gather_probe_units": d_lk,
                          "icosahedral_mc_ms": ico_ms, "icosahedral_mc_units": 32768 * ng3 * 1800,
                          "node_bytes_per_unit": 512}))
        return
    g64 = D.Grid(W["gs"], W["qmax"], D.PRECISION_FP64)

    # builder launch shape: from the shape table the package loads (tuned on MI355X by
    # python -m dplus_mi355x.tune), else the occupancy model
    shape = st.build_shape(g32)
    log(f"[bench] builder launch shape {shape}")
    mb = D.microbench_valu()
    w_t = 2.0 * mb["fma32"] / mb["sin32"]  # flop-equivalents of one v_sin/v_cos
    log(f"[bench] VALU microbenchmark ops/s: {mb}; transcendental weight {w_t:.2f} flop-eq")

    # ---- mixed precision: warmup, then timed steps
    for i in range(args.warmup):
        curve(D, st, g32, W, seed=100 + i)
    D.timing_reset()
    D.timing_enable(True)
    lat = []
    sync()
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        I32 = curve(D, st, g32, W, seed=7)
        lat.append(time.perf_counter() - t0)
    sync()
    elapsed = time.perf_counter() - t_all
    D.timing_enable(False)
    nb, ms_build = D.timing_get("build")
    nv, ms_vegas = D.timing_get("vegas")
    ns, ms_spl = D.timing_get("splines")
    log(f"[bench] mixed: {args.steps} curves in {elapsed:.3f} s; build {ms_build / max(nb, 1):.2f} ms/launch, "
        f"splines {ms_spl / max(ns, 1):.2f} ms, vegas {ms_vegas / max(args.steps, 1):.2f} ms/curve")

    # ---- FP64 path on the same request: the paired builder's algebra in double
    # (like for like), then the plain FP64 builder (one node at a time)
    lat64 = []
    for i in range(args.fp64_steps):
        t0 = time.perf_counter()
        I64 = curve(D, st, g64, W, seed=7)
        lat64.append(time.perf_counter() - t0)
    err = float(np.max(np.abs(I32 - I64) / np.abs(I64)))
    log(f"[bench] fp64: {np.mean(lat64):.3f} s/curve; max rel I(q) mixed vs fp64 = {err:.2e}")
    lat64p = []
    D.set_fp64_builder(True)
    try:
        for i in range(args.fp64_plain_steps):
            t0 = time.perf_counter()
            I64p = curve(D, st, g64, W, seed=7)
            lat64p.append(time.perf_counter() - t0)
    finally:
        D.set_fp64_builder(False)
    err_plain = float(np.max(np.abs(I64p - I64) / np.abs(I64))) if lat64p else None
    if lat64p:
        log(f"[bench] fp64 plain builder: {np.mean(lat64p):.3f} s/curve; vs paired FP64 {err_plain:.2e}")

    # ---- the same grids averaged by adaptive Gauss-Kronrod (relative tolerance 1e-4, depth 8)
    # (one untimed average first: the first one grows the device memory pool by the GK arenas)
    gk = {}
    D.orientation_average(D.Docking.single(g32), W["q"], method=D.OA_GK, eps=1e-4, gk_max_depth=8)
    for name, g in (("mixed", g32), ("fp64", g64)):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            Igk, stg = D.orientation_average(D.Docking.single(g), W["q"], method=D.OA_GK, eps=1e-4, gk_max_depth=8)
            ts.append(time.perf_counter() - t0)
        gk[name] = (Igk, float(np.median(ts)), stg.evaluations)
    gk_err = float(np.max(np.abs(gk["mixed"][0] - gk["fp64"][0]) / np.abs(gk["fp64"][0])))
    gk_out = {"tolerance": 1e-4, "max_depth": 8, "mixed_s": gk["mixed"][1], "fp64_s": gk["fp64"][1],
              "timing": "median of 3 per mode after one untimed average",
              "evaluations": gk["mixed"][2], "evaluations_fp64": gk["fp64"][2], "curve_s_mixed": gk["mixed"][1] + ms_build / max(nb, 1) * 1e-3,
              "max_rel_err_vs_fp64": gk_err}
    log(f"[bench] GK (tol 1e-4): mixed {gk['mixed'][1]:.3f} s, fp64 {gk['fp64'][1]:.3f} s per average "
        f"({gk['mixed'][2]} integrand evaluations); max rel I(q) mixed vs fp64 = {gk_err:.2e}")

    # ---- the VEGAS kernels against the gather ceiling of the same lookups on the same grid
    ngroups = len(np.unique(np.maximum(np.ceil(W["q"] / g32.step_size) - 1, 0)))
    v_lookups = float(W["iters"]) * 1024 * W["ppb"] * ngroups
    v_rate = v_lookups * 512 / (ms_vegas / max(args.steps, 1) * 1e-3)
    p_ms, p_lk = D.gather_probe(g32, 1 << 18, 0, ngroups)
    ceiling = p_lk * 512 / (p_ms * 1e-3)
    oa_roof = {"bound": "gather", "kernel": "vegas_sample<float, float4, true> + vegas_reduce",
               "achieved": v_rate / 1e9, "peak": ceiling / 1e9, "unit": "GB/s of nodes", "frac": v_rate / ceiling,
               "lookups_per_curve": v_lookups, "bytes_per_lookup": 512,
               "ceiling": "dplus_gather_probe: the same lookups' node loads alone at full occupancy"}
    log(f"[bench] VEGAS lookups {v_rate / 1e12:.2f} TB/s of nodes, gather ceiling {ceiling / 1e12:.2f} TB/s "
        f"({v_rate / ceiling:.2f})")

    # ---- roofline of the dominant kernel (the amplitude builder, VALU-bound)
    N = W["gs"] // 2 + 3
    nhalf = 3 * (N + 1) ** 2 * N
    sources = len(W["xyz"]) + len(W["dims"])
    terms = float(nhalf) * sources
    # executed per atom and unit (three phi / phi + pi node pairs 120 degrees apart: six terms):
    # q_x x + q_y y of two pairs (6 flops), their sin/cos (4 transcendentals), the third pair's
    # e^{i u2} = conj(e^{i u0} e^{i u1}) (6 flops) and twelve FMAs (24 flops); sin/cos of q_z z
    # come from LDS, evaluated once per (row, atom) per workgroup, whose units span a few rows
    # (under 2 % of the work, not counted)
    flop_eq = 36.0 / 6.0 + (4.0 / 6.0) * w_t
    t_launch = ms_build / max(nb, 1) * 1e-3
    model_achieved = terms * flop_eq / t_launch / 1e12
    # counted: the builder launch's VALU wave instructions from rocprofv3 --pmc of this
    # workload (profiles/build_atoms_pairs_c2_valu.json), 64 lanes each: FP32 flops
    # (FMA 2, MUL and ADD 1) plus w_T per transcendental, over the live launch time;
    # issue_frac = their issue time at the microbenchmarked rates / the launch time
    counted = None
    vjs = os.path.join(ROOT, "profiles", "build_atoms_pairs_c2_valu.json")
    if os.path.exists(vjs):
        wi = json.load(open(vjs))["wave_instructions"]
        flops32 = 64.0 * (2.0 * wi["valu_fma_f32"] + wi["valu_mul_f32"] + wi["valu_add_f32"])
        trans = 64.0 * wi["valu_trans_f32"]
        issue_s = 64.0 * (wi["valu"] - wi["valu_trans_f32"]) / mb["fma32"] + trans / mb["sin32"]
        counted = {"flops_fp32": flops32, "transcendentals": trans, "valu_instructions_per_term": 64.0 * wi["valu"] / terms,
                   "achieved": (flops32 + w_t * trans) / t_launch / 1e12, "issue_frac": issue_s / t_launch}
    achieved = counted["achieved"] if counted else model_achieved
    # HBM bytes per builder launch from the two rocprofv3 --pmc passes of this workload
    # (FETCH_SIZE, WRITE_SIZE, each scaled by a known-size copy), kept in profiles/
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "build_atoms_pairs_c2_pmc.json")
    if os.path.exists(pmc):
        traffic = json.load(open(pmc)).get("traffic_bytes")

    cpu = None
    if not args.no_cpu_baseline:
        g64.download()
        cpu = cpu_baseline(W, g64.data.copy(), g64.interp.copy())
        log(f"[bench] cpu baseline: {cpu}")

    docking = None if args.no_docking else docking_curve(D, S, mb)
    plain = plain_curve(D, S)

    serving = sessions = None
    if not args.no_serving:
        serving = serving_stream(D, S, n_requests=args.serving_requests, streams=args.serving_streams,
                                 graphs=not args.serving_eager)
        sessions = fit_sessions(D, S, evals=args.session_evals)

    value = args.steps / elapsed
    fp64_value = 1.0 / float(np.mean(lat64))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "curves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "50k-atom SAW ellipsoid 8/6/5 nm + 0.3 nm hydration layer; grid 250, qmax 7.5, "
                               "1000 q; VEGAS 10 x 100352; mixed precision",
                   "atoms": len(W["xyz"]), "hydration_boxes": len(W["dims"]), "grid_size": W["gs"],
                   "q_points": len(W["q"]), "oa": "VEGAS 10x(1024x98)", "build_shape": shape},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "fp64": {"value": fp64_value, "unit": "curves/s", "p50_ms": 1e3 * float(np.percentile(lat64, 50)),
                 "p99_ms": 1e3 * float(np.percentile(lat64, 99)), "steps": args.fp64_steps},
        "speedup_vs_fp64": value / fp64_value,
        "max_rel_err_vs_fp64": err,
        # the plain FP64 builder (one node at a time, one double sincos per term) on the same curve
        "fp64_plain": None if not lat64p else {
            "value": 1.0 / float(np.mean(lat64p)), "unit": "curves/s", "steps": len(lat64p),
            "max_rel_err_vs_fp64": err_plain},
        "speedup_vs_fp64_plain": None if not lat64p else value * float(np.mean(lat64p)),
        "gauss_kronrod": gk_out,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": f"build_atoms_pairs<false, {shape['wg']}>", "launch_ms": t_launch * 1e3,
                     "terms_per_launch": terms, "flop_eq_per_term": flop_eq,
                     "transcendental_weight": w_t,
                     "achieved_from": "counted VALU instructions (rocprofv3 --pmc)" if counted else "flop model",
                     "counted": counted, "model_achieved": model_achieved,
                     # the direct sum's unit: 8 flops + 2 transcendentals per (node, source) term over
                     # all P nodes (the builder does half the nodes by Friedel symmetry and shares sin/cos
                     # pairs between the unit's nodes, so this exceeds the peak)
                     "direct_sum_unit_achieved": float(O_total_points(W["gs"])) * sources * (8.0 + 2.0 * w_t) / t_launch / 1e12,
                     # the same launch against the HBM roofline: PMC bytes per launch over its time
                     "hbm": None if traffic is None else {
                         "achieved": traffic / t_launch / 1e9, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                         "frac": traffic / t_launch / 1e9 / PEAK_HBM_GBPS}},
        "oa_roofline": oa_roof,
        "cpu_baseline": cpu,
        "docking": docking,
        "plain_2000_atoms": plain,
        "serving": serving,
        "fit_sessions": sessions,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
