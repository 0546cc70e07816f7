"""Serving on the GPU: the request engine gives every request the same result
alone and in a concurrent mix, admits by HBM budget, and matches the FP64
oracle; fit sessions replay captured graphs with the same intensities as the
eager orientation average and chi^2 reduced on the device."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pdb(n, seed, radius):
    import dplus_mi355x.synth as S
    xyz, el = S.walk_structure(seed, n, (radius, radius, radius))
    coef9, rad, ion, _ = S.structure_tables(el)
    return dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, sol_ed=333.0, c1=1.0, flags=3, scale=1.0)


def requests(D):
    import dplus_mi355x.synth as S
    out = []
    rng = np.random.default_rng(5)
    for i in range(10):
        n = int(rng.integers(150, 2500))
        qmax = float(rng.uniform(2.0, 6.0))
        gs = int(2 * round(rng.uniform(12, 36) / 2))
        q = np.linspace(0.02, qmax, int(rng.integers(40, 160)))
        kw = dict(pdb=make_pdb(n, 100 + i, 1.0 + n / 2500.0), grid_size=gs, qmax=qmax,
                  precision=D.PRECISION_FP64 if i % 4 == 3 else D.PRECISION_MIXED, seed=7 + i)
        if i % 2:
            kw.update(method=D.OA_VEGAS, vegas_iterations=3, vegas_points_per_bin=4)
        else:
            kw.update(method=D.OA_MC, max_dirs=8192, chunk=4096)
        if i % 3 == 1:
            # request 4 (mixed MC): 40 copies turned about z, the copy-major kernel
            kw["placements"] = S.rows_to_docking(S.helix_rows(40 if i == 4 else 6, 2.0, 0.7, 31.0))
        if i % 3 == 2:
            kw.update(data=np.ones(len(q)), sigma=np.full(len(q), 0.5), fit_scale=1e-6, fit_constant=0.1)
        out.append((q, kw))
    return out


def test_engine_same_alone_and_in_the_mix(dplus):
    from dplus_mi355x.serving import Engine
    say("[case] engine: 10 mixed requests alone (1 stream, one at a time) vs concurrently (4 streams)")
    reqs = requests(dplus)
    alone = []
    e1 = Engine(streams=1)
    for q, kw in reqs:
        alone.append(e1.wait(e1.submit(q, **kw)))
    e1.close()
    e4 = Engine(streams=4)
    tickets = [e4.submit(q, **kw) for q, kw in reqs]
    mix = [e4.wait(t) for t in tickets]
    st = e4.stats()
    e4.close()
    say(f"    engine stats {st}")
    assert st["completed"] == len(reqs) and st["max_concurrent"] > 1
    for (I1, r1), (I4, r4), (q, kw) in zip(alone, mix, reqs):
        assert np.array_equal(I1, I4), kw.get("method")
        assert r1["chi2"] == r4["chi2"]
        if "data" in kw:
            model = kw["fit_scale"] * I1 + kw["fit_constant"]
            ref = float(np.sum(((model - kw["data"]) / kw["sigma"]) ** 2))
            assert r1["chi2"] == pytest.approx(ref, rel=1e-12)
            assert r1["chi2_on_device"] == 1


def test_engine_graph_requests_match_eager(dplus):
    from dplus_mi355x.serving import Engine
    say("[case] engine: 10 requests eager vs captured graphs (1 stream twice, then 4 streams)")
    reqs = requests(dplus)
    e1 = Engine(streams=1, graphs=False)
    eager = [e1.wait(e1.submit(q, **kw)) for q, kw in reqs]
    e1.close()
    assert not any(r["graph"] for _, r in eager)
    # one slot: the second pass updates the instantiations the first pass kept
    g1 = Engine(streams=1, graphs=True)
    passes = [[g1.wait(g1.submit(q, **kw)) for q, kw in reqs] for _ in range(2)]
    st1 = g1.stats()
    g1.close()
    g4 = Engine(streams=4, graphs=True)
    tickets = [g4.submit(q, **kw) for q, kw in reqs]
    passes.append([g4.wait(t) for t in tickets])
    st4 = g4.stats()
    g4.close()
    say(f"    graph stats one slot {st1}, four slots {st4}")
    assert st1["graph_launches"] == 2 * len(reqs) and st1["graph_updates"] >= len(reqs)
    assert st4["graph_launches"] == len(reqs)
    for res in passes:
        for (Ie, re), (Ig, rg), (q, kw) in zip(eager, res, reqs):
            assert rg["graph"] == 1
            assert np.array_equal(Ie, Ig), (kw.get("method"), kw["grid_size"], kw["precision"])
            assert re["chi2"] == rg["chi2"] and re["chi2_on_device"] == rg["chi2_on_device"]
            assert (re["dirs_used"], re["iterations_used"]) == (rg["dirs_used"], rg["iterations_used"])


def test_engine_matches_oracle(dplus, oracle):
    from dplus_mi355x.serving import Engine
    say("[case] engine request against the FP64 oracle")
    import dplus_mi355x.synth as S
    pdb = make_pdb(400, 3, 1.2)
    gs, qmax = 20, 4.0
    q = np.linspace(0.05, qmax, 60)
    rows = S.helix_rows(5, 1.5, 0.6, 40.0)
    rot, nt, tr = S.rows_to_docking(rows)
    ions = oracle.Ions(pdb["coef9"], pdb["rad"], np.bincount(pdb["ion"], minlength=len(pdb["coef9"])), 333.0, 1.0,
                       oracle.CALC_ATOMIC_FORMFACTORS | oracle.CALC_DUMMY_SOLVENT)
    data = oracle.grid_fill_pdb(gs, qmax, pdb["xyz"], pdb["ion"], ions)
    co = oracle.grid_splines(gs, data)
    ref, _ = oracle.mc_oa(oracle.Docked([(gs, qmax, data, co, rot, nt, tr)]), q, 11, 8192, 4096)
    e = Engine(streams=2)
    for prec, tol in ((dplus.PRECISION_FP64, 1e-6), (dplus.PRECISION_MIXED, 1e-4)):
        I, _ = e.wait(e.submit(q, pdb=pdb, grid_size=gs, qmax=qmax, precision=prec, placements=(rot, nt, tr),
                               method=dplus.OA_MC, seed=11, max_dirs=8192, chunk=4096))
        assert np.max(np.abs(I - ref) / ref) < tol
    e.close()


def test_engine_admission_budget_and_resident_grid(dplus):
    from dplus_mi355x.serving import Engine
    say("[case] engine admission by HBM budget; resident grid request")
    pdb = make_pdb(300, 9, 1.0)
    q = np.linspace(0.05, 3.0, 50)
    kw = dict(pdb=pdb, grid_size=24, qmax=3.0, method=dplus.OA_MC, seed=2, max_dirs=4096, chunk=4096)
    probe = Engine(streams=1)
    I0, r0 = probe.wait(probe.submit(q, **kw))
    probe.close()
    need = r0["reserved_bytes"]
    tight = Engine(hbm_budget_bytes=need + need // 2, streams=4)  # room for one request at a time
    ts = [tight.submit(q, **kw) for _ in range(6)]
    outs = [tight.wait(t)[0] for t in ts]
    st = tight.stats()
    assert st["max_concurrent"] == 1 and st["peak_reserved"] <= st["budget"]
    assert all(np.array_equal(o, I0) for o in outs)
    with pytest.raises(dplus.DplusError) as err:
        tight.submit(q, **dict(kw, grid_size=200))
    assert err.value.error_code == 13
    tight.close()
    # a resident grid gives the same curve as the direct average of that grid
    st_ = dplus.Structure(pdb["xyz"], pdb["ion"], pdb["coef9"], pdb["rad"], sol_ed=333.0, flags=3)
    g = dplus.Grid(24, 3.0, dplus.PRECISION_MIXED)
    st_.build(g)
    direct, _ = dplus.orientation_average(dplus.Docking.single(g), q, seed=2, max_dirs=4096, chunk=4096)
    e = Engine(streams=2)
    Ig, _ = e.wait(e.submit(q, grid=g, method=dplus.OA_MC, seed=2, max_dirs=4096, chunk=4096))
    e.close()
    assert np.array_equal(Ig, direct)
    assert np.array_equal(Ig, I0)


def test_fit_session_graph_replays(dplus):
    from dplus_mi355x.serving import FitSession
    import dplus_mi355x.synth as S
    say("[case] fit session: captured graph replays vs eager averages, chi^2 on the device")
    pdb = make_pdb(500, 21, 1.2)
    g = dplus.Grid(24, 4.0, dplus.PRECISION_MIXED)
    dplus.Structure(pdb["xyz"], pdb["ion"], pdb["coef9"], pdb["rad"], sol_ed=333.0, flags=3).build(g)
    q = np.linspace(0.05, 4.0, 80)
    y = np.linspace(2.0, 1.0, 80)
    sig = np.full(80, 0.1)
    dirs = 16384
    sess = FitSession([g], q, y, sig, seed=4, directions=dirs)
    # 48 copies turned about z alone run the copy-major MC kernel; the same count
    # tilted about x goes back to the general kernel (a capture of its own)
    cases = [(12, 2.0, 0.8, 30.0, 0.0), (12, 2.2, 0.8, 31.5, 0.0), (12, 2.0, 0.8, 30.0, 0.0),
             (20, 1.8, 0.7, 25.0, 0.0), (48, 2.0, 0.8, 30.0, 0.0), (48, 2.1, 0.8, 29.0, 0.0),
             (48, 2.0, 0.8, 30.0, 5.0)]
    prev = {}
    for n, rad, rise, twist, tilt in cases:
        rows = S.helix_rows(n, rad, rise, twist)
        rows[:, 3] = tilt
        rot, nt, tr = S.rows_to_docking(rows)
        I, c2 = sess.evaluate([(rot, nt, tr)], fit_scale=1e-5, fit_constant=0.2)
        eager, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=4, max_dirs=dirs,
                                             chunk=8192)
        assert np.array_equal(I, eager)
        ref = float(np.sum(((1e-5 * eager + 0.2 - y) / sig) ** 2))
        assert c2 == pytest.approx(ref, rel=1e-12)
        key = (n, rad, rise, twist, tilt)
        if key in prev:
            assert np.array_equal(prev[key], I)
        prev[key] = I
    info = sess.info()
    say(f"    session {info}")
    assert info["evaluations"] == 7 and info["graph_captures"] == 4
    sess.close()


def test_fit_session_rejects_q_outside_the_grid(dplus):
    """A session checks q like the orientation average does (q >= 0, q up to shell
    grid_size/2 + 1) before it allocates or captures anything."""
    from dplus_mi355x.serving import FitSession
    say("[case] fit session: q outside the grid is rejected at creation")
    pdb = make_pdb(200, 22, 1.0)
    g = dplus.Grid(24, 4.0, dplus.PRECISION_MIXED)
    dplus.Structure(pdb["xyz"], pdb["ion"], pdb["coef9"], pdb["rad"], sol_ed=333.0, flags=3).build(g)
    step = 4.0 / 12
    for q in (np.linspace(0.05, 13 * step * 1.001, 40), np.linspace(-0.1, 3.0, 40), np.linspace(0.05, 4.3 * 1.5, 40)):
        with pytest.raises(dplus.DplusError) as ei:
            FitSession([g], q, np.ones(40), np.ones(40), seed=1, directions=1024)
        assert ei.value.error_code == 21
    sess = FitSession([g], np.linspace(0.0, 13 * step, 40), np.ones(40), np.ones(40), seed=1, directions=1024)
    I, _ = sess.evaluate([(np.array([[0.0, 0.0, 0.0, 1.0]]), np.array([1]), np.zeros((1, 3)))])
    assert np.all(np.isfinite(I)) and np.all(I > 0)
    sess.close()


def test_concurrent_sessions_capture_while_others_allocate(dplus):
    """Eight fit sessions in threads, each re-capturing its graph as the copy
    count grows, while the main thread keeps allocating grids: every capture
    and allocation succeeds and every session still equals its eager average."""
    import threading
    from dplus_mi355x.serving import FitSession
    import dplus_mi355x.synth as S
    say("[case] 8 sessions capturing concurrently with allocations on other threads")
    pdb = make_pdb(300, 31, 1.0)
    g = dplus.Grid(20, 3.0, dplus.PRECISION_MIXED)
    dplus.Structure(pdb["xyz"], pdb["ion"], pdb["coef9"], pdb["rad"], sol_ed=333.0, flags=3).build(g)
    q = np.linspace(0.05, 3.0, 40)
    y, sig = np.linspace(2.0, 1.0, 40), np.full(40, 0.1)
    results, errors = {}, []

    def run(k):
        try:
            sess = FitSession([g], q, y, sig, seed=50 + k, directions=2048)
            for n in (3, 7, 12, 20):
                rot, nt, tr = S.rows_to_docking(S.helix_rows(n + k, 1.5, 0.5, 30.0 + k))
                results[(k, n)] = (sess.evaluate([(rot, nt, tr)])[0], (rot, nt, tr))
            sess.close()
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    scratch = []
    while any(t.is_alive() for t in threads):
        scratch.append(dplus.Grid(16, 2.0, dplus.PRECISION_MIXED))
        if len(scratch) > 6:
            scratch.pop(0)
    for t in threads:
        t.join()
    assert not errors, errors
    for (k, n), (I, (rot, nt, tr)) in sorted(results.items())[::5]:
        eager, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=50 + k, max_dirs=2048,
                                             chunk=2048)
        assert np.array_equal(I, eager)


def test_engine_structure_cache(dplus):
    """Requests with the same PDB content share one resident structure (a cache
    hit), a one-coordinate change is a different structure (a miss), and cached
    structures give the same curves as a fresh engine."""
    from dplus_mi355x.serving import Engine
    say("[case] engine structure cache: hits, misses, results")
    pdb = make_pdb(700, 41, 1.2)
    moved = dict(pdb, xyz=pdb["xyz"].copy())
    moved["xyz"][3, 0] += 0.05
    q = np.linspace(0.05, 3.0, 40)
    kw = dict(method=dplus.OA_MC, seed=3, max_dirs=4096, chunk=4096)
    e = Engine(streams=2)
    outs = [e.wait(e.submit(q, pdb=p, grid_size=gs, qmax=3.0, **kw))[0]
            for p, gs in ((pdb, 20), (pdb, 26), (moved, 20), (pdb, 20))]
    st = e.stats()
    e.close()
    say(f"    {st}")
    assert st["cache_misses"] == 2 and st["cache_hits"] == 2 and 0 < st["cache_bytes"] <= st["cache_capacity"]
    fresh = Engine(streams=1)
    for (p, gs), I in zip(((pdb, 20), (pdb, 26), (moved, 20)), outs):
        assert np.array_equal(fresh.wait(fresh.submit(q, pdb=p, grid_size=gs, qmax=3.0, **kw))[0], I)
    fresh.close()
    assert np.array_equal(outs[0], outs[3]) and not np.array_equal(outs[0], outs[2])


def test_engine_grouped_builds_same_as_alone(dplus):
    """With grouped builds on, every mixed PDB build goes through the batcher's
    shared launches and each request keeps the bits it has alone."""
    from dplus_mi355x.serving import Engine
    say("[case] engine: grouped builder launches vs one request at a time")
    reqs = requests(dplus)
    e1 = Engine(streams=1)
    alone = [e1.wait(e1.submit(q, **kw)) for q, kw in reqs]
    e1.close()
    eb = Engine(streams=8, batch_builds=True)
    tickets = [eb.submit(q, **kw) for q, kw in reqs]
    mix = [eb.wait(t) for t in tickets]
    st = eb.stats()
    eb.close()
    say(f"    engine stats {st}")
    nmixed = sum(1 for _, kw in reqs if kw["precision"] == dplus.PRECISION_MIXED)
    assert st["batched_builds"] == nmixed and 1 <= st["batch_launches"] <= nmixed
    for (I1, r1), (Ib, rb) in zip(alone, mix):
        assert np.array_equal(I1, Ib)
        assert r1["chi2"] == rb["chi2"]


def test_engine_autotunes_new_build_classes(dplus):
    """autotune=True: the first mixed build of a (grid, tile class) runs the
    autotuner and records its shape; results agree with the untuned engine."""
    from dplus_mi355x.serving import Engine
    say("[case] engine: autotune on first sight")
    pdb = make_pdb(1500, 21, 1.8)
    q = np.linspace(0.05, 5.0, 80)
    kw = dict(pdb=pdb, grid_size=40, qmax=5.0, method=dplus.OA_MC, seed=3, max_dirs=8192, chunk=4096)
    dplus.build_shapes_clear()
    try:
        e0 = Engine(streams=1)
        I0, _ = e0.wait(e0.submit(q, **kw))
        e0.close()
        assert dplus.build_shapes_save() == ""
        et = Engine(streams=2, autotune=True)
        I1, _ = et.wait(et.submit(q, **kw))
        I2, _ = et.wait(et.submit(q, **kw))
        et.close()
        table = dplus.build_shapes_save().split()
        say(f"    shape table {table}")
        assert len(table) == 4 and int(table[0]) == 40 // 2 + 3
        assert np.max(np.abs(I1 - I0) / I0) < 1e-6
        assert np.array_equal(I1, I2)
    finally:
        dplus.build_shapes_clear()
