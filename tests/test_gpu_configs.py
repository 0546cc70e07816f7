"""The benchmark workloads at their stated sizes on the GPU, against the FP64
oracle on the same seeded inputs.

* 50,000 atoms + hydration layer, grid 250, 1,000 q: sampled grid nodes against
  the oracle's direct atom and box sums, VEGAS 10 x 100,352 and Gauss-Kronrod
  (tolerance 1e-4) mixed against FP64, and the oracle's VEGAS on the FP64 grid.
* 5,000-atom subunit on a grid of 200 docked as a 1,000-copy helix, 800 q, MC
  over 2 x 10^6 directions: mixed against FP64 and the closed form at q = 0;
  the oracle's docked MC on the same grids.
* 512-request serving stream, every request its own structure (500..200,000
  atoms): alone against in the mix (bit-identical), mixed against FP64, and
  eight requests (largest atom count, largest grid) against the oracle.
* 16 fit sessions (2,000..20,000-atom subunits on grids of 150..250, 60..1,980
  helix or icosahedral copies, 600 q): one evaluation each against the oracle's
  docked MC on the downloaded grids, chi^2 against the host's.

Directions are keyed by (seed, direction, shell layer), so the oracle can check
the q values of a few shell layers of a full-size average exactly.
Tolerances: FP64 mode 1e-6 relative on I(q) (grid nodes 1e-9 of the shell's RMS
amplitude); mixed 1e-4 on I(q) (grid nodes 2e-5 of the shell's RMS amplitude).
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-6
MIXED_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


def sample_nodes(O, gs, qmax, per_shell=48, nshells=40, seed=0):
    """Node indices: the origin, and per_shell nodes in each of ~nshells shells
    (always the first three and the last two), among them both poles, the
    equator row (when 3i is even) and random (theta row, phi) pairs."""
    N = gs // 2 + 3
    rng = np.random.default_rng(seed)
    shells = sorted(set([1, 2, 3, N - 1, N] + [int(x) for x in np.linspace(1, N, nshells)]))
    idx, shell_of = [0], [0]
    for i in shells:
        picks = [(0, 0), (0, 3 * i), (3 * i, 0), (3 * i, 5)]
        if (3 * i) % 2 == 0:
            picks += [(3 * i // 2, 0), (3 * i // 2, 3 * i), (3 * i // 2, 1)]
        while len(picks) < per_shell:
            picks.append((int(rng.integers(0, 3 * i + 1)), int(rng.integers(0, 6 * i))))
        for t, p in picks:
            idx.append(O.index_from_indices(i, t, p % (6 * i)))
            shell_of.append(i)
    idx = np.array(idx, dtype=np.int64)
    qv = np.array([O.qvector_from_index(gs, qmax, int(k)) for k in idx])
    return idx, np.array(shell_of), qv


def node_errors(vals, ref, shell_of):
    """Per shell: max |vals - ref| over the sampled nodes / the RMS of |ref| there."""
    worst = 0.0
    for i in np.unique(shell_of):
        m = shell_of == i
        scale = np.sqrt(np.mean(np.abs(ref[m]) ** 2))
        worst = max(worst, float(np.max(np.abs(vals[m] - ref[m])) / scale))
    return worst


def grid_values(g, idx):
    d = g.data
    return d[2 * idx] + 1j * d[2 * idx + 1]


def layer_subset(q, step, layers):
    """Indices of the q values that fall in the given shell layers (q in
    (l step, (l+1) step], q = 0 in layer 0)."""
    lay = np.maximum(np.ceil(q / step) - 1, 0).astype(int)
    return np.nonzero(np.isin(lay, layers))[0]


def oracle_ions(O, coef9, rad, ion, sol=333.0):
    return O.Ions(coef9, rad, np.bincount(ion, minlength=len(coef9)), sol, 1.0,
                  O.CALC_ATOMIC_FORMFACTORS | O.CALC_DUMMY_SOLVENT)


# ---------------------------------------------------------------- 50,000 atoms + hydration, grid 250
def test_50k_atoms_hydration_grid250_vs_oracle(dplus, oracle):
    import dplus_mi355x.synth as S
    O = oracle
    say("[case] 50k atoms + 0.3 nm hydration layer, grid 250, qmax 7.5, 1000 q")
    t0 = time.time()
    xyz, el = S.walk_structure(2024, 50000, (8.0, 6.0, 5.0))
    coef9, rad, ion, _ = S.structure_tables(el)
    dims, com = dplus.hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.3, vox=0.05)
    vox, contrast, gs, qmax = 0.05, 366.3 - 333.0, 250, 7.5
    say(f"  {len(xyz)} atoms, {len(dims)} boxes, host prep {time.time() - t0:.1f} s")
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, c1=1.0, flags=3, box_dims=dims, box_com=com,
                         vox_step=vox, box_contrast=contrast)
    grids = {}
    for prec in (0, 1):
        say(f"  [step] build precision {prec}")
        g = dplus.Grid(gs, qmax, prec)
        st.build(g, host_copy=True)
        grids[prec] = g
    # sampled nodes against the oracle's direct sums
    idx, shell_of, qv = sample_nodes(O, gs, qmax)
    assert len(idx) >= 2000
    say(f"  [step] oracle direct sums at {len(idx)} nodes")
    ref = O.amp_atoms(xyz, ion, oracle_ions(O, coef9, rad, ion), qv) + O.amp_boxes(dims, com, vox, contrast, qv)
    for prec, tol in ((0, 1e-9), (1, 2e-5)):
        vals = grid_values(grids[prec], idx)
        e = node_errors(vals, ref, shell_of)
        e0 = abs(vals[0] - ref[0]) / abs(ref[0])
        say(f"  precision {prec}: nodes max err / shell RMS {e:.3e}, origin {e0:.3e}")
        assert e < tol and e0 < tol
    # VEGAS 10 x 100,352 and Gauss-Kronrod, mixed against FP64
    q = np.linspace(0.0, qmax, 1000)
    curves, gks = {}, {}
    for prec in (0, 1):
        say(f"  [step] VEGAS and GK precision {prec}")
        curves[prec], _ = dplus.orientation_average(dplus.Docking.single(grids[prec]), q, method=dplus.OA_VEGAS,
                                                    seed=7, vegas_iterations=10, vegas_points_per_bin=98)
        gks[prec], _ = dplus.orientation_average(dplus.Docking.single(grids[prec]), q, method=dplus.OA_GK,
                                                 eps=1e-4, gk_max_depth=8)
    ev, eg = rel(curves[1], curves[0]), rel(gks[1], gks[0])
    say(f"  mixed vs FP64: VEGAS {ev:.3e}, GK {eg:.3e}")
    assert ev < MIXED_TOL and eg < MIXED_TOL
    # the oracle's VEGAS on the downloaded FP64 grid, on the q of five shell layers
    step = grids[0].step_size
    sub = layer_subset(q, step, [0, 1, 40, 90, 124])
    say(f"  [step] oracle VEGAS on {len(sub)} q")
    d64 = O.Docked.single(gs, qmax, grids[0].data, grids[0].interp)
    orc, _ = O.vegas_oa(d64, q[sub], 7, 10, 98)
    e64, e32 = rel(curves[0][sub], orc), rel(curves[1][sub], orc)
    say(f"  VEGAS vs oracle: FP64 {e64:.3e}, mixed {e32:.3e}")
    assert e64 < FP64_TOL and e32 < MIXED_TOL


# ---------------------------------------------------------------- 1,000-copy helix, 2 x 10^6 directions
def test_helix_of_1000_copies_2e6_directions(dplus, oracle):
    """A 5,000-atom subunit (2.5 nm sphere) on a grid of 200 (qmax 6) docked as a
    1,000-copy helix (radius 12 nm, rise 0.9 nm, twist 27.7 deg; translations to
    ~450 nm), 800 q on [0, 6], MC over 2 x 10^6 directions in both precisions:
    mixed within 1e-4 of FP64 and I(0) = (1000 A(0))^2; the oracle's docked MC on
    the downloaded grids (8 shell layers, 4,096 directions)."""
    import dplus_mi355x.synth as S
    O = oracle
    say("[case] 5000-atom subunit gs 200, helix of 1000 copies, 800 q, MC 2e6 directions")
    xyz, el = S.walk_structure(3, 5000, (2.5, 2.5, 2.5))
    coef9, rad, ion, _ = S.structure_tables(el)
    rot, nt, tr = S.rows_to_docking(S.helix_rows(1000, 12.0, 0.9, 27.7))
    q = np.linspace(0.0, 6.0, 800)
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3)
    out, grids = {}, {}
    for prec in (0, 1):
        g = dplus.Grid(200, 6.0, prec)
        st.build(g, host_copy=True)
        grids[prec] = g
        say(f"  [step] MC 2e6 directions precision {prec}")
        t0 = time.time()
        I, stt = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=5, max_dirs=2_000_000,
                                           chunk=8192)
        say(f"    {stt.dirs_used} directions in {time.time() - t0:.1f} s")
        assert stt.dirs_used >= 2_000_000
        out[prec] = I
        closed = (1000.0 * float(g.data[0])) ** 2
        assert I[0] == pytest.approx(closed, rel=1e-12 if prec == 0 else 1e-4)
    err = rel(out[1], out[0])
    say(f"  mixed vs FP64 {err:.3e}")
    assert err < MIXED_TOL
    step = grids[0].step_size
    sub = layer_subset(q, step, [0, 1, 5, 20, 45, 70, 95, 99])
    say(f"  [step] oracle docked MC on {len(sub)} q, 4096 directions")
    for prec, tol in ((0, FP64_TOL), (1, MIXED_TOL)):
        g = grids[prec]
        mine, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q[sub], seed=9, max_dirs=4096,
                                            chunk=4096)
        ref, _ = O.mc_oa(O.Docked([(200, 6.0, g.data, g.interp, rot, nt, tr)]), q[sub], 9, 4096, 4096)
        e = rel(mine, ref)
        say(f"    precision {prec} vs oracle {e:.3e}")
        assert e < tol


# ---------------------------------------------------------------- 512-request serving stream
def test_serving_stream_512_distinct_structures(dplus, oracle):
    from dplus_mi355x.serving import Engine
    import dplus_mi355x.synth as S
    O = oracle
    t0 = time.time()
    _, reqs, _ = S.serving_workload(512)
    say(f"[case] serving: 512 requests, {sum(n for *_, n in reqs)} atoms in 512 distinct structures "
        f"(prepared in {time.time() - t0:.1f} s)")
    say("  [step] mixed, one request at a time on one stream, eager launches")
    e1 = Engine(hbm_budget_bytes=200 * 10 ** 9, streams=1, graphs=False)
    alone = [e1.wait(e1.submit(q, precision=dplus.PRECISION_MIXED, **kw))[0] for q, kw, _ in reqs]
    e1.close()
    say("  [step] mixed, all submitted at once on 16 streams, one captured graph per request")
    e16 = Engine(hbm_budget_bytes=200 * 10 ** 9, streams=16, graphs=True)
    tickets = [e16.submit(q, precision=dplus.PRECISION_MIXED, **kw) for q, kw, _ in reqs]
    mix = [e16.wait(t)[0] for t in tickets]
    st = e16.stats()
    e16.close()
    say(f"    max concurrent {st['max_concurrent']}, cache hits {st['cache_hits']}, graph launches "
        f"{st['graph_launches']} (updated {st['graph_updates']})")
    assert st["max_concurrent"] > 1 and st["graph_launches"] == len(reqs)
    for i, (a, m) in enumerate(zip(alone, mix)):
        assert np.array_equal(a, m), i
    say("  [step] FP64, 16 streams")
    e64 = Engine(hbm_budget_bytes=200 * 10 ** 9, streams=16)
    tickets = [e64.submit(q, precision=dplus.PRECISION_FP64, **kw) for q, kw, _ in reqs]
    fp64 = [e64.wait(t)[0] for t in tickets]
    e64.close()
    errs = [rel(m, f) for m, f in zip(mix, fp64)]
    say(f"  mixed vs FP64: max {max(errs):.3e}")
    assert max(errs) < MIXED_TOL
    # eight requests against the oracle: the largest atom count, the largest grid,
    # and six more of both averaging methods
    n_at = np.array([n for *_, n in reqs])
    gsz = np.array([kw["grid_size"] for _, kw, _ in reqs])
    chosen = [int(np.argmax(n_at)), int(np.argmax(gsz))]
    for k in np.argsort(n_at)[::64]:
        if int(k) not in chosen and len(chosen) < 8:
            chosen.append(int(k))
    for k in chosen:
        q, kw, n = reqs[k]
        p = kw["pdb"]
        gs, qmax = kw["grid_size"], kw["qmax"]
        say(f"  [step] oracle: request {k}, {n} atoms, grid {gs}, qmax {qmax:.2f}, {len(q)} q, "
            f"{'MC' if kw['method'] == dplus.OA_MC else 'VEGAS'}")
        g = dplus.Grid(gs, qmax, dplus.PRECISION_FP64)
        dplus.Structure(p["xyz"], p["ion"], p["coef9"], p["rad"], sol_ed=p["sol_ed"], flags=p["flags"]).build(
            g, host_copy=True)
        idx, shell_of, qv = sample_nodes(O, gs, qmax, per_shell=12, nshells=12, seed=k)
        ref = O.amp_atoms(p["xyz"], p["ion"], oracle_ions(O, p["coef9"], p["rad"], p["ion"], p["sol_ed"]), qv)
        assert node_errors(grid_values(g, idx), ref, shell_of) < 1e-9
        step = g.step_size
        nl = int(np.ceil(q[-1] / step))
        sub = layer_subset(q, step, sorted({0, 1, nl // 3, (2 * nl) // 3, nl - 1}))
        dock = O.Docked.single(gs, qmax, g.data, g.interp)
        if kw["method"] == dplus.OA_MC:
            orc, _ = O.mc_oa(dock, q[sub], kw["seed"], kw["max_dirs"], kw["chunk"])
        else:
            orc, _ = O.vegas_oa(dock, q[sub], kw["seed"], kw["vegas_iterations"], kw["vegas_points_per_bin"])
        e64, e32 = rel(fp64[k][sub], orc), rel(mix[k][sub], orc)
        say(f"    vs oracle on {len(sub)} q: FP64 {e64:.3e}, mixed {e32:.3e}")
        assert e64 < FP64_TOL and e32 < MIXED_TOL


# ---------------------------------------------------------------- 16 fit sessions
def test_16_fit_sessions_vs_oracle(dplus, oracle):
    from dplus_mi355x.serving import FitSession
    import dplus_mi355x.synth as S
    O = oracle
    t0 = time.time()
    plans = S.session_plans(16, 5)
    say(f"[case] 16 fit sessions (prepared in {time.time() - t0:.1f} s): grids "
        f"{[p['gs'] for p in plans]}, copies {[p['copies'] for p in plans]}")
    qmax, nq, dirs = 5.0, 600, 2048
    q = np.linspace(0.01, qmax, nq)
    rng = np.random.default_rng(3)
    for si, p in enumerate(plans):
        say(f"  [step] session {si}: {p['kind']}, {len(p['leaves'])} subunits "
            f"({[len(l[0]) for l in p['leaves']]} atoms), grid {p['gs']}, {p['copies']} copies")
        grids = []
        for xyz, ion, coef9, rad in p["leaves"]:
            g = dplus.Grid(p["gs"], qmax, dplus.PRECISION_MIXED)
            dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3).build(g, host_copy=True)
            grids.append(g)
        y = rng.uniform(0.5, 2.0, nq)
        sig = rng.uniform(0.05, 0.2, nq)
        sess = FitSession(grids, q, y, sig, seed=p["seed"], directions=dirs)
        pl = S.session_placements(p, 1.03, 1.5)
        I, c2 = sess.evaluate(pl, fit_scale=2e-7, fit_constant=0.3)
        sess.close()
        host = float(np.sum(((2e-7 * I + 0.3 - y) / sig) ** 2))
        assert c2 == pytest.approx(host, rel=1e-12)
        step = grids[0].step_size
        nl = int(np.ceil(q[-1] / step))
        sub = layer_subset(q, step, [1, nl // 2, nl - 1])
        dock = O.Docked([(p["gs"], qmax, g.data, g.interp, rot, nt, tr) for g, (rot, nt, tr) in zip(grids, pl)])
        ref, _ = O.mc_oa(dock, q[sub], p["seed"], dirs, dirs)
        e = rel(I[sub], ref)
        say(f"    vs oracle on {len(sub)} q: {e:.3e}")
        assert e < MIXED_TOL
        del grids, dock
