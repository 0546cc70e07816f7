"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same seeded inputs.  Tolerances: FP64 mode 1e-9 relative on grids and 1e-6 on
I(q) (the FP64-mode bar); mixed mode 1e-4 relative on I(q)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-6
MIXED_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def structure(n, seed, radius=1.5):
    import dplus_mi355x.synth as S
    xyz, el = S.walk_structure(seed, n, (radius, radius, radius))
    coef9, rad, ion, order = S.structure_tables(el)
    return xyz, ion, coef9, rad


def oracle_ions(O, coef9, rad, ion, sol=334.0, c1=1.0):
    flags = O.CALC_ATOMIC_FORMFACTORS | (O.CALC_DUMMY_SOLVENT if sol else 0)
    return O.Ions(coef9, rad, np.bincount(ion, minlength=len(coef9)), sol, c1, flags)


def gpu_grid(D, gs, qmax, prec, xyz, ion, coef9, rad, sol=334.0, **kw):
    g = D.Grid(gs, qmax, prec)
    flags = D.CALC_ATOMIC_FORMFACTORS | (D.CALC_DUMMY_SOLVENT if sol else 0)
    g.build_pdb(xyz, ion, coef9, rad, sol_ed=sol, flags=flags, **kw)
    return g


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def small(dplus, oracle):
    say("[case] small structure: 300 atoms, gs 24, qmax 5")
    xyz, ion, coef9, rad = structure(300, 7)
    gs, qmax = 24, 5.0
    ions = oracle_ions(oracle, coef9, rad, ion)
    data = oracle.grid_fill_pdb(gs, qmax, xyz, ion, ions)
    co = oracle.grid_splines(gs, data)
    return dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, gs=gs, qmax=qmax, ions=ions, data=data, co=co)


def test_device_is_gfx950(dplus):
    info = dplus.device_info()
    say("[device]", info)
    assert info["ok"] and info["arch"].startswith("gfx950")


@pytest.mark.parametrize("prec", [0, 1])
def test_grid_build_matches_oracle(dplus, oracle, small, prec):
    say(f"[case] grid build precision={prec}")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    tol = 1e-10 if prec == 0 else 2e-5
    assert relmax(g.data, s["data"]) < tol
    assert relmax(g.interp, s["co"]) < (1e-10 if prec == 0 else 2e-5)


@pytest.mark.parametrize("gs,natoms", [(2, 1), (4, 3), (6, 7), (10, 1), (16, 40), (40, 5)])
def test_grid_build_edge_sizes(dplus, oracle, gs, natoms):
    """The smallest grids (shells 1..4, where a pair unit's rows span the
    equator and a workgroup spans more rows than the shared-row cache holds),
    a single atom and a handful of atoms: every node of the mixed and FP64
    builds against the oracle's direct sum."""
    say(f"[case] grid build at grid size {gs} with {natoms} atoms")
    xyz, ion, coef9, rad = structure(natoms, 20 + gs, radius=0.6)
    ions = oracle_ions(oracle, coef9, rad, ion)
    ref = oracle.grid_fill_pdb(gs, 3.0, xyz, ion, ions)
    for prec, tol in ((0, 1e-10), (1, 2e-5)):
        g = gpu_grid(dplus, gs, 3.0, prec, xyz, ion, coef9, rad)
        assert relmax(g.data, ref) < tol


@pytest.mark.parametrize("gs,natoms", [(6, 7), (24, 300), (70, 4000)])
def test_fp64_paired_builder_matches_plain(dplus, gs, natoms):
    """FP64 grids: the paired builder (the default FP64 form: Friedel halving,
    phi / phi + pi node pairs sharing the sincos of q_x x + q_y y, the row's
    sincos of q_z z through LDS, atom slices) and the plain builder (one node at
    a time, one double sincos per term) give the same nodes and spline
    coefficients to rounding, with hydration-layer and excluded-volume boxes."""
    say(f"[case] FP64 paired vs plain builder, grid {gs}, {natoms} atoms")
    xyz, ion, coef9, rad = structure(natoms, 30 + gs, radius=1.2 if natoms < 1000 else 2.6)
    ld, lc = dplus.hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.28, vox=0.1)
    ed, ec = dplus.excluded_volume_boxes(xyz, rad[ion], probe=0.14, thickness=0.28, vox=0.1)
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=dplus.CALC_ATOMIC_FORMFACTORS, box_dims=ld,
                         box_com=lc, vox_step=0.1, box_contrast=33.3, excl_dims=ed, excl_com=ec,
                         excl_contrast=-333.0)
    g1 = dplus.Grid(gs, 5.0, 0)
    st.build(g1, host_copy=True)
    dplus.set_fp64_builder(True)
    try:
        g2 = dplus.Grid(gs, 5.0, 0)
        st.build(g2, host_copy=True)
    finally:
        dplus.set_fp64_builder(False)
    shape = st.build_shape(g1)
    err, err_i = relmax(g1.data, g2.data), relmax(g1.interp, g2.interp)
    say(f"    {len(ld)} layer and {len(ed)} excluded-volume boxes, shape {shape}; "
        f"nodes {err:.2e}, splines {err_i:.2e}")
    assert err < 1e-13 and err_i < 1e-12


def test_splines_of_uploaded_data(dplus, oracle):
    say("[case] splines of uploaded random data")
    gs = 40
    rng = np.random.default_rng(1)
    g = dplus.Grid(gs, 5.0, 0)
    d = rng.normal(size=2 * g.points)
    g.data[:] = d
    g.upload()
    g.download()
    assert np.array_equal(g.data, d)
    assert relmax(g.interp, oracle.grid_splines(gs, d)) < 1e-13


@pytest.mark.parametrize("gs", [12, 250, 350])
def test_splines_wave_rows_match_serial_sweeps(dplus, gs, monkeypatch):
    """The one-wave-per-row spline kernel (affine-map scans of the forward and
    backward sweeps across a row's lanes) against the serial one-thread-per-row
    sweeps on random data: FP64 to 1e-13 of the largest coefficient, mixed to
    1e-6 (the serial form rounds its forward values to FP32 between the sweeps).
    Grid 350 has shells beyond the wave kernel's 1,024-node rows, which keep the
    serial sweeps."""
    say(f"[case] spline rows by waves vs serial sweeps, grid {gs}")
    rng = np.random.default_rng(gs)
    for prec, tol in ((0, 1e-13), (1, 1e-6)):
        out = []
        for serial in (False, True):
            if serial:
                monkeypatch.setenv("DPLUS_SPLINES_SERIAL", "1")
            else:
                monkeypatch.delenv("DPLUS_SPLINES_SERIAL", raising=False)
            g = dplus.Grid(gs, 5.0, prec)
            if not out:
                d = rng.normal(size=2 * g.points)
            g.data[:] = d
            g.upload()
            g.download()
            out.append(g.interp.copy())
            del g
        monkeypatch.delenv("DPLUS_SPLINES_SERIAL", raising=False)
        err = relmax(out[0], out[1])
        say(f"    precision {prec}: max |wave - serial| / max {err:.2e}")
        assert err < tol


def test_get_sphr_matches_oracle(dplus, oracle, small):
    say("[case] get_sphr")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], 0, s["xyz"], s["ion"], s["coef9"], s["rad"])
    rng = np.random.default_rng(2)
    n = 400
    q = np.concatenate([rng.uniform(0, s["qmax"], n - 4), [0.0, s["qmax"], 1.0, 2.5]])
    th = np.concatenate([rng.uniform(0, np.pi, n - 4), [0.0, np.pi, 0.0, np.pi]])
    ph = np.concatenate([rng.uniform(0, 2 * np.pi, n - 4), [0.0, 2 * np.pi, 0.0, 3.0]])
    a = g.get_sphr(q, th, ph)
    b = np.array([oracle.get_sphr(s["gs"], s["qmax"], s["data"], s["co"], *x) for x in zip(q, th, ph)])
    assert relmax(a, b) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_interpolate_theta_phi_plane_matches_oracle(dplus, oracle, small, prec):
    """InterpolateThetaPhiPlane (Grid.cpp:61-134) through the C ABI against the oracle:
    FP64 grid 1e-10, mixed (FP32 storage) 1e-5 of the shell's largest value; the
    CJacobianSphereGrid mirror gives the same; out-of-range shells are rejected."""
    say(f"[case] interpolate_theta_phi_plane precision={prec}")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    N = s["gs"] // 2 + 3
    rng = np.random.default_rng(11)
    n = 600
    ri = np.concatenate([rng.integers(1, N + 1, n - 8), [1, 1, N, N, 2, 5, N, 1]])
    th = np.concatenate([rng.uniform(0, np.pi, n - 8), [0.0, np.pi, 0.0, np.pi, np.pi / 6, 1.0, 3.1, 0.0]])
    ph = np.concatenate([rng.uniform(0, 2 * np.pi, n - 8), [0.0, 2 * np.pi, 6.0, 0.0, np.pi, 2 * np.pi, 0.1, 3.0]])
    a = g.interpolate_theta_phi_plane(ri, th, ph)
    b = np.array([oracle.interp_plane(s["gs"], s["data"], s["co"], int(r), t, p) for r, t, p in zip(ri, th, ph)])
    assert relmax(a, b) < (1e-10 if prec == 0 else 1e-5)
    import dplus_mi355x.wrappers as W
    cg = W.CJacobianSphereGrid._wrap(g)
    assert cg.interpolate_theta_phi_plane(int(ri[0]), th[0], ph[0]) == a[0]
    assert cg.extras == 3
    for bad in ((0, 1.0, 1.0), (N + 1, 1.0, 1.0), (1, -0.1, 1.0), (1, 1.0, 7.0)):
        with pytest.raises(dplus.DplusError):
            g.interpolate_theta_phi_plane(*bad)


def test_seeded_intensity_matches_oracle(dplus, oracle, small):
    say("[case] seeded CalculateIntensity")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], 0, s["xyz"], s["ion"], s["coef9"], s["rad"])
    for q, eps, seed, n in ((2.3, 1e-3, 0, 20000), (0.7, 0.0, 5, 4000), (0.0, 1e-3, 1, 100)):
        mine = g.calculate_intensity(q, None, eps, seed, n)
        ref, _ = oracle.grid_intensity_seeded(s["gs"], s["qmax"], s["data"], s["co"], q, eps, seed, n)
        assert mine == pytest.approx(ref, rel=1e-10)


@pytest.mark.parametrize("prec", [0, 1])
def test_mc_matches_oracle(dplus, oracle, small, prec):
    say(f"[case] MC precision={prec}")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    q = np.linspace(0, s["qmax"], 61)
    mine, st = dplus.orientation_average(dplus.Docking.single(g), q, seed=9, max_dirs=4096, chunk=1024)
    ref, _ = oracle.mc_oa(oracle.Docked.single(s["gs"], s["qmax"], s["data"], s["co"]), q, 9, 4096, 1024)
    assert np.max(np.abs(mine - ref) / ref) < (FP64_TOL * 1e-3 if prec == 0 else MIXED_TOL)


def test_mc_convergence_path_matches_oracle(dplus, oracle, small):
    say("[case] MC with convergence")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], 0, s["xyz"], s["ion"], s["coef9"], s["rad"])
    q = np.linspace(0.05, s["qmax"], 40)
    mine, st = dplus.orientation_average(dplus.Docking.single(g), q, seed=3, max_dirs=20 * 512, chunk=512,
                                         eps=3e-3)
    ref, used = oracle.mc_oa(oracle.Docked.single(s["gs"], s["qmax"], s["data"], s["co"]), q, 3, 20 * 512, 512,
                             3e-3)
    assert st.dirs_used == used
    assert np.max(np.abs(mine - ref) / ref) < FP64_TOL


@pytest.mark.parametrize("prec", [0, 1])
def test_vegas_matches_oracle(dplus, oracle, small, prec):
    say(f"[case] VEGAS precision={prec}")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    q = np.linspace(0.02, s["qmax"], 45)
    mine, st = dplus.orientation_average(dplus.Docking.single(g), q, method=dplus.OA_VEGAS, seed=4,
                                         vegas_iterations=4, vegas_points_per_bin=2)
    ref, _ = oracle.vegas_oa(oracle.Docked.single(s["gs"], s["qmax"], s["data"], s["co"]), q, 4, 4, 2)
    assert np.max(np.abs(mine - ref) / ref) < (FP64_TOL if prec == 0 else MIXED_TOL)


def test_gk_matches_oracle(dplus, oracle, small):
    say("[case] Gauss-Kronrod")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], 0, s["xyz"], s["ion"], s["coef9"], s["rad"])
    q = np.array([0.0, 0.3, 1.1, 2.5, 4.0])
    mine, st = dplus.orientation_average(dplus.Docking.single(g), q, method=dplus.OA_GK, eps=1e-4)
    ref, ev = oracle.gk_oa(oracle.Docked.single(s["gs"], s["qmax"], s["data"], s["co"]), q, 1e-4)
    assert st.evaluations == ev
    assert np.max(np.abs(mine - ref) / ref) < FP64_TOL


@pytest.mark.parametrize("prec", [0, 1])
def test_gk_device_recursion_matches_host_replay(dplus, small, prec, monkeypatch):
    """The Gauss-Kronrod phi recursion grown on the device (accept / split and the
    left + right tree sums in gk_decide) against the host replay of the same
    recursion: identical intensities and evaluation counts, on a single grid and
    on a docked 12-copy helix, with depth limits and a q = 0 entry."""
    say(f"[case] GK device recursion vs host replay precision={prec}")
    import dplus_mi355x.synth as S
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    rot, nt, tr = S.rows_to_docking(S.helix_rows(12, 3.0, 0.9, 27.7))
    q = np.concatenate([[0.0], np.linspace(0.05, 4.5, 40)])
    for dock in (dplus.Docking.single(g), dplus.Docking([(g, rot, nt, tr)])):
        for eps, depth in ((1e-4, 1 << 20), (1e-6, 5)):
            monkeypatch.delenv("DPLUS_GK_HOST", raising=False)
            a, sa = dplus.orientation_average(dock, q, method=dplus.OA_GK, eps=eps, gk_max_depth=depth)
            monkeypatch.setenv("DPLUS_GK_HOST", "1")
            b, sb = dplus.orientation_average(dock, q, method=dplus.OA_GK, eps=eps, gk_max_depth=depth)
            monkeypatch.delenv("DPLUS_GK_HOST", raising=False)
            say(f"    eps {eps} depth {depth}: {sa.evaluations} evaluations, max |device - host| "
                f"{np.max(np.abs(a - b)):.2e}")
            assert sa.evaluations == sb.evaluations
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", [0, 1])
def test_hybrid_docking_mc_matches_oracle(dplus, oracle, small, prec):
    say(f"[case] hybrid docking MC precision={prec}")
    import dplus_mi355x.synth as S
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    rows = S.helix_rows(12, 3.0, 0.9, 27.7)
    rot, nt, tr = S.rows_to_docking(rows)
    q = np.linspace(0.05, 4.0, 50)
    mine, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=2, max_dirs=2048, chunk=1024)
    ref, _ = oracle.mc_oa(oracle.Docked([(s["gs"], s["qmax"], s["data"], s["co"], rot, nt, tr)]), q, 2, 2048, 1024)
    assert np.max(np.abs(mine - ref) / ref) < (FP64_TOL if prec == 0 else MIXED_TOL)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("case", ["two_leaves", "sparse", "one_copy_chunk"])
def test_zcopy_docking_mc_matches_oracle(dplus, oracle, small, prec, case, monkeypatch):
    """Copies that all turn about z run the copy-major MC kernel: against the
    oracle's docked MC, against the general kernel on the same input
    (DPLUS_ZCOPY=0), and in chunked launches.  two_leaves: a 70-copy helix whose
    rows repeat at a second height (two translations per rotation) and a
    37-copy helix of another twist, 177 copies (not a multiple of the wave, one
    chunk spanning both leaves); every shell group takes the collapsed lookup.
    sparse: 33 copies 137.5 deg apart, one chunk around the whole circle, so
    only the low shell groups fit the collapse window and the others take the
    direct lookup.  one_copy_chunk: 65 copies, the second chunk holding one
    copy (a segment of two nodes, 63 lanes repeating it with scale 0)."""
    say(f"[case] z-copy docking MC {case} precision={prec}")
    import dplus_mi355x.synth as S
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    if case == "two_leaves":
        a = S.helix_rows(70, 3.0, 0.9, 27.7)
        b = a.copy()
        b[:, 2] += 0.45
        rotA, ntA, trA = S.rows_to_docking(np.concatenate([a, b]))
        assert (ntA == 2).all()
        rotB, ntB, trB = S.rows_to_docking(S.helix_rows(37, 2.0, 0.6, 41.0), scale=0.7)
        leaves = [(g, rotA, ntA, trA), (g, rotB, ntB, trB)]
    elif case == "sparse":
        rot, nt, tr = S.rows_to_docking(S.helix_rows(33, 2.5, 0.3, 137.5))
        leaves = [(g, rot, nt, tr)]
    else:
        rot, nt, tr = S.rows_to_docking(S.helix_rows(65, 2.5, 0.3, 27.7))
        leaves = [(g, rot, nt, tr)]
    q = np.linspace(0.05, 4.0, 50)
    mine, _ = dplus.orientation_average(dplus.Docking(leaves), q, seed=2, max_dirs=2048, chunk=1024)
    ref, _ = oracle.mc_oa(oracle.Docked([(s["gs"], s["qmax"], s["data"], s["co"], r, n, t)
                                         for _, r, n, t in leaves]), q, 2, 2048, 1024)
    e = np.max(np.abs(mine - ref) / ref)
    monkeypatch.setenv("DPLUS_ZCOPY", "0")
    gen, _ = dplus.orientation_average(dplus.Docking(leaves), q, seed=2, max_dirs=2048, chunk=1024)
    monkeypatch.delenv("DPLUS_ZCOPY")
    # chunked MC (eps > 0): launches with a direction offset
    ch, st = dplus.orientation_average(dplus.Docking(leaves), q, seed=2, max_dirs=2048, chunk=1024, eps=1e-12)
    ab = np.max(np.abs(mine - gen) / gen)
    say(f"    vs oracle {e:.2e}, vs general kernel {ab:.2e}, chunked {np.max(np.abs(ch - mine) / mine):.2e}")
    assert e < (FP64_TOL if prec == 0 else MIXED_TOL)
    assert ab < (1e-12 if prec == 0 else 2e-6)
    assert st.dirs_used == 2048
    np.testing.assert_allclose(ch, mine, rtol=1e-12 if prec == 0 else 1e-6)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("method", ["mc", "vegas"])
def test_grouped_translations_mc_matches_oracle(dplus, oracle, small, prec, method):
    """An icosahedral shell of 180 copies (60 rotations x 3 translations): with
    equally spaced q the mixed kernels (docked MC and VEGAS) step each
    translation's phase from q to q (two range-reduced sin/cos per translation),
    with unevenly spaced q they take one sin/cos per q; both against the
    oracle's docked MC / VEGAS."""
    say(f"[case] grouped translations {method} precision={prec}")
    import dplus_mi355x.synth as S
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    rot, nt, tr = S.rows_to_docking(S.icosahedral_rows(4.0, 180, seed=1))
    assert len(rot) == 60 and (nt == 3).all()
    rng = np.random.default_rng(4)
    dock = oracle.Docked([(s["gs"], s["qmax"], s["data"], s["co"], rot, nt, tr)])
    for q in (np.linspace(0.05, 4.0, 60), np.sort(rng.uniform(0.05, 4.0, 60))):
        if method == "mc":
            mine, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=3, max_dirs=2048,
                                                chunk=1024)
            ref, _ = oracle.mc_oa(dock, q, 3, 2048, 1024)
        else:
            mine, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, method=dplus.OA_VEGAS,
                                                seed=3, vegas_iterations=3, vegas_points_per_bin=2)
            ref, _ = oracle.vegas_oa(dock, q, 3, 3, 2)
        e = np.max(np.abs(mine - ref) / ref)
        say(f"    {'uniform' if np.allclose(np.diff(q), np.diff(q)[0]) else 'uneven'} q: vs oracle {e:.2e}")
        assert e < (FP64_TOL if prec == 0 else MIXED_TOL)


def test_dock_grid_matches_oracle(dplus, oracle, small):
    say("[case] grid-level manual symmetry")
    import dplus_mi355x.synth as S
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], 0, s["xyz"], s["ion"], s["coef9"], s["rad"])
    rot, nt, tr = S.rows_to_docking(S.helix_rows(4, 1.2, 0.5, 40.0))
    out = dplus.Grid(20, 4.0, 0)
    dplus.dock_grid(out, dplus.Docking([(g, rot, nt, tr)]), 1.0)
    out.download()
    ref = oracle.dock_grid(20, 4.0, oracle.Docked([(s["gs"], s["qmax"], s["data"], s["co"], rot, nt, tr)]), 1.0)
    assert relmax(out.data, ref) < 1e-10


def test_single_atom_fixture_on_gpu(dplus, oracle):
    """The reference's golden single-atom curve (o_explicit_SS334) from the GPU path."""
    import json, os
    say("[case] golden o_explicit_SS334 on the GPU")
    rec = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "o_explicit_SS334.json")))
    import dplus_mi355x.elements as E
    coef9, rad, ion, _ = E.ion_tables(["O"])
    for prec in (0, 1):
        g = dplus.Grid(rec["domain"]["GridSize"], rec["domain"]["qMax"], prec)
        g.build_pdb([[0, 0, 0]], ion, coef9, rad, sol_ed=334.0,
                    flags=dplus.CALC_ATOMIC_FORMFACTORS | dplus.CALC_DUMMY_SOLVENT)
        I, _ = dplus.orientation_average(dplus.Docking.single(g), rec["q"], seed=1, max_dirs=256, chunk=256)
        ref = np.array(rec["intensity"])
        assert np.max(np.abs(I - ref)) <= 2e-5 * ref.max()


def test_friedel_and_debye_at_larger_size(dplus, oracle):
    """2,000 atoms in a 3 nm sphere, gs 100: the builder output is
    Friedel-symmetric and the MC curve follows the exact Debye sum within the
    grid's interpolation error."""
    say("[case] 2000 atoms gs 100 Debye")
    xyz, ion, coef9, rad = structure(2000, 1, radius=3.0)
    g = gpu_grid(dplus, 100, 5.0, 1, xyz, ion, coef9, rad)
    q = np.linspace(0.0, 5.0, 101)
    I, _ = dplus.orientation_average(dplus.Docking.single(g), q, seed=1, max_dirs=1 << 16, chunk=8192)
    deb = oracle.debye(xyz, ion, oracle_ions(oracle, coef9, rad, ion), q)
    err = np.abs(I - deb) / deb
    say(f"  max rel err vs Debye {err.max():.3e} median {np.median(err):.3e}")
    assert np.median(err) < 0.02 and err.max() < 0.2


def test_helix_of_1000_copies_of_a_5000_atom_grid200(dplus):
    """A 5,000-atom subunit on a grid of size 200 (qmax 6),
    docked as a 1,000-copy helix (radius 12 nm, rise 0.9 nm, twist 27.7 deg; the
    translations reach 900 nm), 800 q, MC with 8,192 directions.  Mixed and FP64
    agree within the mixed bar, and at q = 0 every copy adds the subunit's
    forward amplitude in phase: I(0) = (1000 A(0))^2 exactly (FP64 1e-12)."""
    import dplus_mi355x.synth as S
    say("[case] docking: 5000 atoms gs 200, helix of 1000 copies, 800 q")
    xyz, ion, coef9, rad = structure(5000, 3, radius=2.5)
    rot, nt, tr = S.rows_to_docking(S.helix_rows(1000, 12.0, 0.9, 27.7))
    q = np.linspace(0.0, 6.0, 800)
    out = {}
    for prec in (0, 1):
        g = gpu_grid(dplus, 200, 6.0, prec, xyz, ion, coef9, rad)
        I, _ = dplus.orientation_average(dplus.Docking([(g, rot, nt, tr)]), q, seed=5, max_dirs=8192, chunk=8192)
        g.download()
        a0 = float(np.real(g.data[0]))
        out[prec] = I
        closed = (1000.0 * a0) ** 2
        say(f"  precision {prec}: I(0) {I[0]:.9e} closed form {closed:.9e}")
        assert I[0] == pytest.approx(closed, rel=1e-12 if prec == 0 else 1e-4)
    err = np.max(np.abs(out[1] - out[0]) / out[0])
    say(f"  max rel mixed vs FP64 {err:.3e}")
    assert err < MIXED_TOL


@pytest.mark.parametrize("prec", [0, 1])
def test_hydration_boxes_match_oracle(dplus, oracle, prec):
    """Atoms plus a hydration layer of solvent boxes (the builder's box tiles)
    against the oracle's direct box sums, through the resident-structure API."""
    say(f"[case] atoms + hydration boxes precision={prec}")
    xyz, ion, coef9, rad = structure(400, 5, radius=1.2)
    dims, com = dplus.hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.3, vox=0.1)
    assert len(dims) > 10
    gs, qmax, contrast = 24, 5.0, 33.3
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3, box_dims=dims, box_com=com, vox_step=0.1,
                         box_contrast=contrast)
    g = dplus.Grid(gs, qmax, prec)
    st.build(g, host_copy=True)
    ions = oracle_ions(oracle, coef9, rad, ion, sol=333.0)
    ref = oracle.grid_fill_pdb(gs, qmax, xyz, ion, ions, 1.0, dims, com, 0.1, contrast)
    assert relmax(g.data, ref) < (1e-10 if prec == 0 else 2e-5)


def test_kernel_timing_and_microbench(dplus):
    say("[case] timing hooks and VALU microbenchmark")
    mb = dplus.microbench_valu()
    say("  ", mb)
    assert all(v > 1e12 for v in mb.values())
    xyz, ion, coef9, rad = structure(200, 2)
    st = dplus.Structure(xyz, ion, coef9, rad)
    g = dplus.Grid(30, 4.0, 1)
    dplus.timing_reset()
    dplus.timing_enable(True)
    st.build(g)
    st.build(g)
    dplus.timing_enable(False)
    n, ms = dplus.timing_get("build")
    assert n == 2 and ms > 0


def test_rebuild_is_idempotent(dplus):
    """Building the same grid twice gives the same values (regression: the
    origin node was accumulated across builds)."""
    say("[case] rebuild idempotence")
    xyz, ion, coef9, rad = structure(150, 4)
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3)
    for prec in (0, 1):
        g = dplus.Grid(20, 4.0, prec)
        st.build(g, host_copy=True)
        first = g.data.copy()
        st.build(g, host_copy=True)
        st.build(g, host_copy=True)
        assert np.array_equal(g.data, first)


def test_grouped_build_matches_single_builds(dplus, oracle):
    """One grouped builder launch over grids of different structures, atom
    counts, hydration boxes, grid sizes and phase forms gives every grid the
    bits of its own single build, and the oracle's values."""
    say("[case] grouped build: 6 grids (5 mixed + 1 FP64) in one call vs one by one")
    specs = [(120, 3, 1.0, 16, 4.0, 1), (900, 4, 1.6, 40, 6.0, 1), (2500, 5, 2.2, 30, 5.0, 1),
             (60, 6, 0.8, 24, 7.5, 0), (700, 7, 1.4, 50, 3.0, 1), (400, 8, 1.5, 44, 7.0, 1)]
    sts, grids, singles, refs = [], [], [], []
    for i, (n, seed, radius, gs, qmax, prec) in enumerate(specs):
        xyz, ion, coef9, rad = structure(n, seed, radius)
        if i == 5:
            xyz = xyz + np.array([100.0, 0.0, 0.0])  # phases beyond 128 revolutions: the reduced form
        kw = {}
        if i == 2:
            dims, com = dplus.hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.3, vox=0.1)
            kw = dict(box_dims=dims, box_com=com, vox_step=0.1, box_contrast=33.3)
        st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3, **kw)
        g1 = dplus.Grid(gs, qmax, prec)
        st.build(g1, host_copy=True)
        sts.append(st)
        grids.append(dplus.Grid(gs, qmax, prec))
        singles.append(g1.data.copy())
        if i == 1:
            ions = oracle_ions(oracle, coef9, rad, ion, sol=333.0)
            refs.append((i, oracle.grid_fill_pdb(gs, qmax, xyz, ion, ions)))
    dplus.build_grids(sts, grids, host_copy=True)
    for i, (g, one) in enumerate(zip(grids, singles)):
        assert np.array_equal(g.data, one), i
    for i, ref in refs:
        assert relmax(grids[i].data, ref) < 2e-5, i


def test_autotune_shape_table(dplus):
    """The autotuner times every launch shape, records the fastest for the
    grid's class, and leaves the grid with the values of a default build."""
    say("[case] autotune: 3,000 atoms at grid 60")
    xyz, ion, coef9, rad = structure(3000, 9, 2.0)
    st = dplus.Structure(xyz, ion, coef9, rad, sol_ed=333.0, flags=3)
    dplus.build_shapes_clear()
    try:
        g0 = dplus.Grid(60, 6.0, 1)
        st.build(g0, host_copy=True)
        g = dplus.Grid(60, 6.0, 1)
        shape = st.autotune(g)
        say("   ", shape)
        assert shape["wg"] in (64, 128, 256, 512) and 1 <= shape["slices"] <= 8
        assert 0 < shape["best_ms"] <= shape["model_ms"]
        g.download()
        assert relmax(g.data, g0.data) < 1e-6
        table = dplus.build_shapes_save()
        lines = table.split()
        assert len(lines) == 4 and int(lines[0]) == g.actual_size
        assert int(lines[2]) == shape["wg"] and int(lines[3]) == shape["slices"]
        # a build of the same class takes the recorded shape: same values as the default build
        g2 = dplus.Grid(60, 6.0, 1)
        st.build(g2, host_copy=True)
        assert relmax(g2.data, g0.data) < 1e-6
        dplus.build_shapes_clear()
        dplus.build_shapes_load(table)
        assert dplus.build_shapes_save() == table
    finally:
        dplus.build_shapes_clear()


@pytest.mark.parametrize("prec,gs,plain", [(0, 40, False), (0, 42, False), (1, 40, False), (1, 42, False),
                                           (1, 40, True)])
def test_build_writes_every_node(dplus, prec, gs, plain):
    """A build writes every node of its grid and every spline coefficient (the
    request engine builds into unzeroed grids): a grid whose nodes all hold NaN
    before the build gives the bits of a build into a zeroed grid."""
    say(f"[case] build over NaN nodes: precision {prec}, grid {gs}, plain FP64 builder {plain}")
    xyz, ion, coef9, rad = structure(1500, 9)
    dplus.set_fp64_builder(plain)
    try:
        ref = gpu_grid(dplus, gs, 4.0, prec, xyz, ion, coef9, rad)
        ref.download()
        g = dplus.Grid(gs, 4.0, prec)
        g.data[:] = np.nan
        g.upload()  # every node value NaN, so every spline coefficient but the origin's (always 0)
        g.download()
        assert np.isnan(g.data).all() and np.isnan(g.interp[2:]).all()
        g.build_pdb(xyz, ion, coef9, rad, sol_ed=334.0, flags=dplus.CALC_ATOMIC_FORMFACTORS | dplus.CALC_DUMMY_SOLVENT)
        g.download()
    finally:
        dplus.set_fp64_builder(False)
    assert np.array_equal(g.data, ref.data) and np.array_equal(g.interp, ref.interp)


def test_2000_atoms_grid100_500q_mc_and_gk_vs_debye(dplus, oracle):
    """The 2,000-atom workload as specified (3 nm sphere, solvent 333 e/nm^3,
    grid 100, qmax 5, 500 q, MC over 10^6 directions in chunks of 8,192, fixed
    seed) against the FP64 Debye sum: the MC curve within its sampling error,
    Gauss-Kronrod (no sampling error) within the grid's interpolation error,
    and mixed within 1e-4 of FP64."""
    say("[case] 2000 atoms, grid 100, 500 q, MC 10^6 directions and GK vs Debye")
    xyz, ion, coef9, rad = structure(2000, 1, radius=3.0)
    q = np.linspace(0.0, 5.0, 500)
    deb = oracle.debye(xyz, ion, oracle_ions(oracle, coef9, rad, ion, sol=333.0), q)
    curves = {}
    for prec in (0, 1):
        g = gpu_grid(dplus, 100, 5.0, prec, xyz, ion, coef9, rad, sol=333.0)
        curves[prec], st = dplus.orientation_average(dplus.Docking.single(g), q, seed=1, max_dirs=10 ** 6,
                                                     chunk=8192)
        assert st.dirs_used == 123 * 8192  # 10^6 rounded up to whole chunks
        if prec == 0:
            sub = q[::25]
            gk, _ = dplus.orientation_average(dplus.Docking.single(g), sub, method=dplus.OA_GK, eps=1e-6)
            egk = np.abs(gk - deb[::25]) / deb[::25]
            say(f"  GK (eps 1e-6, {len(sub)} q) vs Debye: max {egk.max():.3e} median {np.median(egk):.3e}")
    err = np.abs(curves[0] - deb) / deb
    say(f"  MC FP64 vs Debye: max {err.max():.3e} median {np.median(err):.3e}")
    mix = np.max(np.abs(curves[1] - curves[0]) / curves[0])
    say(f"  MC mixed vs FP64: max {mix:.3e}")
    assert mix < 1e-4
    # measured on MI355X: GK median 9.4e-4, max 2.9e-3 (the grid's interpolation error);
    # MC median 1.1e-3, max 3.3e-3
    assert np.median(egk) < 3e-3 and egk.max() < 1e-2
    assert np.median(err) < 3e-3 and err.max() < 1e-2


@pytest.mark.parametrize("prec", [0, 1])
def test_azimuthal_average_matches_oracle(dplus, oracle, small, prec):
    """dplus_azimuthal_average (the 2-D path's phi average) against the oracle on the
    same stratified Philox azimuths, a single grid and a docked helix."""
    import dplus_mi355x.synth as S
    say(f"[case] azimuthal average precision={prec}")
    s = small
    g = gpu_grid(dplus, s["gs"], s["qmax"], prec, s["xyz"], s["ion"], s["coef9"], s["rad"])
    g.download()
    rng = np.random.default_rng(4)
    q = np.concatenate([rng.uniform(0, s["qmax"], 60), [0.0, s["qmax"]]])
    th = np.concatenate([rng.uniform(0, np.pi, 60), [0.3, np.pi]])
    rot, nt, tr = S.rows_to_docking(S.helix_rows(6, 1.5, 0.4, 33.0))
    for dock, odock in ((dplus.Docking.single(g), oracle.Docked.single(s["gs"], s["qmax"], g.data, g.interp)),
                        (dplus.Docking([(g, rot, nt, tr)]),
                         oracle.Docked([(s["gs"], s["qmax"], g.data, g.interp, rot, nt, tr)]))):
        mine = dplus.azimuthal_average(dock, q, th, nphi=1000, seed=3)
        ref = oracle.azimuthal(odock, q, th, 1000, 3)
        assert np.max(np.abs(mine - ref) / ref) < (1e-10 if prec == 0 else 1e-4)
