"""The allocation-free orientation average (dplus_orientation_average_noalloc)
against dplus_orientation_average: the same kernels on the same seeded inputs,
so the intensities must be identical, with no device allocation during the
call; two calls in a row, the second smaller, for Monte Carlo (fixed count and
with the convergence test), VEGAS and docked leaves; sizes past the scratch
maximum and Gauss-Kronrod are rejected before any work."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def allocations(D):
    return D.device_info()["device_allocations"]


@pytest.fixture(scope="module")
def grid(dplus):
    import dplus_mi355x.synth as S
    say("[case] no-alloc: 400-atom grid, gs 24, qmax 5")
    xyz, el = S.walk_structure(17, 400, (1.4, 1.4, 1.4))
    coef9, rad, ion, _ = S.structure_tables(el)
    g = dplus.Grid(24, 5.0, dplus.PRECISION_MIXED)
    g.build_pdb(xyz, ion, coef9, rad, sol_ed=333.0, flags=dplus.CALC_ATOMIC_FORMFACTORS | dplus.CALC_DUMMY_SOLVENT)
    return g


def both(D, dock, q, **kw):
    from dplus_mi355x.grid import orientation_average_noalloc
    ref, st_ref = D.orientation_average(dock, q, **kw)
    a0 = allocations(D)
    mine, st = orientation_average_noalloc(dock, q, **kw)
    assert allocations(D) == a0, "the allocation-free call allocated device memory"
    return ref, mine, st_ref, st


@pytest.mark.parametrize("case", ["mc", "mc_converge", "vegas"])
def test_noalloc_equals_orientation_average_twice(dplus, grid, case):
    dock = dplus.Docking.single(grid)
    calls = {
        "mc": [dict(q=np.linspace(0, 5.0, 90), max_dirs=32768, chunk=4096),
               dict(q=np.linspace(0.1, 3.0, 25), max_dirs=8192, chunk=2048)],
        "mc_converge": [dict(q=np.linspace(0, 5.0, 70), max_dirs=65536, chunk=4096, eps=1e-3),
                        dict(q=np.linspace(0.2, 2.0, 11), max_dirs=16384, chunk=2048, eps=1e-3)],
        "vegas": [dict(q=np.linspace(0, 5.0, 80), method=dplus.OA_VEGAS, vegas_iterations=4,
                       vegas_points_per_bin=8),
                  dict(q=np.linspace(0.3, 2.5, 17), method=dplus.OA_VEGAS, vegas_iterations=2,
                       vegas_points_per_bin=4)],
    }[case]
    for k, kw in enumerate(calls):
        q = kw.pop("q")
        say(f"[case] no-alloc {case} call {k}: {len(q)} q, {kw}")
        ref, mine, st_ref, st = both(dplus, dock, q, seed=3 + k, **kw)
        assert np.array_equal(ref, mine)
        assert st.dirs_used == st_ref.dirs_used and st.iterations_used == st_ref.iterations_used
        assert np.all(np.isfinite(mine)) and np.all(mine > 0)


def test_noalloc_docked_leaves_twice(dplus, grid):
    import dplus_mi355x.synth as S
    for k, (n, rad) in enumerate(((24, 2.0), (6, 1.2))):
        say(f"[case] no-alloc docked helix of {n} copies")
        rot, nt, tr = S.rows_to_docking(S.helix_rows(n, rad, 0.6, 29.0))
        dock = dplus.Docking([(grid, rot, nt, tr)])
        q = np.linspace(0.05, 4.0, 60 - 30 * k)
        ref, mine, _, _ = both(dplus, dock, q, seed=11, max_dirs=8192 >> k, chunk=2048)
        assert np.array_equal(ref, mine)


def test_noalloc_rejects_oversize_and_gk(dplus, grid):
    from dplus_mi355x.grid import orientation_average_noalloc
    say("[case] no-alloc: scratch maximum and Gauss-Kronrod rejected")
    dock = dplus.Docking.single(grid)
    q = np.linspace(0, 5.0, 400)
    a0 = allocations(dplus)
    with pytest.raises(dplus.DplusError) as ei:
        orientation_average_noalloc(dock, q, max_dirs=1 << 24, chunk=8192)
    assert ei.value.error_code == 21 and "maximum" in str(ei.value)
    with pytest.raises(dplus.DplusError) as ei:
        orientation_average_noalloc(dock, q, method=dplus.OA_GK)
    assert ei.value.error_code == 21
    assert allocations(dplus) == a0
    # still usable afterwards
    ref, mine, _, _ = both(dplus, dock, q[:50], seed=2, max_dirs=4096, chunk=4096)
    assert np.array_equal(ref, mine)
