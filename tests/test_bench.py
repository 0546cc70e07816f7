"""The measurement harness on the CPU: bench.py's one-GPU contract, the seeded
serving workload it replays, and the autotuner's coverage of that workload's
builder classes (no GPU is touched)."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_serving_workload_is_seeded_and_follows_the_grid_rule():
    import dplus_mi355x.synth as S
    pool, reqs, arr = S.serving_workload(n_requests=64, pool_size=6, seed=4)
    pool2, reqs2, arr2 = S.serving_workload(n_requests=64, pool_size=6, seed=4)
    assert np.array_equal(arr, arr2) and np.all(np.diff(arr) > 0)
    for (q, kw, n), (q2, kw2, n2) in zip(reqs, reqs2):
        assert n == n2 and np.array_equal(q, q2) and kw["grid_size"] == kw2["grid_size"]
        assert np.array_equal(kw["pdb"]["xyz"], kw2["pdb"]["xyz"])
    for pdb, L, n in pool:
        assert 500 <= n <= 200000 and len(pdb["xyz"]) == n
        assert abs(L - 2.0 * np.max(np.linalg.norm(pdb["xyz"], axis=1))) < 1e-12
    for q, kw, n in reqs:
        L = next(Lp for p, Lp, _ in pool if p is kw["pdb"])
        # grid size: 2 qmax L / pi + 3 rounded up to a multiple of 10 (CalculationInput.py:301-320)
        assert kw["grid_size"] % 10 == 0 and kw["grid_size"] >= 2 * kw["qmax"] * L / np.pi + 3
        assert kw["grid_size"] - 10 < 2 * kw["qmax"] * L / np.pi + 3
        assert 200 <= len(q) <= 1000 and 2.0 <= kw["qmax"] <= 8.0
        assert kw["method"] in (S.OA_MC, S.OA_VEGAS)


def test_autotuner_covers_every_serving_class():
    import dplus_mi355x.synth as S
    from dplus_mi355x import tune
    _, reqs, _ = S.serving_workload()
    classes = {(kw["grid_size"] // 2 + 3, tune._tile_class(tune._ntiles(kw["pdb"]["ion"]))) for _, kw, _ in reqs}
    specs = tune._serving_specs(S)
    tuned = {(gs // 2 + 3, tune._tile_class(tune._ntiles(p["ion"]))) for p, gs, _ in specs}
    assert tuned == classes and len(specs) == len(classes)
