"""Autotuning of the mixed builder's launch shapes for a set of workloads.

The builder's launch shape (workgroup size 64/128/256/512 units, atom slices
1..8) is chosen per (grid shells, tile-count class) from the library's shape
table, or by an occupancy model where the table has no entry.  This module
runs dplus_structure_autotune on one representative (structure, grid) of
every class the named workloads build and writes the table, which the package
loads at import from build_shapes_mi355x.txt:

    python -m dplus_mi355x.tune --out dplus-dev_amd/dplus_mi355x/build_shapes_mi355x.txt

Workloads: "serving" (the seeded 512-request serving stream, synth.serving_workload),
"subunits" (the 2,000-atom grid-100, 5,000-atom grid-200 and 2k..20k-atom
grid 150..250 subunits of the other benchmark configurations) and "large"
(the 50,000-atom grid-250 structure with its hydration layer).
"""
import argparse
import sys
import time

import numpy as np


def _tile_class(ntiles):
    return int(np.floor(2.0 * np.log2(max(ntiles, 1))))


def _ntiles(ion, nboxes_tiles=0):
    counts = np.bincount(np.asarray(ion))
    return int(np.sum((counts + 255) // 256)) + nboxes_tiles


def _serving_specs(S):
    pool, reqs, _ = S.serving_workload()
    out = {}
    for _, kw, n in reqs:
        pdb = kw["pdb"]
        N = kw["grid_size"] // 2 + 3
        key = (N, _tile_class(_ntiles(pdb["ion"])))
        out.setdefault(key, (pdb, kw["grid_size"], kw["qmax"]))
    return list(out.values())


def _subunit_specs(S):
    specs = []
    for n, radius, gs, qmax, seed in ((2000, 3.0, 100, 5.0, 11), (5000, 2.5, 200, 6.0, 12)):
        xyz, el = S.walk_structure(seed, n, (radius, radius, radius))
        coef9, rad, ion, _ = S.structure_tables(el)
        specs.append((dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, sol_ed=333.0, flags=3), gs, qmax))
    for n in (2000, 5000, 10000, 20000):
        R = (3.0 * n / (4.0 * np.pi * 50.0)) ** (1.0 / 3.0)
        xyz, el = S.walk_structure(300 + n, n, (R, R, R))
        coef9, rad, ion, _ = S.structure_tables(el)
        for gs in (150, 200, 250):
            specs.append((dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, sol_ed=333.0, flags=3), gs, 5.0))
    return specs


def _large_specs(D, S):
    xyz, el = S.walk_structure(2024, 50000, (8.0, 6.0, 5.0))
    coef9, rad, ion, _ = S.structure_tables(el)
    dims, com = D.hydration_boxes(xyz, rad[ion], probe=0.14, thickness=0.3, vox=0.05)
    return [(dict(xyz=xyz, ion=ion, coef9=coef9, rad=rad, sol_ed=333.0, flags=3, box_dims=dims, box_com=com,
                  vox_step=0.05, box_contrast=33.3), 250, 7.5)]


def tune(D, S, workloads=("serving", "subunits", "large"), log=print):
    """Autotunes one representative of every (grid, tile class) of the workloads;
    returns [(grid_size, natoms, shape dict)]."""
    specs = []
    if "serving" in workloads:
        specs += _serving_specs(S)
    if "subunits" in workloads:
        specs += _subunit_specs(S)
    if "large" in workloads:
        specs += _large_specs(D, S)
    out = []
    for pdb, gs, qmax in specs:
        st = D.Structure(**pdb)
        g = D.Grid(gs, qmax, D.PRECISION_MIXED)
        t0 = time.time()
        r = st.autotune(g)
        out.append((gs, len(pdb["xyz"]), r))
        log(f"[tune] grid {gs} atoms {len(pdb['xyz'])}: wg {r['wg']} slices {r['slices']} "
            f"{r['best_ms']:.3f} ms (model {r['model_ms']:.3f} ms), sweep {time.time() - t0:.1f} s")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="shape table file to write")
    ap.add_argument("--workloads", default="serving,subunits,large")
    args = ap.parse_args(argv)
    import dplus_mi355x as D
    import dplus_mi355x.synth as S
    D.build_shapes_clear()
    res = tune(D, S, tuple(args.workloads.split(",")), log=lambda *a: print(*a, flush=True))
    gain = [r["model_ms"] / r["best_ms"] for _, _, r in res if r["best_ms"] > 0]
    with open(args.out, "w") as f:
        f.write("# builder launch shapes measured on one MI355X by python -m dplus_mi355x.tune\n")
        f.write(f"# workloads {args.workloads}: {len(res)} classes, model/best time ratio median "
                f"{np.median(gain):.3f}, max {np.max(gain):.3f}\n")
        f.write("# grid_shells tile_class wg slices\n")
        f.write(D.build_shapes_save())
    print(f"[tune] {len(res)} classes written to {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
