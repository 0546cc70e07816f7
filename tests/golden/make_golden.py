"""Extract the reference's reviewer-test fixtures for the hot path into small
JSON files under tests/golden/.  Reads data files only (PDB text, .state JSON
and the <name>TestStandard.out mean/sigma curves) from the reference tree and
stores them with the full state document:

  PythonInterface/tests/reviewer_tests/files_for_tests/generate/cpu/short/<name>/

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os

REF = "/root/reference/PythonInterface/tests/reviewer_tests/files_for_tests/generate/cpu/short"
OUT = os.path.dirname(os.path.abspath(__file__))
CASES = ([f"{e}_explicit_{s}" for e in ("h", "c", "n", "o", "points") for s in ("Vacuum", "SS334")] +
         [f"{g}_implicit_{s}" for g in ("ch", "ch2", "ch3", "nh", "nh2", "nh3", "oh") for s in ("Vacuum", "SS334")])


def read_pdb(path):
    atoms = []
    for line in open(path):
        if line.startswith(("ATOM", "HETATM")):
            x, y, z = float(line[30:38]), float(line[38:46]), float(line[46:54])
            el = line[76:78].strip() or line[12:16].strip()[0]
            atoms.append((el, x / 10.0, y / 10.0, z / 10.0))  # Angstrom -> nm (PDBReaderLib.cpp:872-874)
    return atoms


def read_standard(path):
    q, I, s, meta = [], [], [], {}
    for line in open(path):
        if line.startswith("#"):
            if ":" in line:
                k, v = line[1:].split(":", 1)
                meta[k.strip()] = v.strip()
            continue
        parts = line.split()
        try:
            vals = [float(p) for p in parts]
        except ValueError:
            continue
        if len(vals) >= 3:
            q.append(vals[0]); I.append(vals[1]); s.append(vals[2])
    return q, I, s, meta


def main():
    for name in CASES:
        d = os.path.join(REF, name)
        state = json.load(open(os.path.join(d, name + ".state")))
        model = state["Domain"]["Populations"][0]["Models"][0]
        pdb = [f for f in os.listdir(d) if f.endswith(".pdb")][0]
        q, I, s, meta = read_standard(os.path.join(d, name + "TestStandard.out"))
        rec = {
            "name": name,
            "source": f"reviewer_tests/files_for_tests/generate/cpu/short/{name}",
            "domain": {k: state["DomainPreferences"][k] for k in
                       ("GridSize", "qMax", "OrientationMethod", "OrientationIterations", "Convergence")},
            "extra_parameters": model["ExtraParameters"],
            "centered": model["Centered"],
            "atoms": read_pdb(os.path.join(d, pdb)),
            "q": q, "intensity": I, "sigma": s,
            "meta": meta,
            # the whole state and PDB text, so the backend can be driven as the reference's tests do
            "state": state,
            "pdb_name": pdb,
            "pdb_text": open(os.path.join(d, pdb)).read(),
        }
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump(rec, f)
        print(name, len(q), rec["atoms"][:2])


GPU_REF = "/root/reference/PythonInterface/tests/reviewer_tests/files_for_tests/generate/gpu/short"
SYM_CASES = [f"Man_Symm_Impl_No_Charges_{h}_Hydr_{m}" for h in ("wo", "With")
             for m in ("MC", "VEGAS", "Hybrid_MC", "Hybrid_VEGAS", "Ga_Kr")] + \
            [f"Space_fill_Symm_Impl_No_Charges_{h}_Hydr_{m}" for h in ("wo", "With")
             for m in ("MC", "Hybrid_MC", "Hybrid_VEGAS", "Ga_Kr", "Hybrid_Ga_Kr")]


def symmetry_cases():
    """Tubulin-dimer symmetry fixtures (manual and space-filling, grid and
    hybrid, MC/VEGAS/GK): the state, the mean/sigma curve, and the one PDB they
    share, gzip-compressed, under tests/golden/symmetry/."""
    import gzip
    import hashlib
    out = os.path.join(OUT, "symmetry")
    os.makedirs(out, exist_ok=True)
    pdb_digest = None
    for name in SYM_CASES:
        d = os.path.join(GPU_REF, name)
        state = json.load(open(os.path.join(d, name + ".state")))
        pdb = [f for f in os.listdir(d) if f.endswith(".pdb")][0]
        raw = open(os.path.join(d, pdb), "rb").read()
        digest = hashlib.sha256(raw).hexdigest()
        if pdb_digest is None:
            with gzip.GzipFile(os.path.join(out, "tubulin_dimer_no_charges.pdb.gz"), "wb", mtime=0) as f:
                f.write(raw)
            pdb_digest = digest
        assert digest == pdb_digest, name
        std = [f for f in os.listdir(d) if f.startswith(name + "TestStandard")][0]
        q, I, s, meta = read_standard(os.path.join(d, std))
        rec = {"name": name, "source": f"reviewer_tests/files_for_tests/generate/gpu/short/{name}",
               "pdb": "tubulin_dimer_no_charges.pdb.gz", "pdb_sha256": digest, "state": state,
               "q": q, "intensity": I, "sigma": s, "meta": meta}
        with open(os.path.join(out, name + ".json"), "w") as f:
            json.dump(rec, f)
        print(name, len(q))


LONG_REF = "/root/reference/PythonInterface/tests/reviewer_tests/files_for_tests/generate/cpu/long"
UNIT_REF = "/root/reference/PythonInterface/tests/unit_tests/files_for_tests"


def read_curve(path):
    """(q, I, header lines) of a two-column .out signal file."""
    q, I, head = [], [], []
    for line in open(path):
        if line.startswith("#"):
            head.append(line.rstrip("\n"))
            continue
        parts = line.split()
        if len(parts) >= 2:
            try:
                a, b = float(parts[0]), float(parts[1])
            except ValueError:
                continue
            q.append(a)
            I.append(b)
    return q, I, head


def voxel_cases():
    """Voxelised-solvent fixtures under tests/golden/voxel/: the hollow-sphere
    hydration layer with and without filled holes (radius type 5, solvent only),
    and 1jff with its excluded volume subtracted at 334 e/nm^3 (radius type 5,
    probe 0.14 nm, voxel 0.2 nm), each with its state, curve and gzip'ed PDB."""
    import gzip
    import hashlib
    out = os.path.join(OUT, "voxel")
    os.makedirs(out, exist_ok=True)

    def store_pdb(src, name):
        raw = open(src, "rb").read()
        with gzip.GzipFile(os.path.join(out, name), "wb", mtime=0) as f:
            f.write(raw)
        return hashlib.sha256(raw).hexdigest()

    for name in ("Hollow Sphere PDB Fill Holes", "Hollow Sphere PDB No Fill Holes"):
        d = os.path.join(LONG_REF, name)
        state = json.load(open(os.path.join(d, name + ".state")))
        digest = store_pdb(os.path.join(d, "Sphere.pdb"), "hollow_sphere.pdb.gz")
        q, I, s, meta = read_standard(os.path.join(d, name + "TestStandard.out"))
        rec = {"name": name, "source": f"reviewer_tests/files_for_tests/generate/cpu/long/{name}",
               "pdb": "hollow_sphere.pdb.gz", "pdb_sha256": digest, "state": state, "q": q, "intensity": I,
               "sigma": s, "meta": meta}
        with open(os.path.join(out, name.replace(" ", "_") + ".json"), "w") as f:
            json.dump(rec, f)
        print(name, len(q))
    # 1jff: the fit test's state (test_fit.py:98-115) at the parameters its target
    # curve was generated with (the curve's header: solvent 334, probe 0.14)
    state = json.load(open(os.path.join(UNIT_REF, "1jff_ED_350_probe_0.125_voxel.state")))
    model = state["Domain"]["Populations"][0]["Models"][0]
    ep = model["ExtraParameters"]
    ep[1] = 334.0
    ep[4] = 0.14
    digest = store_pdb(os.path.join(UNIT_REF, "1jff.pdb"), "1jff.pdb.gz")
    q, I, head = read_curve(os.path.join(UNIT_REF, "1jff_ED_334_probe_0.14_voxel.out"))
    rec = {"name": "1jff_ED_334_probe_0.14_voxel", "source": "unit_tests/files_for_tests/1jff_ED_334_probe_0.14_voxel.out",
           "pdb": "1jff.pdb.gz", "pdb_sha256": digest, "state": state, "q": q, "intensity": I,
           "header": [h for h in head if h.startswith("#")]}
    with open(os.path.join(out, "1jff_ED_334_probe_0.14_voxel.json"), "w") as f:
        json.dump(rec, f)
    print("1jff", len(q))


SCRIPTED_CASES = [f"Scripted_Symm_Impl_No_Charges_{h}_Hydr_{m}" for h in ("wo", "With")
                  for m in ("MC", "VEGAS", "Hybrid_MC", "Hybrid_VEGAS", "Ga_Kr")]
RING_CASES = [f"{e}_Ring{k}_{h}Solution_{s}_Solvation" for e in ("Impl", "Expl") for k in ("", "_DOL")
              for h in ("", "Hybrid_") for s in ("with", "without")]
MT_EXAMPLE = "/root/reference/Example Files/Examples/GPU/Slow_GPU/4_Microtubule"
# the reference's symmetry scripts and the scripts under tests/golden/scripted/
# that take the same layer parameters to the same rows
SCRIPT_FOR = {"Space filling symmetry - Cartesian.lua": "cartesian_lattice.lua",
              "Super Helical Left Hand MT Helix with in-ring twist.lua": "helical_ring.lua",
              "Left_Hand_Helix.lua": "left_hand_helix.lua"}


def scripted_cases():
    """Scripted-symmetry fixtures under tests/golden/scripted/: the tubulin dimer
    in the Cartesian-lattice script (MC, VEGAS, hybrid, GK, with and without the
    hydration layer), the 13-unit tubulin ring as a helix script and as the
    equivalent manual symmetry ("DOL") with implicit and explicit hydrogens, and
    the microtubule example (672 copies from the left-handed helix script, VEGAS,
    its model curve).  Each record holds the state, the curve, the PDB it names
    (gzip'ed, shared when identical) and the script that replaces the
    reference's (the scripts themselves are not copied)."""
    import gzip
    import hashlib
    out = os.path.join(OUT, "scripted")
    os.makedirs(out, exist_ok=True)
    pdbs = {}

    def store_pdb(path):
        raw = open(path, "rb").read()
        digest = hashlib.sha256(raw).hexdigest()
        if digest not in pdbs:
            name = {"453ecc065e91": "../symmetry/tubulin_dimer_no_charges.pdb.gz"}.get(digest[:12])
            if name is None:
                name = os.path.basename(path)[:-4].lower() + ".pdb.gz"
                with gzip.GzipFile(os.path.join(out, name), "wb", mtime=0) as f:
                    f.write(raw)
            pdbs[digest] = name
        return pdbs[digest], digest

    def tree_files(m, acc):
        if "Filename" in m:
            acc.append((m["Type"], os.path.basename(m["Filename"].replace("\\", "/"))))
        for c in m.get("Children", []):
            tree_files(c, acc)
        return acc

    for name in SCRIPTED_CASES + RING_CASES:
        d = os.path.join(GPU_REF, name)
        state = json.load(open(os.path.join(d, name + ".state")))
        files = tree_files(state["Domain"]["Populations"][0]["Models"][0], [])
        pdb = [f for t, f in files if t == "PDB"][0]
        script = [f for t, f in files if t == "Scripted Symmetry"]
        pdb_name, digest = store_pdb(os.path.join(d, pdb))
        std = [f for f in os.listdir(d) if f.startswith(name + "TestStandard")][0]
        q, I, s, meta = read_standard(os.path.join(d, std))
        rec = {"name": name, "source": f"reviewer_tests/files_for_tests/generate/gpu/short/{name}",
               "pdb": pdb_name, "pdb_sha256": digest, "state": state, "q": q, "intensity": I, "sigma": s,
               "meta": meta}
        if script:
            rec["reference_script"] = script[0]
            rec["script"] = SCRIPT_FOR[script[0]]
        with open(os.path.join(out, name + ".json"), "w") as f:
            json.dump(rec, f)
        print(name, len(q), pdb_name, rec.get("script"))
    # the microtubule example: model curve of MT_State.state (program revision 4.0.5)
    state = json.load(open(os.path.join(MT_EXAMPLE, "MT_State.state")))
    pdb_name, digest = store_pdb(os.path.join(MT_EXAMPLE, "3j6f_Dimer_2_2_Added_H_GCentered.pdb"))
    q, I, head = read_curve(os.path.join(MT_EXAMPLE, "MT_Model.out"))
    rec = {"name": "Microtubule_Left_Hand_Helix", "source": "Example Files/Examples/GPU/Slow_GPU/4_Microtubule",
           "pdb": pdb_name, "pdb_sha256": digest, "state": state, "q": q, "intensity": I,
           "header": [h for h in head if h.startswith("#") and not h.startswith("###")],
           "reference_script": "Left_Hand_Helix.lua", "script": SCRIPT_FOR["Left_Hand_Helix.lua"]}
    with open(os.path.join(out, "Microtubule_Left_Hand_Helix.json"), "w") as f:
        json.dump(rec, f)
    print("microtubule", len(q), pdb_name)


if __name__ == "__main__":
    import sys
    if "--voxel" in sys.argv:
        voxel_cases()
    elif "--scripted" in sys.argv:
        scripted_cases()
    else:
        main()
        symmetry_cases()
        voxel_cases()
        scripted_cases()
