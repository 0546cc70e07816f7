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
CASES = [f"{e}_explicit_{s}" for e in ("h", "c", "n", "o", "points") for s in ("Vacuum", "SS334")]


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


if __name__ == "__main__":
    main()
