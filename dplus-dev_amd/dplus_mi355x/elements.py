"""Element data the amplitude builder needs, in the reference's units.

* X-ray form factors: Cromer-Mann 4 Gaussians + constant, International
  Tables Vol IV Table 2.2B, as tabulated in PDBReaderLib/PDBReaderLib.cpp:2270-2293
  (row order a1, b1, a2, b2, a3, b3, a4, b4, c; q in nm^-1 with
  s^2 = q^2 / (100 * 16 pi^2), Backend/GPU/Atomic Form Factor.cpp:151).
* Dummy-atom ("svg") radii in nm used with radius type 4
  (PDBReaderLib.cpp:204-216, selected by SetRadiusType :1216-1219).
* Atomic weights for centring on the centre of mass (PDBReaderLib.cpp:77-93, 271-292).
"""
import numpy as np

XRAY = {
    "H": (0.49300, 10.51090, 0.32290, 26.1257, 0.14020, 3.14240, 0.04080, 57.79980, 0.0030),
    "C": (2.31000, 20.8439, 1.02000, 10.2075, 1.58860, 0.56870, 0.86500, 51.65120, 0.2156),
    "N": (12.2126, 0.00570, 3.13220, 9.89330, 2.01250, 28.9975, 1.16630, 0.58260, -11.5290),
    "O": (3.04850, 13.2771, 2.28680, 5.70110, 1.54630, 0.32390, 0.86700, 32.90890, 0.2508),
    "P": (6.43450, 1.90670, 4.17910, 27.1570, 1.78000, 0.52600, 1.49080, 68.16450, 1.1149),
    "S": (6.90530, 1.46790, 5.20340, 22.2151, 1.43790, 0.25360, 1.58630, 56.1720, 0.86690),
}

SVG_RADIUS = {"H": 0.107, "C": 0.1577, "N": 0.08414, "O": 0.130, "P": 0.111, "S": 0.168}

WEIGHT = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "P": 30.97376, "S": 32.06}


def ion_tables(elements):
    """coef9 [nion, 9], radius [nion], index of every atom into those rows.

    Rows follow the element order of first appearance sorted by the reference's
    ion index order (H < C < N < O < P < S), matching its ion-sorted atoms."""
    order = [e for e in ("H", "C", "N", "O", "P", "S") if e in set(elements)]
    unknown = set(elements) - set(order)
    if unknown:
        raise ValueError(f"no X-ray form factor for {sorted(unknown)}")
    row = {e: i for i, e in enumerate(order)}
    coef9 = np.array([XRAY[e] for e in order], dtype=np.float64)
    rad = np.array([SVG_RADIUS[e] for e in order], dtype=np.float64)
    ion = np.array([row[e] for e in elements], dtype=np.int32)
    return coef9, rad, ion, order
