"""Element data the amplitude builder needs, in the reference's units.

* X-ray form factors: Cromer-Mann 4 Gaussians + constant, International
  Tables Vol IV Table 2.2B, as tabulated in PDBReaderLib/PDBReaderLib.cpp:2270-2489
  (row order a1, b1, a2, b2, a3, b3, a4, b4, c; q in nm^-1 with
  s^2 = q^2 / (100 * 16 pi^2), Backend/GPU/Atomic Form Factor.cpp:151).
* Dummy-atom ("svg") radii in nm used with radius types 4 and 5
  (PDBReaderLib.cpp:204-264, selected by SetRadiusType :1216-1219), and the van der
  Waals, empirical and calculated radii of types 1, 2 and 3.
* Atomic weights for centring on the centre of mass (PDBReaderLib.cpp:77-264, 271-292).
"""
import numpy as np

XRAY = {
    "H": (0.49300, 10.51090, 0.32290, 26.1257, 0.14020, 3.14240, 0.04080, 57.79980, 0.0030),
    "C": (2.31000, 20.8439, 1.02000, 10.2075, 1.58860, 0.56870, 0.86500, 51.65120, 0.2156),
    "N": (12.2126, 0.00570, 3.13220, 9.89330, 2.01250, 28.9975, 1.16630, 0.58260, -11.5290),
    "O": (3.04850, 13.2771, 2.28680, 5.70110, 1.54630, 0.32390, 0.86700, 32.90890, 0.2508),
    "P": (6.43450, 1.90670, 4.17910, 27.1570, 1.78000, 0.52600, 1.49080, 68.16450, 1.1149),
    "S": (6.90530, 1.46790, 5.20340, 22.2151, 1.43790, 0.25360, 1.58630, 56.1720, 0.86690),
    "Na": (4.76260, 3.28500, 3.17360, 8.84220, 1.26740, 0.31360, 1.11280, 129.42400, 0.676),
    "Mg": (5.42040, 2.82750, 2.17350, 79.2611, 1.22690, 0.38080, 2.30730, 7.19370, 0.8584),
    "Cl": (11.46040, 0.01040, 7.19640, 1.16620, 6.25560, 18.5194, 1.64550, 47.77840, -9.5574),
    "K": (8.21860, 12.79490, 7.43980, 0.77480, 1.05190, 213.187, 0.86590, 41.68410, 1.4228),
    "Ca": (8.62660, 10.44210, 7.38730, 0.65990, 1.58990, 85.7484, 1.02110, 178.43700, 1.3751),
    "Mn": (11.28190, 5.34090, 7.35730, 0.34320, 3.01930, 17.86740, 2.24410, 83.75430, 1.0896),
    "Fe": (11.76950, 4.76110, 7.35730, 0.30720, 3.52220, 15.35350, 2.30450, 76.88050, 1.0369),
    "Cu": (13.33800, 3.58280, 7.16760, 0.24700, 5.61580, 11.39660, 1.67350, 64.81260, 1.191),
    "Zn": (14.07430, 3.26550, 7.03180, 0.23330, 5.16520, 10.31630, 2.41000, 58.70970, 1.3041),
    # ions (PDBReaderLib.cpp:2270-2489, the rows of ion indices 11, 16, 18, 26, 29, 31,
    # 46, 50, 51, 59, 60, 62; Cu2+'s b2 as tabulated there, International Tables 0.244078)
    "O1-": (4.19160, 12.8573, 1.63969, 4.17236, 1.52673, 47.0179, -20.307, -0.01404, 21.9412),
    "Na1+": (3.25650, 2.66710, 3.93620, 6.11530, 1.39980, 0.20010, 1.00320, 14.03900, 0.404),
    "Mg2+": (3.49880, 2.16760, 3.83780, 4.75420, 1.32840, 0.18500, 0.84970, 10.14110, 0.4853),
    "Cl1-": (18.29150, 0.00660, 7.40840, 1.17170, 6.53370, 19.5424, 2.33860, 60.44860, -16.378),
    "K1+": (7.95780, 12.63310, 7.49170, 0.76740, 6.35900, -0.0020, 1.19150, 31.91280, -4.9978),
    "Ca2+": (15.63480, -0.00740, 7.95180, 0.60890, 8.43720, 10.3116, 0.85370, 25.99050, -14.875),
    "Mn2+": (10.80610, 5.27960, 7.36200, 0.34350, 3.52680, 14.34300, 0.21840, 41.32350, 1.0874),
    "Fe2+": (11.04240, 4.65380, 7.37400, 0.30530, 4.13460, 12.05460, 0.43990, 31.28090, 1.0097),
    "Fe3+": (11.17640, 4.61470, 7.38630, 0.30050, 3.39480, 11.67290, 0.07240, 38.55660, 0.9707),
    "Cu1+": (11.94750, 3.36690, 7.35730, 0.22740, 6.24550, 8.66250, 1.55780, 25.84870, 0.89),
    "Cu2+": (11.81680, 3.37484, 7.11181, 244078, 5.78135, 7.98760, 1.14523, 19.89700, 1.14431),
    "Zn2+": (11.97190, 2.99460, 7.38620, 0.20310, 6.46680, 7.08260, 1.39400, 18.09950, 0.7807),
    # implicit-hydrogen groups (fitted coefficients, PDBReaderLib.cpp:2479-2486)
    "CH": (0.894937, 55.7145, 0.894429, 4.03158, 3.78824, 24.8323, 3.14683e-6, 956.628, 1.42149),
    "CH2": (1.61908, 52.1451, 2.27205, 24.6589, 2.1815, 24.6587, 0.0019254, 152.165, 1.92445),
    "CH3": (12.5735, 38.7341, -0.456658, -6.28167, 5.71547, 54.955, -11.711, 47.898, 2.87762),
    "NH": (0.00506991, 108.256, 2.03147, 14.6199, 1.82122, 14.628, 2.06506, 35.4102, 2.07168),
    "NH2": (3.00872, 28.3717, 0.288137, 63.9637, 3.39248, 3.51866, 2.03511, 28.3675, 0.269952),
    "NH3": (0.294613, 67.4408, 6.48379, 29.1576, 5.67182, 0.54735, 6.57164, 0.547493, -9.02757),
    "OH": (-2.73406, 22.1288, 0.00966263, 94.3428, 6.64439, 13.9044, 2.67949, 32.7607, 2.39981),
    "SH": (-127.811, 7.19935, 62.5514, 12.1591, 160.747, 1.88979, 2.34822, 55.952, -80.836),
}

SVG_RADIUS = {"H": 0.107, "C": 0.1577, "N": 0.08414, "O": 0.130, "P": 0.111, "S": 0.168, "Na": 0.180,
              "Mg": 0.160, "Cl": 0.100, "K": 0.220, "Ca": 0.197, "Mn": 0.130, "Fe": 0.124, "Cu": 0.128,
              "Zn": 0.133, "CH": 0.173, "CH2": 0.185, "CH3": 0.197, "NH": 0.122, "NH2": 0.145, "NH3": 0.162,
              "OH": 0.150, "SH": 0.181}

# van der Waals, empirical and calculated radii (nm) of radius types 1, 2 and 3
# (PDBReaderLib.cpp:77-111, pm / 1000); a group takes its heavy atom's radius (the
# reference adds entry 0's radius, which is zero, per hydrogen, :213-267)
_ATOM_RADII = {"H": (0.120, 0.025, 0.053), "C": (0.170, 0.070, 0.067), "N": (0.155, 0.065, 0.056),
               "O": (0.152, 0.060, 0.048), "P": (0.180, 0.100, 0.098), "S": (0.180, 0.100, 0.088),
               "Na": (0.227, 0.180, 0.190), "Mg": (0.173, 0.150, 0.145), "Cl": (0.175, 0.100, 0.079),
               "K": (0.275, 0.220, 0.243), "Ca": (0.231, 0.180, 0.194), "Mn": (0.186, 0.140, 0.161),
               "Fe": (0.181, 0.140, 0.156), "Cu": (0.140, 0.135, 0.145), "Zn": (0.139, 0.135, 0.142)}
_HEAVY = {"CH": "C", "CH2": "C", "CH3": "C", "NH": "N", "NH2": "N", "NH3": "N", "OH": "O", "SH": "S"}
# an ion takes its element's radii and weight
_ION_OF = {k: k.rstrip("+-0123456789") for k in XRAY if k[-1] in "+-"}
_HEAVY.update(_ION_OF)
for _ion, _el in _ION_OF.items():
    SVG_RADIUS[_ion] = SVG_RADIUS[_el]
VDW_RADIUS = {e: _ATOM_RADII[_HEAVY.get(e, e)][0] for e in SVG_RADIUS}
EMP_RADIUS = {e: _ATOM_RADII[_HEAVY.get(e, e)][1] for e in SVG_RADIUS}
CALC_RADIUS = {e: _ATOM_RADII[_HEAVY.get(e, e)][2] for e in SVG_RADIUS}

# a group weighs what its heavy atom weighs (the reference adds atmWt[0] = 0 per hydrogen)
WEIGHT = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "P": 30.97376, "S": 32.06, "Na": 22.98976,
          "Mg": 24.3050, "Cl": 35.45, "K": 39.0983, "Ca": 40.078, "Mn": 54.93804, "Fe": 55.845, "Cu": 63.546,
          "Zn": 65.38, "CH": 12.011, "CH2": 12.011, "CH3": 12.011, "NH": 14.007, "NH2": 14.007, "NH3": 14.007,
          "OH": 15.999, "SH": 32.06}
WEIGHT.update({_ion: WEIGHT[_el] for _ion, _el in _ION_OF.items()})

# table order, the order of ion rows: elements by atomic number, each followed by its ions
SPECIES = tuple(sorted(XRAY, key=lambda e: (list(XRAY).index(_HEAVY.get(e, e)) if e in _ION_OF else list(XRAY).index(e),
                                            e in _ION_OF, e)))


def ion_tables(elements):
    """coef9 [nion, 9], radius [nion], index of every atom into those rows.

    Rows follow the species present, in table order (atoms by atomic number,
    then the implicit-hydrogen groups), matching the reference's ion-sorted atoms."""
    names, inverse = np.unique(np.asarray(elements, dtype=str), return_inverse=True)
    present = set(names.tolist())
    order = [e for e in SPECIES if e in present]
    unknown = present - set(order)
    if unknown:
        raise ValueError(f"no X-ray form factor for {sorted(unknown)}")
    row = {e: i for i, e in enumerate(order)}
    coef9 = np.array([XRAY[e] for e in order], dtype=np.float64)
    rad = np.array([SVG_RADIUS[e] for e in order], dtype=np.float64)
    ion = np.array([row[e] for e in names.tolist()], dtype=np.int32)[inverse.ravel()]
    return coef9, rad, ion, order
