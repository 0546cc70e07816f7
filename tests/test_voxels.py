"""The voxelised solvent on the CPU (host code of the library, no kernel): the
excluded-volume and hydration-layer voxel sets of PDBAmplitude::CalculateSolventSpace
(Backend/Backend/VoxelizedSolvent.cpp:4-194) against brute force.

Boxes are decoded back into voxel centres on the voxel lattice (a box of n_a
voxels centred at c covers c + (k - (n_a - 1)/2) vox, k < n_a), then compared
with sets built independently here: the voxels whose corner point lies within an
atom radius (MarkVoxelsNeighboringAtoms, :365-396), the cavities that are not
26-connected to the outside (Floodfill3D, :398-441), and the layer/excluded
partition.
"""
import numpy as np
import pytest
from scipy import ndimage

VOX = 0.05


def structure(seed, n, radius):
    import dplus_mi355x.synth as S
    xyz, el = S.walk_structure(seed, n, (radius, radius, radius))
    coef9, rad, ion, _ = S.structure_tables(el)
    return xyz, rad[ion]


def hollow_shell(n=2500, R=1.2):
    """Atoms on a Fibonacci lattice of a sphere, 0.085 nm apart: a closed shell."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    rho = np.sqrt(1.0 - z * z)
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1) * R, np.full(n, 0.1577)


def decode(dims, com, vox=VOX):
    """Voxel centres (as integer lattice keys relative to the first box) of a box list."""
    dims, com = np.asarray(dims), np.asarray(com)
    if len(dims) == 0:
        return set(), np.zeros(3)
    lo = com - 0.5 * (dims - 1) * vox  # centre of each box's first voxel
    origin = lo[0]
    keys = set()
    for d, l in zip(dims, lo):
        base = np.rint((l - origin) / vox).astype(int)
        for i in range(d[0]):
            for j in range(d[1]):
                for k in range(d[2]):
                    keys.add((base[0] + i, base[1] + j, base[2] + k))
    return keys, origin


def brute_atoms(xyz, rad, origin, lo_key, hi_key, vox=VOX, extra=0.0):
    """Lattice keys in [lo_key, hi_key] whose point lies within rad + extra of an atom
    (coordinates and radii in float, as the reference keeps them)."""
    x = xyz.astype(np.float32).astype(np.float64)
    r = rad.astype(np.float32).astype(np.float64) + extra
    ii = np.arange(lo_key[0], hi_key[0] + 1)
    jj = np.arange(lo_key[1], hi_key[1] + 1)
    kk = np.arange(lo_key[2], hi_key[2] + 1)
    P = np.stack(np.meshgrid(ii, jj, kk, indexing="ij"), axis=-1).reshape(-1, 3)
    pts = origin + P * vox
    inside = np.zeros(len(P), bool)
    for a in range(len(x)):
        d2 = np.sum((pts - x[a]) ** 2, axis=1)
        inside |= d2 <= r[a] ** 2
    return {tuple(p) for p in P[inside]}, P, inside


def test_excluded_volume_without_probe_is_the_atom_voxels():
    import dplus_mi355x as D
    xyz, rad = structure(5, 120, 1.0)
    dims, com = D.excluded_volume_boxes(xyz, rad, probe=0.0, thickness=0.0, vox=VOX, fill_holes=False)
    keys, origin = decode(dims, com)
    assert len(keys) == int(np.prod(dims, axis=1).sum())  # the boxes are disjoint
    K = np.array(sorted(keys))
    lo, hi = K.min(axis=0) - 3, K.max(axis=0) + 3
    ref, _, _ = brute_atoms(xyz, rad, origin, lo, hi)
    assert keys == ref
    # the hydration layer needs a probe radius
    ld, _ = D.hydration_boxes(xyz, rad, probe=0.0, thickness=0.3, vox=VOX)
    assert len(ld) == 0


def test_excluded_volume_fills_cavities():
    """A hollow shell of atoms: with fill_holes the enclosed voxels join the
    excluded volume, exactly the empty voxels not 26-connected to the outside."""
    import dplus_mi355x as D
    xyz, rad = hollow_shell()
    d0, c0 = D.excluded_volume_boxes(xyz, rad, probe=0.0, thickness=0.0, vox=VOX, fill_holes=False)
    d1, c1 = D.excluded_volume_boxes(xyz, rad, probe=0.0, thickness=0.0, vox=VOX, fill_holes=True)
    k0, o0 = decode(d0, c0)
    k1, o1 = decode(d1, c1)
    shift = np.rint((o1 - o0) / VOX).astype(int)
    k1 = {(a + shift[0], b + shift[1], c + shift[2]) for a, b, c in k1}
    assert k0 < k1
    K = np.array(sorted(k0))
    lo, hi = K.min(axis=0) - 2, K.max(axis=0) + 2
    _, P, inside = brute_atoms(xyz, rad, o0, lo, hi)
    shape = tuple(hi - lo + 1)
    occ = inside.reshape(shape)
    lab, _ = ndimage.label(~occ, structure=np.ones((3, 3, 3)))
    outside = lab[0, 0, 0]
    cavity = (~occ) & (lab != outside)
    assert cavity.sum() > 1000
    filled = {tuple(p) for p in P[(occ | cavity).ravel()]}
    assert k1 == filled


def test_excluded_volume_and_layer_partition_the_probe_shell():
    """With a probe radius the excluded volume holds the atom voxels and the part of
    the probe shell the layer does not take; the layer and the excluded volume are
    disjoint, and their union covers every voxel within radius + probe of an atom
    (away from the corner rows the reference blanks)."""
    import dplus_mi355x as D
    xyz, rad = structure(7, 150, 1.1)
    probe, thick = 0.14, 0.3
    de, ce = D.excluded_volume_boxes(xyz, rad, probe=probe, thickness=thick, vox=VOX, fill_holes=True)
    dl, cl = D.hydration_boxes(xyz, rad, probe=probe, thickness=thick, vox=VOX, fill_holes=True)
    ke, oe = decode(de, ce)
    kl, ol = decode(dl, cl)
    shift = np.rint((ol - oe) / VOX).astype(int)
    kl = {(a + shift[0], b + shift[1], c + shift[2]) for a, b, c in kl}
    assert ke and kl and not (ke & kl)
    K = np.array(sorted(ke | kl))
    lo, hi = K.min(axis=0) - 2, K.max(axis=0) + 2
    atoms, _, _ = brute_atoms(xyz, rad, oe, lo, hi)
    assert atoms <= ke
    shell, _, _ = brute_atoms(xyz, rad, oe, lo, hi, extra=probe)
    assert shell <= (ke | kl)
    # the layer reaches beyond the probe shell when thickness > probe
    assert len(kl - shell) > 0


def test_radius_types_change_the_excluded_volume():
    """Radius types pick the radii of the marking (PDBReaderLib.cpp:1207-1236):
    van der Waals radii are the largest, so their excluded volume is the largest."""
    import dplus_mi355x.elements as E
    import dplus_mi355x as D
    xyz, _ = structure(9, 80, 0.9)
    vols = {}
    for name, table in (("dummy", E.SVG_RADIUS), ("vdw", E.VDW_RADIUS), ("emp", E.EMP_RADIUS),
                        ("calc", E.CALC_RADIUS)):
        r = np.full(len(xyz), table["C"])
        d, _ = D.excluded_volume_boxes(xyz, r, probe=0.0, thickness=0.0, vox=VOX, fill_holes=False)
        vols[name] = int(np.prod(d, axis=1).sum())
    assert vols["vdw"] > vols["dummy"] > vols["emp"] > vols["calc"] > 0
