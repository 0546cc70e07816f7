"""Jacobian spherical grids, PDB amplitude leaves, docking and orientation
averages over the C ABI of libdplus_mi355x."""
import ctypes

import numpy as np

from . import _lib as L


class Grid:
    """A Jacobian spherical grid resident on the GPU (JacobianSphereGrid,
    Backend/Backend/Grid.h:231-327).  ``data``/``interp`` are zero-copy numpy
    views of the host mirror in the reference layout (2*P doubles)."""

    def __init__(self, grid_size, q_max, precision=L.PRECISION_MIXED):
        h = ctypes.c_void_p()
        L.check(L.lib().dplus_grid_create(int(grid_size), float(q_max), int(precision), ctypes.byref(h)))
        self._init_from_handle(h)

    def _init_from_handle(self, h):
        self._h = h
        gs, qm, st, n, pts, pr = (ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int(),
                                  ctypes.c_longlong(), ctypes.c_int())
        L.check(L.lib().dplus_grid_info(h, ctypes.byref(gs), ctypes.byref(qm), ctypes.byref(st), ctypes.byref(n),
                                        ctypes.byref(pts), ctypes.byref(pr)))
        self.grid_size, self.q_max, self.step_size = gs.value, qm.value, st.value
        self.actual_size, self.points, self.precision = n.value, pts.value, pr.value

    @property
    def handle(self):
        return self._h

    @property
    def data(self):
        p = L.lib().dplus_grid_data(self._h)
        return np.ctypeslib.as_array(p, shape=(2 * self.points,))

    @property
    def interp(self):
        p = L.lib().dplus_grid_interp(self._h)
        return np.ctypeslib.as_array(p, shape=(2 * self.points,))

    def upload(self):
        L.check(L.lib().dplus_grid_upload(self._h))

    def download(self):
        L.check(L.lib().dplus_grid_download(self._h))

    def calculate_splines(self):
        L.check(L.lib().dplus_grid_calculate_splines(self._h))

    def get_sphr(self, q, theta, phi):
        q, theta, phi = (np.atleast_1d(L.f64(x)) for x in (q, theta, phi))
        n = len(q)
        out = np.zeros(2 * n)
        L.check(L.lib().dplus_grid_get_sphr(self._h, n, L.dp(q), L.dp(theta), L.dp(phi), L.dp(out)))
        return out[0::2] + 1j * out[1::2]

    def interpolate_theta_phi_plane(self, ri, theta, phi):
        """InterpolateThetaPhiPlane (Grid.cpp:61-134) inside shell ri at (theta, phi), many points."""
        ri = np.atleast_1d(L.i32(ri))
        theta, phi = (np.atleast_1d(L.f64(x)) for x in (theta, phi))
        ri, theta, phi = (np.ascontiguousarray(x) for x in np.broadcast_arrays(ri, theta, phi))
        n = len(ri)
        out = np.zeros(2 * n)
        L.check(L.lib().dplus_grid_interpolate_theta_phi_plane(self._h, n, L.ip(ri), L.dp(theta), L.dp(phi),
                                                                 L.dp(out)))
        return out[0::2] + 1j * out[1::2]

    def calculate_intensity(self, q, theta=None, epsilon=1e-3, seed=0, iterations=1000000, phi_min=0.0,
                            phi_max=2 * np.pi):
        r = ctypes.c_double()
        th = -1.0 if theta is None else float(theta)
        L.check(L.lib().dplus_grid_calculate_intensity(self._h, float(q), th, float(epsilon), int(seed),
                                                       int(iterations), float(phi_min), float(phi_max),
                                                       ctypes.byref(r)))
        return r.value

    def build_pdb(self, xyz, ion, coef9, rad, sol_ed=0.0, c1=1.0, flags=L.CALC_ATOMIC_FORMFACTORS, scale=1.0,
                  box_dims=None, box_com=None, vox_step=0.0, box_contrast=0.0, host_copy=True, excl_dims=None,
                  excl_com=None, excl_contrast=0.0):
        """PDBAmplitude::calculateGrid for atoms (nm) of ion types ion[i]."""
        keep = []
        d = _pdb_desc(xyz, ion, coef9, rad, sol_ed, c1, flags, scale, box_dims, box_com, vox_step, box_contrast, keep,
                      excl_dims, excl_com, excl_contrast)
        L.check(L.lib().dplus_pdb_build_grid(self._h, ctypes.byref(d), 1 if host_copy else 0))

    def __del__(self):
        try:
            if self._h:
                L.lib().dplus_grid_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _boxes(dims, com):
    if dims is None or len(dims) == 0:
        return np.zeros((1, 3), np.int32), np.zeros((1, 3)), 0
    return L.i32(dims).reshape(-1, 3), L.f64(com).reshape(-1, 3), len(dims)


def _pdb_desc(xyz, ion, coef9, rad, sol_ed, c1, flags, scale, box_dims, box_com, vox_step, box_contrast, keep,
              excl_dims=None, excl_com=None, excl_contrast=0.0):
    xyz, ion = L.f64(xyz).reshape(-1, 3), L.i32(ion)
    coef9, rad = L.f64(coef9).reshape(-1, 9), L.f64(rad)
    bd, bc, nb = _boxes(box_dims, box_com)
    ed, ec, ne = _boxes(excl_dims, excl_com)
    keep += [xyz, ion, coef9, rad, bd, bc, ed, ec]
    return L.PdbDesc(len(ion), L.dp(xyz), L.ip(ion), len(coef9), L.dp(coef9), L.dp(rad), sol_ed, c1, flags, scale,
                     nb, L.ip(bd), L.dp(bc), vox_step, box_contrast, ne, L.ip(ed), L.dp(ec), excl_contrast)


class Structure:
    """Atoms and solvent boxes of a PDB leaf resident in HBM (sorted and tiled once)."""

    def __init__(self, xyz, ion, coef9, rad, sol_ed=0.0, c1=1.0, flags=L.CALC_ATOMIC_FORMFACTORS, scale=1.0,
                 box_dims=None, box_com=None, vox_step=0.0, box_contrast=0.0, excl_dims=None, excl_com=None,
                 excl_contrast=0.0):
        keep = []
        d = _pdb_desc(xyz, ion, coef9, rad, sol_ed, c1, flags, scale, box_dims, box_com, vox_step, box_contrast, keep,
                      excl_dims, excl_com, excl_contrast)
        h = ctypes.c_void_p()
        L.check(L.lib().dplus_structure_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.natoms = len(L.i32(ion))
        self.nboxes = 0 if box_dims is None else len(box_dims)

    def build(self, grid, host_copy=False):
        L.check(L.lib().dplus_structure_build_grid(self._h, grid.handle, 1 if host_copy else 0))

    def build_shape(self, grid):
        """dplus_structure_build_shape: {"wg", "slices", "tuned"} of a mixed build of grid."""
        wg, sl, tu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().dplus_structure_build_shape(self._h, grid.handle, ctypes.byref(wg), ctypes.byref(sl),
                                                    ctypes.byref(tu)))
        return {"wg": wg.value, "slices": sl.value, "tuned": bool(tu.value)}

    def autotune(self, grid):
        """dplus_structure_autotune: times every builder launch shape on this grid,
        records the fastest for the grid's class and builds the grid; returns
        {"wg", "slices", "best_ms", "model_ms"}."""
        wg, sl = ctypes.c_int(), ctypes.c_int()
        tb, tm = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().dplus_structure_autotune(self._h, grid.handle, ctypes.byref(wg), ctypes.byref(sl),
                                                 ctypes.byref(tb), ctypes.byref(tm)))
        return {"wg": wg.value, "slices": sl.value, "best_ms": tb.value, "model_ms": tm.value}

    def __del__(self):
        try:
            if self._h:
                L.lib().dplus_structure_destroy(self._h)
                self._h = None
        except Exception:
            pass


def index_from_indices(qi, ti, pi):
    return int(L.lib().dplus_grid_index_from_indices(qi, ti, pi))


def indices_from_index(idx):
    qi, ti, pi = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
    L.lib().dplus_grid_indices_from_index(idx, ctypes.byref(qi), ctypes.byref(ti), ctypes.byref(pi))
    return qi.value, ti.value, pi.value


class Docking:
    """Leaves of a hybrid orientation average: each leaf is a Grid seen under
    rotations rot[r] = (alpha, beta, gamma [rad], scale) with translations."""

    def __init__(self, leaves):
        self._keep = []
        arr = (L.Leaf * len(leaves))()
        for i, (grid, rot, ntrans, trans) in enumerate(leaves):
            rot, ntrans, trans = L.f64(rot).reshape(-1, 4), L.i32(ntrans), L.f64(trans).reshape(-1, 3)
            self._keep += [grid, rot, ntrans, trans]
            arr[i] = L.Leaf(grid.handle, len(rot), L.dp(rot), L.ip(ntrans), L.dp(trans))
        self.arr, self.n = arr, len(leaves)

    @staticmethod
    def single(grid):
        return Docking([(grid, [[0, 0, 0, 1]], [1], [[0, 0, 0]])])


def orientation_average(dock, q, method=L.OA_MC, seed=1, max_dirs=1 << 20, chunk=8192, eps=0.0,
                        vegas_iterations=10, vegas_points_per_bin=4, gk_max_depth=1 << 40):
    q = L.f64(q)
    out = np.zeros(len(q))
    p = L.OAParams(method, seed, max_dirs, chunk, eps, vegas_iterations, vegas_points_per_bin, gk_max_depth)
    st = L.OAStats()
    L.check(L.lib().dplus_orientation_average(dock.n, dock.arr, len(q), L.dp(q), ctypes.byref(p), L.dp(out),
                                              ctypes.byref(st)))
    return out, st


def azimuthal_average(dock, q, theta, nphi=4096, seed=1):
    """dplus_azimuthal_average: mean over nphi stratified phi of |A(q_i, theta_i, phi)|^2."""
    q, theta = L.f64(q), L.f64(theta)
    out = np.zeros(len(q))
    L.check(L.lib().dplus_azimuthal_average(dock.n, dock.arr, len(q), L.dp(q), L.dp(theta), int(nphi), int(seed),
                                            L.dp(out)))
    return out


def orientation_average_noalloc(dock, q, method=L.OA_MC, seed=1, max_dirs=1 << 20, chunk=8192, eps=0.0,
                                vegas_iterations=10, vegas_points_per_bin=4):
    """orientation_average (Monte Carlo or VEGAS) with every device buffer taken
    from the library's fixed device scratch array: no device allocation."""
    q = L.f64(q)
    out = np.zeros(len(q))
    p = L.OAParams(method, seed, max_dirs, chunk, eps, vegas_iterations, vegas_points_per_bin, 0)
    st = L.OAStats()
    L.check(L.lib().dplus_orientation_average_noalloc(dock.n, dock.arr, len(q), L.dp(q), ctypes.byref(p),
                                                      L.dp(out), ctypes.byref(st)))
    return out, st


def dock_grid(out, dock, scale=1.0):
    L.check(L.lib().dplus_dock_grid(out.handle, dock.n, dock.arr, float(scale)))


def dock_grid_space_filling(out, dock, distances, angles_deg, reps, scale=1.0):
    """GridSymmetry::calculateGrid: dock the leaves, times the lattice sum of the
    basis built from distances (nm), angles (degrees) and repetitions."""
    d, a, r = L.f64(distances), L.f64(angles_deg), L.f64(reps)
    L.check(L.lib().dplus_dock_grid_space_filling(out.handle, dock.n, dock.arr, float(scale), L.dp(d), L.dp(a),
                                                  L.dp(r)))


def _voxel_boxes(fn, xyz, rad, probe, thickness, vox, fill_holes):
    xyz, rad = L.f64(xyz).reshape(-1, 3), L.f64(rad)
    n = ctypes.c_int()
    fh = 1 if fill_holes else 0
    L.check(fn(len(rad), L.dp(xyz), L.dp(rad), probe, thickness, vox, fh, 0, ctypes.byref(n), None, None))
    dims, com = np.zeros((max(1, n.value), 3), np.int32), np.zeros((max(1, n.value), 3))
    L.check(fn(len(rad), L.dp(xyz), L.dp(rad), probe, thickness, vox, fh, n.value, ctypes.byref(n), L.ip(dims),
               L.dp(com)))
    return dims[:n.value], com[:n.value]


def hydration_boxes(xyz, rad, probe=0.14, thickness=0.3, vox=0.05, fill_holes=True):
    """Hydration-layer boxes (dims [n,3] voxels, centres [n,3] nm) around atoms of radius rad,
    with the reference's voxel marking (cavities count as atoms when fill_holes)."""
    return _voxel_boxes(L.lib().dplus_hydration_boxes, xyz, rad, probe, thickness, vox, fill_holes)


def excluded_volume_boxes(xyz, rad, probe=0.14, thickness=0.3, vox=0.05, fill_holes=True):
    """Excluded-volume boxes of the same voxel marking (atoms, filled cavities and the
    probe shell the hydration layer leaves), subtracted with rho_0 by radius types 1, 2, 3, 5."""
    return _voxel_boxes(L.lib().dplus_excluded_volume_boxes, xyz, rad, probe, thickness, vox, fill_holes)


def set_fp64_builder(plain):
    """Form of later FP64 grid builds: False the paired builder (default), True the
    plain form, one node at a time (dplus_set_fp64_builder)."""
    L.check(L.lib().dplus_set_fp64_builder(1 if plain else 0))


def lua_symmetry_rows(script, params):
    """Rows [k, 6] = (x, y, z, alpha, beta, gamma) (nm, degrees) that a symmetry
    script's Populate(p, nlayers) returns for the layer parameters params
    [nlayers][nlp] (LuaSymmetry::PreCalculate, Backend/Backend/Symmetry.cpp:463-521).
    Runs on the host in the library's own interpreter; script errors raise DplusError.
    Ragged layers are padded with zeros."""
    layers = [np.atleast_1d(np.asarray(r, dtype=np.float64)) for r in params]
    nl, nlp = len(layers), max((len(r) for r in layers), default=0)
    P = np.zeros((nl, nlp))
    for i, r in enumerate(layers):
        P[i, :len(r)] = r
    text = script.encode() if isinstance(script, str) else bytes(script)
    n = ctypes.c_int()
    rows = np.zeros(6 * 64)
    for _ in range(2):
        cap = len(rows) // 6
        L.check(L.lib().dplus_lua_symmetry_rows(text, nl, nlp, L.dp(P) if P.size else None, cap, ctypes.byref(n),
                                                L.dp(rows)))
        if n.value <= cap:
            return rows[:6 * n.value].reshape(-1, 6).copy()
        rows = np.zeros(6 * n.value)
    raise RuntimeError("the script's row count changed between calls")


def synth_walk(seed, n, step=0.38, min_sep=0.12, axes=(3.0, 3.0, 3.0)):
    out = np.zeros(3 * n)
    L.check(L.lib().dplus_synth_walk(int(seed), int(n), float(step), float(min_sep), float(axes[0]),
                                     float(axes[1]), float(axes[2]), L.dp(out)))
    return out.reshape(-1, 3)


def build_shapes_save():
    """The autotuner's shape table as text ("grid_shells tile_class wg slices" lines)."""
    p, n = ctypes.c_void_p(), ctypes.c_longlong()
    L.check(L.lib().dplus_build_shapes_save(ctypes.byref(p), ctypes.byref(n)))
    try:
        return ctypes.string_at(p, n.value).decode()
    finally:
        L.lib().dplus_free(p)


def build_grids(structures, grids, host_copy=False):
    """dplus_structure_build_grids: grid i from structure i, mixed grids in grouped launches."""
    n = len(structures)
    if len(grids) != n:
        raise ValueError("one grid per structure")
    S = (ctypes.c_void_p * max(n, 1))(*[s._h.value for s in structures])
    G = (ctypes.c_void_p * max(n, 1))(*[g.handle.value for g in grids])
    L.check(L.lib().dplus_structure_build_grids(n, S, G, 1 if host_copy else 0))


def build_shapes_clear():
    """Empties the autotuner's shape table (builds use the occupancy model again)."""
    L.check(L.lib().dplus_build_shapes_clear())


def build_shapes_load(text):
    """Adds the lines of a saved shape table to this process's table."""
    L.check(L.lib().dplus_build_shapes_load(text.encode()))
