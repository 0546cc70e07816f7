"""ctypes view of the FP64 oracle (oracle/oracle.c) and of the reference grid
build (oracle/_ref/libgridref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py to compare results with; the product package never
imports it.  Each wrapper names the oracle function it calls; the reference citations
live beside those functions in oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgridref.so")

CALC_ATOMIC_FORMFACTORS = 1
CALC_DUMMY_SOLVENT = 2

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_LL = ctypes.POINTER(ctypes.c_longlong)


def build(with_ref=True):
    """Compile the oracle (and the reference grid when the reference tree exists)."""
    targets = ["all"]
    if with_ref and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class Leaf(ctypes.Structure):
    _fields_ = [
        ("grid_size", ctypes.c_int),
        ("qmax", ctypes.c_double),
        ("data", _D),
        ("coeffs", _D),
        ("nrot", ctypes.c_int),
        ("rot", _D),
        ("ntrans", _I),
        ("trans", _D),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_total_points.restype = ctypes.c_longlong
        L.orc_step_size.restype = ctypes.c_double
        L.orc_index_from_indices.restype = ctypes.c_longlong
        L.orc_index_from_indices.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong]
        L.orc_step_size.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_uniform_pair.argtypes = [ctypes.c_uint64] + [ctypes.c_uint32] * 4 + [_D, _D]
        L.orc_mc_stratify.argtypes = [ctypes.c_longlong, ctypes.c_int, _D, _D]
        L.orc_grid_intensity_dirs.restype = ctypes.c_double
        L.orc_mc_oa.argtypes = [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D, ctypes.c_uint64,
                                ctypes.c_longlong, ctypes.c_int, ctypes.c_double, _D, _LL]
        L.orc_vegas_oa.argtypes = [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _I]
        L.orc_gk_oa.argtypes = [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D, ctypes.c_double,
                                ctypes.c_longlong, _D, _LL]
        L.orc_azimuthal.argtypes = [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D, _D, ctypes.c_int,
                                    ctypes.c_uint64, _D]
        L.orc_mt19937_pairs.argtypes = [ctypes.c_uint32, ctypes.c_longlong] + [ctypes.c_double] * 4 + [_D]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- geometry
def total_points(grid_size):
    return int(lib().orc_total_points(grid_size))


def step_size(grid_size, qmax):
    return float(lib().orc_step_size(grid_size, qmax))


def index_from_indices(qi, ti, pi):
    return int(lib().orc_index_from_indices(qi, ti, pi))


def indices_from_index(idx):
    qi, ti, pi = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
    lib().orc_indices_from_index(ctypes.c_longlong(idx), ctypes.byref(qi), ctypes.byref(ti), ctypes.byref(pi))
    return qi.value, ti.value, pi.value


def qvector_from_index(grid_size, qmax, idx):
    out = np.zeros(3)
    lib().orc_qvector_from_index(grid_size, ctypes.c_double(qmax), ctypes.c_longlong(idx), _dp(out))
    return out


# ---------------------------------------------------------------- rng
def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def mc_stratify(d, chunk, u, v):
    """(u, v) of direction d of an MC chunk mapped into its run's patch."""
    cu, cv = ctypes.c_double(u), ctypes.c_double(v)
    lib().orc_mc_stratify(d, chunk, ctypes.byref(cu), ctypes.byref(cv))
    return cu.value, cv.value


def uniform_pair(seed, c0, c1, c2, c3):
    u, v = ctypes.c_double(), ctypes.c_double()
    lib().orc_uniform_pair(seed, c0, c1, c2, c3, ctypes.byref(u), ctypes.byref(v))
    return u.value, v.value


def mt19937_pairs(seed, n, a0, b0, a1, b1):
    out = np.zeros(2 * n)
    lib().orc_mt19937_pairs(seed, n, a0, b0, a1, b1, _dp(out))
    return out[0::2].copy(), out[1::2].copy()


# ---------------------------------------------------------------- form factors / amplitudes
class Ions:
    """Per-ion description: coef9 rows (a1,b1,..,a4,b4,c), radius [nm], atom counts."""

    def __init__(self, coef9, rad, count, sol_ed=0.0, c1=1.0, flags=CALC_ATOMIC_FORMFACTORS):
        self.coef9 = f64(coef9).reshape(-1, 9)
        self.rad = f64(rad)
        self.count = i32(count)
        self.sol_ed = float(sol_ed)
        self.c1 = float(c1)
        self.flags = int(flags)

    @property
    def nion(self):
        return self.coef9.shape[0]

    def args(self):
        return (self.nion, _dp(self.coef9), _dp(self.rad), _ip(self.count), ctypes.c_double(self.sol_ed),
                ctypes.c_double(self.c1), self.flags)


def form_factors(ions, q):
    out = np.zeros(ions.nion)
    lib().orc_form_factors(*ions.args(), ctypes.c_double(q), _dp(out))
    return out


def amp_atoms(xyz, ion, ions, qvec, scale=1.0):
    xyz, ion, qvec = f64(xyz), i32(ion), f64(qvec).reshape(-1, 3)
    out = np.zeros(2 * len(qvec))
    lib().orc_amp_atoms(len(ion), _dp(xyz), _ip(ion), *ions.args(), ctypes.c_double(scale), len(qvec),
                        _dp(qvec), _dp(out))
    return out[0::2] + 1j * out[1::2]


def amp_boxes(dims, com, vox, contrast, qvec, scale=1.0):
    dims, com, qvec = i32(dims), f64(com), f64(qvec).reshape(-1, 3)
    out = np.zeros(2 * len(qvec))
    lib().orc_amp_boxes(len(dims), _ip(dims), _dp(com), ctypes.c_double(vox), ctypes.c_double(contrast),
                        ctypes.c_double(scale), len(qvec), _dp(qvec), _dp(out))
    return out[0::2] + 1j * out[1::2]


def grid_fill_pdb(grid_size, qmax, xyz, ion, ions, scale=1.0, box_dims=None, box_com=None, vox=0.0,
                  contrast=0.0):
    xyz, ion = f64(xyz), i32(ion)
    P = total_points(grid_size)
    data = np.zeros(2 * P)
    if box_dims is None:
        bd, bc, nb = np.zeros(3, np.int32), np.zeros(3), 0
    else:
        bd, bc, nb = i32(box_dims), f64(box_com), len(box_dims)
    lib().orc_grid_fill_pdb(grid_size, ctypes.c_double(qmax), len(ion), _dp(xyz), _ip(ion), *ions.args(),
                            ctypes.c_double(scale), nb, _ip(bd), _dp(bc), ctypes.c_double(vox),
                            ctypes.c_double(contrast), _dp(data))
    return data


# ---------------------------------------------------------------- splines / lookup
def periodic_spline_ref(pts):
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.complex128))
    out = np.zeros(len(pts), dtype=np.complex128)
    lib().orc_periodic_spline_ref(len(pts), pts.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return out


def periodic_spline_exact(pts):
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.complex128))
    out = np.zeros(len(pts), dtype=np.complex128)
    lib().orc_periodic_spline_exact(len(pts), pts.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return out


def grid_splines(grid_size, data):
    data = f64(data)
    out = np.zeros_like(data)
    lib().orc_grid_splines(grid_size, _dp(data), _dp(out))
    return out


def get_sphr(grid_size, qmax, data, coeffs, q, th, ph):
    out = np.zeros(2)
    lib().orc_get_sphr(grid_size, ctypes.c_double(qmax), _dp(data), _dp(coeffs), ctypes.c_double(q),
                       ctypes.c_double(th), ctypes.c_double(ph), _dp(out))
    return complex(out[0], out[1])


def interp_plane(grid_size, data, coeffs, ri, th, ph):
    out = np.zeros(2)
    lib().orc_interp_plane(grid_size, _dp(data), _dp(coeffs), ri, ctypes.c_double(th), ctypes.c_double(ph),
                           _dp(out))
    return complex(out[0], out[1])


def amp_group(grid_size, qmax, data, coeffs, layer, th, ph, qs):
    qs = f64(qs)
    out = np.zeros(2 * len(qs))
    lib().orc_amp_group(grid_size, ctypes.c_double(qmax), _dp(data), _dp(coeffs), layer, ctypes.c_double(th),
                        ctypes.c_double(ph), len(qs), _dp(qs), _dp(out))
    return out[0::2] + 1j * out[1::2]


# ---------------------------------------------------------------- docking / orientation averages
class Docked:
    """Leaves for the docked-amplitude functions.  Each leaf: (grid_size, qmax, data,
    coeffs, rot[nrot,4] = (alpha, beta, gamma [rad], scale), ntrans[nrot], trans[sum ntrans,3])."""

    def __init__(self, leaves):
        self._keep = []
        arr = (Leaf * len(leaves))()
        for i, (gs, qmax, data, coeffs, rot, ntrans, trans) in enumerate(leaves):
            data, coeffs = f64(data), f64(coeffs)
            rot, ntrans, trans = f64(rot).reshape(-1, 4), i32(ntrans), f64(trans).reshape(-1, 3)
            self._keep += [data, coeffs, rot, ntrans, trans]
            arr[i] = Leaf(gs, qmax, _dp(data), _dp(coeffs), len(rot), _dp(rot), _ip(ntrans), _dp(trans))
        self.arr = arr
        self.n = len(leaves)

    @staticmethod
    def single(grid_size, qmax, data, coeffs):
        return Docked([(grid_size, qmax, data, coeffs, [[0, 0, 0, 1]], [1], [[0, 0, 0]])])


def docked_group(dock, layer, th, ph, qs):
    qs = f64(qs)
    out = np.zeros(2 * len(qs))
    lib().orc_docked_group(dock.n, dock.arr, layer, ctypes.c_double(th), ctypes.c_double(ph), len(qs), _dp(qs),
                           _dp(out))
    return out[0::2] + 1j * out[1::2]


def docked_cart(dock, q3):
    out = np.zeros(2)
    lib().orc_docked_cart(dock.n, dock.arr, ctypes.c_double(q3[0]), ctypes.c_double(q3[1]),
                          ctypes.c_double(q3[2]), _dp(out))
    return complex(out[0], out[1])


def azimuthal(dock, q, theta, nphi, seed):
    q, theta = f64(q), f64(theta)
    out = np.zeros(len(q))
    lib().orc_azimuthal(dock.n, dock.arr, len(q), _dp(q), _dp(theta), int(nphi), ctypes.c_uint64(seed), _dp(out))
    return out


def mc_oa(dock, qs, seed, max_dirs, chunk=8192, eps=0.0):
    qs = f64(qs)
    out = np.zeros(len(qs))
    used = ctypes.c_longlong()
    lib().orc_mc_oa(dock.n, dock.arr, len(qs), _dp(qs), seed, max_dirs, chunk, eps, _dp(out), ctypes.byref(used))
    return out, used.value


def vegas_oa(dock, qs, seed, iterations, points_per_bin=4, eps=0.0):
    qs = f64(qs)
    out = np.zeros(len(qs))
    used = ctypes.c_int()
    lib().orc_vegas_oa(dock.n, dock.arr, len(qs), _dp(qs), seed, iterations, points_per_bin, eps, _dp(out),
                       ctypes.byref(used))
    return out, used.value


def gk_oa(dock, qs, eps, max_depth=1 << 40):
    qs = f64(qs)
    out = np.zeros(len(qs))
    ev = ctypes.c_longlong()
    lib().orc_gk_oa(dock.n, dock.arr, len(qs), _dp(qs), eps, max_depth, _dp(out), ctypes.byref(ev))
    return out, ev.value


def dock_grid(out_grid_size, out_qmax, dock, scale=1.0):
    P = total_points(out_grid_size)
    out = np.zeros(2 * P)
    lib().orc_dock_grid(out_grid_size, ctypes.c_double(out_qmax), dock.n, dock.arr, ctypes.c_double(scale),
                        _dp(out))
    return out


def grid_intensity_seeded(grid_size, qmax, data, coeffs, q, eps, seed, iters):
    """JacobianSphereGrid::CalculateIntensity(q, eps, seed, iters) with its mt19937 stream."""
    phi, v2 = mt19937_pairs(seed, iters, 0.0, 2 * np.pi, 0.0, 2.0)
    ct = v2 - 1.0
    used = ctypes.c_longlong()
    r = lib().orc_grid_intensity_dirs(grid_size, ctypes.c_double(qmax), _dp(data), _dp(coeffs),
                                      ctypes.c_double(q), ctypes.c_double(eps), ctypes.c_longlong(iters),
                                      _dp(phi), _dp(ct), ctypes.byref(used))
    return r, used.value


def debye(xyz, ion, ions, qs, scale=1.0):
    xyz, ion, qs = f64(xyz), i32(ion), f64(qs)
    out = np.zeros(len(qs))
    lib().orc_debye(len(ion), _dp(xyz), _ip(ion), *ions.args(), ctypes.c_double(scale), len(qs), _dp(qs),
                    _dp(out))
    return out


# ---------------------------------------------------------------- reference grid (oracle/_ref)
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.refg_new.restype = ctypes.c_void_p
        R.refg_new.argtypes = [ctypes.c_int, ctypes.c_double]
        R.refg_free.argtypes = [ctypes.c_void_p]
        R.refg_data.restype = _D
        R.refg_data.argtypes = [ctypes.c_void_p]
        R.refg_interp.restype = _D
        R.refg_interp.argtypes = [ctypes.c_void_p]
        R.refg_real_size.restype = ctypes.c_longlong
        R.refg_real_size.argtypes = [ctypes.c_void_p]
        R.refg_splines.argtypes = [ctypes.c_void_p]
        R.refg_get_sphr.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
        R.refg_get_cart.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
        R.refg_intensity.restype = ctypes.c_double
        R.refg_intensity.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_uint,
                                     ctypes.c_ulonglong]
        R.refg_amps_at_points.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, _D]
        _ref = R
    return _ref


class RefGrid:
    """The reference's JacobianSphereGrid (compiled from the reference sources)."""

    def __init__(self, grid_size, qmax, data=None):
        self.h = ref().refg_new(grid_size, qmax)
        n = ref().refg_real_size(self.h) // 8
        self.n = n
        if data is not None:
            self.data[:] = data
            ref().refg_splines(self.h)

    @property
    def data(self):
        return np.ctypeslib.as_array(ref().refg_data(self.h), shape=(self.n,))

    @property
    def coeffs(self):
        return np.ctypeslib.as_array(ref().refg_interp(self.h), shape=(self.n,))

    def get_sphr(self, q, th, ph):
        out = np.zeros(2)
        ref().refg_get_sphr(self.h, q, th, ph, _dp(out))
        return complex(out[0], out[1])

    def get_cart(self, x, y, z):
        out = np.zeros(2)
        ref().refg_get_cart(self.h, x, y, z, _dp(out))
        return complex(out[0], out[1])

    def intensity(self, q, eps, seed, iters):
        return ref().refg_intensity(self.h, q, eps, seed, iters)

    def amps_at_points(self, qs, th, ph):
        qs = f64(qs)
        out = np.zeros(2 * len(qs))
        ref().refg_amps_at_points(self.h, len(qs), _dp(qs), th, ph, _dp(out))
        return out[0::2] + 1j * out[1::2]

    def __del__(self):
        try:
            ref().refg_free(self.h)
        except Exception:
            pass
