"""Drop-in counterparts of the reference's compiled Python extension
``dplus.wrappers`` (PythonInterface/dplus/wrappers/wrappers.pyx), backed by
libdplus_mi355x through its C ABI.

``BackendWrapper`` mirrors BackendWrapper.pyx (the PythonBackendWrapper job
interface); failures raise ``RuntimeError`` whose text is the JSON object
``{"code": int, "message": str}``, as PythonBackendWrapper::ConvertException
does (Backend/Backend/PythonBackendWrapper.cpp:49-60), so the reference's
``dplus.Backend`` layer (PythonInterface/dplus/Backend.py:4-8) parses it
unchanged.

``CJacobianSphereGrid`` mirrors Grid.pyx:10-200 (JacobianSphereGrid): a grid
whose host mirror ``get_data()`` is filled on the host and then splined on the
GPU by ``calculate_splines()``.
"""
import ctypes
import json

import numpy as np

from . import _lib as L


def _raise_runtime(rc):
    """RuntimeError carrying the library's {"code", "message"} JSON."""
    msg = L.lib().dplus_last_error().decode()
    try:
        json.loads(msg)
    except ValueError:
        msg = json.dumps({"code": int(rc), "message": msg})
    raise RuntimeError(msg)


def _call(rc):
    if rc != 0:
        _raise_runtime(rc)


def _text(s):
    return s.encode("utf-8") if isinstance(s, str) else bytes(s)


class BackendWrapper:
    """BackendWrapper.pyx:4-56 over dplus_backend_* (one job at a time)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _call(L.lib().dplus_backend_create(ctypes.byref(h)))
        self._h = h

    def check_capabilities(self, tdrLevel=True):
        _call(L.lib().dplus_backend_check_capabilities(self._h, 1 if tdrLevel else 0))

    def get_all_model_metadata(self):
        p = ctypes.c_void_p()
        _call(L.lib().dplus_backend_get_all_model_metadata(self._h, ctypes.byref(p)))
        return L.take_string(p).decode()

    def initialize_cache(self, cache_dir=""):
        _call(L.lib().dplus_backend_initialize_cache(self._h, _text(cache_dir)))

    def start_generate(self, state, useGPU=True):
        _call(L.lib().dplus_backend_start_generate(self._h, _text(state), 1 if useGPU else 0))

    def start_generate_2D(self, state, useGPU=True):
        """BackendWrapper.pyx:22-23: the (q_z, q_perp) intensity matrix job."""
        _call(L.lib().dplus_backend_start_generate_2d(self._h, _text(state), 1 if useGPU else 0))

    def get_generate_2D_results(self):
        """BackendWrapper.pyx:31-32: {"Headers", "2DGraph"} JSON."""
        p = ctypes.c_void_p()
        _call(L.lib().dplus_backend_get_generate_2d_results(self._h, ctypes.byref(p)))
        return L.take_string(p).decode()

    def get_job_status(self):
        p = ctypes.c_void_p()
        _call(L.lib().dplus_backend_get_job_status(self._h, ctypes.byref(p)))
        return L.take_string(p).decode()

    def get_generate_results(self):
        p = ctypes.c_void_p()
        _call(L.lib().dplus_backend_get_generate_results(self._h, ctypes.byref(p)))
        return L.take_string(p).decode()

    def save_amp(self, modelptr, path):
        _call(L.lib().dplus_backend_save_amplitude(self._h, int(modelptr), _text(path)))

    def get_amp(self, modelptr):
        p, n = ctypes.c_void_p(), ctypes.c_longlong()
        _call(L.lib().dplus_backend_get_amplitude(self._h, int(modelptr), ctypes.byref(p), ctypes.byref(n)))
        return L.take_string(p, n.value)

    def get_pdb(self, modelptr):
        """GetPDB decoded as BackendWrapper.get_pdb does (wrappers/BackendWrapper.pyx:40-43)."""
        p, n = ctypes.c_void_p(), ctypes.c_longlong()
        _call(L.lib().dplus_backend_get_pdb(self._h, int(modelptr), ctypes.byref(p), ctypes.byref(n)))
        return L.take_string(p, n.value).decode("utf-8")

    def get_model_ptrs(self):
        n = ctypes.c_int()
        _call(L.lib().dplus_backend_get_model_ptrs(self._h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_uint * max(1, n.value))()
        _call(L.lib().dplus_backend_get_model_ptrs(self._h, buf, n.value, ctypes.byref(n)))
        return list(buf[:n.value])

    def stop(self):
        _call(L.lib().dplus_backend_stop(self._h))

    def wait(self, poll=0.01):
        """Poll get_job_status until the job ends (EmbeddedLocalRunner.generate,
        PythonInterface/dplus/CalculationRunner.py:724-738); returns the last status."""
        import time
        while True:
            st = json.loads(self.get_job_status())
            if not (st["isRunning"] and st["code"] == -1):
                return st
            time.sleep(poll)

    def __del__(self):
        try:
            if self._h:
                L.lib().dplus_backend_destroy(self._h)
                self._h = None
        except Exception:
            pass


class CJacobianSphereGrid:
    """Grid.pyx:10-200 (JacobianSphereGrid) on the GPU.  The host mirror is in
    the reference layout (2*P doubles, complex interleaved); FP64 storage keeps
    it exact through calculate_splines()."""

    theta_divisions = 3  # Grid.h: thetaDivisions
    phi_divisions = 6    # Grid.h: phiDivisions

    def __init__(self, gridSize=0, qMax=0, precision=L.PRECISION_FP64):
        from .grid import Grid
        self._g = Grid(gridSize, qMax, precision)
        self._g.download()

    @classmethod
    def _wrap(cls, grid):
        obj = cls.__new__(cls)
        obj._g = grid
        return obj

    @property
    def grid(self):
        return self._g

    @property
    def q_max(self):
        return self._g.q_max

    @property
    def grid_size(self):
        return self._g.grid_size

    @property
    def step_size(self):
        return self._g.step_size

    @property
    def actual_size(self):
        return self._g.actual_size

    @property
    def extras(self):
        """GetExtras (Grid.h:203): shells past gridSize/2 (actualGridSize = gridSize/2 + Extras, Grid.cpp:178)."""
        return self.actual_size - self.grid_size // 2

    @property
    def real_size(self):
        """GetRealSize: bytes of the value array (Grid.cpp GetRealSize)."""
        return 2 * self._g.points * 8

    def index_from_indices(self, qi, ti, pi):
        return int(L.lib().dplus_grid_index_from_indices(int(qi), int(ti), int(pi)))

    def indices_from_index(self, index):
        qi, ti, pi = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        L.lib().dplus_grid_indices_from_index(int(index), ctypes.byref(qi), ctypes.byref(ti), ctypes.byref(pi))
        return qi.value, ti.value, pi.value

    def get_data(self):
        return self._g.data

    def get_interpolant_coeffs(self):
        return self._g.interp

    def calculate_splines(self):
        """Host values -> GPU, CalculateSplines there, coefficients back to the host mirror."""
        self._g.upload()
        self._g.download()

    def get_param_json_string(self):
        """GetParamJsonString (Grid.cpp:238-269)."""
        return json.dumps({"qmax": self.q_max, "qmin": 0.0, "gridSize": self.grid_size,
                           "stepSize": self.step_size, "Params": []})

    def get_dim(self, x, y):
        """GetDimX/GetDimY/GetDimZ (Grid.cpp:366-392)."""
        dim_x = self.actual_size
        dim_y = 1 if x == 0 else self.theta_divisions * x + 1
        dim_z = 1 if x == 0 else self.phi_divisions * x
        return [dim_x, dim_y, dim_z]

    def _node_angles(self):
        n = self._g.points
        idx = np.arange(n)
        qi = np.zeros(n, np.int64)
        ti = np.zeros(n, np.int64)
        pi = np.zeros(n, np.int64)
        # shells in node order: shell i holds (3i+1)*6i nodes after 1 + 6(i-1)i^2
        for i in range(1, self.actual_size + 1):
            base = 1 + 6 * (i - 1) * i * i
            cnt = (3 * i + 1) * 6 * i
            sl = slice(base, base + cnt)
            k = idx[sl] - base
            qi[sl] = i
            ti[sl] = k // (6 * i)
            pi[sl] = k % (6 * i)
        q = qi * self.step_size
        with np.errstate(invalid="ignore", divide="ignore"):
            th = np.where(qi > 0, np.pi * ti / np.maximum(3 * qi, 1), 0.0)
            ph = np.where(qi > 0, 2 * np.pi * pi / np.maximum(6 * qi, 1), 0.0)
        return q, th, ph

    def fill(self, calcAmplitude, args=()):
        """Grid.pyx fill: data[node] = calcAmplitude(q, theta, phi, *args), then splines."""
        data = self.get_data()
        q, th, ph = self._node_angles()
        for ind in range(self._g.points):
            rs = complex(calcAmplitude(q[ind], th[ind], ph[ind], *args))
            data[2 * ind], data[2 * ind + 1] = rs.real, rs.imag
        self.calculate_splines()

    def fill_cart(self, calcAmplitude_cart):
        """Grid.pyx fill_cart: data[node] = calcAmplitude_cart(qx, qy, qz), then splines."""
        data = self.get_data()
        q, th, ph = self._node_angles()
        x, y, z = q * np.sin(th) * np.cos(ph), q * np.sin(th) * np.sin(ph), q * np.cos(th)
        for ind in range(self._g.points):
            rs = complex(calcAmplitude_cart(x[ind], y[ind], z[ind]))
            data[2 * ind], data[2 * ind + 1] = rs.real, rs.imag
        self.calculate_splines()

    def get_sphr(self, q, theta, phi):
        """GetSphr (Grid.cpp:487-548) at one point."""
        return complex(self._g.get_sphr(q, theta, phi)[0])

    def interpolate_theta_phi_plane(self, ri, theta, phi):
        """InterpolateThetaPhiPlane (Grid.pyx:139-140, Grid.cpp:61-134): the value inside shell ri."""
        return complex(self._g.interpolate_theta_phi_plane(ri, theta, phi)[0])

    def get_sphr_array(self, q, theta, phi):
        """GetSphr at many points in one GPU launch."""
        return self._g.get_sphr(q, theta, phi)

    def get_intensity(self, q, theta, epsi, seed, iterations, phi_min, phi_max):
        """CalculateIntensity (Grid.cpp:393-484); theta None draws theta at random."""
        return self._g.calculate_intensity(q, theta, epsi, seed, iterations, phi_min, phi_max)


def amp_read(path, precision=L.PRECISION_FP64):
    """(CJacobianSphereGrid, header text) of a .amp/.ampj amplitude file."""
    from .grid import Grid
    h, hdr = ctypes.c_void_p(), ctypes.c_void_p()
    L.check(L.lib().dplus_amp_read(_text(path), int(precision), ctypes.byref(h), ctypes.byref(hdr)))
    g = Grid.__new__(Grid)
    g._init_from_handle(h)
    g.download()
    return CJacobianSphereGrid._wrap(g), L.take_string(hdr).decode("utf-8", "replace")


def amp_write(grid, path, header=""):
    """Write a grid as .ampj (zip) or legacy .amp, by the path's extension."""
    g = grid.grid if isinstance(grid, CJacobianSphereGrid) else grid
    L.check(L.lib().dplus_amp_write(g.handle, _text(path), _text(header)))
