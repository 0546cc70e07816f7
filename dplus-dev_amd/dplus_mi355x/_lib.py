"""ctypes binding of libdplus_mi355x (include/dplus_mi355x.h).

The shared library is built in-tree by ``make -C dplus-dev_amd``; loading
fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdplus_mi355x.so")

PRECISION_FP64 = 0
PRECISION_MIXED = 1
OA_MC, OA_VEGAS, OA_GK = 0, 1, 2
CALC_ATOMIC_FORMFACTORS = 1
CALC_DUMMY_SOLVENT = 2

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.POINTER(ctypes.c_void_p)  # char ** outputs, released with dplus_free


class DplusError(RuntimeError):
    """Error raised by the library: ``code`` follows the reference's ErrorCode enum
    (Common/Common.h:184-228), the message is the library's."""

    def __init__(self, code, message):
        self.error_code = code
        super().__init__(message)


class PdbDesc(ctypes.Structure):
    _fields_ = [
        ("natoms", ctypes.c_int), ("xyz", _D), ("ion", _I), ("nion", ctypes.c_int), ("coef9", _D),
        ("rad", _D), ("sol_ed", ctypes.c_double), ("c1", ctypes.c_double), ("flags", ctypes.c_int),
        ("scale", ctypes.c_double), ("nbox", ctypes.c_int), ("box_dims", _I), ("box_com", _D),
        ("vox_step", ctypes.c_double), ("box_contrast", ctypes.c_double), ("nexcl", ctypes.c_int),
        ("excl_dims", _I), ("excl_com", _D), ("excl_contrast", ctypes.c_double),
    ]


class Leaf(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_void_p), ("nrot", ctypes.c_int), ("rot", _D), ("ntrans", _I), ("trans", _D)]


class OAParams(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("max_dirs", ctypes.c_longlong),
        ("chunk", ctypes.c_int), ("eps", ctypes.c_double), ("vegas_iterations", ctypes.c_int),
        ("vegas_points_per_bin", ctypes.c_int), ("gk_max_depth", ctypes.c_longlong),
    ]


class OAStats(ctypes.Structure):
    _fields_ = [("dirs_used", ctypes.c_longlong), ("iterations_used", ctypes.c_int),
                ("evaluations", ctypes.c_longlong), ("seconds", ctypes.c_double)]


class EngineConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_longlong), ("streams", ctypes.c_int), ("batch_builds", ctypes.c_int),
                ("batch_terms", ctypes.c_longlong), ("autotune", ctypes.c_int), ("graphs", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [
        ("pdb", ctypes.POINTER(PdbDesc)), ("grid", ctypes.c_void_p), ("grid_size", ctypes.c_int),
        ("qmax", ctypes.c_double), ("precision", ctypes.c_int), ("nrot", ctypes.c_int), ("rot", _D),
        ("ntrans", _I), ("trans", _D), ("nq", ctypes.c_int), ("q", _D), ("oa", OAParams), ("data", _D),
        ("sigma", _D), ("fit_scale", ctypes.c_double), ("fit_constant", ctypes.c_double),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
        ("chi2", ctypes.c_double), ("chi2_on_device", ctypes.c_int), ("slot", ctypes.c_int),
        ("reserved_bytes", ctypes.c_longlong), ("oa", OAStats), ("graph", ctypes.c_int),
    ]


# every symbol of include/dplus_mi355x.h with its ctypes signature
SIGNATURES = {
    "dplus_last_error": (ctypes.c_char_p, []),
    "dplus_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "dplus_grid_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "dplus_grid_destroy": (None, [ctypes.c_void_p]),
    "dplus_grid_info": (ctypes.c_int, [ctypes.c_void_p, _I, _D, _D, _I, ctypes.POINTER(ctypes.c_longlong), _I]),
    "dplus_grid_data": (_D, [ctypes.c_void_p]),
    "dplus_grid_interp": (_D, [ctypes.c_void_p]),
    "dplus_grid_index_from_indices": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong]),
    "dplus_grid_indices_from_index": (None, [ctypes.c_longlong, _I, ctypes.POINTER(ctypes.c_longlong),
                                             ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_grid_upload": (ctypes.c_int, [ctypes.c_void_p]),
    "dplus_grid_calculate_splines": (ctypes.c_int, [ctypes.c_void_p]),
    "dplus_grid_download": (ctypes.c_int, [ctypes.c_void_p]),
    "dplus_grid_get_sphr": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D]),
    "dplus_gather_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_longlong, ctypes.c_int,
                                          ctypes.c_int, _D, _D]),
    "dplus_grid_interpolate_theta_phi_plane": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I, _D, _D, _D]),
    "dplus_grid_calculate_intensity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_uint, ctypes.c_ulonglong,
                                                      ctypes.c_double, ctypes.c_double, _D]),
    "dplus_pdb_build_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PdbDesc), ctypes.c_int]),
    "dplus_orientation_average": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D,
                                                 ctypes.POINTER(OAParams), _D, ctypes.POINTER(OAStats)]),
    "dplus_orientation_average_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D,
                                                 ctypes.POINTER(OAParams), _D, ctypes.POINTER(OAStats)]),
    "dplus_dock_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_double]),
    "dplus_dock_grid_space_filling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Leaf),
                                                     ctypes.c_double, _D, _D, _D]),
    "dplus_synth_walk": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]),
    "dplus_hydration_boxes": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _I, _I, _D]),
    "dplus_excluded_volume_boxes": (ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, _I, _I, _D]),
    "dplus_set_fp64_builder": (ctypes.c_int, [ctypes.c_int]),
    "dplus_lua_symmetry_rows": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _I,
                                               _D]),
    "dplus_structure_create": (ctypes.c_int, [ctypes.POINTER(PdbDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "dplus_structure_destroy": (None, [ctypes.c_void_p]),
    "dplus_structure_build_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "dplus_structure_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _I, _I, _D, _D]),
    "dplus_structure_build_grids": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                                   ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "dplus_build_shapes_clear": (ctypes.c_int, []),
    "dplus_structure_build_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _I, _I, _I]),
    "dplus_build_shapes_save": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_build_shapes_load": (ctypes.c_int, [ctypes.c_char_p]),
    "dplus_timing_enable": (ctypes.c_int, [ctypes.c_int]),
    "dplus_timing_get": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong), _D]),
    "dplus_timing_reset": (ctypes.c_int, []),
    "dplus_microbench_valu": (ctypes.c_int, [_D, _D, _D, _D]),
    "dplus_copy_probe": (ctypes.c_int, [ctypes.c_longlong, _D]),
    "dplus_amp_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _VP]),
    "dplus_amp_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]),
    "dplus_free": (None, [ctypes.c_void_p]),
    "dplus_backend_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "dplus_backend_destroy": (None, [ctypes.c_void_p]),
    "dplus_backend_check_capabilities": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dplus_backend_get_all_model_metadata": (ctypes.c_int, [ctypes.c_void_p, _VP]),
    "dplus_backend_initialize_cache": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dplus_backend_start_generate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "dplus_backend_start_generate_2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "dplus_backend_get_generate_2d_results": (ctypes.c_int, [ctypes.c_void_p, _VP]),
    "dplus_azimuthal_average": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Leaf), ctypes.c_int, _D, _D, ctypes.c_int,
                                               ctypes.c_ulonglong, _D]),
    "dplus_backend_get_job_status": (ctypes.c_int, [ctypes.c_void_p, _VP]),
    "dplus_backend_get_generate_results": (ctypes.c_int, [ctypes.c_void_p, _VP]),
    "dplus_backend_get_model_ptrs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint), ctypes.c_int, _I]),
    "dplus_backend_get_amplitude": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _VP,
                                                   ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_backend_get_pdb": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _VP, ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_backend_save_amplitude": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_char_p]),
    "dplus_backend_stop": (ctypes.c_int, [ctypes.c_void_p]),
    "dplus_engine_create": (ctypes.c_int, [ctypes.POINTER(EngineConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "dplus_engine_destroy": (None, [ctypes.c_void_p]),
    "dplus_engine_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_engine_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _I]),
    "dplus_engine_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _D, ctypes.POINTER(RequestResult)]),
    "dplus_engine_stats": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 5),
    "dplus_engine_cache_stats": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 4),
    "dplus_engine_batch_stats": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 2),
    "dplus_engine_graph_stats": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 3 + [_D, _D]),
    "dplus_session_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _D, _D, _D,
                                            ctypes.c_ulonglong, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p)]),
    "dplus_session_evaluate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Leaf), ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, _D, _D]),
    "dplus_session_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                          ctypes.POINTER(ctypes.c_longlong)]),
    "dplus_session_destroy": (None, [ctypes.c_void_p]),
}

_lib = None


def lib():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C dplus-dev_amd` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().dplus_last_error().decode()
        try:
            d = json.loads(msg)
            raise DplusError(d.get("code", rc), d.get("message", msg))
        except ValueError:
            raise DplusError(rc, msg)


def take_string(p, length=None):
    """Bytes of a library-allocated buffer, then release it with dplus_free."""
    if not p.value:
        return b""
    try:
        return ctypes.string_at(p.value, length) if length is not None else ctypes.string_at(p.value)
    finally:
        lib().dplus_free(p.value)


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def device_info():
    buf = ctypes.create_string_buffer(512)
    check(lib().dplus_device_info(buf, 512))
    return json.loads(buf.value.decode())


def timing_enable(on=True):
    check(lib().dplus_timing_enable(1 if on else 0))


def timing_reset():
    check(lib().dplus_timing_reset())


def timing_get(kernel):
    """(launches, total ms) of a kernel family measured with HIP events on the library stream."""
    n, ms = ctypes.c_longlong(), ctypes.c_double()
    check(lib().dplus_timing_get(kernel.encode(), ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def microbench_valu():
    """Device-wide ops/s of v_fma_f32, v_sin_f32, v_cos_f32, v_fma_f64."""
    r = [ctypes.c_double() for _ in range(4)]
    check(lib().dplus_microbench_valu(*[ctypes.byref(x) for x in r]))
    return dict(zip(("fma32", "sin32", "cos32", "fma64"), (x.value for x in r)))


def copy_probe(nbytes):
    """Time a plain device copy of nbytes (ms); the known-size kernel for byte counters."""
    ms = ctypes.c_double()
    check(lib().dplus_copy_probe(int(nbytes), ctypes.byref(ms)))
    return ms.value


def gather_probe(grid, directions, layer0, nlayers, rot=None):
    """dplus_gather_probe: (ms, lookups) of the MC lookups' node loads alone."""
    ms, n = ctypes.c_double(), ctypes.c_double()
    r = None if rot is None else f64(rot).reshape(-1, 4)
    check(lib().dplus_gather_probe(grid.handle, 0 if r is None else len(r), None if r is None else dp(r),
                                   int(directions), int(layer0), int(nlayers), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value
