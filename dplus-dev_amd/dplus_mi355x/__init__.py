"""dplus_mi355x: MI355X-native scattering hot path of D+ (amplitude grids,
symmetry docking and orientation averages) behind a C ABI
(include/dplus_mi355x.h) and the reference's backend interface."""
from ._lib import (CALC_ATOMIC_FORMFACTORS, CALC_DUMMY_SOLVENT, OA_GK, OA_MC, OA_VEGAS, PRECISION_FP64,
                   PRECISION_MIXED, DplusError, device_info, lib)
from ._lib import copy_probe, gather_probe, microbench_valu, timing_enable, timing_get, timing_reset
from .grid import (Docking, Grid, Structure, azimuthal_average, build_grids, build_shapes_clear, build_shapes_load, build_shapes_save,
                   dock_grid, dock_grid_space_filling, excluded_volume_boxes, hydration_boxes, index_from_indices, indices_from_index, lua_symmetry_rows, set_fp64_builder,
                   orientation_average, synth_walk)


def _load_shipped_shapes():
    """Loads the builder launch shapes tuned on MI355X (build_shapes_mi355x.txt,
    written by python -m dplus_mi355x.tune) into the library's shape table."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "build_shapes_mi355x.txt")
    if os.path.exists(path) and os.path.exists(_lib.LIB_PATH):
        with open(path) as f:
            build_shapes_load(f.read())


_load_shipped_shapes()
