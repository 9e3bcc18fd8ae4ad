"""ctypes binding of libcrabsaxs_hip.so (the C ABI declared in include/crabsaxs.h).

The library is loaded from this package directory only.  There is no
fallback: if it has not been built, importing the compute API raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcrabsaxs_hip.so"

c_double_p = C.POINTER(C.c_double)
c_size_t_p = C.POINTER(C.c_size_t)


class SolventParamsC(C.Structure):
    _fields_ = [
        ("c1", C.c_double),
        ("c2", C.c_double),
        ("shell_thickness_angstrom", C.c_double),
        ("bulk_density_e_per_a3", C.c_double),
        ("hydration_model", C.c_int32),
        ("hydration_grid_spacing_angstrom", C.c_double),
    ]


class ParseOptionsC(C.Structure):
    _fields_ = [
        ("include_hetatm", C.c_int32),
        ("include_waters", C.c_int32),
        ("include_glycans", C.c_int32),
        ("model_index", C.c_size_t),
        ("allow_unknown_elements", C.c_int32),
        ("occupancy_policy", C.c_int32),
        ("altloc_policy", C.c_int32),
        ("chains", C.c_char_p),
    ]


class AtomRecordC(C.Structure):
    _fields_ = [
        ("x", C.c_double),
        ("y", C.c_double),
        ("z", C.c_double),
        ("occupancy", C.c_double),
        ("element", C.c_int32),
        ("group", C.c_int32),
        ("is_hetatm", C.c_int32),
        ("residue_id", C.c_int64),
        ("atom_name", C.c_char * 16),
        ("residue_name", C.c_char * 16),
        ("chain_id", C.c_char * 16),
    ]


class ComputeOptionsC(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("implicit_hydrogen", C.c_int32),
        ("has_solvent", C.c_int32),
        ("solvent", SolventParamsC),
        ("l_max", C.c_int32),
        ("precision", C.c_int32),
    ]


class FitParamsC(C.Structure):
    _fields_ = [("scale", C.c_double), ("background", C.c_double), ("c1", C.c_double),
                ("c2", C.c_double)]


class FitOptionsC(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("implicit_hydrogen", C.c_int32),
        ("c1_bounds", C.c_double * 2),
        ("c2_bounds", C.c_double * 2),
        ("max_iterations", C.c_int32),
        ("q_sampling_stride", C.c_int32),
        ("solvent", SolventParamsC),
        ("background_mode", C.c_int32),
        ("background_value", C.c_double),
        ("scale_mode", C.c_int32),
        ("scale_value", C.c_double),
        ("optimizer", C.c_int32),
        ("precision", C.c_int32),
    ]


class FitResultC(C.Structure):
    _fields_ = [("params", FitParamsC), ("chi2", C.c_double), ("iterations", C.c_int64),
                ("converged", C.c_int32), ("termination", C.c_char * 96)]


_fo = C.POINTER(FitOptionsC)
_fp = C.POINTER(FitParamsC)

# name -> (restype, argtypes); every symbol declared in include/crabsaxs.h.
SIGNATURES = {
    "csx_last_error": (C.c_char_p, []),
    "csx_solvent_default": (None, [C.POINTER(SolventParamsC)]),
    "csx_compute_options_default": (None, [C.POINTER(ComputeOptionsC)]),
    "csx_device_count": (C.c_int, []),
    "csx_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "csx_ctx_destroy": (None, [C.c_void_p]),
    "csx_ctx_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64),
                                      c_double_p]),
    "csx_ctx_reset_timing": (C.c_int, [C.c_void_p]),
    "csx_ctx_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "csx_element_from_symbol": (C.c_int, [C.c_char_p]),
    "csx_united_atom_group": (C.c_int, [C.c_char_p, C.c_char_p]),
    "csx_form_factor": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]),
    "csx_excluded_volume": (C.c_double, [C.c_int, C.c_int, C.c_int]),
    "csx_water_form_factor": (C.c_double, [C.c_double]),
    "csx_excluded_volume_scale": (C.c_double, [C.c_double, C.c_double]),
    "csx_structure_create": (C.c_int, [C.c_void_p, C.c_size_t, c_double_p,
                                       C.POINTER(C.c_uint8), C.POINTER(C.c_int8), c_double_p,
                                       C.POINTER(C.c_void_p)]),
    "csx_structure_destroy": (None, [C.c_void_p]),
    "csx_structure_atom_count": (C.c_size_t, [C.c_void_p]),
    "csx_structure_set_hydration_sites": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                                    C.c_double]),
    "csx_residue_kind": (C.c_int, [C.c_char_p]),
    "csx_parse_options_default": (None, [C.c_void_p]),
    "csx_read_structure_file": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "csx_parse_structure": (C.c_int, [C.c_char_p, C.c_size_t, C.c_int, C.c_void_p,
                                      C.POINTER(C.c_void_p)]),
    "csx_atoms_count": (C.c_size_t, [C.c_void_p]),
    "csx_atoms_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "csx_atoms_arrays": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(C.c_uint8),
                                   C.POINTER(C.c_int8), c_double_p]),
    "csx_atoms_destroy": (None, [C.c_void_p]),
    "csx_read_dat_file": (C.c_int, [C.c_char_p, c_double_p, c_double_p, c_double_p, C.c_size_t,
                                    c_size_t_p, C.POINTER(C.c_int)]),
    "csx_parse_dat": (C.c_int, [C.c_char_p, C.c_size_t, c_double_p, c_double_p, c_double_p,
                                C.c_size_t, c_size_t_p, C.POINTER(C.c_int)]),
    "csx_build_adaptive_hydration_grid": (C.c_int, [C.c_void_p, C.c_double, C.c_double,
                                                    C.c_double, c_double_p, C.c_size_t,
                                                    c_size_t_p, c_double_p]),
    "csx_build_hydration_grid": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                           c_double_p, C.c_size_t, c_size_t_p, c_double_p]),
    "csx_debye_intensity": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, C.c_int, C.c_int,
                                      c_double_p]),
    "csx_debye_solvent_basis": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                          C.POINTER(SolventParamsC), C.c_int, C.c_int,
                                          c_double_p]),
    "csx_solvent_corrected_intensity": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                                  C.POINTER(SolventParamsC), C.c_int, C.c_int,
                                                  c_double_p]),
    "csx_debye_pair_tables": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                        C.POINTER(SolventParamsC), c_double_p, c_double_p,
                                        c_size_t_p]),
    "csx_combine_solvent_basis": (C.c_int, [c_double_p, c_double_p, C.c_size_t, C.c_double,
                                            C.c_double, c_double_p]),
    "csx_multipole_intensity": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, C.c_int,
                                          C.c_int, c_double_p]),
    "csx_multipole_solvent_intensity": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, C.c_int,
                                                  C.c_int, C.POINTER(SolventParamsC),
                                                  c_double_p]),
    "csx_multipole_solvent_basis": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, C.c_int,
                                              C.c_int, C.POINTER(SolventParamsC), c_double_p]),
    "csx_compute_curve": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                    C.POINTER(ComputeOptionsC), c_double_p,
                                    C.POINTER(C.c_int)]),
    "csx_ctx_set_persistence": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_set_vjp_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_device_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "csx_ctx_set_workspace_limit": (C.c_int, [C.c_void_p, C.c_size_t]),
    "csx_ctx_set_cache_limit": (C.c_int, [C.c_void_p, C.c_size_t]),
    "csx_pcg32": (C.c_int, [C.c_uint64, C.c_uint64, C.c_size_t, C.c_int, C.POINTER(C.c_uint32)]),
    "csx_ctx_set_vjp_split": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_set_multipole_group": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_set_multipole_radial": (C.c_int, [C.c_void_p, C.c_int]),
    "csx_ctx_last_multipole": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "csx_ctx_last_vjp_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "csx_debye_pair_tables_noalloc": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, C.c_void_p,
                                                C.c_void_p]),
    "csx_ctx_last_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "csx_ctx_last_pair_items": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "csx_batch_debye_intensity": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, c_double_p,
                                            C.c_size_t, C.c_int, c_double_p]),
    "csx_batch_debye_solvent_basis": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, c_double_p,
                                                C.c_size_t, C.POINTER(SolventParamsC), C.c_int,
                                                c_double_p]),
    "csx_debye_grad_x": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, c_double_p,
                                   C.POINTER(SolventParamsC), C.c_int, c_double_p]),
    "csx_debye_jacobian": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t,
                                     C.POINTER(SolventParamsC), C.c_int, c_double_p]),
    "csx_fit_options_default": (None, [_fo]),
    "csx_solve_scale_background": (C.c_int, [c_double_p, c_double_p, c_double_p, C.c_size_t,
                                             C.c_int, C.c_double, C.c_int, C.c_double,
                                             c_double_p]),
    "csx_cubic_hermite_interpolate": (C.c_int, [c_double_p, c_double_p, C.c_size_t, c_double_p,
                                                C.c_size_t, c_double_p]),
    "csx_fit_basis": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, _fo, c_double_p]),
    "csx_fit_to_experimental": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p,
                                          C.c_size_t, _fp, _fo, C.POINTER(FitResultC),
                                          c_double_p]),
    "csx_fit_from_basis": (C.c_int, [c_double_p, c_double_p, c_double_p, c_double_p, C.c_size_t,
                                     _fp, _fo, C.POINTER(FitResultC), c_double_p]),
    "csx_chi2_grad": (C.c_int, [c_double_p, c_double_p, c_double_p, c_double_p, C.c_size_t, _fp,
                                c_double_p, c_double_p]),
    "csx_project_simplex": (C.c_int, [c_double_p, C.c_size_t, c_double_p]),
    "csx_optimize_weights": (C.c_int, [c_double_p, C.c_size_t, c_double_p, c_double_p,
                                       C.c_size_t, _fo, C.c_int, C.c_double, c_double_p,
                                       c_double_p, c_double_p]),
    "csx_nnls_weights": (C.c_int, [c_double_p, C.c_size_t, c_double_p, c_double_p, C.c_size_t,
                                   _fo, c_double_p, c_double_p, c_double_p]),
    "csx_ensemble_fit_core": (C.c_int, [c_double_p, C.c_size_t, c_double_p, c_double_p,
                                        c_double_p, C.c_size_t, _fo, C.c_int, C.c_double,
                                        C.c_int, c_double_p, _fp, c_double_p, c_double_p]),
    "csx_ensemble_fit_core_device": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t, c_double_p,
                                               c_double_p, c_double_p, C.c_size_t, _fo, C.c_int,
                                               C.c_double, C.c_int, c_double_p, _fp, c_double_p,
                                               c_double_p]),
    "csx_coordinate_features": (C.c_int, [C.c_void_p, C.c_size_t, c_double_p, c_double_p,
                                          c_double_p, c_double_p, C.POINTER(C.c_size_t),
                                          c_double_p, c_double_p, C.POINTER(C.c_int)]),
    "csx_bootstrap_indices": (C.c_int, [C.c_uint64, C.c_size_t, C.c_size_t,
                                        C.POINTER(C.c_uint64)]),
    "csx_chacha_block": (C.c_int, [C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64, C.c_int,
                                   C.POINTER(C.c_uint32)]),
    "csx_generate_globule": (C.c_int, [C.c_size_t, C.c_double, C.c_double, C.c_uint64,
                                       c_double_p, C.POINTER(C.c_uint8)]),
    "csx_generate_domain_chain": (C.c_int, [C.c_size_t, C.c_int, C.c_uint64, c_double_p,
                                            C.POINTER(C.c_uint8)]),
    "csx_generate_shell_sites": (C.c_int, [c_double_p, C.c_size_t, C.c_double, C.c_double,
                                           C.c_double, c_double_p, C.c_size_t, c_size_t_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m crabsaxs_amd.build` first")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(c_double_p)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
