"""ctypes binding of libchimera_hip.so (declared in include/chimera_hip.h).

The library is built in-tree (``python -m chimeracl_amd.build``).  There is no
fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

from .build import LIB as LIBPATH

CHM_MAXM = 4
CHM_DHT_F32 = 0
CHM_DHT_BF16X3 = 1

P = ctypes.c_void_p
i32, i64, f64, f32, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_float, ctypes.c_size_t


class Grid(ctypes.Structure):
    _fields_ = [("Nx", i32), ("Nr", i32), ("M", i32), ("pad_", i32),
                ("xmin", f64), ("dx_inv", f64), ("rmin", f64), ("dr_inv", f64)]


class Planes(ctypes.Structure):
    _fields_ = [("m0", P), ("mc", P * CHM_MAXM)]


class Fb(ctypes.Structure):
    _fields_ = [("m", P * (CHM_MAXM + 1))]


_MP = P * (CHM_MAXM + 1)


class Spectral(ctypes.Structure):
    _fields_ = [("Nx", i32), ("Nr", i32), ("M", i32), ("ndamp", i32),
                ("dht_kind", i32), ("pad_", i32), ("dt_inv", f64),
                ("kx", P), ("phase_minus", P), ("phase_plus", P), ("damp", P),
                ("twiddle", P),
                ("DHT", _MP), ("DHT_inv", _MP), ("dDp", _MP), ("dDm", _MP),
                ("DHT_lo", _MP), ("DHT_inv_lo", _MP), ("dDp_lo", _MP), ("dDm_lo", _MP),
                ("c1", _MP), ("c2", _MP), ("c3", _MP), ("smooth", _MP),
                ("c1d", _MP), ("c2d", _MP), ("c3d", _MP)]


class TablesOut(ctypes.Structure):
    _fields_ = [("kx", P), ("kr", P), ("phase_minus", P), ("phase_plus", P),
                ("damp", P), ("twiddle", P),
                ("w", _MP), ("c1", _MP), ("c2", _MP), ("c3", _MP), ("smooth", _MP),
                ("DHT", _MP), ("DHT_inv", _MP), ("dDp", _MP), ("dDm", _MP)]


class Beam(ctypes.Structure):
    _fields_ = [("c", f64 * 3), ("L", f64 * 3), ("pc", f64 * 3), ("dp", f64 * 3), ("w", f64)]


R = ctypes.POINTER
_SIGS = {
    "chm_push_x_f64": [P, P, P, P, P, P, P, f64, i64, P],
    "chm_push_x_f32": [P, P, P, P, P, P, P, f32, i64, P],
    "chm_sort_workspace_bytes": ([i64, i64], sz),
    "chm_index_sort_f64": [P, P, P, i64, R(Grid), P, P, P, P, P, sz, P],
    "chm_index_sort_f32": [P, P, P, i64, R(Grid), P, P, P, P, P, sz, P],
    "chm_push_x_keys_f64": [P, P, P, P, P, P, P, f64, i64, R(Grid), P, P, sz, P],
    "chm_push_x_keys_f32": [P, P, P, P, P, P, P, f32, i64, R(Grid), P, P, sz, P],
    "chm_sort_from_keys": [i64, R(Grid), P, P, P, P, P, sz, P],
    "chm_index_sort_f64_noalloc": [P, P, P, i64, R(Grid), P, P, P, P, P],
    "chm_align_elems": [P, i64, P, P, i32, P],
    "chm_align_multi": [P, i64, i32, P, P, P, P],
    "chm_dens_profile_f64": [P, P, i64, P, P, P, i32, P],
    "chm_dens_profile_f32": [P, P, i64, P, P, P, i32, P],
    "chm_deposit_workspace_bytes": ([i32, i32, i32, i32], sz),
    "chm_deposit_scalar_f64": [P, P, P, P, P, P, i32, R(Grid), R(Planes), P, sz, P],
    "chm_deposit_scalar_f32": [P, P, P, P, P, P, i32, R(Grid), R(Planes), P, sz, P],
    "chm_deposit_vector_f64": [P, P, P, P, P, P, P, P, P, P, i32, R(Grid), R(Planes), P, sz, P],
    "chm_deposit_vector_f32": [P, P, P, P, P, P, P, P, P, P, i32, R(Grid), R(Planes), P, sz, P],
    "chm_deposit_vector_push_f64": [P, P, P, P, P, P, P, P, P, P, P, P, P, i32, f64, i64, P,
                                    R(Grid), R(Planes), P, sz, P],
    "chm_deposit_vector_push_f32": [P, P, P, P, P, P, P, P, P, P, P, P, P, i32, f32, i64, P,
                                    R(Grid), R(Planes), P, sz, P],
    "chm_deposit_add_staged_f64": [P, R(Grid), R(Planes), i32, P, sz, P],
    "chm_deposit_add_staged_f32": [P, R(Grid), R(Planes), i32, P, sz, P],
    "chm_postproc_deposit_f64": [R(Planes), i32, P, i32, i32, i32, P],
    "chm_postproc_deposit_f32": [R(Planes), i32, P, i32, i32, i32, P],
    "chm_zero_planes_f64": [R(Planes), i32, i32, i32, i32, P],
    "chm_zero_planes_f32": [R(Planes), i32, i32, i32, i32, P],
    "chm_warp_axis_f64": [R(Planes), i32, i32, i32, i32, P],
    "chm_warp_axis_f32": [R(Planes), i32, i32, i32, i32, P],
    "chm_gather_push_f64": [P, P, P, P, P, P, P, i64, P, P, P, f64, R(Grid), R(Planes), R(Planes),
                            P],
    "chm_gather_push_f32": [P, P, P, P, P, P, P, i64, P, P, P, f32, R(Grid), R(Planes), R(Planes),
                            P],
    "chm_gather_push_ids_f64": [P, P, P, P, P, P, P, i64, P, P, P, P, f64, R(Grid), R(Planes),
                                R(Planes), P],
    "chm_gather_push_ids_f32": [P, P, P, P, P, P, P, i64, P, P, P, P, f32, R(Grid), R(Planes),
                                R(Planes), P],
    "chm_gather_push_pre_f64": [P, P, P, P, P, P, P, i64, P, P, P, P, P, f64, f64, P, P, P, P,
                                R(Grid), R(Planes), R(Planes), P],
    "chm_gather_push_pre_f32": [P, P, P, P, P, P, P, i64, P, P, P, P, P, f32, f32, P, P, P, P,
                                R(Grid), R(Planes), R(Planes), P],
    "chm_sort_sorted_keys": ([P, i64, R(Grid)], P),
    "chm_transform_workspace_bytes": ([i32, i32, i32, i32, i32], sz),
    "chm_transform_f64": [R(Spectral), R(Planes), R(Fb), i32, i32, i32, i32, P, sz, P],
    "chm_transform_f32": [R(Spectral), R(Planes), R(Fb), i32, i32, i32, i32, P, sz, P],
    "chm_mult_table_f64": [R(Spectral), R(Fb), i32, i32, P],
    "chm_mult_table_f32": [R(Spectral), R(Fb), i32, i32, P],
    "chm_spectral_workspace_bytes": ([i32, i32, i32, i32], sz),
    "chm_field_grad_f64": [R(Spectral), R(Fb), R(Fb), R(Fb), P, sz, P],
    "chm_field_grad_f32": [R(Spectral), R(Fb), R(Fb), R(Fb), P, sz, P],
    "chm_field_div_f64": [R(Spectral), R(Fb), R(Fb), P, sz, P],
    "chm_field_div_f32": [R(Spectral), R(Fb), R(Fb), P, sz, P],
    "chm_field_rot_f64": [R(Spectral), R(Fb), R(Fb), i32, P, sz, P],
    "chm_field_rot_f32": [R(Spectral), R(Fb), R(Fb), i32, P, sz, P],
    "chm_advance_fields_f64": [R(Spectral), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), P],
    "chm_advance_fields_f32": [R(Spectral), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), P],
    "chm_damp_fields_f64": [R(Spectral), R(Fb), R(Fb), i32, R(Fb), R(Fb), R(Fb), P],
    "chm_damp_fields_f32": [R(Spectral), R(Fb), R(Fb), i32, R(Fb), R(Fb), R(Fb), P],
    "chm_tables_workspace_bytes": ([i32, i32, i32], sz),
    "chm_build_tables_f64": [i32, i32, i32, f64, f64, f64, f64, i32, R(TablesOut), P, sz, P],
    "chm_gemm_f64": [P, P, P, i32, i32, i32, P],
    "chm_dht_timing": [i32],
    "chm_dht_timing_collect": ([P, P, P, i64], i64),
    "chm_dht_table_elems": ([i32, i32], sz),
    "chm_pad_table": [P, i32, i32, P, P, P, P, P],
    "chm_convert_table": [P, i64, P, P, P, P],
    "chm_laser_pulse_f64": [P, i32, i32, f64, f64, f64, f64, f64, f64, f64, f64, f64, P],
    "chm_laser_spectral_f64": [R(Spectral), P, P, P, f64, P],
    "chm_fill_domain_f64": [P, P, P, P, P, P, P, P, P, P, i32, i64, i32, i32, i32, f64, f64, f64,
                            f64, f64, f64, f64, ctypes.c_uint64, ctypes.c_uint64, P],
    "chm_fill_domain_f32": [P, P, P, P, P, P, P, P, P, P, i32, i64, i32, i32, i32, f64, f64, f64,
                            f64, f64, f64, f64, ctypes.c_uint64, ctypes.c_uint64, P],
    "chm_phase_tables_f64": [P, i32, f64, P, P, P],
    "chm_fill_beam_f64": [P, P, P, P, P, P, P, P, i64, R(Beam), ctypes.c_uint64, ctypes.c_uint64, P],
    "chm_fill_beam_f32": [P, P, P, P, P, P, P, P, i64, R(Beam), ctypes.c_uint64, ctypes.c_uint64, P],
    # adjoints (differentiable form)
    "chm_push_x_adj_f64": [P, P, P, P, P, P, P, P, P, P, P, f64, i64, P],
    "chm_push_x_adj_f32": [P, P, P, P, P, P, P, P, P, P, P, f32, i64, P],
    "chm_deposit_scalar_adj_f64": [P, P, P, P, i32, i64, R(Grid), R(Planes), P, P, P, P, P],
    "chm_deposit_scalar_adj_f32": [P, P, P, P, i32, i64, R(Grid), R(Planes), P, P, P, P, P],
    "chm_deposit_vector_adj_f64": [P, P, P, P, P, P, P, P, i32, i64, R(Grid), R(Planes),
                                   P, P, P, P, P, P, P, P, P],
    "chm_deposit_vector_adj_f32": [P, P, P, P, P, P, P, P, i32, i64, R(Grid), R(Planes),
                                   P, P, P, P, P, P, P, P, P],
    "chm_gather_push_adj_f64": [P, P, P, P, P, P, i64, P, f64, R(Grid), R(Planes), R(Planes),
                                P, P, P, P, P, P, P, P, P],
    "chm_gather_push_adj_f32": [P, P, P, P, P, P, i64, P, f32, R(Grid), R(Planes), R(Planes),
                                P, P, P, P, P, P, P, P, P],
    "chm_gather_adj_scatter_f64": [P, P, P, P, P, P, P, P, R(Grid), R(Planes), P, sz, P],
    "chm_gather_adj_scatter_f32": [P, P, P, P, P, P, P, P, R(Grid), R(Planes), P, sz, P],
    "chm_postproc_deposit_adj_f64": [R(Planes), i32, P, i32, i32, i32, P],
    "chm_postproc_deposit_adj_f32": [R(Planes), i32, P, i32, i32, i32, P],
    "chm_warp_axis_adj_f64": [R(Planes), i32, i32, i32, i32, P],
    "chm_warp_axis_adj_f32": [R(Planes), i32, i32, i32, i32, P],
    "chm_profile_adj_workspace_bytes": ([i32], sz),
    "chm_dens_profile_adj_f64": [P, P, P, i64, P, P, i32, P, P, sz, P],
    "chm_dens_profile_adj_f32": [P, P, P, i64, P, P, i32, P, P, sz, P],
    "chm_field_grad_adj_f64": [R(Spectral), R(Fb), R(Fb), i32, P, sz, P],
    "chm_field_grad_adj_f32": [R(Spectral), R(Fb), R(Fb), i32, P, sz, P],
    "chm_field_rot_adj_f64": [R(Spectral), R(Fb), R(Fb), i32, i32, P, sz, P],
    "chm_field_rot_adj_f32": [R(Spectral), R(Fb), R(Fb), i32, i32, P, sz, P],
    "chm_damp_fields_adj_f64": [R(Spectral), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), P],
    "chm_damp_fields_adj_f32": [R(Spectral), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), R(Fb), P],
    "chm_laser_pulse_dR_f64": [P, i32, i32, f64, f64, f64, f64, f64, f64, f64, f64, f64, P],
}

ERRORS = {-1: "CHM_EINVAL (argument out of range)", -2: "CHM_EWORKSPACE (workspace too small)",
          -3: "CHM_EUNSUPPORTED (size not supported)",
          -4: "CHM_ESCRATCH (input larger than the fixed scratch array allows)"}


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(
            f"{LIBPATH} is missing: build it with `python -m chimeracl_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIBPATH)
    for name, sig in _SIGS.items():
        fn = getattr(lib, name)
        if isinstance(sig, tuple):
            fn.argtypes, fn.restype = sig
        else:
            fn.argtypes, fn.restype = sig, ctypes.c_int
    return lib


lib = _load()


def call(name, *args):
    """Call an int-returning entry point and raise on a nonzero code."""
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed: {ERRORS.get(rc, f'hipError_t {rc}')}")


def exported_symbols():
    return sorted(_SIGS)
