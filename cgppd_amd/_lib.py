"""ctypes binding of libcgppd_hip.so (include/cgppd_hip.h).

The product path has no CPU fallback: if the library is missing, loading
raises.  Build it with ``python -m cgppd_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
HIP_LIB_PATH = _PKG / "libcgppd_hip.so"

_i32p = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_ctxp = C.c_void_p

# name -> (restype, argtypes); mirrors include/cgppd_hip.h
SIGNATURES = {
    "cgppd_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "cgppd_destroy": (C.c_int, [_ctxp]),
    "cgppd_last_error": (C.c_char_p, [_ctxp]),
    "cgppd_device_count": (C.c_int, [_i32p]),
    "cgppd_set_force_field": (C.c_int, [_ctxp, _f32p, _f32p, _f32p]),
    "cgppd_set_box": (C.c_int, [_ctxp, C.c_float]),
    "cgppd_set_system": (C.c_int, [_ctxp, C.c_int, _i32p, _i32p, C.c_int]),
    "cgppd_set_torsion_table": (C.c_int, [_ctxp, _f64p, _f64p]),
    "cgppd_set_topology": (C.c_int, [_ctxp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _i32p, _i32p]),
    "cgppd_energy_internal": (C.c_int, [_ctxp, C.c_int, _f64p]),
    "cgppd_topology_list": (C.c_int, [C.c_int, _i32p, _i32p, C.c_int, C.c_int, _i32p, _i32p, C.c_int, _i32p, _i32p]),
    "cgppd_set_replicas": (C.c_int, [_ctxp, C.c_int, _f32p, _f32p, _f64p, _f32p, _f32p, _i32p, C.c_uint64]),
    "cgppd_upload_positions": (C.c_int, [_ctxp, C.c_int, _f32p]),
    "cgppd_update_molecule": (C.c_int, [_ctxp, C.c_int, C.c_int, _f32p]),
    "cgppd_energy": (C.c_int, [_ctxp, C.c_int, _f64p, _f64p, _f64p]),
    "cgppd_energy_nc": (C.c_int, [_ctxp, C.c_int, _f64p, _f64p, _f64p]),
    "cgppd_energy_nc_all": (C.c_int, [_ctxp, _f64p]),
    "cgppd_energy_async": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_energy_wait": (C.c_int, [_ctxp, C.c_int, _f64p, _f64p, _f64p]),
    "cgppd_energy_all": (C.c_int, [_ctxp, _f64p, _f64p, _f64p, _f64p]),
    "cgppd_energy_all_noalloc": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cgppd_mc_prepare": (C.c_int, [_ctxp]),
    "cgppd_mc_steps": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_exchange": (C.c_int, [_ctxp]),
    "cgppd_run": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int]),
    "cgppd_synchronize": (C.c_int, [_ctxp]),
    "cgppd_get_replica_stats": (C.c_int, [_ctxp, _f64p, _f64p, _i32p, _i64p, _i64p, _i64p]),
    "cgppd_get_pair_evaluations": (C.c_int, [_ctxp, _i64p]),
    "cgppd_get_coordinates": (C.c_int, [_ctxp, C.c_int, _f32p, _f32p, _f32p]),
    "cgppd_get_energy_matrix": (C.c_int, [_ctxp, C.c_int, _f64p]),
    "cgppd_set_launch": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_set_pair_layout": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_set_replica_groups": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_get_replica_groups": (C.c_int, [_ctxp, _i32p]),
    "cgppd_autotune": (C.c_int, [_ctxp, C.c_int, _i32p, _i32p, _f64p]),
    "cgppd_get_pair_launch": (C.c_int, [_ctxp, _i32p, _i32p]),
    "cgppd_set_window": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_get_window": (C.c_int, [_ctxp, _i32p]),
    "cgppd_autotune_table": (C.c_int, [_ctxp, _f64p]),
    "cgppd_set_image_shift": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_set_repulsive_crowding": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_set_rigid_potential_only": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_set_timing": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_last_kernel_ms": (C.c_int, [_ctxp, _f64p, _i64p]),
    "cgppd_model_constants": (C.c_int, [_f32p, _f32p, _f32p]),
    "cgppd_get_sizes": (C.c_int, [_ctxp, _i32p, _i32p, _i32p]),
    "cgppd_surrogate_param_count": (C.c_int, [_i32p]),
    "cgppd_surrogate_init": (C.c_int, [_ctxp, C.c_uint64]),
    "cgppd_surrogate_set_params": (C.c_int, [_ctxp, _f32p]),
    "cgppd_surrogate_get_params": (C.c_int, [_ctxp, _f32p]),
    "cgppd_surrogate_predict": (C.c_int, [_ctxp, _f32p, C.c_int, _f32p]),
    "cgppd_surrogate_gradient": (C.c_int, [_ctxp, _f32p, _f32p, C.c_int, _f32p, _f64p]),
    "cgppd_surrogate_add_samples": (C.c_int, [_ctxp, _f32p, _f32p, C.c_int]),
    "cgppd_surrogate_collect": (C.c_int, [_ctxp]),
    "cgppd_surrogate_sample_count": (C.c_int, [_ctxp, _i64p]),
    "cgppd_surrogate_clear_samples": (C.c_int, [_ctxp]),
    "cgppd_surrogate_train": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_float, C.c_uint64, _f64p]),
    "cgppd_surrogate_features": (C.c_int, [_ctxp, C.c_int, _f32p]),
    "cgppd_set_delayed_acceptance": (C.c_int, [_ctxp, C.c_int]),
    "cgppd_get_stage1_rejections": (C.c_int, [_ctxp, _i64p]),
}

_lib = None


def load(path: Path | None = None):
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("CGPPD_HIP_LIB", HIP_LIB_PATH))
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m cgppd_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


TOPOLOGY_KINDS = {"bonds": (0, 2), "angles": (1, 3), "torsions": (2, 4), "local": (3, 1), "crankshaft": (4, 1),
                  "flex": (5, 1), "segment_bonds": (6, 2), "indirect": (7, 2), "domain_of": (8, 1),
                  "flex_nb": (9, 1), "flex_nb_off": (10, 1), "branch": (11, 1), "branch_off": (12, 1),
                  "segbond_of": (13, 1), "counts": (14, 1)}


def _segments_csr(segments):
    off = np.zeros(len(segments) + 1, dtype=np.int32)
    flat = []
    for k, sg in enumerate(segments):
        flat.extend(int(v) for v in sg)
        off[k + 1] = len(flat)
    return off, np.asarray(flat if flat else [0], dtype=np.int32)


def topology_list(chain_code, chain_index, all_flexible, segments, kind):
    """Host-only query of the library's topology builder
    (cgppd_topology_list): one list of one molecule, shaped (-1, width)."""
    lib = load()
    code, idx = i32(chain_code), i32(chain_index)
    off, res = _segments_csr(segments)
    k, width = TOPOLOGY_KINDS[kind]
    n = C.c_int()
    args = (len(code), ptr(code, C.c_int), ptr(idx, C.c_int), int(bool(all_flexible)), len(segments),
            ptr(off, C.c_int), ptr(res, C.c_int), k)
    rc = lib.cgppd_topology_list(*args, None, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"cgppd_topology_list failed ({rc})")
    out = np.zeros(max(n.value, 1), dtype=np.int32)
    rc = lib.cgppd_topology_list(*args, ptr(out, C.c_int), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"cgppd_topology_list failed ({rc})")
    out = out[:n.value]
    return out.reshape(-1, width) if width > 1 else out
