"""ctypes binding of libcfdhip.so (the C ABI in include/cfdhip.h).

There is no CPU fallback: if the library is missing or no GPU is visible,
every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_NAME = "libcfdhip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

CFD_OK = 0
CFD_EINVAL = 1
CFD_EHIP = 2
CFD_ENOMEM = 3
CFD_ENOTINIT = 4
CFD_EINITIAL = 10
CFD_EINVALID_DT = 11
CFD_EREJECTED = 12

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class CfdReferenceElement(C.Structure):
    _fields_ = [("dr", _dp), ("ds", _dp), ("lift", _dp), ("cell_weights", _dp), ("average_weights", _dp),
                ("face_weights", _dp), ("p1_midpoint", _dp), ("linear_basis", _dp), ("face_nodes", _ip),
                ("midpoint_nodes", _ip)]


class CfdMeshTables(C.Structure):
    _fields_ = [("num_elements", C.c_int64), ("geo", _dp), ("conn", _ip), ("lim", _dp), ("bc_states", _dp),
                ("num_bc_states", C.c_int32)]


class CfdStepReport(C.Structure):
    _fields_ = [("boundary_flux", C.c_double * 4), ("admissible", C.c_int32), ("min_density", C.c_double),
                ("min_pressure", C.c_double), ("troubled_cells", C.c_int64)]


class CfdRunResult(C.Structure):
    _fields_ = [("time", C.c_double), ("steps", C.c_int64), ("rejected_steps", C.c_int64),
                ("boundary_flux_integral", C.c_double * 4), ("min_density", C.c_double),
                ("min_pressure", C.c_double), ("max_troubled_cells", C.c_int64), ("fail_time", C.c_double),
                ("fail_dt", C.c_double)]


class CfdAdjointStats(C.Structure):
    _fields_ = [("num_steps", C.c_int64), ("checkpoints", C.c_int64), ("tape_records", C.c_int64),
                ("tape_peak", C.c_int64), ("tape_regrows", C.c_int64), ("positivity_scaled", C.c_int64),
                ("forward_ms", C.c_double), ("backward_ms", C.c_double), ("recompute_ms", C.c_double),
                ("device_bytes", C.c_int64), ("skipped_passes", C.c_int64)]


class CfdDeviceMesh(C.Structure):
    _fields_ = [("num_elements", C.c_int64), ("geo", C.c_void_p), ("conn", C.c_void_p), ("alpha_beta", C.c_void_p),
                ("dt_geometry", C.c_void_p), ("bc_states", C.c_void_p), ("num_bc_states", C.c_int32)]


CFD_STATIC_MAX_ELEMENTS = 65536


class CfdMeshInput(C.Structure):
    _fields_ = [("num_vertices", C.c_int64), ("vertices", _dp), ("num_elements", C.c_int64), ("elements", _i64p),
                ("num_tags", C.c_int32), ("tag_names", C.c_char_p), ("tag_offsets", _i64p), ("tag_edges", _i64p),
                ("tag_kind", _ip), ("tag_state", _ip), ("tag_partner", _ip), ("tag_shift", _dp)]


class CfdMeshOutput(C.Structure):
    _fields_ = [("geo", _dp), ("conn", _ip), ("lim", _dp), ("neighbor_element", _i64p), ("neighbor_face", _i64p),
                ("face_tag", _i64p), ("centroid", _dp), ("normal", _dp), ("edge_length", _dp),
                ("face_midpoint", _dp), ("periodic_shift", _dp)]

# name -> (restype, argtypes)
SIGNATURES = {
    "cfd_last_error": (C.c_char_p, []),
    "cfd_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "cfd_init": (C.c_int, [C.POINTER(CfdReferenceElement)]),
    "cfd_ctx_create": (C.c_int, [C.POINTER(CfdMeshTables), C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(C.c_void_p)]),
    "cfd_ctx_destroy": (C.c_int, [C.c_void_p]),
    "cfd_ctx_set_stage_block": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd_rhs": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "cfd_rhs_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    "cfd_limit": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _i64p, _u8p]),
    "cfd_time_step": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cfd_step": (C.c_int, [C.c_void_p, _dp, C.c_double, _dp, C.POINTER(CfdStepReport), _u8p]),
    "cfd_step_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.POINTER(CfdStepReport)]),
    "cfd_step_device_static": (C.c_int, [C.POINTER(CfdDeviceMesh), C.c_int, C.c_int, C.c_double, C.c_void_p,
                                         C.c_double, C.c_void_p, C.POINTER(CfdStepReport)]),
    "cfd_cell_averages": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cfd_integrate": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cfd_run": (C.c_int, [C.c_void_p, _dp, C.c_double, C.c_int64, _dp, C.POINTER(CfdRunResult)]),
    "cfd_state_upload": (C.c_int, [C.c_void_p, _dp]),
    "cfd_state_download": (C.c_int, [C.c_void_p, _dp]),
    "cfd_state_advance": (C.c_int, [C.c_void_p, C.c_int64, _dp, C.POINTER(CfdStepReport)]),
    "cfd_synchronize": (C.c_int, [C.c_void_p]),
    "cfd_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd_timing_read": (C.c_int, [C.c_void_p, _dp]),
    "cfd_time_rhs": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "cfd_rhs_vjp": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "cfd_limit_vjp": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "cfd_adjoint_create": (C.c_int, [C.c_void_p, C.c_int64, _dp, C.c_int, C.POINTER(C.c_void_p)]),
    "cfd_adjoint_destroy": (C.c_int, [C.c_void_p]),
    "cfd_adjoint_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cfd_adjoint_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cfd_adjoint_forward_replay": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cfd_adjoint_time_steps": (C.c_int, [C.c_void_p, _dp]),
    "cfd_adjoint_stats": (C.c_int, [C.c_void_p, C.POINTER(CfdAdjointStats)]),
    "cfd_flux_pointwise": (C.c_int, [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, C.c_double, _dp]),
    "cfd_limiter_pointwise": (C.c_int, [C.c_int, C.c_int64, _dp, _dp]),
    "cfd_build_tables": (C.c_int, [C.POINTER(CfdMeshInput), C.POINTER(CfdMeshOutput)]),
}

_lib = None


class NativeError(RuntimeError):
    """A failure inside libcfdhip.so (HIP error, allocation failure, ...)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libcfdhip error {code}: {message}")
        self.code = code
        self.message = message


def load() -> C.CDLL:
    """Load libcfdhip.so (built by ``__graft_entry__.build()``); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("CFDHIP_LIBRARY", LIB_PATH))
    if not path.exists():
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (there is no CPU fallback).")
    try:
        # torch ships its own libamdhip64; loaded first, it is the one HIP
        # runtime of the process (a second runtime loaded before it leaves
        # torch without devices).
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().cfd_last_error().decode(errors="replace")


def check(code: int) -> None:
    if code == CFD_OK:
        return
    msg = last_error()
    if code == CFD_EINVAL:
        raise ValueError(msg)
    raise NativeError(code, msg)


def device_count() -> int:
    n = C.c_int(0)
    load().cfd_device_count(C.byref(n))
    return n.value


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


_initialised = False
_keepalive = []


def ensure_init(reference) -> None:
    """Upload the reference-element tables once per process."""
    global _initialised
    if _initialised:
        return
    lib = load()
    if device_count() < 1:
        raise RuntimeError("No HIP device is visible; libcfdhip.so has no CPU fallback.")
    arrs = dict(
        dr=np.ascontiguousarray(reference.dr, dtype=np.float64),
        ds=np.ascontiguousarray(reference.ds, dtype=np.float64),
        lift=np.ascontiguousarray(reference.lift, dtype=np.float64),
        cell_weights=np.ascontiguousarray(reference.cell_weights, dtype=np.float64),
        average_weights=np.ascontiguousarray(reference.cell_weights / reference.cell_weights.sum()),
        face_weights=np.ascontiguousarray(reference.face_weights.reshape(-1), dtype=np.float64),
        p1_midpoint=np.ascontiguousarray(reference.p1_midpoint, dtype=np.float64),
        linear_basis=np.ascontiguousarray(reference.linear_midpoint_basis, dtype=np.float64),
        face_nodes=np.ascontiguousarray(reference.face_nodes.reshape(-1), dtype=np.int32),
        midpoint_nodes=np.ascontiguousarray(reference.face_midpoint_nodes, dtype=np.int32),
    )
    ref = CfdReferenceElement(**{k: (iptr(v) if v.dtype == np.int32 else dptr(v)) for k, v in arrs.items()})
    _keepalive.append(arrs)
    check(lib.cfd_init(C.byref(ref)))
    _initialised = True
