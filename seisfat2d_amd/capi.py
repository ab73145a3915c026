"""ctypes binding of the C ABI in include/seisfat2d_eik.h.

Loads seisfat2d_amd/libseisfat2d_eik.so (built in-tree by seisfat2d_amd.build);
raises if it is missing — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .build import EIK_LIB

OK, EINVAL, EHIP, ENOMEM, ESTATE, EIO = 0, -1, -2, -3, -4, -5
ISO, ANI = 0, 1
SCHED_PLAIN, SCHED_TILED, SCHED_TILED_DIAG = 0, 1, 2

# Every function the header declares (checked by the CPU test suite).
EXPORTS = [
    "seisfat_last_error", "seisfat_eik_create", "seisfat_eik_destroy", "seisfat_eik_set_slowness",
    "seisfat_eik_set_slowness_device", "seisfat_eik_set_stiffness", "seisfat_eik_solve",
    "seisfat_eik_device_T", "seisfat_eik_copy_T", "seisfat_batch_create", "seisfat_batch_run", "seisfat_batch_picks",
    "seisfat_batch_device_picks", "seisfat_batch_device_T", "seisfat_batch_copy_T",
    "seisfat_batch_num_launches", "seisfat_batch_num_sweep_launches", "seisfat_batch_sweep_ms",
    "seisfat_batch_status", "seisfat_batch_destroy", "seisfat_batch_run_record", "seisfat_batch_vjp_device",
    "seisfat_batch_misfit_gradient", "seisfat_eik_fast_division", "seisfat_eik_set_slowness_padded",
    "seisfat_eik_device_S", "seisfat_eik_solve_noalloc", "seisfat_batch_vjp", "seisfat_batch_run_converged",
    "seisfat_batch_tile_counts", "seisfat_batch_create_opts", "seisfat_batch_get_opts",
    # include/seisfat2d_kdm.h
    "seisfat_kdm_create", "seisfat_kdm_set_tables", "seisfat_kdm_set_traces", "seisfat_kdm_adjoint",
    "seisfat_kdm_forward", "seisfat_kdm_destroy", "seisfat_trace_convolve", "seisfat_ricker",
    "seisfat_kdm_set_table",
]
KDM_IMAGE, KDM_ANGLE = 0, 1


class SeisfatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"seisfat error {code}: {msg}")
        self.code = code


class SchedOpts(C.Structure):
    """seisfat_sched_opts: schedule options of a tiled batch (bit-identical results)."""
    _fields_ = [("pairs", C.c_int32), ("wpcu", C.c_int32), ("skip", C.c_int32), ("priority", C.c_int32)]

    FIELDS = ("pairs", "wpcu", "skip", "priority")

    @classmethod
    def of(cls, d):
        o = cls(1, 0, 1, 1)
        for k, v in (d or {}).items():
            if k not in cls.FIELDS:
                raise ValueError(f"unknown schedule option {k!r} (known: {', '.join(cls.FIELDS)})")
            setattr(o, k, int(v))
        return o

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.FIELDS}


class KdmParams(C.Structure):
    """seisfat_kdm_params (include/seisfat2d_kdm.h)."""
    _fields_ = [("domain", C.c_int32), ("nt", C.c_int32), ("dt", C.c_float), ("aperture", C.c_float),
                ("da", C.c_float), ("nang", C.c_int32), ("ncmp", C.c_int32)]


class Shot(C.Structure):
    _fields_ = [("sx", C.c_float), ("sz", C.c_float), ("x0", C.c_int32), ("width", C.c_int32),
                ("rec_offset", C.c_int32), ("nrec", C.c_int32)]


_lib = None
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(EIK_LIB):
            raise RuntimeError(f"{EIK_LIB} is not built; run __graft_entry__.build()")
        L = C.CDLL(EIK_LIB)
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        L.seisfat_last_error.restype = C.c_char_p
        L.seisfat_last_error.argtypes = []
        L.seisfat_eik_create.argtypes = [i, i, f, f, C.POINTER(vp)]
        L.seisfat_eik_destroy.argtypes = [vp]
        L.seisfat_eik_set_slowness.argtypes = [vp, f32p]
        L.seisfat_eik_set_slowness_device.argtypes = [vp, vp, vp]
        L.seisfat_eik_set_slowness_padded.argtypes = [vp, f32p]
        L.seisfat_eik_device_S.restype = vp
        L.seisfat_eik_device_S.argtypes = [vp]
        L.seisfat_eik_set_stiffness.argtypes = [vp, C.POINTER(vp)]
        L.seisfat_eik_solve.argtypes = [vp, i, i, f, f, vp, vp, vp, i, vp]
        L.seisfat_eik_solve_noalloc.argtypes = [vp, f, f, vp, vp, vp, i, vp, vp, vp]
        L.seisfat_eik_device_T.restype = vp
        L.seisfat_eik_device_T.argtypes = [vp]
        L.seisfat_eik_copy_T.argtypes = [vp, vp]
        L.seisfat_eik_fast_division.argtypes = [vp]
        L.seisfat_batch_create.argtypes = [vp, i, i, i, C.POINTER(Shot), i, vp, vp, C.POINTER(vp)]
        L.seisfat_batch_create_opts.argtypes = [vp, i, i, i, C.POINTER(Shot), i, vp, vp, C.POINTER(SchedOpts),
                                                C.POINTER(vp)]
        L.seisfat_batch_get_opts.argtypes = [vp, C.POINTER(SchedOpts)]
        L.seisfat_batch_run.argtypes = [vp, vp]
        L.seisfat_batch_picks.argtypes = [vp, vp]
        L.seisfat_batch_device_picks.restype = vp
        L.seisfat_batch_device_picks.argtypes = [vp]
        L.seisfat_batch_device_T.restype = vp
        L.seisfat_batch_device_T.argtypes = [vp, i, C.POINTER(i), C.POINTER(i)]
        L.seisfat_batch_copy_T.argtypes = [vp, i, vp]
        L.seisfat_batch_num_launches.argtypes = [vp]
        L.seisfat_batch_num_sweep_launches.argtypes = [vp]
        L.seisfat_batch_sweep_ms.argtypes = [vp, C.POINTER(f)]
        L.seisfat_batch_tile_counts.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.seisfat_batch_destroy.argtypes = [vp]
        L.seisfat_batch_status.argtypes = [vp]
        L.seisfat_batch_run_record.argtypes = [vp, vp]
        L.seisfat_batch_vjp_device.argtypes = [vp, vp, vp, vp]
        L.seisfat_batch_misfit_gradient.argtypes = [vp, vp, vp, vp, vp]
        L.seisfat_batch_vjp.argtypes = [vp, vp, vp]
        L.seisfat_batch_run_converged.argtypes = [vp, i, i, C.POINTER(i), vp]
        L.seisfat_kdm_create.argtypes = [vp, C.POINTER(KdmParams), i, i, C.POINTER(vp)]
        L.seisfat_kdm_set_tables.argtypes = [vp, vp, vp, vp]
        L.seisfat_kdm_set_traces.argtypes = [vp, i, vp, vp, vp, vp]
        L.seisfat_kdm_set_table.argtypes = [vp, i, i, vp, vp]
        L.seisfat_kdm_adjoint.argtypes = [vp, vp, vp, vp]
        L.seisfat_kdm_forward.argtypes = [vp, vp, vp, vp]
        L.seisfat_kdm_destroy.argtypes = [vp]
        L.seisfat_trace_convolve.argtypes = [i, i, i, vp, i, vp, vp, vp]
        L.seisfat_ricker.argtypes = [f, f, vp, i]
        ptr_ret = {"seisfat_last_error", "seisfat_eik_device_T", "seisfat_batch_device_picks",
                   "seisfat_batch_device_T", "seisfat_eik_device_S"}
        for name in EXPORTS:
            if name not in ptr_ret:
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise SeisfatError(rc, lib().seisfat_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)
