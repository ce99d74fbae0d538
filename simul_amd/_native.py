"""ctypes binding of libpdx.so (include/pdx.h).

The library is the only compute path: importing this module raises if the
in-tree build is missing, and every call checks the returned status and maps
it onto the reference's exception types (FEAValidationError / ValueError for
bad input, RuntimeError for ordering and device errors).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .validation import FEAValidationError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SIMUL_AMD_LIB") or os.path.join(_HERE, "_lib", "libpdx.so")

OK, EINVAL, ESTATE, EDEVICE, ENOMEM, EUNSUP = 0, 1, 2, 3, 4, 5

MODEL_LINEAR_ELASTIC, MODEL_NEO_HOOKEAN, MODEL_J2 = 0, 1, 2
FORM_NOSB, FORM_BOND = 0, 1
DR_ADAPTIVE, DR_KINETIC = 0, 1
SING_ABSOLUTE, SING_RELATIVE = 0, 1

FIELD = {
    "X": 0, "x": 1, "u": 2, "v": 3, "a": 4, "f": 5, "volume": 6, "mass": 7, "bulk": 8,
    "shear": 9, "cstab": 10, "K": 11, "K_inv": 12, "F": 13, "P": 14, "damage": 15,
    "fixed": 16, "fext": 17, "ep": 18, "epbar": 19, "back": 20, "nnb": 21, "sing": 22,
    "drmass": 23, "density": 24,
}
# per-particle shape kind: "vec" n*dim, "scal" n, "ten" n*dim*dim, "sym9" n*9
FIELD_KIND = {
    "X": ("vec", np.float64), "x": ("vec", np.float64), "u": ("vec", np.float64),
    "v": ("vec", np.float64), "a": ("vec", np.float64), "f": ("vec", np.float64),
    "volume": ("scal", np.float64), "mass": ("scal", np.float64), "bulk": ("scal", np.float64),
    "shear": ("scal", np.float64), "cstab": ("scal", np.float64), "K": ("ten", np.float64),
    "K_inv": ("ten", np.float64), "F": ("ten", np.float64), "P": ("ten", np.float64),
    "damage": ("scal", np.float64), "fixed": ("scal", np.uint8), "fext": ("vec", np.float64),
    "ep": ("sym9", np.float64), "epbar": ("scal", np.float64), "back": ("sym9", np.float64),
    "nnb": ("scal", np.int32), "sing": ("scal", np.uint8), "drmass": ("scal", np.float64),
    "density": ("scal", np.float64),
}


class DRParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("use_graph", C.c_int32), ("chunk", C.c_int32),
                ("resume", C.c_int32), ("max_iterations", C.c_int64), ("tol", C.c_double),
                ("ref", C.c_double), ("dt", C.c_double), ("viscous_damping", C.c_double),
                ("cmax", C.c_double), ("refs", C.POINTER(C.c_double))]


class DRResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("reserved", C.c_int32), ("iterations", C.c_int64),
                ("residual", C.c_double), ("relative_residual", C.c_double),
                ("reference_residual", C.c_double), ("damping", C.c_double),
                ("resets", C.c_int64), ("elapsed_s", C.c_double)]


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); the exported surface of include/pdx.h
SIGNATURES = {
    "pdx_create": [C.c_int32, C.c_int32, C.POINTER(_vp)],
    "pdx_create_batch": [C.c_int32, _i64p, C.c_int32, C.POINTER(_vp)],
    "pdx_destroy": [_vp],
    "pdx_num_cases": [_vp, _i32p],
    "pdx_set_particles": [_vp, _dp, _dp, _dp],
    "pdx_set_materials": [_vp, C.c_int32, _dp, _dp, _dp, _dp, _dp],
    "pdx_set_stabilization": [_vp, C.c_double, _dp, C.c_double],
    "pdx_set_formulation": [_vp, C.c_int32],
    "pdx_set_bc": [_vp, _u8p, _dp],
    "pdx_count_neighbors": [_vp, C.c_double, _i32p],
    "pdx_build_bonds": [_vp, C.c_double, _i64p],
    "pdx_get_bonds": [_vp, _i64p, _i32p, _u8p],
    "pdx_get_bond_data": [_vp, _dp, _dp, _dp],
    "pdx_set_broken": [_vp, _u8p],
    "pdx_set_layout_builder": [_vp, C.c_int32],
    "pdx_get_halo_tables": [_vp, _i64p, _i32p, _i32p, _vp, _vp, _vp, _dp],
    "pdx_shape_tensor": [_vp, C.c_int32, C.c_double],
    "pdx_compute_F_stress": [_vp],
    "pdx_compute_forces": [_vp],
    "pdx_residual_norm": [_vp, _dp],
    "pdx_cold_residual": [_vp, _dp],
    "pdx_cold_residual_noalloc": [_vp, _dp],
    "pdx_break_bonds": [_vp, C.c_double, _i64p],
    "pdx_local_damage": [_vp],
    "pdx_commit_state": [_vp],
    "pdx_estimate_dt": [_vp, C.c_double, C.c_double, C.c_double, _dp],
    "pdx_adr_mass": [_vp, C.c_double],
    "pdx_kd_step": [_vp, C.c_double, C.c_double, _dp, _i32p, _dp, _dp, _i32p],
    "pdx_dr_solve": [_vp, C.POINTER(DRParams), C.POINTER(DRResult)],
    "pdx_explicit_run": [_vp, C.c_double, C.c_double, C.c_int64, C.c_double, _i64p, _dp],
    "pdx_time_kernels": [_vp, C.c_int32, _dp],
    "pdx_stats": [_vp, _i64p],
    "pdx_lattice_spacing": [_vp, _dp],
    "pdx_autotune": [_vp, C.c_int32, _i32p],
    "pdx_set_group": [_vp, C.c_int32],
    "pdx_set_variants": [_vp, C.c_int32, C.c_int32],
    "pdx_get_variants": [_vp, _i32p, _i32p],
    "pdx_neighbor_mean": [_vp, C.c_int32, _dp, _dp],
    "pdx_adr_mass_auto": [_vp, C.c_int32, C.c_double, _dp],
    "pdx_neighbor_features": [_vp, C.c_int32, C.c_int32, _dp, _vp, C.c_int32],
    "pdx_neighbor_features_device": [_vp, C.c_int32, C.c_int32, _vp, _vp, C.c_int32],
    "pdx_warm_raw_features": [_vp, _dp, _dp, _dp, _u8p, _i32p, C.c_double, _vp],
    "pdx_nn_linear_fwd": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp],
    "pdx_nn_linear_dgrad": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp],
    "pdx_nn_linear_wgrad": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp],
    "pdx_nn_mse": [_vp, C.c_int, _vp, _vp, C.c_int, C.c_float, _vp, _vp, _vp],
    "pdx_nn_adam": [_vp, _vp, _vp, _vp, _vp, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int64,
                    C.c_float, _vp],
    "pdx_nn_to_bf16": [_vp, C.c_int64, _vp, _vp],
    "pdx_get": [_vp, C.c_int32, _vp],
    "pdx_get_u_device": [_vp, _vp],
    "pdx_set_u_device": [_vp, _vp],
    "pdx_set": [_vp, C.c_int32, _vp],
}

_lib = None


def load():
    """Loads libpdx.so; raises ImportError when the in-tree build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libpdx.so not found at {LIB_PATH}; build it with `python -m simul_amd._build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    lib.pdx_last_error.restype = C.c_char_p
    lib.pdx_last_error.argtypes = []
    _lib = lib
    return lib


def check(status: int, what: str = ""):
    if status == OK:
        return
    msg = (_lib.pdx_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == EINVAL:
        raise FEAValidationError(text)
    if status == EUNSUP:
        raise NotImplementedError(text)
    if status == ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)


def ptr(a, t=None):
    if a is None:
        return None
    return a.ctypes.data_as(t) if t is not None else C.c_void_p(a.ctypes.data)


class Context:
    """Owns one pdx_ctx (all device memory of a particle system)."""

    def __init__(self, n: int, dim: int, case_offsets=None):
        self.lib = load()
        self.n, self.dim = int(n), int(dim)
        h = C.c_void_p()
        if case_offsets is None:
            check(self.lib.pdx_create(self.n, self.dim, C.byref(h)), "pdx_create")
            self.ncase = 1
            self.case_offsets = np.array([0, self.n], np.int64)
        else:
            off = np.ascontiguousarray(case_offsets, np.int64)
            check(self.lib.pdx_create_batch(len(off) - 1, ptr(off, _i64p), self.dim, C.byref(h)),
                  "pdx_create_batch")
            self.ncase = len(off) - 1
            self.case_offsets = off
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.pdx_destroy(h)
            self.h = None

    def call(self, name, *args):
        check(getattr(self.lib, name)(self.h, *args), name)

    def shape_of(self, field):
        kind, dt = FIELD_KIND[field]
        n, d = self.n, self.dim
        return {"vec": (n, d), "scal": (n,), "ten": (n, d, d), "sym9": (n, 3, 3)}[kind], dt

    def get(self, field) -> np.ndarray:
        shape, dt = self.shape_of(field)
        out = np.zeros(shape, dt)
        self.call("pdx_get", FIELD[field], ptr(out))
        return out

    def set(self, field, value):
        shape, dt = self.shape_of(field)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dt), shape))
        self.call("pdx_set", FIELD[field], ptr(a))
