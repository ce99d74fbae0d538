"""ctypes front end of the FP64 CPU oracle (oracle/pd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(simul_amd) never imports this module.

`Problem` owns every array the C oracle touches, in caller particle order,
with bonds as CSR rows sorted by ascending j (the neighbour *set* semantics of
reference peridynamics/core/neighbor.py:154-206).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class _OrcPD(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("dim", C.c_int32),
        ("model", C.c_int32), ("formulation", C.c_int32),
        ("per_bond_recompute", C.c_int32), ("nthreads", C.c_int32),
        ("horizon", C.c_double), ("G_s", C.c_double),
        ("X", _dp), ("x", _dp), ("vol", _dp), ("mass", _dp),
        ("bulk", _dp), ("shear", _dp), ("cstab", _dp),
        ("sy0", _dp), ("h_iso", _dp), ("h_kin", _dp),
        ("ep", _dp), ("epbar", _dp), ("beta", _dp),
        ("ep_t", _dp), ("epbar_t", _dp), ("beta_t", _dp),
        ("fixed", _u8p), ("fext", _dp),
        ("row", _i64p), ("col", _i32p), ("broken", _u8p),
        ("xi", _dp), ("xlen", _dp), ("omega", _dp),
        ("K", _dp), ("Kinv", _dp), ("sing", _u8p),
        ("F", _dp), ("P", _dp), ("A", _dp),
        ("f", _dp), ("v", _dp), ("a", _dp),
    ]


def build() -> str:
    """Compiles liboracle.so with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.POINTER(_OrcPD)
        L.orc_neighbor_count.restype = C.c_int64
        L.orc_neighbor_count.argtypes = [C.c_int32, C.c_int32, _dp, C.c_double, _i32p, C.c_int32]
        L.orc_neighbor_fill.restype = None
        L.orc_neighbor_fill.argtypes = [C.c_int32, C.c_int32, _dp, C.c_double, _i64p, _i32p, C.c_int32]
        L.orc_neighbor_count_brute.restype = C.c_int64
        L.orc_neighbor_count_brute.argtypes = [C.c_int32, C.c_int32, _dp, C.c_double, _i32p]
        for name in ("orc_bond_data", "orc_deformation_gradient", "orc_stress", "orc_commit",
                     "orc_force", "orc_compute_forces"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [P]
        L.orc_shape_tensor.restype = None
        L.orc_shape_tensor.argtypes = [P, C.c_int32, C.c_double]
        L.orc_residual_norm.restype = C.c_double
        L.orc_residual_norm.argtypes = [P]
        L.orc_estimate_dt.restype = C.c_double
        L.orc_estimate_dt.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.orc_kd_solve.restype = C.c_int64
        L.orc_kd_solve.argtypes = [P, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_double, _dp]
        L.orc_adr_mass.restype = None
        L.orc_adr_mass.argtypes = [P, C.c_double, _dp]
        L.orc_adr_solve_mode.restype = C.c_int64
        L.orc_adr_solve_mode.argtypes = [P, _dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int32, _dp]
        L.orc_adr_solve.restype = C.c_int64
        L.orc_adr_solve.argtypes = [P, _dp, C.c_int64, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_explicit_run.restype = C.c_int64
        L.orc_explicit_run.argtypes = [P, C.c_double, C.c_double, C.c_int64, C.c_double, _dp]
        L.orc_update_damage.restype = C.c_int64
        L.orc_update_damage.argtypes = [P, C.c_double]
        L.orc_local_damage.restype = None
        L.orc_local_damage.argtypes = [P, _i32p, _dp]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def neighbors(X: np.ndarray, horizon: float, nthreads: int | None = None):
    """CSR neighbour sets (row, col), rows sorted by ascending j."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, dim = X.shape
    counts = np.zeros(n, dtype=np.int32)
    nt = nthreads or default_threads()
    total = lib().orc_neighbor_count(n, dim, _ptr(X, _dp), horizon, _ptr(counts, _i32p), nt)
    row = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row[1:])
    assert row[-1] == total
    col = np.zeros(max(int(total), 1), dtype=np.int32)
    lib().orc_neighbor_fill(n, dim, _ptr(X, _dp), horizon, _ptr(row, _i64p), _ptr(col, _i32p), nt)
    return row, col[:total]


def neighbor_counts_brute(X: np.ndarray, horizon: float) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, dim = X.shape
    counts = np.zeros(n, dtype=np.int32)
    lib().orc_neighbor_count_brute(n, dim, _ptr(X, _dp), horizon, _ptr(counts, _i32p))
    return counts


MODELS = {"linear_elastic": 0, "neo_hookean": 1, "j2_plasticity": 2}


class Problem:
    """One peridynamic body for the oracle.  All arrays are FP64, caller order."""

    def __init__(self, X, vol, horizon, bulk, shear, *, mass=None, density=1.0,
                 model="linear_elastic", formulation="nosb", G_s=0.1, cstab=None,
                 fixed=None, fext=None, sy0=None, h_iso=None, h_kin=None,
                 nthreads=None, per_bond_recompute=False, row=None, col=None):
        X = np.ascontiguousarray(X, dtype=np.float64)
        n, dim = X.shape
        self.n, self.dim = n, dim
        full = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (n,)))
        self.X = X
        self.x = X.copy()
        self.vol = full(vol)
        self.mass = full(mass) if mass is not None else self.vol * density
        self.bulk, self.shear = full(bulk), full(shear)
        if cstab is None:
            cstab = 18.0 * self.bulk / (np.pi * horizon ** 4) if dim == 3 else \
                9.0 * (4.0 * self.bulk * self.shear / (self.bulk + self.shear)) / (np.pi * horizon ** 3)
        self.cstab = full(cstab)
        self.model = MODELS[model] if isinstance(model, str) else int(model)
        self.formulation = {"nosb": 0, "bond": 1}[formulation] if isinstance(formulation, str) else int(formulation)
        self.horizon, self.G_s = float(horizon), float(G_s)
        self.nthreads = nthreads or default_threads()
        self.per_bond_recompute = int(bool(per_bond_recompute))
        self.fixed = np.zeros(n, np.uint8) if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        self.fext = None if fext is None else np.ascontiguousarray(fext, np.float64).reshape(n, dim)
        if self.model == 2:
            self.sy0, self.h_iso = full(sy0), full(h_iso)
            self.h_kin = full(0.0 if h_kin is None else h_kin)
        else:
            self.sy0 = self.h_iso = self.h_kin = None
        z9 = lambda: np.zeros((n, 9))
        self.ep, self.epbar, self.beta = z9(), np.zeros(n), z9()
        self.ep_t, self.epbar_t, self.beta_t = z9(), np.zeros(n), z9()
        if row is None:
            row, col = neighbors(X, horizon, self.nthreads)
        self.row, self.col = np.ascontiguousarray(row, np.int64), np.ascontiguousarray(col, np.int32)
        nb = len(self.col)
        self.nb = nb
        self.broken = np.zeros(max(nb, 1), np.uint8)
        self.xi = np.zeros((max(nb, 1), dim))
        self.xlen = np.zeros(max(nb, 1))
        self.omega = np.zeros(max(nb, 1))
        dd = dim * dim
        self.K, self.Kinv = np.zeros((n, dim, dim)), np.zeros((n, dim, dim))
        self.sing = np.zeros(n, np.uint8)
        self.F, self.P, self.A = np.zeros((n, dim, dim)), np.zeros((n, dim, dim)), np.zeros((n, dim, dim))
        self.f, self.v, self.a = np.zeros((n, dim)), np.zeros((n, dim)), np.zeros((n, dim))
        self._s = _OrcPD()
        self._bind()
        lib().orc_bond_data(C.byref(self._s))
        del dd

    def _bind(self):
        s = self._s
        s.n, s.dim, s.model, s.formulation = self.n, self.dim, self.model, self.formulation
        s.per_bond_recompute, s.nthreads = self.per_bond_recompute, self.nthreads
        s.horizon, s.G_s = self.horizon, self.G_s
        for name in ("X", "x", "vol", "mass", "bulk", "shear", "cstab", "sy0", "h_iso", "h_kin",
                     "ep", "epbar", "beta", "ep_t", "epbar_t", "beta_t", "fext", "xi", "xlen",
                     "omega", "K", "Kinv", "F", "P", "A", "f", "v", "a"):
            setattr(s, name, _ptr(getattr(self, name), _dp))
        s.fixed, s.broken, s.sing = _ptr(self.fixed, _u8p), _ptr(self.broken, _u8p), _ptr(self.sing, _u8p)
        s.row, s.col = _ptr(self.row, _i64p), _ptr(self.col, _i32p)

    # -- kernels -------------------------------------------------------
    def shape_tensor(self, mode: str = "relative", tol: float = 1e-12):
        lib().orc_shape_tensor(C.byref(self._s), {"absolute": 0, "relative": 1}[mode], tol)

    def deformation_gradient(self):
        lib().orc_deformation_gradient(C.byref(self._s))

    def stress(self):
        lib().orc_stress(C.byref(self._s))

    def force(self):
        lib().orc_force(C.byref(self._s))

    def compute_forces(self):
        lib().orc_compute_forces(C.byref(self._s))

    def commit(self):
        lib().orc_commit(C.byref(self._s))

    def residual_norm(self) -> float:
        return lib().orc_residual_norm(C.byref(self._s))

    def estimate_dt(self, bulk, shear, rho_mean) -> float:
        return lib().orc_estimate_dt(C.byref(self._s), bulk, shear, rho_mean)

    def kd_solve(self, dt, visc=0.0, max_it=100000, tol=1e-6, ref=0.0):
        out = np.zeros(4)
        it = lib().orc_kd_solve(C.byref(self._s), dt, visc, int(max_it), tol, ref, _ptr(out, _dp))
        return {"converged": it > 0, "iterations": abs(int(it)), "residual": out[0],
                "relative_residual": out[1], "reference_residual": out[2], "resets": int(out[3])}

    def adr_mass(self, beta: float = 1.0) -> np.ndarray:
        lam = np.zeros(self.n)
        lib().orc_adr_mass(C.byref(self._s), beta, _ptr(lam, _dp))
        return lam

    def adr_solve(self, lam, max_it=100000, tol=1e-6, ref=0.0, cmax=1.9, damping="velocity"):
        lam = np.ascontiguousarray(lam, np.float64)
        out = np.zeros(4)
        mode = {"local": 0, "velocity": 1}[damping]
        it = lib().orc_adr_solve_mode(C.byref(self._s), _ptr(lam, _dp), int(max_it), tol, ref, cmax, mode,
                                      _ptr(out, _dp))
        return {"converged": it >= 0 and out[1] < tol, "iterations": abs(int(it)), "residual": out[0],
                "relative_residual": out[1], "reference_residual": out[2], "damping": out[3]}

    def explicit_run(self, dt, damping=0.0, n_steps=1, critical_stretch=0.0):
        """Velocity-Verlet explicit dynamics (+ per-step bond breaking when
        critical_stretch > 0) from the current x, v, a; returns
        (bonds broken, kinetic energy after the last step)."""
        out = np.zeros(1)
        nb = lib().orc_explicit_run(C.byref(self._s), dt, damping, int(n_steps), critical_stretch,
                                    _ptr(out, _dp))
        return int(nb), float(out[0])

    def update_damage(self, s_c: float) -> int:
        return int(lib().orc_update_damage(C.byref(self._s), s_c))

    def local_damage(self) -> np.ndarray:
        initial = np.diff(self.row).astype(np.int32)
        dmg = np.zeros(self.n)
        lib().orc_local_damage(C.byref(self._s), _ptr(initial, _i32p), _ptr(dmg, _dp))
        return dmg

    # -- helpers -------------------------------------------------------
    @property
    def u(self) -> np.ndarray:
        return self.x - self.X

    def set_u(self, u):
        self.x[...] = self.X + np.asarray(u, np.float64).reshape(self.n, self.dim)

    def bond_pairs(self):
        """(i, j) for every bond in CSR order."""
        i = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.row))
        return i, self.col[: self.nb].astype(np.int64)

    def free_mask(self) -> np.ndarray:
        bits = (self.fixed[:, None] >> np.arange(self.dim)[None, :]) & 1
        return bits == 0
