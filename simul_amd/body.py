"""PDBody: one peridynamic body (or a batch of independent bodies) on the GPU.

Thin host-side object over the libpdx C ABI (include/pdx.h).  It is the common
core under the reference-API mirror (simul_amd.peridynamics, simul_amd.framework),
the tests and the benchmark.  Arrays are numpy, caller particle order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

MODELS = {"linear_elastic": N.MODEL_LINEAR_ELASTIC, "neo_hookean": N.MODEL_NEO_HOOKEAN,
          "j2_plasticity": N.MODEL_J2}
FORMS = {"nosb": N.FORM_NOSB, "bond": N.FORM_BOND}


@dataclass
class DRResult:
    converged: bool
    iterations: int
    residual: float
    relative_residual: float
    reference_residual: float
    damping: float
    resets: int
    elapsed_s: float


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.asarray(a, np.float64)
    if shape is not None:
        a = np.broadcast_to(a, shape)
    return np.ascontiguousarray(a)



class _Diverged(Exception):
    """A solve with the spectral mass left the stable range."""

class PDBody:
    def __init__(self, X, vol, horizon, bulk, shear, *, mass=None, model="linear_elastic",
                 formulation="nosb", G_s=0.1, cstab=None, fixed=None, fext=None, u0=None,
                 yield_stress=None, h_iso=None, h_kin=None, singular_test="relative",
                 singular_tol=1e-12, case_offsets=None, build=True):
        X = np.ascontiguousarray(X, np.float64)
        n, dim = X.shape
        self.n, self.dim = n, dim
        self.ctx = N.Context(n, dim, case_offsets)
        self.ncase = self.ctx.ncase
        self.horizon = float(horizon)
        vol = _f64(vol, (n,))
        mass = _f64(mass, (n,)) if mass is not None else None
        self.ctx.call("pdx_set_particles", N.ptr(X, N._dp), N.ptr(vol, N._dp), N.ptr(mass, N._dp))
        self.set_materials(model, bulk, shear, yield_stress, h_iso, h_kin)
        if cstab is None:
            # NOSBMaterial.set_stabilization_modulus (nosb.py:335-346): 18 K / (pi d^4)
            # in 3-D, 9 E / (pi t d^3) in 2-D (t = 1), with the plane-stress
            # E = 4 K mu / (K + mu) recovered from K = E / (2 (1 - nu)), mu = E / (2 (1 + nu))
            b = _f64(bulk, (n,))
            if dim == 3:
                cstab = 18.0 * b / (np.pi * self.horizon ** 4)
            else:
                s = _f64(shear, (n,))
                cstab = 9.0 * (4.0 * b * s / (b + s)) / (np.pi * self.horizon ** 3)
        self.set_stabilization(G_s, cstab)
        self.ctx.call("pdx_set_formulation", FORMS[formulation])
        self.formulation = formulation
        self.set_bc(fixed, fext)
        self.singular_test, self.singular_tol = singular_test, singular_tol
        self.n_bonds = 0
        if build:
            self.build_bonds(self.horizon)
            if formulation == "nosb":
                self.shape_tensor(singular_test, singular_tol)
        if u0 is not None:
            self.u = u0

    # -- inputs ----------------------------------------------------------
    def set_materials(self, model, bulk, shear, yield_stress=None, h_iso=None, h_kin=None):
        n = self.n
        self.model = model
        m = MODELS[model] if isinstance(model, str) else int(model)
        self.ctx.call("pdx_set_materials", m, N.ptr(_f64(bulk, (n,)), N._dp), N.ptr(_f64(shear, (n,)), N._dp),
                      N.ptr(_f64(yield_stress, (n,)), N._dp), N.ptr(_f64(h_iso, (n,)), N._dp),
                      N.ptr(_f64(h_kin, (n,)), N._dp))

    def set_stabilization(self, G_s, cstab):
        c = _f64(cstab, (self.n,))
        self.G_s = float(G_s)
        self.ctx.call("pdx_set_stabilization", float(G_s), N.ptr(c, N._dp), 0.0)

    def set_bc(self, fixed=None, fext=None):
        fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        fe = None if fext is None else _f64(fext, (self.n, self.dim))
        self.ctx.call("pdx_set_bc", N.ptr(fx, N._u8p), N.ptr(fe, N._dp))

    def set_formulation(self, formulation):
        self.ctx.call("pdx_set_formulation", FORMS[formulation])
        self.formulation = formulation

    def build_bonds(self, horizon):
        nb = C.c_int64()
        self.ctx.call("pdx_build_bonds", float(horizon), C.byref(nb))
        self.horizon = float(horizon)
        self.n_bonds = nb.value
        return nb.value

    def count_neighbors(self, horizon):
        out = np.zeros(self.n, np.int32)
        self.ctx.call("pdx_count_neighbors", float(horizon), N.ptr(out, N._i32p))
        return out

    def shape_tensor(self, test="relative", tol=1e-12):
        self.ctx.call("pdx_shape_tensor", {"absolute": N.SING_ABSOLUTE, "relative": N.SING_RELATIVE}[test], tol)

    # -- state -------------------------------------------------------------
    def get(self, name):
        return self.ctx.get(name)

    def set(self, name, value):
        self.ctx.set(name, value)

    @property
    def u(self):
        return self.ctx.get("u")

    @u.setter
    def u(self, value):
        self.ctx.set("u", value)

    def get_u_device(self):
        """The displacement as a (n, dim) float64 torch tensor on the GPU
        (pdx_get_u_device: no host round trip)."""
        import torch
        t = torch.empty((self.n, self.dim), dtype=torch.float64, device="cuda")
        torch.cuda.current_stream().synchronize()
        self.ctx.call("pdx_get_u_device", C.c_void_p(t.data_ptr()))
        return t

    def set_u_device(self, t):
        """Sets the displacement from a (n, dim) float64 GPU tensor
        (pdx_set_u_device)."""
        import torch
        if not (t.is_cuda and t.dtype == torch.float64 and tuple(t.shape) == (self.n, self.dim)):
            raise ValueError(f"expected a ({self.n}, {self.dim}) float64 CUDA tensor")
        t = t.contiguous()
        torch.cuda.current_stream().synchronize()
        self.ctx.call("pdx_set_u_device", C.c_void_p(t.data_ptr()))

    @property
    def x(self):
        return self.ctx.get("x")

    @x.setter
    def x(self, value):
        self.ctx.set("x", value)

    def bonds(self):
        """(row, col, broken) in caller order, rows sorted by j."""
        row = np.zeros(self.n + 1, np.int64)
        nb = self.n_bonds
        col = np.zeros(max(nb, 1), np.int32)
        br = np.zeros(max(nb, 1), np.uint8)
        self.ctx.call("pdx_get_bonds", N.ptr(row, N._i64p), N.ptr(col, N._i32p), N.ptr(br, N._u8p))
        return row, col[:nb], br[:nb]

    def bond_data(self):
        nb = self.n_bonds
        xi = np.zeros((max(nb, 1), self.dim))
        L = np.zeros(max(nb, 1))
        w = np.zeros(max(nb, 1))
        self.ctx.call("pdx_get_bond_data", N.ptr(xi, N._dp), N.ptr(L, N._dp), N.ptr(w, N._dp))
        return xi[:nb], L[:nb], w[:nb]

    def set_broken(self, broken):
        b = np.ascontiguousarray(broken, np.uint8)
        self.ctx.call("pdx_set_broken", N.ptr(b, N._u8p))

    # -- kernels ------------------------------------------------------------
    def compute_F_stress(self):
        self.ctx.call("pdx_compute_F_stress")

    def compute_forces(self, download=True):
        """Force density of the current state (host copy unless download is
        False; residual_norm() reads the per-case norms either way)."""
        self.ctx.call("pdx_compute_forces")
        return self.ctx.get("f") if download else None

    def residual_norm(self):
        out = np.zeros(self.ncase)
        self.ctx.call("pdx_residual_norm", N.ptr(out, N._dp))
        return out if self.ncase > 1 else float(out[0])

    def cold_residual(self):
        out = np.zeros(self.ncase)
        self.ctx.call("pdx_cold_residual", N.ptr(out, N._dp))
        return out if self.ncase > 1 else float(out[0])

    def break_bonds(self, s_c):
        n = C.c_int64()
        self.ctx.call("pdx_break_bonds", float(s_c), C.byref(n))
        return n.value

    def neighbor_mean(self, values):
        """Weighted mean of a per-particle vector (n x width, width <= 64) over
        each particle's intact bonds (pdx_neighbor_mean)."""
        v = np.ascontiguousarray(values, np.float64)
        w = 1 if v.ndim == 1 else v.shape[1]
        out = np.zeros_like(v)
        self.ctx.call("pdx_neighbor_mean", int(w), N.ptr(v, N._dp), N.ptr(out, N._dp))
        return out

    def local_damage(self):
        self.ctx.call("pdx_local_damage")
        return self.ctx.get("damage")

    def commit_state(self):
        self.ctx.call("pdx_commit_state")

    def estimate_dt(self, bulk, shear, rho_mean):
        dt = C.c_double()
        self.ctx.call("pdx_estimate_dt", float(bulk), float(shear), float(rho_mean), C.byref(dt))
        return dt.value

    def adr_mass(self, beta=1.0):
        self.ctx.call("pdx_adr_mass", float(beta))

    def tune_mass(self, iters=20, safety=1.2):
        """Adaptive-DR mass scaled to the spectrum of each case
        (pdx_adr_mass_auto); returns beta per case (a float for one case).
        safety 1.2: a 1.2 margin on the estimated stability limit (DR
        converged with every factor down to 1.1 on the dx 0.4 vertebral body,
        the 64-case batch and 400 cases of the warm start's families, and
        diverges on the vertebra near 0.9, beta 0.07); solve(mass="spectral")
        falls back to the Gershgorin mass should a case diverge all the same."""
        b = np.zeros(self.ncase)
        self.ctx.call("pdx_adr_mass_auto", int(iters), float(safety), N.ptr(b, N._dp))
        return b if self.ncase > 1 else float(b[0])

    def kd_step(self, dt, visc, prev_ke, ke_increasing):
        pk = np.array(np.broadcast_to(prev_ke, (self.ncase,)), np.float64)
        inc = np.array(np.broadcast_to(ke_increasing, (self.ncase,)), np.int32)
        ke = np.zeros(self.ncase)
        res = np.zeros(self.ncase)
        rs = np.zeros(self.ncase, np.int32)
        self.ctx.call("pdx_kd_step", float(dt), float(visc), N.ptr(pk, N._dp), N.ptr(inc, N._i32p),
                      N.ptr(ke, N._dp), N.ptr(res, N._dp), N.ptr(rs, N._i32p))
        return {"kinetic_energy": ke, "residual": res, "velocity_reset": rs.astype(bool),
                "prev_ke": pk, "ke_increasing": inc.astype(bool)}

    def explicit_run(self, dt, damping=0.0, n_steps=1, critical_stretch=0.0):
        """Velocity-Verlet explicit dynamics from the current u, v and a
        (pdx_explicit_run), with bonds breaking after every step when
        critical_stretch > 0.  Returns (bonds broken during the run, kinetic
        energy per case after the last step; a float for one case)."""
        nb = C.c_int64()
        ke = np.zeros(self.ncase)
        self.ctx.call("pdx_explicit_run", float(dt), float(damping), int(n_steps), float(critical_stretch),
                      C.byref(nb), N.ptr(ke, N._dp))
        return nb.value, (ke if self.ncase > 1 else float(ke[0]))

    def solve(self, method="adaptive", tol=1e-6, max_iterations=100000, ref=0.0, dt=0.0,
              viscous_damping=0.0, cmax=1.9, use_graph=True, chunk=64, resume=False, refs=None,
              progress=None, slice_iterations=None, mass=None):
        """Dynamic relaxation to |f_free| / ref < tol per case.

        ref > 0 fixes the reference residual of every case, refs (ncase) of
        each case; otherwise each case's residual at the start state is used.
        With ``slice_iterations`` the iterations run in slices of that many
        (resumed on the device, so the trajectory is the same) and
        ``progress(done, result)`` is called after each slice with the slice's
        result (a list for a batch); the returned iteration counts are totals
        over the slices.

        mass (adaptive DR): None keeps the context's mass (the Gershgorin mass
        unless set otherwise); "gershgorin" recomputes it; "spectral" scales it
        to each case's spectrum first (tune_mass, its beta kept as
        ``last_beta``) and, should a case diverge (relative residual above 1 or
        not finite at the end), solves again from the same start with the
        Gershgorin mass; a number beta: the Gershgorin mass times beta (a beta
        an earlier tuning of the same body found, reused without tuning),
        and, should a case diverge, solves again from the same start with
        mass="spectral"."""
        if mass is not None and method == "adaptive" and not resume:
            reuse = not isinstance(mass, str)
            if not reuse and mass == "gershgorin":
                self.adr_mass(1.0)
            elif reuse or mass == "spectral":
                u_start = self.get_u_device()
                if reuse:
                    if not (float(mass) > 0.0):
                        raise ValueError("a reused mass factor beta must be positive")
                    self.adr_mass(float(mass))
                else:
                    self.last_beta = self.tune_mass()

                def watch(done, rs):  # a diverging case ends the attempt at the slice
                    for x in (rs if isinstance(rs, list) else [rs]):
                        if not x.converged and not (np.isfinite(x.relative_residual)
                                                    and x.relative_residual <= 100.0):
                            raise _Diverged()
                    if progress is not None:
                        progress(done, rs)

                try:
                    out = self.solve(method, tol, max_iterations, ref, dt, viscous_damping, cmax, use_graph,
                                     chunk, False, refs, watch, slice_iterations or 2048)
                    lst = out if isinstance(out, list) else [out]
                    if all(x.converged or (np.isfinite(x.relative_residual) and x.relative_residual <= 1.0)
                           for x in lst):
                        return out
                except _Diverged:
                    pass
                self.set_u_device(u_start)
                if reuse:
                    return self.solve(method, tol, max_iterations, ref, dt, viscous_damping, cmax, use_graph,
                                      chunk, False, refs, progress, slice_iterations, mass="spectral")
                self.adr_mass(1.0)
                return self.solve(method, tol, max_iterations, ref, dt, viscous_damping, cmax, use_graph, chunk,
                                  False, refs, progress, slice_iterations)
            else:
                raise ValueError(f"mass must be None, 'gershgorin', 'spectral' or a beta, not {mass!r}")
        if slice_iterations is None:
            return self._solve_once(method, tol, max_iterations, ref, dt, viscous_damping, cmax, use_graph,
                                    chunk, resume, refs)
        if refs is None:
            if ref > 0.0:
                refs = np.full(self.ncase, float(ref))
            else:
                self.compute_forces(download=False)
                refs = np.atleast_1d(self.residual_norm()).astype(np.float64)
                refs = np.where(refs < 1e-20, 1.0, refs)
        refs = np.ascontiguousarray(np.broadcast_to(np.asarray(refs, np.float64), (self.ncase,)))
        total = np.zeros(self.ncase, np.int64)
        done = np.zeros(self.ncase, bool)
        last = None
        launched = 0
        sl = max(2 * int(chunk), int(slice_iterations) // (2 * int(chunk)) * (2 * int(chunk)))
        while launched < max_iterations and not done.all():
            n = int(min(sl, max_iterations - launched))
            out = self._solve_once(method, tol, n, 0.0, dt, viscous_damping, cmax, use_graph, chunk,
                                   resume or launched > 0, refs)
            out = out if isinstance(out, list) else [out]
            for k, r in enumerate(out):
                if not done[k]:
                    total[k] += r.iterations
                    done[k] = r.converged
            launched += n
            last = out
            if progress is not None:
                progress(launched, last if self.ncase > 1 else last[0])
        res = [DRResult(bool(done[k]), int(total[k]), r.residual, r.relative_residual, r.reference_residual,
                        r.damping, r.resets, r.elapsed_s) for k, r in enumerate(last)]
        return res if self.ncase > 1 else res[0]

    def _solve_once(self, method, tol, max_iterations, ref, dt, viscous_damping, cmax, use_graph, chunk,
                    resume, refs):
        p = N.DRParams()
        p.method = {"adaptive": N.DR_ADAPTIVE, "kinetic": N.DR_KINETIC}[method]
        p.use_graph, p.chunk, p.resume = int(bool(use_graph)), int(chunk), int(bool(resume))
        p.max_iterations, p.tol, p.ref = int(max_iterations), float(tol), float(ref)
        p.dt, p.viscous_damping, p.cmax = float(dt), float(viscous_damping), float(cmax)
        if refs is not None:
            refs = np.ascontiguousarray(np.broadcast_to(np.asarray(refs, np.float64), (self.ncase,)))
            p.refs = N.ptr(refs, N._dp)
        res = (N.DRResult * self.ncase)()
        self.ctx.call("pdx_dr_solve", C.byref(p), res)
        out = [DRResult(bool(r.converged), int(r.iterations), r.residual, r.relative_residual,
                        r.reference_residual, r.damping, int(r.resets), r.elapsed_s) for r in res]
        return out if self.ncase > 1 else out[0]

    def time_kernels(self, iters=20):
        ms = np.zeros(3)
        self.ctx.call("pdx_time_kernels", int(iters), N.ptr(ms, N._dp))
        return {"fpass_ms": ms[0], "force_ms": ms[1], "total_ms": ms[2]}

    def autotune(self, iters=10):
        g = C.c_int32()
        self.ctx.call("pdx_autotune", int(iters), C.byref(g))
        return g.value

    def set_group(self, g):
        self.ctx.call("pdx_set_group", int(g))

    def set_variants(self, fpass, force):
        """Hot-pass variants of the F-pass and the force pass (pdx_set_variants)."""
        self.ctx.call("pdx_set_variants", int(fpass), int(force))

    def variants(self):
        """(F-pass variant, force-pass variant) in use."""
        f, g = C.c_int32(), C.c_int32()
        self.ctx.call("pdx_get_variants", C.byref(f), C.byref(g))
        return f.value, g.value

    def lattice_spacing(self):
        """Lattice spacing h of the PDX_VARIANT_LATTICE kernels, or 0.0 when the
        context does not qualify for them (pdx_lattice_spacing)."""
        h = C.c_double()
        self.ctx.call("pdx_lattice_spacing", C.byref(h))
        return h.value

    def stats(self):
        o = np.zeros(4, np.int64)
        self.ctx.call("pdx_stats", N.ptr(o, N._i64p))
        return {"n_particles": int(o[0]), "n_bonds": int(o[1]), "n_intact": int(o[2]), "sell_slots": int(o[3])}


def from_specimen(spec, *, model=None, G_s=0.1, case_offsets=None, **kw) -> PDBody:
    bulk, shear = spec.bulk, spec.shear
    cst = 18.0 * bulk / (np.pi * spec.horizon ** 4)
    mdl = model or spec.model
    extra = {}
    if mdl == "j2_plasticity":
        extra = dict(yield_stress=spec.extra["yield"], h_iso=spec.extra["h_iso"],
                     h_kin=spec.extra.get("h_kin"))
    return PDBody(spec.X, spec.vol, spec.horizon, bulk, shear, mass=spec.vol * spec.rho, model=mdl,
                  G_s=G_s, cstab=cst, fixed=spec.fixed, fext=spec.fext, u0=spec.u0,
                  case_offsets=case_offsets, **extra, **kw)
