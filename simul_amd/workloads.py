"""Synthetic specimens of the benchmark workloads (mm - N - MPa - t/mm^3).

* cube: the 20x20x20 linear-elastic cube at 1 mm spacing (8,000 particles),
  bottom layer fixed, top layer prescribed -0.5 % axial strain.
* vertebra: elliptic cylinder x^2/20^2 + y^2/15^2 <= 1, 0 <= z <= 25 mm,
  voxel centres at (i + 1/2) dx; cortical shell (distance to the outer surface
  <= 1.2 mm, E = 12 GPa) around a cancellous core with E drawn log-normal
  (median 300 MPa, sigma_log 0.4) from a seeded generator; nu = 0.3.
  Bottom endplate fixed, top endplate loaded by a uniform pressure applied as
  the force density p / dx on the top voxel layer.
* vertebra_j2: the same geometry with J2 plasticity under a prescribed 2 %
  endplate compression in 20 increments, bonds breaking at stretch 0.02.
* batch: several random box / cylinder / truncated-ellipsoid specimens at dx = 0.5 mm.
* random_vertebra / random_vertebra_j2 / vertebra_family: the learned warm
  start's families of vertebral bodies (random shapes, pressure-loaded
  Neo-Hookean; random shapes, J2 compressed between platens in one
  increment; the headline's shape with random cancellous fields and loads).

The lateral distance to the ellipse uses the first-order estimate
(1 - r) / |grad r| with r = sqrt(x^2/a^2 + y^2/b^2).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

HORIZON_FACTOR = 3.015
RHO_BONE = 1.9e-9  # t/mm^3


def bulk_shear(E, nu):
    E = np.asarray(E, np.float64)
    return E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def stab_modulus(bulk, horizon, dim=3):
    """c = 18 K / (pi delta^4) (reference nosb.py:343-346, 3-D)."""
    return 18.0 * np.asarray(bulk, np.float64) / (math.pi * horizon ** 4)


@dataclass
class Specimen:
    name: str
    X: np.ndarray               # (n, 3) positions [mm]
    dx: float
    horizon: float
    E: np.ndarray               # (n,) Young's modulus [MPa]
    nu: float
    rho: float
    fixed: np.ndarray           # (n,) uint8 DOF mask
    u0: np.ndarray              # (n, 3) prescribed displacement (fixed DOFs)
    fext: np.ndarray            # (n, 3) external force density
    model: str = "linear_elastic"
    material_id: np.ndarray = None  # 0 cortical / 1 cancellous
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.X)

    @property
    def vol(self):
        return np.full(self.n, self.dx ** 3)

    @property
    def bulk(self):
        return bulk_shear(self.E, self.nu)[0]

    @property
    def shear(self):
        return bulk_shear(self.E, self.nu)[1]


def cube(n_side=20, dx=1.0, E=15000.0, nu=0.3, strain=-0.005):
    """20^3 cube: z = 0 layer fixed, top layer fixed with u_z = strain * (n-1) dx."""
    g = np.arange(n_side, dtype=np.float64) * dx
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(X)
    fixed = np.zeros(n, np.uint8)
    bot = X[:, 2] < 0.5 * dx
    top = X[:, 2] > (n_side - 1.5) * dx
    fixed[bot | top] = 7
    u0 = np.zeros((n, 3))
    u0[top, 2] = strain * (n_side - 1) * dx
    return Specimen("cube", X, dx, HORIZON_FACTOR * dx, np.full(n, E), nu, RHO_BONE, fixed, u0,
                    np.zeros((n, 3)), "linear_elastic", np.zeros(n, np.int32))


def _ellipse_lateral_distance(x, y, a, b):
    r = np.sqrt((x / a) ** 2 + (y / b) ** 2)
    rs = np.maximum(r, 1e-12)
    g = np.sqrt((x / a ** 2) ** 2 + (y / b ** 2) ** 2) / rs
    return (1.0 - r) / np.maximum(g, 1e-12)


def vertebra(dx=0.4, a=20.0, b=15.0, height=25.0, shell=1.2, E_cort=12000.0, E_canc_median=300.0,
             sigma_log=0.4, nu=0.3, pressure=2.0, seed=2000, model="neo_hookean", direction=None):
    nx, ny, nz = int(round(2 * a / dx)), int(round(2 * b / dx)), int(math.floor(height / dx + 1e-9))
    xs = -a + (np.arange(nx) + 0.5) * dx
    ys = -b + (np.arange(ny) + 0.5) * dx
    zs = (np.arange(nz) + 0.5) * dx
    zs = zs[zs <= height]
    XX, YY = np.meshgrid(xs, ys, indexing="ij")
    inside = (XX / a) ** 2 + (YY / b) ** 2 <= 1.0
    px, py = XX[inside], YY[inside]
    nl = len(px)
    X = np.empty((nl * len(zs), 3))
    for k, z in enumerate(zs):
        X[k * nl:(k + 1) * nl, 0] = px
        X[k * nl:(k + 1) * nl, 1] = py
        X[k * nl:(k + 1) * nl, 2] = z
    n = len(X)
    d_lat = _ellipse_lateral_distance(X[:, 0], X[:, 1], a, b)
    d_end = np.minimum(X[:, 2], height - X[:, 2])
    cortical = np.minimum(d_lat, d_end) <= shell
    rng = np.random.default_rng(seed)
    E = np.where(cortical, E_cort, np.exp(np.log(E_canc_median) + sigma_log * rng.standard_normal(n)))
    fixed = np.zeros(n, np.uint8)
    bot = X[:, 2] < zs[0] + 0.5 * dx
    top = X[:, 2] > zs[-1] - 0.5 * dx
    fixed[bot] = 7
    fext = np.zeros((n, 3))
    d = np.array([0.0, 0.0, -1.0]) if direction is None else np.asarray(direction, np.float64)
    fext[top] = d / np.linalg.norm(d) * pressure / dx
    return Specimen(f"vertebra_dx{dx}", X, dx, HORIZON_FACTOR * dx, E, nu, RHO_BONE, fixed,
                    np.zeros((n, 3)), fext, model, np.where(cortical, 0, 1).astype(np.int32),
                    {"pressure": pressure, "seed": seed, "top": top, "bottom": bot})


def vertebra_j2(dx=0.25, yield_cort=80.0, yield_canc=4.0, hardening=0.05, strain=-0.02, n_increments=20,
                critical_stretch=0.02, seed=3000, **kw):
    """The vertebral body with J2 plasticity (linear isotropic hardening H =
    hardening x E, yield 80 MPa cortical / 4 MPa cancellous), both endplates
    held by rigid platens (all DOFs prescribed): bottom fixed, top moved to
    u_z = strain x height in ``n_increments`` load increments, bonds breaking
    at ``critical_stretch``."""
    spec = vertebra(dx=dx, pressure=0.0, seed=seed, model="j2_plasticity", **kw)
    top, bot = spec.extra["top"], spec.extra["bottom"]
    spec.fixed[top | bot] = 7
    height = kw.get("height", 25.0)
    cort = spec.material_id == 0
    spec.name = f"vertebra_j2_dx{dx}"
    spec.extra.update({"yield": np.where(cort, yield_cort, yield_canc), "h_iso": hardening * spec.E,
                       "top_idx": np.where(top)[0], "u_final": np.array([0.0, 0.0, strain * height]),
                       "n_increments": n_increments, "critical_stretch": critical_stretch})
    return spec


def random_specimen(seed, n_target, dx=0.5, shell=1.2, E_cort=12000.0, E_canc_median=300.0,
                    sigma_log=0.4, nu=0.3, p_range=(0.5, 3.0), cone_deg=20.0, model="neo_hookean"):
    """One batched load case: a box, cylinder or ellipsoid scaled to about n_target
    particles, two-phase material, pressure of random magnitude and direction
    within cone_deg of the axis applied on the top layer, bottom layer fixed."""
    rng = np.random.default_rng(seed)
    shape = ["box", "cylinder", "ellipsoid"][int(rng.integers(3))]
    asp = rng.uniform(0.7, 1.4, size=3)
    frac = {"box": 1.0, "cylinder": math.pi / 4, "ellipsoid": 0.8 * math.pi / 4 * (1 - 0.64 / 3)}[shape]
    s = (n_target * dx ** 3 / (frac * np.prod(asp))) ** (1.0 / 3.0)
    L = asp * s
    cnt = np.maximum(np.round(L / dx).astype(int), 3)
    g = [(np.arange(c) + 0.5) * dx - c * dx / 2 for c in cnt]
    X = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3)
    hx, hy, hz = cnt * dx / 2
    if shape == "cylinder":
        X = X[(X[:, 0] / hx) ** 2 + (X[:, 1] / hy) ** 2 <= 1.0]
    elif shape == "ellipsoid":
        # truncated at |z| <= 0.8 hz so that the specimen has flat endplates to
        # hold and to load (an ellipsoid supported and loaded at its tips has
        # near-rigid modes)
        X = X[((X[:, 0] / hx) ** 2 + (X[:, 1] / hy) ** 2 + (X[:, 2] / hz) ** 2 <= 1.0)
              & (np.abs(X[:, 2]) <= 0.8 * hz)]
    n = len(X)
    if shape == "box":
        d = np.min(np.stack([hx - np.abs(X[:, 0]), hy - np.abs(X[:, 1]), hz - np.abs(X[:, 2])]), 0)
    elif shape == "cylinder":
        d = np.minimum(_ellipse_lateral_distance(X[:, 0], X[:, 1], hx, hy), hz - np.abs(X[:, 2]))
    else:
        r = np.sqrt((X[:, 0] / hx) ** 2 + (X[:, 1] / hy) ** 2 + (X[:, 2] / hz) ** 2)
        d = np.minimum((1.0 - r) * min(hx, hy, hz), 0.8 * hz - np.abs(X[:, 2]))
    cortical = d <= shell
    E = np.where(cortical, E_cort, np.exp(np.log(E_canc_median) + sigma_log * rng.standard_normal(n)))
    z = X[:, 2]
    fixed = np.zeros(n, np.uint8)
    bot = z < z.min() + 0.5 * dx
    top = z > z.max() - 0.5 * dx
    fixed[bot] = 7
    p = rng.uniform(*p_range)
    th = math.radians(cone_deg) * math.sqrt(rng.uniform())
    ph = rng.uniform(0, 2 * math.pi)
    dirv = -np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    fext = np.zeros((n, 3))
    fext[top] = dirv * p / dx
    return Specimen(f"batch{seed}_{shape}", X, dx, HORIZON_FACTOR * dx, E, nu, RHO_BONE, fixed,
                    np.zeros((n, 3)), fext, model, np.where(cortical, 0, 1).astype(np.int32),
                    {"shape": shape, "pressure": p, "direction": dirv})


def batch(n_cases=64, seed=4000, n_lo=10_000, n_hi=200_000, dx=0.5, model="neo_hookean"):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_lo, n_hi + 1, size=n_cases)
    return [random_specimen(seed + 1 + k, int(sizes[k]), dx=dx, model=model) for k in range(n_cases)]


def batch_family(seed, dx=0.5, n_lo=10_000, n_hi=200_000):
    """One case of the 64-case batch's family: random_specimen with a size
    drawn uniformly from [n_lo, n_hi] particles (the batch itself is
    batch(seed=4000): cases 4001-4064)."""
    n = int(np.random.default_rng([seed, 7]).integers(n_lo, n_hi + 1))
    return random_specimen(seed, n, dx=dx)


def pack(specs):
    """Several specimens of one spacing packed caller-order into one batched
    input: the concatenated specimen and the case offsets (ncase + 1)."""
    dx = specs[0].dx
    if any(s.dx != dx for s in specs):
        raise ValueError("packed specimens must share the particle spacing")
    off = np.zeros(len(specs) + 1, np.int64)
    off[1:] = np.cumsum([s.n for s in specs])
    cat = lambda name: np.concatenate([getattr(s, name) for s in specs])
    extra = {"cases": [s.name for s in specs]}
    if specs[0].model == "j2_plasticity":
        # the per-particle J2 parameters travel with their particles
        for key in ("yield", "h_iso"):
            extra[key] = np.concatenate([np.broadcast_to(np.asarray(s.extra[key], np.float64), (s.n,))
                                         for s in specs])
    packed = Specimen(f"batch{len(specs)}_dx{dx}", cat("X"), dx, specs[0].horizon, cat("E"), specs[0].nu,
                      specs[0].rho, cat("fixed"), cat("u0"), cat("fext"), specs[0].model, cat("material_id"),
                      extra)
    return packed, off


def random_vertebra(seed, dx=0.8, p_range=(0.5, 3.0), cone_deg=20.0, model="neo_hookean"):
    """A vertebral body of random size (semi-axes 14-24 mm x 0.6-0.9 of that,
    height 18-30 mm), random cancellous field and a pressure of random magnitude
    and direction within cone_deg of the axis: the training / held-out family of
    the learned warm start."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(14.0, 24.0)
    b = a * rng.uniform(0.6, 0.9)
    h = rng.uniform(18.0, 30.0)
    p = rng.uniform(*p_range)
    th = math.radians(cone_deg) * math.sqrt(rng.uniform())
    ph = rng.uniform(0, 2 * math.pi)
    d = -np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    spec = vertebra(dx=dx, a=a, b=b, height=h, pressure=p, seed=int(rng.integers(1 << 31)), model=model,
                    direction=d)
    spec.name = f"rvert{seed}_dx{dx}"
    return spec


def vertebra_family(seed, dx=0.4, p_range=(0.5, 3.0), cone_deg=20.0, model="neo_hookean", fixed_load=False):
    """The headline vertebral body's geometry (20 x 15 x 25 mm) with a random
    cancellous field and a pressure of random magnitude and direction within
    cone_deg of the axis: the family of the headline's learned warm start
    (the headline itself is seed 2000 at 2 MPa along -z).  fixed_load: the
    headline's own load (2 MPa along -z), only the cancellous field random
    (a Monte Carlo study over the material)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(*p_range)
    th = math.radians(cone_deg) * math.sqrt(rng.uniform())
    ph = rng.uniform(0, 2 * math.pi)
    d = -np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    if fixed_load:
        p, d = 2.0, None
    spec = vertebra(dx=dx, pressure=p, seed=int(rng.integers(1 << 31)), model=model, direction=d)
    spec.name = f"vfam{seed}_dx{dx}"
    return spec


def random_vertebra_j2(seed, dx=0.8, strain_range=(1e-3, 2e-2), **kw):
    """A vertebral body of random size (as random_vertebra) with the J2
    material and the platens of vertebra_j2, compressed in one increment to a
    nominal strain drawn log-uniform from ``strain_range``: u0 holds the
    prescribed top displacement (0, 0, -strain x height).  The J2 family of
    the learned warm start."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(14.0, 24.0)
    b = a * rng.uniform(0.6, 0.9)
    h = rng.uniform(18.0, 30.0)
    eps = math.exp(rng.uniform(math.log(strain_range[0]), math.log(strain_range[1])))
    spec = vertebra_j2(dx=dx, a=a, b=b, height=h, strain=-eps, n_increments=1, seed=int(rng.integers(1 << 31)),
                       **kw)
    spec.u0[spec.extra["top_idx"]] = spec.extra["u_final"]
    spec.name = f"rvertj2_{seed}_dx{dx}"
    spec.extra["strain"] = eps
    return spec
