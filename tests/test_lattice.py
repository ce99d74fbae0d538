"""Lattice kernels (PDX_VARIANT_LATTICE) against the halo kernels and the FP64 oracle.

On a voxel lattice every bond vector is h o for an offset o of the 122-site
stencil, so the lattice kernels take xi, |xi| and the influence weight as
constants of the offset.  Their bond arithmetic is the halo kernels' (FP32
bond terms, FP64 sums) with exact bond vectors instead of FP32 differences of
anchored positions, so forces agree with the halo kernels to FP32 rounding
(2e-6 relative-L2) and with the oracle to the same bound as the halo kernels
(5e-5); solves reach the same displacements (1e-5 relative-L2).
"""
import numpy as np
import pytest

from helpers import lattice, pair, rel_l2, smooth_u
from simul_amd import workloads as W
from simul_amd.body import PDBody

pytestmark = pytest.mark.gpu

LAT = (256, 256 | 64, 256 | 128)          # PDX_VARIANT_LATTICE [| SPLIT2 / SPLIT4]
HALO_REF = 32 | 1                          # PDX_VARIANT_HALO, group 1, no split


def _lattice_batch(model="neo_hookean", dx=0.5, seed=1, **kw):
    """Two exact lattice cases (the second ends in padding particles) with
    log-normal moduli, a fixed bottom layer and loads on top."""
    A = lattice(11, 9, 8, dx=dx)
    B = lattice(6, 7, 9, dx=dx) + 40.0
    X = np.concatenate([A, B])
    off = np.array([0, len(A), len(X)])
    n = len(X)
    E = 15000.0 * np.exp(0.4 * np.random.default_rng(seed).standard_normal(n))
    bulk, shear = W.bulk_shear(E, 0.3)
    fixed = np.zeros(n, np.uint8)
    fixed[X[:, 2] < 0.25 * dx] = 7
    fixed[(X[:, 2] > 40.0 - 0.25 * dx) & (X[:, 2] < 40.0 + 0.25 * dx)] = 7
    fext = np.zeros((n, 3))
    fext[(X[:, 2] > 6.5 * dx) & (X[:, 2] < 20.0), 2] = -30.0
    fext[X[:, 2] > 40.0 + 7.5 * dx, 0] = 20.0
    h = 3.015 * dx
    body = PDBody(X, dx ** 3, h, bulk, shear, cstab=18 * bulk / (np.pi * h ** 4), fixed=fixed, fext=fext,
                  model=model, case_offsets=off, **kw)
    return X, body


def _pass_forces(body, u):
    """Internal forces at u through the solver's own F-pass and force pass (a
    kinetic-DR step from rest leaves u in place and stores f), where
    compute_forces runs the diagnostic F-pass."""
    body.u = u
    body.set("v", np.zeros_like(u))
    body.set("a", np.zeros_like(u))
    body.kd_step(1e-3, 0.0, 0.0, 1)
    assert np.array_equal(body.u, u)
    return body.get("f")


def test_lattice_qualification(gpu):
    """An exact lattice at 3.015 h qualifies with its spacing; jitter, two
    volumes, a horizon at or below 3 h, a later change of volumes and 2-D do
    not."""
    X = lattice(9, 8, 7, dx=0.4)
    _, shear = W.bulk_shear(np.full(len(X), 1000.0), 0.3)
    bulk = shear * 2.0
    body = PDBody(X, 0.4 ** 3, 3.015 * 0.4, bulk, shear)
    assert abs(body.lattice_spacing() - 0.4) < 1e-12
    vol = np.full(len(X), 0.4 ** 3)
    vol[5] *= 1.01
    body.set("volume", vol)
    assert body.lattice_spacing() == 0.0
    assert PDBody(X, vol, 3.015 * 0.4, bulk, shear).lattice_spacing() == 0.0
    Xj = lattice(9, 8, 7, dx=0.4, jitter=0.01, seed=3)
    assert PDBody(Xj, 0.4 ** 3, 3.015 * 0.4, bulk, shear).lattice_spacing() == 0.0
    assert PDBody(X, 0.4 ** 3, 2.9 * 0.4, bulk, shear).lattice_spacing() == 0.0
    g = np.arange(20.0)
    X2 = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    b2, s2 = W.bulk_shear(np.full(len(X2), 1000.0), 0.3)
    assert PDBody(X2, 1.0, 3.015, b2, s2).lattice_spacing() == 0.0


@pytest.mark.parametrize("model", ["linear_elastic", "neo_hookean"])
def test_lattice_forces_match_oracle(gpu, model):
    """F-pass and force pass of every lattice split against the FP64 oracle on
    a heterogeneous lattice cube under a smooth non-affine displacement, with
    the oracle's own bond vectors (FP64 position differences)."""
    X = lattice(10, 9, 8, dx=0.5)
    n = len(X)
    fext = np.zeros((n, 3))
    fext[X[:, 2] > 3.0, 2] = -3.0
    body, prob = pair(X, 3.015 * 0.5, model=model, hetero=True, fext=fext, dx=0.5)
    assert abs(body.lattice_spacing() - 0.5) < 1e-12
    u = smooth_u(X, amp=2e-3)
    body.u = u
    prob.set_u(u)
    prob.compute_forces()
    for v in LAT:
        body.set_variants(v, v)
        f = body.compute_forces()
        assert rel_l2(f, prob.f) < 5e-5, (v, rel_l2(f, prob.f))
        f = _pass_forces(body, u)
        assert rel_l2(f, prob.f) < 5e-5, (v, rel_l2(f, prob.f))


def test_lattice_forces_match_halo(gpu):
    """Every lattice split of each pass against the halo kernels, also mixed
    (a lattice F-pass with a halo force pass and the reverse), with broken
    bonds inside rows (from the damage kernel, then rewritten by
    pdx_set_broken) and a batch whose cases end in padding particles."""
    X, body = _lattice_batch()
    assert abs(body.lattice_spacing() - 0.5) < 1e-12
    body.u = smooth_u(X, amp=0.02, seed=7)
    assert body.break_bonds(0.02) > 0
    u1 = smooth_u(X, amp=2e-3, seed=8)
    for step in ("damage", "set_broken"):
        if step == "set_broken":
            _, _, br = body.bonds()
            br = br.copy()
            br[::7] = 1
            br[3::11] = 0
            body.set_broken(br)
        body.u = u1
        body.set_variants(HALO_REF, HALO_REF)
        f0 = body.compute_forces().copy()
        g0 = _pass_forces(body, u1).copy()
        assert rel_l2(g0, f0) < 2e-6, step
        for v in LAT:
            body.set_variants(v, v)
            f = body.compute_forces()
            assert rel_l2(f, f0) < 2e-6, (step, v, rel_l2(f, f0))
            # the solver's passes: lattice F-pass writing the force pass's
            # records, and each lattice pass beside a halo pass
            assert rel_l2(_pass_forces(body, u1), g0) < 2e-6, (step, v, "both")
            body.set_variants(v, HALO_REF)
            assert rel_l2(_pass_forces(body, u1), g0) < 2e-6, (step, v, "F-pass")
            body.set_variants(HALO_REF, v)
            assert rel_l2(_pass_forces(body, u1), g0) < 2e-6, (step, v, "force")
            assert rel_l2(body.compute_forces(), f0) < 2e-6, (step, v, "force after the diagnostic F-pass")


def test_lattice_j2_forces_match_oracle(gpu):
    """J2 (radial return in the lattice F-pass's FP64 finish) against the
    oracle, yielding under a large smooth displacement."""
    X = lattice(9, 9, 8, dx=0.5)
    body, prob = pair(X, 3.015 * 0.5, model="j2_plasticity", hetero=True, dx=0.5, yield_stress=60.0,
                      h_iso=750.0, h_kin=100.0)
    u = smooth_u(X, amp=8e-3, seed=5)
    body.u = u
    prob.set_u(u)
    prob.compute_forces()
    body.set_variants(HALO_REF, HALO_REF)
    fh = _pass_forces(body, u).copy()
    for v in LAT:
        body.set_variants(v, v)
        f = _pass_forces(body, u)
        assert rel_l2(f, prob.f) < 5e-5, (v, rel_l2(f, prob.f))
        assert rel_l2(f, fh) < 2e-6, (v, rel_l2(f, fh))


def test_lattice_solves_match_halo_and_oracle(gpu):
    """Adaptive DR to 1e-6 with the lattice kernels: the iteration counts of
    the halo kernels (+-2%) and their displacements (1e-5); 40 kinetic-DR
    steps agree to 1e-5; the first case's solution matches the oracle's
    adaptive-DR solution of that case alone (1e-5)."""
    X, body = _lattice_batch(model="linear_elastic")
    ref = None
    for v in (HALO_REF,) + LAT:
        body.set_variants(v, v)
        body.u = np.zeros_like(X)
        cold = body.cold_residual()
        r = body.solve("adaptive", tol=1e-6, max_iterations=40000, refs=cold)
        assert all(x.converged for x in r), v
        uu = body.u.copy()
        u1 = smooth_u(X, amp=2e-3, seed=8)
        body.u = u1
        body.set("v", np.zeros_like(X))
        body.set("a", np.zeros_like(X))
        body.solve("kinetic", tol=0.0, max_iterations=40, dt=1e-3, use_graph=False)
        uk = body.u - u1
        if ref is None:
            ref = {"it": [x.iterations for x in r], "u": uu, "uk": uk}
            continue
        for a_, b_ in zip([x.iterations for x in r], ref["it"]):
            assert abs(a_ - b_) <= max(3, 0.02 * b_), (v, a_, b_)
        assert rel_l2(uu, ref["u"]) < 1e-5, (v, rel_l2(uu, ref["u"]))
        assert rel_l2(uk, ref["uk"]) < 1e-5, (v, rel_l2(uk, ref["uk"]))
    # the first case alone against the oracle
    n0 = 11 * 9 * 8
    Xa = X[:n0]
    fixed = np.zeros(n0, np.uint8)
    fixed[Xa[:, 2] < 0.125] = 7
    fext = np.zeros((n0, 3))
    fext[Xa[:, 2] > 3.25, 2] = -30.0
    a, prob = pair(Xa, 3.015 * 0.5, hetero=True, fixed=fixed, fext=fext, dx=0.5, seed=4)
    a.set_variants(256 | 128, 256 | 128)
    cold = a.cold_residual()
    r = a.solve("adaptive", tol=1e-6, max_iterations=40000, ref=cold)
    ro = prob.adr_solve(prob.adr_mass(1.0), max_it=40000, tol=1e-6, ref=cold)
    assert r.converged and ro["converged"]
    assert rel_l2(a.u, prob.u) < 1e-5, rel_l2(a.u, prob.u)


def test_lattice_graph_and_eager_bitwise(gpu):
    """hipGraph-captured chunks of lattice iterations give the eager bits."""
    X, body = _lattice_batch()
    body.set_variants(256 | 64, 256 | 128)
    r1 = body.solve("adaptive", tol=1e-6, use_graph=True, chunk=16)
    u1 = body.u
    body.u = np.zeros_like(X)
    r2 = body.solve("adaptive", tol=1e-6, use_graph=False, chunk=16)
    assert [x.iterations for x in r1] == [x.iterations for x in r2]
    assert np.array_equal(u1, body.u)


def test_autotune_considers_lattice(gpu):
    """The autotuner times the lattice variants on a qualifying context and
    keeps a valid variant for each pass."""
    X, body = _lattice_batch()
    body.autotune(iters=4)
    f, g = body.variants()
    ok = {1, 2, 4, 16 | 1, 16 | 2} | {32 | s | q for s in (0, 64, 128) for q in (1, 2, 4)} | set(LAT)
    assert f in ok and g in ok


def test_lattice_force_pass_with_a_wide_halo(gpu):
    """A lattice body one of whose blocks reaches more particles than the
    80-byte staged records fit in LDS (a random J2 vertebral body at dx 0.75,
    2 170 slots in its widest block) still runs the lattice force pass, with
    64-byte records: its forces match the gathering kernels (2e-6) for both
    the F-pass's records and the force pass's own staging, and a few adaptive
    DR iterations agree."""
    import ctypes as C

    from simul_amd import _native as NN
    from simul_amd.body import from_specimen
    spec = W.random_vertebra_j2(105229, dx=0.75)
    body = from_specimen(spec)
    assert abs(body.lattice_spacing() - 0.75) < 1e-12
    sizes = np.zeros(5, np.int64)
    body.ctx.call("pdx_get_halo_tables", NN.ptr(sizes, NN._i64p), None, None, None, None, None, None)
    cap = int(sizes[4])
    assert cap * 80 > 160 * 1024 - 1024 >= cap * 64, cap  # only 64-byte records fit
    u = smooth_u(spec.X, amp=2e-3, seed=3)
    u[spec.fixed != 0] = spec.u0[spec.fixed != 0]
    body.set_variants(1, 1)
    g0 = _pass_forces(body, u).copy()
    f0 = body.compute_forces().copy()
    for v in LAT:
        body.set_variants(v, v)
        assert rel_l2(_pass_forces(body, u), g0) < 2e-6, v
        body.set_variants(HALO_REF, v)  # the force pass stages the records itself
        assert rel_l2(_pass_forces(body, u), g0) < 2e-6, v
        assert rel_l2(body.compute_forces(), f0) < 2e-6, v
