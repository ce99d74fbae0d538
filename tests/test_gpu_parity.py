"""HIP path (libpdx through the C ABI) against the FP64 oracle on identical inputs.

Tolerances: integer/index work (neighbour sets, broken flags) bit-exact; FP64
setup (bond data, K, K^-1) to 1e-12 relative; FP32 bond arithmetic (F, P, f)
to the relative-L2 bounds written in each test; solved displacements to 1e-5
relative-L2 of the oracle solved to the same residual.
"""
import numpy as np
import pytest

from helpers import lattice, pair, rel_l2, smooth_u
from oracle import pdoracle as po
from simul_amd import workloads as W
from simul_amd.body import PDBody

pytestmark = pytest.mark.gpu


def _oracle_csr_equal(body, prob):
    row, col, br = body.bonds()
    assert np.array_equal(row, prob.row)
    assert np.array_equal(col, prob.col[: prob.nb])
    assert not br.any()


@pytest.mark.parametrize("jitter", [0.0, 0.23])
def test_neighbor_sets_match_oracle(gpu, jitter):
    X = lattice(9, 8, 7, dx=1.0, jitter=jitter, seed=5)
    body, prob = pair(X, 3.015)
    _oracle_csr_equal(body, prob)
    assert body.n_bonds == prob.nb


def test_neighbor_counts_reference_kats(gpu):
    # backend/fea/peridynamics/tests/test_3d.py:67-138
    X = lattice(3, 3, 3)
    for h, centre, corner in ((1.1, 6, 3), (1.5, 18, 6)):
        body = PDBody(X, 1.0, h, 1.0, 1.0, build=False)
        c = body.count_neighbors(h)
        assert c[13] == centre and c[0] == corner
    # 2-D: test_particles.py:84-146
    g = np.arange(3.0)
    X2 = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    b2 = PDBody(X2, 1.0, 1.5, 1.0, 1.0, build=False)
    c2 = b2.count_neighbors(1.5)
    assert c2[4] == 8 and all(c2[k] == 3 for k in (0, 2, 6, 8))
    X22 = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0]])
    assert np.all(PDBody(X22, 1.0, 1.1, 1.0, 1.0, build=False).count_neighbors(1.1) == 2)


def test_config1_bond_count(gpu):
    spec = W.cube()
    body = PDBody(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, build=False)
    assert body.build_bonds(spec.horizon) == 816728
    nnb = body.get("nnb")
    assert nnb.min() == 28 and nnb.max() == 122


def test_bond_data_and_shape_tensor(gpu):
    X = lattice(8, 7, 6, jitter=0.15, seed=2)
    body, prob = pair(X, 3.015, hetero=True)
    xi, L, w = body.bond_data()
    nb = prob.nb
    assert np.array_equal(xi, prob.xi[:nb])
    assert np.array_equal(L, prob.xlen[:nb])
    assert np.array_equal(w, prob.omega[:nb])
    K, Ki = body.get("K"), body.get("K_inv")
    assert rel_l2(K, prob.K) < 1e-13
    assert rel_l2(Ki, prob.Kinv) < 1e-12
    assert np.array_equal(body.get("sing"), prob.sing)


@pytest.mark.parametrize("model", ["linear_elastic", "neo_hookean"])
def test_patch_test_affine(gpu, model):
    """u = G X gives F = I + G at every particle, surfaces included."""
    X = lattice(7, 7, 7, jitter=0.1, seed=7)
    body, _ = pair(X, 3.015, model=model)
    G = np.random.default_rng(3).uniform(-2e-3, 2e-3, (3, 3))
    body.u = X @ G.T
    body.compute_F_stress()
    F = body.get("F")
    err = np.abs(F - (np.eye(3) + G)[None]).max() / np.abs(G).max()
    assert err < 2e-5, err


@pytest.mark.parametrize("model", ["linear_elastic", "neo_hookean"])
def test_F_P_force_match_oracle(gpu, model):
    X = lattice(10, 9, 8, jitter=0.1, seed=11)
    n = len(X)
    fext = np.zeros((n, 3))
    fext[X[:, 2] > 6.5, 2] = -3.0
    body, prob = pair(X, 3.015, model=model, hetero=True, fext=fext)
    u = smooth_u(X, amp=2e-3)
    body.u = u
    prob.set_u(u)
    f = body.compute_forces()
    prob.compute_forces()
    I = np.eye(3)[None]
    assert rel_l2(body.get("F") - I, prob.F - I) < 2e-5
    assert rel_l2(body.get("P"), prob.P) < 2e-5
    assert rel_l2(f, prob.f) < 5e-5, rel_l2(f, prob.f)


def test_bond_formulation_force(gpu):
    X = lattice(9, 8, 7, jitter=0.1, seed=4)
    body, prob = pair(X, 3.015, formulation="bond")
    u = smooth_u(X, amp=2e-3, seed=9)
    body.u = u
    prob.set_u(u)
    f = body.compute_forces()
    prob.compute_forces()
    assert rel_l2(f, prob.f) < 5e-5


def test_linear_momentum(gpu):
    X = lattice(9, 9, 9, jitter=0.2, seed=8)
    body, _ = pair(X, 3.015, model="neo_hookean", hetero=True)
    body.u = smooth_u(X, amp=3e-3, seed=2)
    f = body.compute_forces()
    # both bond terms are antisymmetric in (i, j): sum_i V_i f_i = 0 (uniform V)
    tot = np.abs(f.sum(0)).max()
    assert tot < 1e-5 * np.abs(f).sum(0).max(), tot


def _cube_pair(n_side=12, model="linear_elastic"):
    spec = W.cube(n_side=n_side)
    body, prob = pair(spec.X, spec.horizon, model=model, fixed=spec.fixed)
    body.u = spec.u0
    prob.set_u(spec.u0)
    return spec, body, prob


@pytest.mark.parametrize("model", ["linear_elastic", "neo_hookean"])
def test_adaptive_dr_matches_oracle(gpu, model):
    spec, body, prob = _cube_pair(12, model)
    cold = body.cold_residual()
    r = body.solve("adaptive", tol=1e-6, max_iterations=20000, ref=cold)
    lam = prob.adr_mass(1.0)
    # FP32 volume/modulus records feed the GPU mass: agreement to FP32 rounding
    assert np.allclose(body.get("drmass"), lam, rtol=1e-6)
    ro = prob.adr_solve(lam, max_it=20000, tol=1e-6, ref=cold)
    assert r.converged and ro["converged"]
    # same algorithm; FP32 forces and FP32-rounded mass inputs shift the trajectory slightly
    assert abs(r.iterations - ro["iterations"]) <= max(5, 0.1 * ro["iterations"])
    assert rel_l2(body.u, prob.u) < 1e-5, rel_l2(body.u, prob.u)
    # the returned state really is at the reported residual
    body.compute_forces()
    assert body.residual_norm() / cold < 1.01e-6


def test_graph_and_eager_are_bitwise_identical(gpu):
    spec, body, _ = _cube_pair(10)
    r1 = body.solve("adaptive", tol=1e-6, use_graph=True, chunk=16)
    u1 = body.u
    body.u = spec.u0
    r2 = body.solve("adaptive", tol=1e-6, use_graph=False, chunk=16)
    assert r1.iterations == r2.iterations
    assert np.array_equal(u1, body.u)


def test_kinetic_dr_steps_match_oracle(gpu):
    spec, body, prob = _cube_pair(8)
    dt = body.estimate_dt(spec.bulk[0], spec.shear[0], spec.rho)
    dto = prob.estimate_dt(spec.bulk[0], spec.shear[0], spec.rho)
    assert abs(dt - dto) <= 1e-12 * dto
    r = body.solve("kinetic", tol=0.0, max_iterations=40, dt=dt, use_graph=False)
    ro = prob.kd_solve(dt, 0.0, max_it=40, tol=0.0)
    assert r.iterations == 40 and ro["iterations"] == 40
    assert rel_l2(body.u, prob.u) < 1e-5
    assert r.resets == ro["resets"]


def test_damage_flags_bitexact(gpu):
    X = lattice(9, 8, 7, jitter=0.1, seed=6)
    body, prob = pair(X, 3.015)
    u = smooth_u(X, amp=0.03, seed=4)
    body.u = u
    x = body.x                       # X + u as the library holds it
    prob.x[...] = x
    nb_g = body.break_bonds(0.01)
    nb_o = prob.update_damage(0.01)
    assert nb_g == nb_o > 0
    _, _, br = body.bonds()
    assert np.array_equal(br, prob.broken[: prob.nb])
    assert np.array_equal(body.local_damage(), prob.local_damage())


@pytest.mark.parametrize("h_kin", [None, 0.08])
def test_j2_stress_and_commit(gpu, h_kin):
    X = lattice(7, 7, 7, jitter=0.05, seed=1)
    n = len(X)
    E = 1000.0
    sy = np.full(n, 2.0)
    hi = np.full(n, 0.05 * E)
    hk = None if h_kin is None else np.full(n, h_kin * E)
    body, prob = pair(X, 3.015, E=E, model="j2_plasticity", yield_stress=sy, h_iso=hi, h_kin=hk)
    G = np.diag([-0.004, 0.001, 0.0015]) + 1e-3 * np.random.default_rng(2).uniform(-1, 1, (3, 3))
    u = X @ G.T + smooth_u(X, amp=5e-4, seed=3)
    body.u = u
    prob.set_u(u)
    body.compute_F_stress()
    prob.deformation_gradient()
    prob.stress()
    assert rel_l2(body.get("P"), prob.P) < 2e-5
    assert prob.epbar_t.max() > 0  # yielding happened
    body.commit_state()
    prob.commit()
    assert rel_l2(body.get("epbar"), prob.epbar) < 1e-4
    assert rel_l2(body.get("ep").reshape(-1, 9), prob.ep) < 1e-4
    if hk is None:
        return
    assert rel_l2(body.get("back").reshape(-1, 9), prob.beta) < 1e-4
    # reverse loading from the committed state: the yield test and the stress see
    # the back stress and the plastic strain of the first step
    u2 = -0.6 * u
    body.u = u2
    prob.set_u(u2)
    body.compute_F_stress()
    prob.deformation_gradient()
    prob.stress()
    assert prob.epbar_t.max() > prob.epbar.max()
    assert rel_l2(body.get("P"), prob.P) < 2e-5
    body.commit_state()
    prob.commit()
    assert rel_l2(body.get("back").reshape(-1, 9), prob.beta) < 1e-4


def test_batch_matches_single_cases(gpu):
    specs = [W.cube(n_side=k) for k in (6, 9, 7)]
    off = np.cumsum([0] + [s.n for s in specs])
    Xb = np.concatenate([s.X for s in specs])
    fx = np.concatenate([s.fixed for s in specs])
    u0 = np.concatenate([s.u0 for s in specs])
    bulk, shear = W.bulk_shear(15000.0, 0.3)
    cst = 18 * bulk / (np.pi * 3.015 ** 4)
    b = PDBody(Xb, 1.0, 3.015, bulk, shear, cstab=cst, fixed=fx, u0=u0, case_offsets=off)
    rb = b.solve("adaptive", tol=1e-6)
    ub = b.u
    for k, s in enumerate(specs):
        single = PDBody(s.X, 1.0, 3.015, bulk, shear, cstab=cst, fixed=s.fixed, u0=s.u0)
        rs = single.solve("adaptive", tol=1e-6)
        assert rb[k].iterations == rs.iterations
        assert np.array_equal(ub[off[k]:off[k + 1]], single.u)


def test_large_bond_build_is_intact(gpu):
    """Regression: every stored bond of the 365,056-particle vertebra survives the
    build (no stream-unordered clear overwrites rows) and the sets match the oracle."""
    spec = W.vertebra(0.4)
    body = PDBody(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, build=False)
    nb = body.build_bonds(spec.horizon)
    st = body.stats()
    assert st["n_bonds"] == nb == st["n_intact"] == 42517012
    row, col, br = body.bonds()
    orow, ocol = po.neighbors(spec.X, spec.horizon)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol) and not br.any()


VARIANTS = (1, 2, 4, 16 | 1, 16 | 2)  # per-lane groups 1/2/4, cooperative groups 1/2


def _variant_runs(body, u0, method, iters, variants=VARIANTS, **kw):
    out = {}
    for g in variants:
        body.set_group(g)
        body.u = u0
        # a solve keeps the accelerations of the previous one (the reference's
        # solve() resets velocities only), so every variant starts from a = 0
        body.set("a", np.zeros_like(u0))
        f = body.compute_forces().copy()
        r = body.solve(method, tol=0.0, max_iterations=iters, use_graph=False, **kw)
        out[g] = (f, body.u.copy(), [x.iterations for x in r])
    return out


@pytest.mark.parametrize("formulation", ["nosb", "bond"])
def test_hot_pass_variants_bitwise_identical(gpu, formulation):
    """Every gather variant runs the same bond arithmetic in the same order:
    forces and DR trajectories agree bit for bit, with broken bonds inside
    rows and a batch whose cases end in padding particles."""
    A = lattice(9, 8, 7, jitter=0.2, seed=11)
    B = lattice(6, 7, 5, jitter=0.2, seed=12) + 50.0
    X = np.concatenate([A, B])
    off = np.array([0, len(A), len(X)])
    n = len(X)
    bulk, shear = W.bulk_shear(15000.0 * np.exp(0.3 * np.random.default_rng(1).standard_normal(n)), 0.3)
    fixed = np.zeros(n, np.uint8)
    fixed[X[:, 2] < 0.5] = 7
    fixed[(X[:, 2] > 49.5) & (X[:, 2] < 50.5)] = 7
    body = PDBody(X, 1.0, 3.015, bulk, shear, cstab=18 * bulk / (np.pi * 3.015 ** 4), fixed=fixed,
                  formulation=formulation, model="neo_hookean", case_offsets=off)
    u = smooth_u(X, amp=0.02, seed=7)
    body.u = u
    assert body.break_bonds(0.02) > 0          # dead slots in the middle of rows
    u0 = smooth_u(X, amp=2e-3, seed=8)
    base = None
    for method, kw in (("adaptive", {}), ("kinetic", {"dt": 1e-3})):
        runs = _variant_runs(body, u0, method, 25, **kw)
        base = runs[1]
        for g, (f, uu, it) in runs.items():
            assert it == base[2] == [25, 25]
            assert np.array_equal(f, base[0]), (method, g)
            assert np.array_equal(uu, base[1]), (method, g)


HALO = tuple(32 | s | g for s in (0, 64, 128) for g in (1, 2, 4))  # PDX_VARIANT_HALO [| SPLIT2/4] | group


def _halo_body(formulation):
    A = lattice(9, 8, 7, jitter=0.2, seed=11)
    B = lattice(6, 7, 5, jitter=0.2, seed=12) + 50.0
    X = np.concatenate([A, B])
    off = np.array([0, len(A), len(X)])
    n = len(X)
    bulk, shear = W.bulk_shear(15000.0 * np.exp(0.3 * np.random.default_rng(1).standard_normal(n)), 0.3)
    fixed = np.zeros(n, np.uint8)
    fixed[X[:, 2] < 0.5] = 7
    fixed[(X[:, 2] > 49.5) & (X[:, 2] < 50.5)] = 7
    fext = np.zeros((n, 3))
    fext[(X[:, 2] > 5.5) & (X[:, 2] < 10.0), 2] = -20.0
    fext[X[:, 2] > 53.5, 0] = 15.0
    body = PDBody(X, 1.0, 3.015, bulk, shear, cstab=18 * bulk / (np.pi * 3.015 ** 4), fixed=fixed, fext=fext,
                  formulation=formulation, model="neo_hookean", case_offsets=off)
    return X, body


@pytest.mark.parametrize("formulation", ["nosb", "bond"])
def test_halo_variants_match_gathering(gpu, formulation):
    """The LDS-staged variants stage positions relative to a per-block anchor
    in FP32, so their bond vectors differ from the gathering kernels' FP64
    differences by ~1e-7 relative.  Forces agree to 2e-6 with broken bonds
    inside rows (set by the damage kernel, then rewritten by pdx_set_broken)
    and a batch whose cases end in padding particles; groups 2 and 4 (NOSB
    force terms summed per group in FP32) agree with group 1 to 1e-7, split
    rows (other FP64 summation order) agree to 1e-12.
    Adaptive DR to 1e-6 converges in the same number of iterations (+-2%) to
    the same displacements (1e-5), and 40 kinetic-DR steps agree to 1e-5."""
    X, body = _halo_body(formulation)
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
        body.set_group(16 | 2)
        f0 = body.compute_forces().copy()
        fs = {}
        for g in HALO:
            body.set_group(g)
            fs[g] = body.compute_forces().copy()
            assert rel_l2(fs[g], f0) < 2e-6, (step, g, rel_l2(fs[g], f0))
        for sp in (0, 64, 128):
            for g in (2, 4):
                assert rel_l2(fs[32 | sp | g], fs[32 | sp | 1]) < 1e-7, (step, sp, g)
            assert rel_l2(fs[32 | sp | 1], fs[32 | 1]) < 1e-12, (step, sp)
    X, body = _halo_body(formulation)
    ref = {}
    for g in (16 | 2,) + HALO:
        body.set_group(g)
        body.u = np.zeros_like(X)
        cold = body.cold_residual()
        r = body.solve("adaptive", tol=1e-6, max_iterations=40000, refs=cold)
        assert all(x.converged for x in r), g
        uu = body.u.copy()
        body.u = u1
        body.set("v", np.zeros_like(X))
        body.set("a", np.zeros_like(X))
        body.solve("kinetic", tol=0.0, max_iterations=40, dt=1e-3, use_graph=False)
        if not ref:
            ref = {"it": [x.iterations for x in r], "u": uu, "uk": body.u.copy()}
            continue
        for a_, b_ in zip([x.iterations for x in r], ref["it"]):
            assert abs(a_ - b_) <= max(3, 0.02 * b_), (g, a_, b_)
        assert rel_l2(uu, ref["u"]) < 1e-5, (g, rel_l2(uu, ref["u"]))
        assert rel_l2(body.u - u1, ref["uk"] - u1) < 1e-5, (g, rel_l2(body.u - u1, ref["uk"] - u1))


def test_m64_reading_fpass_equals_summing_fpass(gpu):
    """The halo F-pass either sums each row's FP64 m = sum_j w V_j xi itself or
    reads it from the per-row array a solve builds first (in the order the
    F-pass variant sums it).  For every halo variant the two paths give the
    same F and the same forces bit for bit, also after a change of variant
    (which invalidates the array), with broken bonds and batch padding."""
    X, body = _halo_body("nosb")
    body.u = smooth_u(X, amp=0.02, seed=7)
    assert body.break_bonds(0.02) > 0
    u1 = smooth_u(X, amp=2e-3, seed=8)
    for g in HALO + (32 | 1, 32 | 128 | 4):
        body.set_group(g)
        body.u = u1
        f_sum = body.compute_forces().copy()       # no per-row array yet: summed in the pass
        F_sum = body.get("F").copy()
        body.solve("adaptive", tol=1e-6, max_iterations=0)   # builds the per-row array
        body.u = u1
        f_read = body.compute_forces().copy()
        assert np.array_equal(body.get("F"), F_sum), g
        assert np.array_equal(f_read, f_sum), g


def test_sliced_solve_matches_single(gpu):
    """Resumed slices of the device DR loop follow the same trajectory: same
    iteration counts and the same bits as one uninterrupted solve."""
    spec, body, _ = _cube_pair(10)
    r1 = body.solve("adaptive", tol=1e-6)
    u1 = body.u
    body.u = spec.u0
    seen = []
    r2 = body.solve("adaptive", tol=1e-6, slice_iterations=128, progress=lambda d, rs: seen.append(d))
    assert r2.converged and r2.iterations == r1.iterations and len(seen) > 1
    assert np.array_equal(body.u, u1)
    specs = [W.cube(n_side=k) for k in (7, 9)]
    off = np.cumsum([0] + [s.n for s in specs])
    bulk, shear = W.bulk_shear(15000.0, 0.3)
    mk = lambda: PDBody(np.concatenate([s.X for s in specs]), 1.0, 3.015, bulk, shear,
                        fixed=np.concatenate([s.fixed for s in specs]), u0=np.concatenate([s.u0 for s in specs]),
                        case_offsets=off)
    b1, b2 = mk(), mk()
    ra = b1.solve("adaptive", tol=1e-6)
    rb = b2.solve("adaptive", tol=1e-6, slice_iterations=128, progress=lambda d, rs: None)
    assert [r.iterations for r in ra] == [r.iterations for r in rb]
    assert np.array_equal(b1.u, b2.u)


def test_halo_variants_2d_and_lds_fallback(gpu):
    """2-D bodies through the halo kernels, and a horizon of 4 spacings whose
    NOSB force-pass halo (~3 000 slots x 80 B) exceeds the 160 KB of LDS
    while the F-pass halo (40 B per slot) fits: the NOSB force pass then runs
    the cooperative kernel.  Forces agree with the gathering kernels to
    2e-6."""
    g = np.arange(40.0)
    X2 = np.stack(np.meshgrid(g, g[:30], indexing="ij"), -1).reshape(-1, 2)
    X2 = X2 + np.random.default_rng(4).uniform(-0.2, 0.2, X2.shape)
    X3 = lattice(14, 13, 12, jitter=0.2, seed=9)
    for X, h in ((X2, 3.015), (X3, 4.0)):
        n, dim = X.shape
        bulk, shear = W.bulk_shear(15000.0 * np.exp(0.3 * np.random.default_rng(2).standard_normal(n)), 0.3)
        fixed = np.zeros(n, np.uint8)
        fixed[X[:, -1] < 0.5] = (1 << dim) - 1
        for form in ("nosb", "bond"):
            body = PDBody(X, 1.0, h, bulk, shear, cstab=18 * bulk / (np.pi * h ** 4), fixed=fixed,
                          formulation=form, model="neo_hookean")
            body.u = 2e-3 * np.sin(X / 5.0) + 1e-4 * np.cos(X.sum(1))[:, None]
            body.set_group(16 | 2)
            f0 = body.compute_forces().copy()
            for g_ in HALO:
                body.set_group(g_)
                f = body.compute_forces()
                assert rel_l2(f, f0) < 2e-6, (dim, h, form, g_, rel_l2(f, f0))


@pytest.mark.parametrize("model", ["linear_elastic", "neo_hookean"])
def test_2d_forces_and_dr_match_oracle(gpu, model):
    """A 2-D plane-stress plate against the oracle: the default stabilisation
    modulus 9 E / (pi t d^3) (nosb.py:335-346, t = 1), NOSB and bond forces at
    the force tolerance of the 3-D tests, and the adaptive-DR solution of a
    compressed plate to 1e-5 relative-L2."""
    g = np.arange(26.0)
    X = np.stack(np.meshgrid(g, g[:18], indexing="ij"), -1).reshape(-1, 2)
    X = X + np.random.default_rng(21).uniform(-0.15, 0.15, X.shape)
    n = len(X)
    E, nu, h = 15000.0, 0.3, 3.015
    K2, mu = E / (2 * (1 - nu)), E / (2 * (1 + nu))
    mass = np.full(n, W.RHO_BONE)
    fixed = np.zeros(n, np.uint8)
    bot, top = X[:, 1] < 0.5, X[:, 1] > 16.5
    fixed[bot] = 3
    fixed[top] = 3
    u0 = np.zeros_like(X)
    u0[top, 1] = -0.004 * 17.0
    body = PDBody(X, 1.0, h, K2, mu, mass=mass, model=model, fixed=fixed)
    prob = po.Problem(X, 1.0, h, K2, mu, mass=mass, model=model, fixed=fixed)
    prob.shape_tensor("relative", 1e-12)
    c_ref = 9.0 * E / (np.pi * h ** 3)
    assert np.allclose(prob.cstab, c_ref, rtol=1e-12)
    assert np.allclose(body.get("cstab"), c_ref, rtol=1e-6)
    u = smooth_u(X, amp=2e-3, seed=5)
    body.u = u
    prob.set_u(u)
    f = body.compute_forces()
    prob.compute_forces()
    assert rel_l2(f, prob.f) < 5e-5, rel_l2(f, prob.f)
    body.u = u0
    prob.set_u(u0)
    cold = body.cold_residual()
    prob.compute_forces()
    assert abs(prob.residual_norm() - cold) < 1e-5 * cold
    r = body.solve("adaptive", tol=1e-6, max_iterations=50000, ref=cold)
    ro = prob.adr_solve(prob.adr_mass(1.0), max_it=50000, tol=1e-6, ref=cold)
    assert r.converged and ro["converged"]
    assert rel_l2(body.u, prob.u) < 1e-5, rel_l2(body.u, prob.u)
    pb = po.Problem(X, 1.0, h, K2, mu, mass=mass, formulation="bond", fixed=fixed)
    bb = PDBody(X, 1.0, h, K2, mu, mass=mass, formulation="bond", fixed=fixed)
    bb.u = u
    pb.set_u(u)
    fb = bb.compute_forces()
    pb.compute_forces()
    assert rel_l2(fb, pb.f) < 5e-5, rel_l2(fb, pb.f)


def test_spectral_mass_same_solution_fewer_iterations(gpu):
    """pdx_adr_mass_auto: beta in (0, 1] per case, and adaptive DR with the
    spectral mass reaches the Gershgorin-mass solution (1e-5) in fewer
    iterations, for one body and for a batch; the oracle-matched default
    (the Gershgorin mass) is what a solve without mass= uses."""
    spec, body, prob = _cube_pair(12, "neo_hookean")
    cold = body.cold_residual()
    body.adr_mass(1.0)
    lam0 = body.get("drmass").copy()
    r1 = body.solve("adaptive", tol=1e-6, ref=cold)
    u1 = body.u
    assert np.allclose(body.get("drmass"), lam0)
    body.u = spec.u0
    beta = body.tune_mass()
    assert 0.0 < beta <= 1.0
    assert np.allclose(body.get("drmass"), beta * lam0, rtol=1e-12)
    body.u = spec.u0
    r2 = body.solve("adaptive", tol=1e-6, ref=cold, mass="spectral")
    assert r1.converged and r2.converged and r2.iterations < r1.iterations
    assert rel_l2(body.u, u1) < 1e-5
    specs = [W.cube(n_side=k) for k in (7, 10, 13)]
    off = np.cumsum([0] + [s.n for s in specs])
    bulk, shear = W.bulk_shear(15000.0, 0.3)
    X = np.concatenate([s.X for s in specs])
    bb = PDBody(X, 1.0, 3.015, bulk, shear, fixed=np.concatenate([s.fixed for s in specs]),
                u0=np.concatenate([s.u0 for s in specs]), case_offsets=off, model="neo_hookean")
    refs = np.atleast_1d(bb.cold_residual())
    ra = bb.solve("adaptive", tol=1e-6, refs=refs)
    ua = bb.u
    bb.u = np.concatenate([s.u0 for s in specs])
    rb = bb.solve("adaptive", tol=1e-6, refs=refs, mass="spectral")
    assert all(x.converged for x in ra + rb)
    assert all(b.iterations < a.iterations for a, b in zip(ra, rb))
    assert rel_l2(bb.u, ua) < 1e-5
    b3 = bb.tune_mass()
    assert b3.shape == (3,) and np.all((b3 > 0) & (b3 <= 1))


def test_spectral_mass_falls_back_when_a_case_diverges(gpu):
    """solve(mass="spectral") with a mass far below the stability limit (the
    tuning's safety factor forced to 0.1): the attempt diverges numerically,
    is stopped at its first slice, and the solve falls back to the Gershgorin
    mass from the same start, ending on the same solution as a plain solve."""
    spec, body, _ = _cube_pair(10, "neo_hookean")
    cold = body.cold_residual()
    r1 = body.solve("adaptive", tol=1e-6, ref=cold)
    u1 = body.u
    body.u = spec.u0
    tune = body.tune_mass
    body.tune_mass = lambda iters=20, safety=2.0: tune(iters, 0.1)
    r2 = body.solve("adaptive", tol=1e-6, ref=cold, mass="spectral", max_iterations=20000)
    assert r2.converged and r2.iterations == r1.iterations
    assert np.array_equal(body.u, u1)


def _tables(body):
    """The context's halo tables (pdx_get_halo_tables), or None when unbuilt."""
    from simul_amd import _native as NN
    sizes = np.zeros(5, np.int64)
    body.ctx.call("pdx_get_halo_tables", NN.ptr(sizes, NN._i64p), None, None, None, None, None, None)
    nbk, nent, nbl, nbkw, cap = (int(v) for v in sizes)
    if nbk == 0:
        return None
    ptr = np.zeros(nbk + 1, np.int32)
    hl = np.zeros(max(nent, 1), np.int32)
    hs = np.zeros(max(nent, 1), np.uint16)
    bl = np.zeros(nbl, np.uint16)
    bk = np.zeros(nbkw, np.uint16)
    anc = np.zeros(3 * nbk)
    body.ctx.call("pdx_get_halo_tables", NN.ptr(sizes, NN._i64p), NN.ptr(ptr, NN._i32p), NN.ptr(hl, NN._i32p),
                  NN.ptr(hs), NN.ptr(bl), NN.ptr(bk), NN.ptr(anc, NN._dp))
    return {"cap": cap, "ptr": ptr, "hl": hl[:nent], "hs": hs[:nent], "bl": bl, "bk": bk, "anchor": anc}


def _layout_pair(X, h, off=None, **kw):
    n, dim = X.shape
    bulk, shear = W.bulk_shear(np.full(n, 15000.0), 0.3)
    out = []
    for where in (1, 0):  # PDX_LAYOUT_HOST, PDX_LAYOUT_DEVICE
        b = PDBody(X, 1.0, h, bulk, shear, case_offsets=off, build=False, **kw)
        b.ctx.call("pdx_set_layout_builder", where)
        b.build_bonds(h)
        out.append((b, _tables(b)))
    return out


@pytest.mark.parametrize("case", ["lattice", "batch", "plate_2d", "wide_horizon", "vertebra", "batch64"])
def test_device_layout_equals_host_layout(gpu, case):
    """The device builder of the halo tables (k_halo_build) gives the host
    builder's tables bit for bit: halo lists, LDS slots (the colour classes and
    the over-full colours' free slots), bl / bk and the FP64 anchors; after
    bonds break (bl / bk rewritten by the damage kernel) both stay equal, and
    the forces of the two contexts are identical."""
    off = None
    h = 3.015
    if case == "lattice":
        X = lattice(20, 14, 11, jitter=0.2, seed=3)
    elif case == "batch":
        specs = [lattice(9, 8, 7), lattice(6, 7, 5, jitter=0.2, seed=4) + 50.0, lattice(13, 5, 4, jitter=0.1, seed=5)]
        X = np.concatenate(specs)
        off = np.cumsum([0] + [len(s) for s in specs])
    elif case == "plate_2d":
        g = np.arange(60.0)
        X = np.stack(np.meshgrid(g, g[:41], indexing="ij"), -1).reshape(-1, 2)
        X = X + np.random.default_rng(6).uniform(-0.2, 0.2, X.shape)
    elif case == "wide_horizon":
        X = lattice(16, 16, 12, jitter=0.15, seed=7)
        h = 4.5
    elif case == "vertebra":
        spec = W.vertebra(0.4)
        X, h = spec.X, spec.horizon
    else:
        spec, off = W.pack(W.batch())
        X, h = spec.X, spec.horizon
    (bh, th), (bd, td) = _layout_pair(X, h, off)
    assert th is not None and td is not None
    for k in ("cap", "ptr", "hl", "hs", "bl", "bk", "anchor"):
        assert np.array_equal(th[k], td[k]), (case, k)
    if case in ("vertebra", "batch64"):
        return
    u = smooth_u(X, amp=0.03, seed=9)
    for b in (bh, bd):
        b.u = u
        assert b.break_bonds(0.01) > 0
    th, td = _tables(bh), _tables(bd)
    for k in ("bl", "bk"):
        assert np.array_equal(th[k], td[k]), (case, k)
    if X.shape[1] == 3:
        for b in (bh, bd):
            b.shape_tensor()
        assert np.array_equal(bh.compute_forces(), bd.compute_forces())
