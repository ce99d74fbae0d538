"""Incremental loading with J2 plasticity and bond breaking:
the HIP driver (simul_amd.loading) against the same increments run by the
oracle, and the full-size config-3 specimen (dx 0.25 mm, 1,507,600 particles)
through size-independent checks: neighbour sets and broken flags bit-exact
with the oracle on the same displacement field, J2 state of that field within
the J2 tolerance.
"""
import os

import numpy as np
import pytest

from helpers import rel_l2
from oracle import pdoracle as po
from simul_amd import workloads as W
from simul_amd.body import from_specimen
from simul_amd.loading import IncrementalLoading

pytestmark = pytest.mark.gpu


def _oracle(spec, nthreads=None):
    p = po.Problem(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, model="j2_plasticity",
                   sy0=spec.extra["yield"], h_iso=spec.extra["h_iso"], fixed=spec.fixed, G_s=0.1,
                   cstab=W.stab_modulus(spec.bulk, spec.horizon), mass=spec.vol * spec.rho, nthreads=nthreads)
    p.shape_tensor("relative", 1e-12)
    return p


def _oracle_increments(p, idx, u_final, n_inc, tol, critical_stretch=None, break_rounds=3):
    """The driver's increments (simul_amd.loading.IncrementalLoading.step) on the
    oracle: prescribed DOFs at k / n of the final displacement, adaptive DR to
    tol, then up to break_rounds rounds of breaking, shape tensor and
    re-equilibration to tol of the increment's start residual, then the J2
    commit.  Returns (iterations, bonds broken) per increment."""
    out = []
    for k in range(1, n_inc + 1):
        u = p.u
        u[idx] = (k / n_inc) * u_final
        p.set_u(u)
        r = p.adr_solve(p.adr_mass(1.0), max_it=100000, tol=tol)
        assert r["converged"]
        its, nb_now, ref = r["iterations"], 0, r["reference_residual"]
        if critical_stretch is not None:
            for _ in range(break_rounds):
                nb = p.update_damage(critical_stretch)
                if nb == 0:
                    break
                nb_now += nb
                p.shape_tensor("relative", 1e-12)
                r = p.adr_solve(p.adr_mass(1.0), max_it=100000, tol=tol, ref=ref)
                assert r["converged"]
                its += r["iterations"]
        p.deformation_gradient()
        p.stress()
        p.commit()
        out.append((its, nb_now))
    return out


def test_incremental_j2_matches_oracle(gpu):
    """Config 3's J2 loading at dx 1.25 mm (12,080 particles), 4 increments to
    1e-6, against the oracle's increments to the same criterion.  Measured:
    the same iteration count in every increment, u within 1.0e-8 and the
    equivalent plastic strain within 6.9e-8 relative-L2 (the bound below
    leaves 100x; against an oracle solved to 1e-9 instead, u differs by
    4.9e-5: the 1e-6 criterion's own error)."""
    spec = W.vertebra_j2(dx=1.25, n_increments=4)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    body = from_specimen(spec)
    drv = IncrementalLoading(body, idx, uf, n_increments=4, critical_stretch=None, tol=1e-6)
    res = drv.run()
    assert all(r.converged for r in res)
    prob = _oracle(spec)
    its = _oracle_increments(prob, idx, uf, 4, 1e-6)
    for r, (io, _) in zip(res, its):
        assert abs(r.iterations - io) <= max(3, 0.01 * io), (r.iterations, io)
    assert np.allclose(body.u[idx], uf)
    assert rel_l2(body.u, prob.u) < 1e-6, rel_l2(body.u, prob.u)
    assert prob.epbar.max() > 0  # the cancellous core yields
    assert rel_l2(body.get("epbar"), prob.epbar) < 1e-5, rel_l2(body.get("epbar"), prob.epbar)


def test_incremental_breaking_matches_oracle(gpu):
    """Config 3's loading with bond breaking at dx 1.25 mm: the configured 2 %
    compression in 3 increments, critical stretch 0.012, up to 3 breaking
    rounds per increment, against the oracle running the same increments.
    Breaking cascades here (about 173 k of 1.4 M bonds break), so any
    difference in the fields would show in the flags.  Measured: broken flags
    identical, the same iterations and bonds broken in every increment, u
    within 2.1e-8 and the plastic strain within 9.2e-8 relative-L2."""
    spec = W.vertebra_j2(dx=1.25, n_increments=3)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    body = from_specimen(spec)
    drv = IncrementalLoading(body, idx, uf, n_increments=3, critical_stretch=0.012, tol=1e-6, max_break_rounds=3)
    res = drv.run()
    assert all(r.converged for r in res)
    prob = _oracle(spec)
    out = _oracle_increments(prob, idx, uf, 3, 1e-6, critical_stretch=0.012, break_rounds=3)
    print("GPU", [(r.iterations, r.broken_now) for r in res], "oracle", out, flush=True)
    for r, (io, nbo) in zip(res, out):
        assert r.broken_now == nbo
        assert abs(r.iterations - io) <= max(3, 0.01 * io), (r.iterations, io)
    _, _, br = body.bonds()
    assert br.sum() > 1000
    assert np.array_equal(br.astype(bool), prob.broken[: prob.nb].astype(bool))
    assert rel_l2(body.u, prob.u) < 1e-6, rel_l2(body.u, prob.u)
    assert rel_l2(body.get("epbar"), prob.epbar) < 1e-5, rel_l2(body.get("epbar"), prob.epbar)


def test_incremental_breaking_invariants(gpu):
    """Breaking inside the driver: counts only grow, every broken flag is one the
    criterion gives for the state it was taken at, damage = broken / initial."""
    spec = W.vertebra_j2(dx=1.25, n_increments=3)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"] * 3.0
    body = from_specimen(spec)
    drv = IncrementalLoading(body, idx, uf, n_increments=3, critical_stretch=0.01, tol=1e-5)
    res = drv.run()
    totals = [r.broken_total for r in res]
    assert totals == sorted(totals) and totals[-1] > 0
    assert sum(r.broken_now for r in res) == totals[-1]
    row, col, br = body.bonds()
    init = np.diff(row)
    nbr = np.add.reduceat(br.astype(np.int64), row[:-1]) * (init > 0)
    assert np.allclose(body.local_damage(), np.where(init > 0, nbr / np.maximum(init, 1), 0.0))
    # breaking keeps the bond family symmetric: internal forces still sum to zero
    f = body.compute_forces()
    assert np.isfinite(f).all()
    assert np.abs(f.sum(0)).max() < 1e-5 * np.abs(f).sum(0).max()


def test_config3_full_size_flags_and_j2(gpu):
    """dx 0.25 mm: 1,507,600 particles.  300 adaptive-DR iterations towards the
    full 2 % compression, then: neighbour CSR equal to the oracle's, broken flags of the
    GPU field equal to the oracle's on the same field (two thresholds), J2 trial
    state of that field within the J2 tolerance."""
    spec = W.vertebra_j2(dx=0.25)
    assert spec.n == 1507600
    body = from_specimen(spec, build=False)
    nb = body.build_bonds(spec.horizon)
    body.shape_tensor("relative", 1e-12)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    u = body.u
    u[idx] = uf  # the whole 2 % at once: the endplate region yields
    body.u = u
    body.solve("adaptive", tol=0.0, max_iterations=300)
    uu = body.u
    row, col, br = body.bonds()
    orow, ocol = po.neighbors(spec.X, spec.horizon)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol) and not br.any() and nb == len(ocol)
    prob = po.Problem(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, model="j2_plasticity",
                      sy0=spec.extra["yield"], h_iso=spec.extra["h_iso"], fixed=spec.fixed, row=orow, col=ocol)
    del orow, ocol
    prob.set_u(uu)
    for s_c in (spec.extra["critical_stretch"], 0.002):
        nb_g = body.break_bonds(s_c)
        nb_o = prob.update_damage(s_c)
        assert nb_g == nb_o
        _, _, br = body.bonds()
        assert np.array_equal(br, prob.broken[: prob.nb].astype(bool))
    assert nb_g > 0
    body.shape_tensor("relative", 1e-12)
    prob.shape_tensor("relative", 1e-12)
    body.compute_F_stress()
    prob.deformation_gradient()
    prob.stress()
    assert rel_l2(body.get("P"), prob.P) < 2e-5
    body.commit_state()
    prob.commit()
    assert prob.epbar.max() > 0
    assert rel_l2(body.get("epbar"), prob.epbar) < 1e-3


def test_secant_predictor_and_spectral_mass_same_states(gpu):
    """The J2 loading with the secant predictor and the spectral DR mass ends
    every increment at the state of the plain driver (same tolerance against
    the same plain-start reference) in fewer DR iterations; the oracle's
    final state agrees within 1e-4 (another DR path to the same 1e-6
    criterion, whose own error is about 5e-5 here)."""
    spec = W.vertebra_j2(dx=1.25, n_increments=6)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    runs = {}
    for key, kw in (("plain", {}), ("fast", {"predictor": "secant", "mass": "spectral"})):
        body = from_specimen(spec)
        drv = IncrementalLoading(body, idx, uf, n_increments=6, critical_stretch=None, tol=1e-6, **kw)
        res = drv.run()
        assert all(r.converged for r in res), key
        runs[key] = (sum(r.iterations for r in res), body.u.copy(), body.get("epbar").copy())
    assert runs["fast"][0] < runs["plain"][0], (runs["fast"][0], runs["plain"][0])
    assert rel_l2(runs["fast"][1], runs["plain"][1]) < 1e-4
    prob = _oracle(spec)
    _oracle_increments(prob, idx, uf, 6, 1e-6)
    assert rel_l2(runs["fast"][1], prob.u) < 1e-4
    assert rel_l2(runs["fast"][2], prob.epbar) < 1e-3


def test_secant_from_the_start_state_and_learned_first_increment(gpu):
    """The secant predictor takes the state the loading starts from as u_0, so
    the second increment already starts from 2 u_1 - u_0 (cheaper than the
    plain start where the response is elastic).  A ``start`` callable (here
    the learned J2 warm start, loading_start) supplies the first increment's
    guess, the only one without a secant guess: every run ends at the plain
    driver's state (same tolerance against the same plain-start reference)."""
    from simul_amd.learned import WarmStartNet
    from simul_amd.learned.warmstart import loading_start
    spec = W.vertebra_j2(dx=1.25, n_increments=4)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    net = WarmStartNet.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "simul_amd",
                                         "learned", "weights_vertebra_j2.npz"))
    runs = {}
    for key in ("plain", "secant", "secant+start"):
        body = from_specimen(spec)
        calls = []
        start = None
        if key == "secant+start":
            s0 = loading_start(net, body, spec, idx, uf, 4)
            start = lambda k: (calls.append(k), s0(k))[1]
        drv = IncrementalLoading(body, idx, uf, n_increments=4, critical_stretch=None, tol=1e-6,
                                 predictor=None if key == "plain" else "secant", start=start)
        res = drv.run()
        assert all(r.converged for r in res), key
        runs[key] = ([r.iterations for r in res], body.u.copy())
        if key == "secant+start":
            assert calls == [1]
    print({k: v[0] for k, v in runs.items()}, flush=True)
    assert runs["secant"][0][0] == runs["plain"][0][0]
    assert runs["secant"][0][1] < runs["plain"][0][1]
    for key in ("secant", "secant+start"):
        assert rel_l2(runs[key][1], runs["plain"][1]) < 1e-4, key


def test_reused_mass_factor(gpu):
    """solve(mass=beta) with the beta a spectral tuning returned runs the
    spectral solve's iterates bit for bit (beta x the cached Gershgorin mass
    rounds as the tuning's scaling does); a beta far below the stability
    limit diverges and the solve falls back to a fresh spectral tuning, which
    converges to the same state."""
    spec = W.vertebra_j2(dx=1.25, n_increments=4)
    idx, uf = spec.extra["top_idx"], spec.extra["u_final"]
    body = from_specimen(spec)
    u = spec.u0.copy()
    u[idx] = 0.25 * uf
    body.u = u
    cold = body.cold_residual()
    body.u = u
    r1 = body.solve("adaptive", tol=1e-6, ref=cold, mass="spectral", slice_iterations=512)
    u1 = body.u.copy()
    beta = body.last_beta
    assert 0.0 < beta <= 1.0
    body.u = u
    r2 = body.solve("adaptive", tol=1e-6, ref=cold, mass=beta, slice_iterations=512)
    assert r2.iterations == r1.iterations and np.array_equal(body.u, u1)
    body.u = u
    r3 = body.solve("adaptive", tol=1e-6, ref=cold, mass=beta * 1e-3, slice_iterations=512)
    assert r3.converged and rel_l2(body.u, u1) < 1e-4
    with pytest.raises(ValueError):
        body.solve("adaptive", tol=1e-6, ref=cold, mass=-1.0)
