"""Benchmark: PD bond-updates/s and time-to-solve at relative residual 1e-6.

Workload (default): the synthetic vertebral body -- elliptic cylinder 20 x 15 mm,
height 25 mm, voxelised at dx = 0.4 mm (365,056 particles, horizon 3.015 dx),
cortical shell 12 GPa around a log-normal cancellous core, Neo-Hookean, bottom
endplate fixed, 2 MPa pressure on the top endplate.  FP32 bond arithmetic, FP64
shape-tensor inverse, FP64 reductions.

A "step" is one adaptive dynamic-relaxation iteration over the specimen
(F-pass + force pass with the fused DR update).  `value` = intact directed
bonds x timed steps / wall time of the timed steps (inputs resident in HBM).
Time-to-solve is a separate cold-start solve to residual / cold residual < 1e-6.

roofline: the dominant kernel's mean duration from HIP events on the solver's
own stream; algorithmic bytes per launch from the per-unit model
  F-pass     8 B/bond + 112 B/particle
  force pass 8 B/bond + 133 B/particle
(16 B/bond + 245 B/particle per iteration).
cpu_baseline: the FP64 oracle in reference-faithful mode (kinetic damping,
P_j K_j^-1 recomputed per bond) on the host cores, a bounded number of
iterations of the same specimen.

The default run also measures the two other solver workloads and attaches
them to the same JSON line under "configs", each with its own roofline (HIP
event means of its hot kernels against the same per-unit model; J2 adds
52 B/particle of plastic state to the F-pass) and its own CPU baseline:
  vertebra_j2_dx0.25  J2 plasticity, 1,507,600 particles, 2 % compression in
                      20 increments solved to 1e-6 with bond breaking
  batch64_dx0.5       64 random specimens (7,045,942 particles) in one context,
                      every case solved to 1e-6 (setup included in the
                      reported time-to-solve), cold and warm-started
  warmstart_j2_...    the learned warm start: 64 held-out J2 vertebral bodies
                      (dx 0.75) in one context, cold and warm-started solves
                      to 1e-6 with the same DR mass
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PD bond-updates/s; time-to-solve at rel. residual 1e-6 (plain vs warm-start)"
HBM_PEAK_GBS = 8000.0
VARIANT_COOP, VARIANT_HALO = 16, 32  # PDX_VARIANT_COOP / PDX_VARIANT_HALO of include/pdx.h
VARIANT_SPLIT2, VARIANT_SPLIT4 = 64, 128
VARIANT_LATTICE = 256


def variant_names(v):
    """(label, F-pass kernel, force kernel) of a hot-pass variant code."""
    if v & VARIANT_LATTICE:
        split = 4 if v & VARIANT_SPLIT4 else (2 if v & VARIANT_SPLIT2 else 1)
        return f"lattice stencil over the LDS-staged halo, rows split over {split} waves", "k_fpass_l", "k_force_l"
    if v & VARIANT_HALO:
        split = 4 if v & VARIANT_SPLIT4 else (2 if v & VARIANT_SPLIT2 else 1)
        return f"LDS-staged halo, bond group {v & 15}, rows split over {split} waves", "k_fpass_h", "k_force_h"
    if v & VARIANT_COOP:
        return f"cooperative gathers, bond group {v & ~VARIANT_COOP}", "k_fpass_w", "k_force_w"
    return f"per-lane gathers, bond group {v}", "k_fpass", "k_force"


def apply_variant(body, spec, iters):
    """--variant: one code for both passes, 'F,G' per pass, 0 = autotune, or
    auto: the library's default variants (the lattice kernels) where the
    context is a lattice, else autotune."""
    if str(spec) == "auto":
        if body.lattice_spacing() > 0.0:
            return
        spec = "0"
    v = [int(x) for x in str(spec).split(",")]
    if v == [0]:
        body.autotune(iters)
    elif len(v) == 1:
        body.set_group(v[0])
    else:
        body.set_variants(v[0], v[1])


def mass_label(spectral):
    return "spectral (pdx_adr_mass_auto, tuning included)" if spectral else "Gershgorin (beta = 1)"


def pass_variants(body):
    """(label, F-pass kernel, force kernel) of the variants a body runs."""
    fv, gv = body.variants()
    if fv == gv:
        return variant_names(fv)
    return (f"F-pass: {variant_names(fv)[0]}; force: {variant_names(gv)[0]}", variant_names(fv)[1],
            variant_names(gv)[2])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--dx", type=float, default=0.4)
    ap.add_argument("--model", default="neo_hookean")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-solve-iters", type=int, default=200000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=50)
    ap.add_argument("--dr-mass", choices=["spectral", "gershgorin"], default="spectral",
                    help="adaptive-DR mass of the time-to-solve runs (spectral: pdx_adr_mass_auto per solve, with "
                         "the Gershgorin fallback should a case diverge)")
    ap.add_argument("--predictor", choices=["secant", "none"], default="secant",
                    help="start of each load increment of the J2 loading (secant: 2 u_(k-1) - u_(k-2))")
    ap.add_argument("--variant", type=str, default="auto",
                    help="hot-pass variant (1, 2, 4, 17, 18, 33, 34, 36, 97, 98, 100, 161, 162, 164, 256, 320, 384) or 'F,G' for the F-pass "
                         "and the force pass; 0 = autotune; auto = the library's default lattice kernels on a "
                         "lattice specimen, else autotune")
    ap.add_argument("--workload", default="vertebra", choices=["vertebra", "vertebra_j2", "batch", "warmstart", "nn"],
                    help="vertebra: the headline (dx 0.4, Neo-Hookean, pressure); vertebra_j2: J2, 20 "
                         "prescribed increments with breaking (dx 0.25); batch: 64 random cases in one context; "
                         "warmstart: the learned warm start on 64 held-out J2 cases, cold vs warm")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (scales rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--headline-only", action="store_true",
                    help="measure only the dx 0.4 vertebral body (no J2-loading and batch blocks)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "simul_amd", "learned", "weights_vertebra_mc.npz"),
                    help="the headline warm start's network weights (default: the network trained on the headline's "
                         "shape and load with 1,024 other cancellous fields)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1 "
                                   "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from simul_amd import workloads as W
    from simul_amd.body import from_specimen

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.workload == "vertebra_j2":
        _print(run_loading(args, torch))
        return
    if args.workload == "batch":
        _print(run_batch(args, torch))
        return
    if args.workload == "nn":
        _print(run_nn(args, torch))
        return
    if args.workload == "warmstart":
        _print(run_warmstart(args, torch))
        return

    t0 = time.perf_counter()
    spec = W.vertebra(dx=args.dx, model=args.model)
    log(f"specimen {spec.name}: {spec.n} particles ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    body = from_specimen(spec)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    st = body.stats()
    nb = st["n_intact"]
    log(f"setup {setup_s:.2f}s: bonds {st['n_bonds']} (SELL slots {st['sell_slots']})")

    cold = body.cold_residual()
    body.u = spec.u0
    apply_variant(body, args.variant, 10)
    vlabel, kf_name, kg_name = pass_variants(body)
    log(f"hot-pass variant: {vlabel}")
    # warmup + timed steps: fixed-count iterations that never stop early (tol = 0)
    body.u = spec.u0
    body.solve("adaptive", tol=0.0, max_iterations=args.warmup, ref=cold, chunk=64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = body.solve("adaptive", tol=0.0, max_iterations=args.steps, ref=cold, chunk=64)
    torch.cuda.synchronize()
    dt_steps = time.perf_counter() - t0
    ms_per_step = 1e3 * dt_steps / args.steps
    value = nb * args.steps / dt_steps
    log(f"timed {args.steps} steps: {ms_per_step * 1e3:.1f} us/step, {value / 1e9:.1f} G bond-updates/s")

    # per-kernel durations (HIP events on the solver stream)
    body.u = spec.u0
    kt = body.time_kernels(args.kernel_iters)
    n_p = spec.n
    bytes_f = 8.0 * nb + 112.0 * n_p
    bytes_g = 8.0 * nb + 133.0 * n_p
    k_f = {"mean_us": kt["fpass_ms"] * 1e3, "bytes": bytes_f,
           "achieved_GBs": bytes_f / (kt["fpass_ms"] * 1e-3) / 1e9}
    k_g = {"mean_us": kt["force_ms"] * 1e3, "bytes": bytes_g,
           "achieved_GBs": bytes_g / (kt["force_ms"] * 1e-3) / 1e9}
    dom_name, dom = (kg_name, k_g) if k_g["mean_us"] >= k_f["mean_us"] else (kf_name, k_f)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{spec.name}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get(dom_name)
    # VALU issue: wave64 VALU instructions per launch of the running variants
    # (SQ_INSTS_VALU of the same workload, profiles/valu_*.json) over the
    # measured launch time, against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
    valu = None
    vfile = os.path.join(ROOT, "profiles", f"valu_{spec.name}.json")
    fv, gv = body.variants()
    if os.path.exists(vfile):
        with open(vfile) as fh:
            vj = json.load(fh)
        peak = vj["peak_wave_instructions_per_s"]
        valu = {"unit": "G wave64 VALU instructions/s", "peak": round(peak / 1e9, 1)}
        for key, code, kt_ms, name in (("fpass", fv, kt["fpass_ms"], kf_name), ("force", gv, kt["force_ms"], kg_name)):
            n_ins = vj[key].get(str(code))
            if n_ins:
                rate = n_ins / (kt_ms * 1e-3)
                valu[name] = {"variant": code, "instructions": n_ins, "achieved": round(rate / 1e9, 1),
                              "frac": round(rate / peak, 4)}
    # the whole DR iteration (both passes) against the per-iteration model
    # 16 B/bond + 245 B/particle
    it_bytes = 16.0 * nb + 245.0 * n_p
    it_gbs = it_bytes / (ms_per_step * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(dom["achieved_GBs"], 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom["achieved_GBs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": dom["bytes"], "mean_us": round(dom["mean_us"], 2),
                "kernels": {kf_name: {k: round(v, 2) for k, v in k_f.items()},
                            kg_name: {k: round(v, 2) for k, v in k_g.items()}},
                "iteration": {"algorithmic_bytes": it_bytes, "achieved_GBs": round(it_gbs, 1),
                              "frac": round(it_gbs / HBM_PEAK_GBS, 4), "ms": round(ms_per_step, 4)},
                "valu": valu}
    log(f"kernels: fpass {k_f['mean_us']:.1f} us ({k_f['achieved_GBs']:.0f} GB/s), "
        f"force {k_g['mean_us']:.1f} us ({k_g['achieved_GBs']:.0f} GB/s)")

    tts = None
    if not args.no_solve:
        # the Gershgorin mass (beta = 1) for reference, then the spectral mass
        # (its tuning inside the timed region), which the cold and warm
        # time-to-solve figures use
        body.u = spec.u0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rg = body.solve("adaptive", tol=args.tol, max_iterations=args.max_solve_iters, ref=cold, chunk=128,
                        mass="gershgorin")
        torch.cuda.synchronize()
        wall_g = time.perf_counter() - t0
        log(f"time-to-solve, Gershgorin mass: {wall_g:.3f}s, {rg.iterations} iterations, converged={rg.converged}")
        body.u = spec.u0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rs = body.solve("adaptive", tol=args.tol, max_iterations=args.max_solve_iters, ref=cold, chunk=128,
                        mass=args.dr_mass)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        tts = {"cold_start_s": round(wall, 4), "iterations": rs.iterations, "converged": rs.converged,
               "relative_residual": rs.relative_residual, "residual_normalisation": "cold-start residual",
               "dr_mass": mass_label(args.dr_mass == "spectral"),
               "cold_start_gershgorin_s": round(wall_g, 4), "iterations_gershgorin": rg.iterations,
               # what the spectral mass alone gains over the Gershgorin mass, both cold
               "spectral_mass_speedup": round(wall_g / wall, 3),
               "warm_start_s": None, "warm_start": "no trained weights", "setup_s": round(setup_s, 3)}
        log(f"time-to-solve: {wall:.3f}s, {rs.iterations} iterations, converged={rs.converged}, "
            f"rel={rs.relative_residual:.2e}")
        u_cold = body.u
        wfile = args.weights
        if os.path.exists(wfile):
            from simul_amd.learned import WarmStartNet, warm_solve
            from simul_amd.learned.warmstart import predict_u_device
            net = WarmStartNet.load(wfile)
            # one untimed inference loads the GPU code objects of the feature
            # and network path (a first-use cost of the process, as the cold
            # solve's kernels are loaded by the warmup steps); nothing of it
            # is kept
            predict_u_device(net, body, spec)
            torch.cuda.synchronize()
            body.u = spec.u0
            rw, tw = warm_solve(body, spec, net, tol=args.tol, ref=cold, max_iterations=args.max_solve_iters,
                                chunk=128, mass=args.dr_mass)
            du = float(np.linalg.norm(body.u - u_cold) / np.linalg.norm(u_cold))
            tts.update({"warm_start_s": round(tw["total_s"], 4), "warm_iterations": rw.iterations,
                        "warm_converged": rw.converged, "warm_inference_s": round(tw["inference_s"], 4),
                        # the learned start's own gain: warm against cold,
                        # both with the same (spectral) DR mass
                        "warm_speedup": round(wall / tw["total_s"], 3),
                        "warm_iteration_ratio": round(rs.iterations / max(rw.iterations, 1), 3),
                        "warm_vs_cold_u_rel_l2": du,
                        "warm_start": "learned per-particle network (BF16 MFMA kernels; "
                                      + os.path.basename(wfile) + "), features, inference and "
                                      "DR included in warm_start_s; one untimed inference before it loads the "
                                      "process's GPU code for that path"})
            log(f"warm start: {tw['total_s']:.3f}s ({tw['inference_s']:.3f}s features+inference), "
                f"{rw.iterations} iterations, converged={rw.converged}, |u_warm - u_cold|/|u_cold| = {du:.2e}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec, args.cpu_seconds)

    configs = None
    if not args.headline_only:
        import gc
        del body
        gc.collect()
        configs = {}
        for fn in (run_loading, run_batch, run_warmstart):
            blk = fn(args, torch)
            configs[blk["config"]["workload"]] = blk
            gc.collect()

    out = {
        "metric": METRIC, "value": value, "unit": "bond-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": spec.name + "_" + spec.model, "n_particles": spec.n, "n_bonds": nb,
                   "horizon_mm": spec.horizon, "dr": "adaptive", "hot_pass_variant": vlabel,
                   "precision": "fp32 bonds, fp64 K^-1 and reductions"},
        "time_to_solve": tts, "roofline": roofline, "cpu_baseline": cpu,
    }
    if configs is not None:
        out["configs"] = configs
    _print(out)


def _print(out):
    print(json.dumps(out), flush=True)


def _line(metric, value, unit, steps, warmup, ms, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


def hot_roofline(body, n_p, nb, iters, state_bytes=0.0):
    """Roofline of the two hot passes from HIP events on the solver stream:
    algorithmic bytes per launch (F-pass 8 B/bond + (112 + state) B/particle,
    force pass 8 B/bond + 133 B/particle) over each kernel's mean duration;
    the dominant kernel heads the record."""
    kt = body.time_kernels(iters)
    _, kf_name, kg_name = pass_variants(body)
    bytes_f = 8.0 * nb + (112.0 + state_bytes) * n_p
    bytes_g = 8.0 * nb + 133.0 * n_p
    k = {kf_name: (kt["fpass_ms"], bytes_f), kg_name: (kt["force_ms"], bytes_g)}
    dom = max(k, key=lambda name: k[name][0])
    ms, b = k[dom]
    ach = b / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": b,
            "mean_us": round(ms * 1e3, 2),
            "kernels": {name: {"mean_us": round(v[0] * 1e3, 2), "bytes": v[1],
                               "achieved_GBs": round(v[1] / (v[0] * 1e-3) / 1e9, 1)} for name, v in k.items()}}


def run_loading(args, torch):
    """J2 vertebra at dx 0.25 mm (1,507,600 particles): 2 % endplate compression
    in 20 prescribed increments, each solved by adaptive DR to 1e-6, bonds
    breaking at stretch 0.02 with re-equilibration.  value = intact bonds x DR
    iterations / DR wall time over the whole loading."""
    from simul_amd import workloads as W
    from simul_amd.body import from_specimen
    from simul_amd.loading import IncrementalLoading

    dx = args.dx if args.dx != 0.4 else 0.25
    spec = W.vertebra_j2(dx=dx)
    t0 = time.perf_counter()
    body = from_specimen(spec)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    nb0 = body.stats()["n_intact"]
    log(f"{spec.name}: {spec.n} particles, {nb0} bonds, setup {setup:.2f}s")
    apply_variant(body, args.variant, 5)
    drv = IncrementalLoading(body, spec.extra["top_idx"], spec.extra["u_final"], spec.extra["n_increments"],
                             critical_stretch=spec.extra["critical_stretch"], tol=args.tol,
                             max_iterations=args.max_solve_iters,
                             mass=args.dr_mass, predictor=None if args.predictor == "none" else args.predictor,
                             progress=lambda done, r: log(f"    {done} iterations, rel {r.relative_residual:.2e}"))

    def progress(r):
        log(f"increment {r.increment:2d}: {r.iterations} iterations, rel {r.relative_residual:.2e}, "
            f"converged={r.converged}, broken now {r.broken_now} (total {r.broken_total}), {r.elapsed_s:.2f}s")

    t0 = time.perf_counter()
    res = drv.run(callback=progress)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    iters = sum(r.iterations for r in res)
    upd = nb0 * iters  # intact bonds at the start; broken bonds are a small fraction
    nb_end = body.stats()["n_intact"]
    # hot kernels at the final (yielded, partly broken) state; J2 adds 52 B per
    # particle of plastic state (eps_p 6, epbar 1, beta 6 floats) to the F-pass
    roof = hot_roofline(body, spec.n, nb_end, args.kernel_iters, state_bytes=52.0)
    log(f"kernels: {roof['kernels']}")
    variant = pass_variants(body)[0]
    del body
    cpu = None
    if not args.no_cpu_baseline:
        # the first increment's boundary values, kinetic DR from there
        u1 = spec.u0.copy()
        u1[spec.extra["top_idx"]] = spec.extra["u_final"] / spec.extra["n_increments"]
        cpu = cpu_baseline(spec, args.cpu_seconds, u_start=u1,
                           what="kinetic-DR iterations of the first load increment (top endplate at 1/20 of "
                                "the final compression) of the full J2 specimen")
    return _line(METRIC, upd / wall, "bond-updates/s", iters, 0, 1e3 * wall / max(iters, 1),
                 {"workload": spec.name, "n_particles": spec.n, "n_bonds": nb0, "increments": len(res),
                  "critical_stretch": spec.extra["critical_stretch"], "variant": variant},
                 time_to_solve={"loading_s": round(wall, 3), "setup_s": round(setup, 3),
                                "dr_mass": mass_label(args.dr_mass == "spectral"), "predictor": args.predictor,
                                "broken_total": res[-1].broken_total if res else 0,
                                "all_converged": all(r.converged for r in res),
                                "increments": [{"k": r.increment, "iterations": r.iterations,
                                                "relative_residual": r.relative_residual,
                                                "broken_now": r.broken_now,
                                                "seconds": round(r.elapsed_s, 3)} for r in res]},
                 roofline=roof, cpu_baseline=cpu)


def run_batch(args, torch):
    """64 random box / cylinder / ellipsoid specimens (10 k - 200 k particles,
    dx 0.5 mm, two-phase Neo-Hookean, pressure 0.5-3 MPa within 20 deg of the
    axis) packed into one context: every launch covers all cases.  value =
    intact bonds of all cases x iterations / wall time; time-to-solve = wall
    time until every case reaches 1e-6, reported with and without the setup
    (neighbour search, bond layout, shape tensors) of the 64 cases."""
    from simul_amd import workloads as W
    from simul_amd.body import from_specimen

    specs = W.batch()
    packed, off = W.pack(specs)
    t0 = time.perf_counter()
    body = from_specimen(packed, case_offsets=off)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    nb = body.stats()["n_intact"]
    log(f"{packed.name}: {len(specs)} cases, {packed.n} particles, {nb} bonds, setup {setup:.2f}s")
    # ref = 0: each case is normalised by its own residual at the start state,
    # which for u0 is the cold-start residual
    apply_variant(body, args.variant, 5)
    body.u = packed.u0
    body.solve("adaptive", tol=0.0, max_iterations=args.warmup, ref=0.0, chunk=64)
    torch.cuda.synchronize()
    steps = args.steps
    t0 = time.perf_counter()
    body.solve("adaptive", tol=0.0, max_iterations=steps, ref=0.0, chunk=64)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"timed {steps} batched iterations: {1e3 * dt / steps:.2f} ms each, {nb * steps / dt / 1e9:.1f} G bond-updates/s")
    tts = None
    if not args.no_solve:
        body.u = packed.u0
        torch.cuda.synchronize()
        t0 = time.perf_counter()

        def progress(done, rs):
            nconv = sum(r.converged for r in rs)
            worst = max((r.relative_residual for r in rs if not r.converged), default=0.0)
            log(f"  {done} iterations: {nconv}/{len(rs)} cases converged, worst rel. residual {worst:.2e}")

        rs = body.solve("adaptive", tol=args.tol, max_iterations=args.max_solve_iters, ref=0.0, chunk=128,
                        slice_iterations=2048, progress=progress,
                        mass=args.dr_mass)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        its = [r.iterations for r in rs]
        tts = {"all_cases_s": round(wall, 3), "all_cases_with_setup_s": round(wall + setup, 3),
               "dr_mass": mass_label(args.dr_mass == "spectral"),
               "converged": sum(r.converged for r in rs), "cases": len(rs),
               "iterations_min": min(its), "iterations_max": max(its), "setup_s": round(setup, 3)}
        log(f"time-to-solve (all cases): {wall:.2f}s (+ setup {setup:.2f}s), iterations {min(its)}..{max(its)}, "
            f"{tts['converged']}/{len(rs)} converged")
        wfile = os.path.join(ROOT, "simul_amd", "learned", "weights_batch.npz")
        if os.path.exists(wfile):
            # the warm-started batch: the network of the batch's family (other
            # seeds) predicts every case's start, the same DR mass, the same
            # criterion (each case's cold-start residual)
            import numpy as np

            from simul_amd.learned import WarmStartNet
            from simul_amd.learned.warmstart import predict_u_device
            net = WarmStartNet.load(wfile)
            u_cold = body.u
            predict_u_device(net, body, packed, off)  # loads the path's GPU code (untimed, nothing kept)
            refs = np.atleast_1d(body.cold_residual())
            body.u = packed.u0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            body.set_u_device(predict_u_device(net, body, packed, off))
            torch.cuda.synchronize()
            t_inf = time.perf_counter() - t0
            rw = body.solve("adaptive", tol=args.tol, max_iterations=args.max_solve_iters, refs=refs, chunk=128,
                            slice_iterations=2048, progress=progress, mass=args.dr_mass)
            torch.cuda.synchronize()
            wall_w = time.perf_counter() - t0
            iw = [r.iterations for r in rw]
            u_warm = body.u  # one download (each read of body.u copies the whole field)
            du = max(float(np.linalg.norm(u_warm[a:b] - u_cold[a:b]) / np.linalg.norm(u_cold[a:b]))
                     for a, b in zip(off[:-1], off[1:]))
            tts.update({"warm_all_cases_s": round(wall_w, 3), "warm_inference_s": round(t_inf, 4),
                        "warm_speedup": round(wall / wall_w, 3), "warm_iterations_max": max(iw),
                        "warm_iteration_ratio_mean": round(float(np.mean(its)) / float(np.mean(iw)), 3),
                        "warm_converged": sum(r.converged for r in rw), "warm_vs_cold_u_rel_l2_max": du,
                        "warm_start": "network of the batch's family (simul_amd/learned/weights_batch.npz, "
                                      "1,024 other random specimens), features, inference and DR included"})
            log(f"warm-started batch: {wall_w:.2f}s incl. {t_inf:.3f}s features+inference, iterations max {max(iw)} "
                f"({wall / wall_w:.2f}x wall)")
    body.u = packed.u0
    roof = hot_roofline(body, packed.n, nb, args.kernel_iters)
    log(f"kernels: {roof['kernels']}")
    variant = pass_variants(body)[0]
    del body
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_cases(specs, args.cpu_seconds)
    return _line(METRIC, nb * steps / dt, "bond-updates/s", steps, args.warmup, 1e3 * dt / steps,
                 {"workload": packed.name, "cases": len(specs), "n_particles": packed.n, "n_bonds": nb,
                  "variant": variant},
                 time_to_solve=tts, roofline=roof, cpu_baseline=cpu)


def run_warmstart(args, torch):
    """The learned warm start on its J2 family: 64 held-out random J2
    vertebral bodies (dx 0.75 mm, about 49 k particles each, compressed
    between platens in one increment to 0.1-2 %; seeds never used in
    training) packed into one context, each solved by adaptive DR to 1e-6 of
    its cold-start residual from rest (cold) and from the network's
    prediction (warm), the same DR mass for both.  value = bond updates of
    the warm solve (each case's share of the intact bonds times its
    iterations) / its DR wall time."""
    import numpy as np

    from simul_amd.learned import WarmStartNet
    from simul_amd.learned.train import _group_body, _specimen
    from simul_amd.learned.warmstart import predict_u_device

    net = WarmStartNet.load(os.path.join(ROOT, "simul_amd", "learned", "weights_vertebra_j2.npz"))
    dx = 0.75
    specs = [_specimen("j2", s, dx) for s in range(105000, 105064)]
    t0 = time.perf_counter()
    packed, off, body = _group_body(specs)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    nb = body.stats()["n_intact"]
    log(f"warm-start family: {len(specs)} held-out J2 cases, {packed.n} particles, {nb} bonds, setup {setup:.2f}s")
    apply_variant(body, args.variant, 5)
    predict_u_device(net, body, packed, off)  # loads the path's GPU code (untimed, nothing kept)
    cold = np.atleast_1d(body.cold_residual())

    def solve():
        return body.solve("adaptive", tol=args.tol, max_iterations=args.max_solve_iters, refs=cold,
                          slice_iterations=2048, progress=lambda d, r: None, mass=args.dr_mass)

    body.u = packed.u0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc = solve()
    torch.cuda.synchronize()
    t_cold = time.perf_counter() - t0
    u_cold = body.u
    body.u = packed.u0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    body.set_u_device(predict_u_device(net, body, packed, off))
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    rw = solve()
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    u_warm = body.u
    ic, iw = [r.iterations for r in rc], [r.iterations for r in rw]
    du = [float(np.linalg.norm(u_warm[a:b] - u_cold[a:b]) / np.linalg.norm(u_cold[a:b]))
          for a, b in zip(off[:-1], off[1:])]
    tts = {"cold_start_s": round(t_cold, 4), "warm_start_s": round(t_warm, 4), "warm_inference_s": round(t_inf, 4),
           "warm_speedup": round(t_cold / t_warm, 3), "cold_iterations_mean": float(np.mean(ic)),
           "warm_iterations_mean": float(np.mean(iw)), "warm_iteration_ratio": round(np.mean(ic) / np.mean(iw), 3),
           "min_case_iteration_ratio": round(min(c / w for c, w in zip(ic, iw)), 3),
           "cold_iterations_max": max(ic), "warm_iterations_max": max(iw),
           "converged": int(sum(r.converged for r in rc) + sum(r.converged for r in rw)), "solves": 2 * len(specs),
           "warm_vs_cold_u_rel_l2_max": max(du), "dr_mass": mass_label(args.dr_mass == "spectral"),
           "setup_s": round(setup, 3),
           "warm_start": "J2 network (simul_amd/learned/weights_vertebra_j2.npz, 4,096 training cases); features, "
                         "inference and DR included in warm_start_s; one untimed inference before it loads the "
                         "process's GPU code for that path"}
    log(f"warm-start family: cold {t_cold:.3f}s ({np.mean(ic):.0f} it mean), warm {t_warm:.3f}s incl. "
        f"{t_inf:.3f}s features+inference ({np.mean(iw):.0f} it mean): {t_cold / t_warm:.2f}x wall, "
        f"{np.mean(ic) / np.mean(iw):.2f}x iterations")
    body.u = packed.u0
    roof = hot_roofline(body, packed.n, nb, args.kernel_iters, state_bytes=52.0)
    variant = pass_variants(body)[0]
    del body
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_cases(specs, args.cpu_seconds, what="the first held-out J2 cases")
    # bond updates of the warm solve: each case's bonds (its share of the
    # particles, one spacing) times its own iterations, converged cases
    # leaving the batched launches
    sizes = np.diff(off)
    upd = nb * float(np.dot(sizes, iw)) / float(sizes.sum())
    return _line(METRIC, upd / (t_warm - t_inf), "bond-updates/s", max(iw), 0,
                 1e3 * (t_warm - t_inf) / max(max(iw), 1),
                 {"workload": f"warmstart_j2_dx{dx}_64heldout", "cases": len(specs), "n_particles": packed.n,
                  "n_bonds": nb, "variant": variant},
                 time_to_solve=tts, roofline=roof, cpu_baseline=cpu)


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def _nn_traffic():
    """HBM bytes per launch of the 256 x 256 forward at 365 056 rows, from the
    committed FETCH_SIZE / WRITE_SIZE summary (profiles/), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_warmstart_mlp_gemm.json")) as fh:
            t = json.load(fh)["fwd_365056"]
        return int(t["fetch"] + t["write"])
    except (OSError, KeyError, ValueError):
        return None


def run_nn(args, torch):
    """The learned warm start's layer GEMMs on the BF16 MFMA kernels
    (include/pdx_nn.h): forward, data gradient and weight gradient of every
    layer of the 64-256-256-256-32 network at the training batch (65 536
    rows) and at the dx 0.4 vertebral body (365 056 rows, forward only), timed
    with HIP events on the stream the kernels run on (`--steps` launches
    each).  value = FLOP/s of the 256 x 256 forward layer at the training batch."""
    import ctypes as C

    import numpy as np

    from simul_amd import _native as N
    from simul_amd.learned.net import DIMS, wgrad_rows_per_slice

    lib = N.load()
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    gen = torch.Generator(device="cuda").manual_seed(1)
    reps = max(5, args.steps)

    def timed(fn):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3  # us

    def bf(shape):
        x = torch.randn(shape, device="cuda", generator=gen, dtype=torch.float32) * 0.1
        y = torch.empty(shape, dtype=torch.int16, device="cuda")
        assert lib.pdx_nn_to_bf16(p(x), x.numel(), p(y), sp) == 0
        return y

    rows = []
    for M, grads in ((65536, True), (365056, False)):
        for li, (k, n) in enumerate(zip(DIMS[:-1], DIMS[1:])):
            X, W, dY = bf((M, k)), bf((k, n)), bf((M, n))
            b = torch.zeros(n, dtype=torch.float32, device="cuda")
            Y = torch.empty((M, n), dtype=torch.int16, device="cuda")
            flop = 2.0 * M * k * n
            us = timed(lambda: lib.pdx_nn_linear_fwd(p(X), p(W), p(b), M, k, n, 1, p(Y), None, n, sp))
            rec = {"rows": M, "k": k, "n": n, "fwd_us": round(us, 2), "fwd_tflops": round(flop / us / 1e6, 1),
                   "fwd_frac_peak": round(flop / us / 1e6 / BF16_PEAK_TFLOPS, 4),
                   "fwd_GBs": round((M * (k + n) * 2 + k * n * 2) / us / 1e3, 1)}
            if grads:
                dX = torch.empty((M, k), dtype=torch.int16, device="cuda")
                us_d = timed(lambda: lib.pdx_nn_linear_dgrad(p(dY), p(W), M, k, n, n, p(X), p(dX), sp))
                R = wgrad_rows_per_slice(M)
                ns = (M + R - 1) // R
                part = torch.empty((ns * k * n,), dtype=torch.float32, device="cuda")
                gW = torch.empty((k, n), dtype=torch.float32, device="cuda")
                gb = torch.empty((n,), dtype=torch.float32, device="cuda")
                us_w = timed(lambda: lib.pdx_nn_linear_wgrad(p(X), p(dY), M, k, n, n, R, p(part), p(gW), p(gb), sp))
                rec.update({"dgrad_us": round(us_d, 2), "dgrad_tflops": round(flop / us_d / 1e6, 1),
                            "wgrad_us": round(us_w, 2), "wgrad_tflops": round(flop / us_w / 1e6, 1)})
            rows.append(rec)
            log(f"M {M} {k}x{n}: " + ", ".join(f"{a} {b}" for a, b in rec.items() if a not in ("rows", "k", "n")))
    # the value: a 256 x 256 layer's forward at the dx 0.4 vertebral body (the warm start's inference there)
    M = 365056
    head = min((r for r in rows if r["rows"] == M and r["k"] == 256 and r["n"] == 256), key=lambda r: r["fwd_us"])
    flop = 2.0 * M * 256 * 256
    byts = M * (256 + 256) * 2 + 256 * 256 * 2  # BF16 rows in and out, the weight
    # one training step of a 65 536-row batch: forward, data gradients (layers 3..1), weight gradients
    tr = [r for r in rows if r["rows"] == 65536]
    step_us = sum(r["fwd_us"] + r["wgrad_us"] for r in tr) + sum(r["dgrad_us"] for r in tr[1:])
    return _line("BF16 MFMA GEMM FLOP/s (learned warm start, 256x256 layer, 365 056 rows)", flop / head["fwd_us"] * 1e6,
                 "FLOP/s", reps, 3, head["fwd_us"] / 1e3,
                 {"workload": "warmstart_mlp_64-256-256-256-32", "rows": [65536, 365056]},
                 roofline={"bound": "mfma", "kernel": "k_nn_gemm_lds", "achieved": head["fwd_tflops"],
                           "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": head["fwd_frac_peak"],
                           "traffic": _nn_traffic(),
                           "hbm": {"achieved": round(byts / head["fwd_us"] / 1e3, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(byts / head["fwd_us"] / 1e3 / HBM_PEAK_GBS, 3)}},
                 training_step={"rows": 65536, "us": round(step_us, 1),
                                "rows_per_s": round(65536 / step_us * 1e6)},
                 layers=rows, cpu_baseline=None)


def copy_calibration(torch):
    """One elementwise kernel reading 256 MiB and writing 256 MiB: the
    rocprof FETCH_SIZE / WRITE_SIZE readings of this launch give the factor
    between the counters and bytes actually moved."""
    n = (256 << 20) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    log(f"copy calibration: read {4 * n} B, wrote {4 * n} B in one elementwise kernel")


def _oracle_kd(spec, u_start=None):
    """The specimen in the oracle, reference-faithful mode, and its stable dt."""
    import numpy as np

    from oracle import pdoracle as po

    j2 = {}
    if spec.model == "j2_plasticity":
        j2 = dict(sy0=spec.extra["yield"], h_iso=spec.extra["h_iso"], h_kin=spec.extra.get("h_kin"))
    prob = po.Problem(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, mass=spec.vol * spec.rho,
                      model=spec.model, fixed=spec.fixed, fext=spec.fext, nthreads=po.default_threads(),
                      per_bond_recompute=True, **j2)
    prob.shape_tensor("absolute", 1e-12)
    prob.set_u(spec.u0 if u_start is None else u_start)
    dt = prob.estimate_dt(float(np.mean(spec.bulk)), float(np.mean(spec.shear)), spec.rho)
    return prob, dt


def cpu_baseline(spec, seconds, u_start=None, what="kinetic-DR iterations of the full specimen"):
    """Oracle in reference-faithful mode on the host cores, bounded sample."""
    from oracle import pdoracle as po

    cores = po.default_threads()
    t0 = time.perf_counter()
    prob, dt = _oracle_kd(spec, u_start)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob.kd_solve(dt, 0.0, max_it=1, tol=0.0)
    t1 = time.perf_counter() - t0
    iters = max(1, int(seconds / max(t1, 1e-3)))
    t0 = time.perf_counter()
    r = prob.kd_solve(dt, 0.0, max_it=iters, tol=0.0)
    el = time.perf_counter() - t0
    nb = prob.nb
    # kd_solve evaluates the forces once before the loop and once per iteration
    v = nb * (iters + 1) / el
    log(f"cpu baseline ({spec.name}): {cores} threads, {iters} iterations in {el:.1f}s -> {v / 1e6:.1f} M "
        f"bond-updates/s, relative residual reached {r['relative_residual']:.3e}")
    return {"value": v, "unit": "bond-updates/s", "cores": cores, "kind": "port",
            "ms_per_iteration": round(1e3 * el / iters, 2),
            "relative_residual_reached": float(r["relative_residual"]),
            "sample": f"{iters} {what} (plus 1 untimed) from the cold start, reference-faithful FP64 "
                      f"C/OpenMP restatement; setup {setup:.1f}s untimed; the reference's kinetic DR needs "
                      f"about 300 000 iterations for 1e-4, so 1e-6 is not reached within the sample"}


def cpu_baseline_cases(specs, seconds, what="the first cases of the batch"):
    """The oracle on the batch's cases one after another (a case at a time:
    the reference solves one body per solver), about `seconds` of kinetic-DR
    iterations spread over the first cases."""
    from oracle import pdoracle as po

    cores = po.default_threads()
    upd, el_tot, used, its_tot = 0.0, 0.0, [], 0
    per_case = seconds / 4
    for spec in specs:
        if el_tot >= seconds or len(used) >= 4:
            break
        prob, dt = _oracle_kd(spec)
        t0 = time.perf_counter()
        prob.kd_solve(dt, 0.0, max_it=1, tol=0.0)
        t1 = time.perf_counter() - t0
        iters = max(1, int(per_case / max(t1, 1e-3)))
        t0 = time.perf_counter()
        prob.kd_solve(dt, 0.0, max_it=iters, tol=0.0)
        el = time.perf_counter() - t0
        upd += prob.nb * (iters + 1)
        el_tot += el
        its_tot += iters
        used.append(f"{spec.name} ({spec.n} particles, {iters} iterations)")
    v = upd / el_tot
    log(f"cpu baseline (batch): {cores} threads, {len(used)} cases in {el_tot:.1f}s -> {v / 1e6:.1f} M bond-updates/s")
    return {"value": v, "unit": "bond-updates/s", "cores": cores, "kind": "port",
            "sample": f"kinetic-DR iterations from the cold start of {what}, one case at a "
                      "time, reference-faithful FP64 C/OpenMP restatement (setup untimed): " + "; ".join(used)}


if __name__ == "__main__":
    main()
