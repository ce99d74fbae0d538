"""Training data, training and evaluation of the learned warm start.

    python -m simul_amd.learned.train --train 1024 --heldout 64 --dx 0.8 \
        --out simul_amd/learned/weights_vertebra.npz --report warmstart_report.json

Data: random vertebral bodies (simul_amd.workloads.random_vertebra: size,
cancellous field, pressure magnitude and direction within 20 deg), solved by
the plain cold-start adaptive DR to 1e-6 in batched contexts of ``group``
cases.  The network learns u / s (s = p H / 1000 MPa) on the free particles.
``--family j2``: J2 vertebral bodies compressed between platens in one
increment to a nominal strain of 0.1-2 % (workloads.random_vertebra_j2); the
features see the prescribed displacement as the load (features.load_view)
and s is the prescribed |u|.
Evaluation: the held-out cases solved cold and warm-started (same batched
context layout, same criterion: residual relative to each case's cold-start
residual), iterations and wall time including feature building and inference.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _group_body(specs):
    from ..body import from_specimen
    from ..workloads import pack
    packed, off = pack(specs)
    body = from_specimen(packed, case_offsets=off)
    return packed, off, body


FAMILIES = ("pressure", "j2", "vertebra", "vertebra_mc", "batch")


def _specimen(family, seed, dx):
    """One case of a training / held-out family: "pressure" is
    workloads.random_vertebra (Neo-Hookean, pressure on the top layer), "j2"
    workloads.random_vertebra_j2 (J2, compressed between platens in one
    increment), "vertebra" workloads.vertebra_family (the headline's shape,
    random cancellous field and load), "vertebra_mc" the same with the
    headline's load (only the cancellous field random), "batch"
    workloads.batch_family (the random boxes, cylinders and truncated
    ellipsoids of the 64-case batch, other seeds)."""
    from ..workloads import batch_family, random_vertebra, random_vertebra_j2, vertebra_family
    if family == "j2":
        return random_vertebra_j2(int(seed), dx=dx)
    if family == "vertebra":
        return vertebra_family(int(seed), dx=dx)
    if family == "vertebra_mc":
        return vertebra_family(int(seed), dx=dx, fixed_load=True)
    if family == "batch":
        return batch_family(int(seed), dx=dx)
    if family == "pressure":
        return random_vertebra(int(seed), dx=dx)
    raise ValueError(f"unknown family {family!r}")


def generate(seeds, dx, group, tol, max_iterations, to_bf16, mass=None, hops=2, family="pressure"):
    """Training data of the cases ``seeds`` of ``family`` at spacing dx, kept
    on the device: BF16 features (int16 bits), FP32 targets u / s and FP32
    weights (1 on free particles), as lists of CUDA tensors per group of
    ``group`` cases.  The cases are solved by the plain cold-start adaptive DR
    to tol (``mass``: PDBody.solve's DR mass)."""
    import torch
    from .features import features_and_scale
    Xs, Ts, Ws, stats = [], [], [], []
    for g0 in range(0, len(seeds), group):
        t0 = time.perf_counter()
        specs = [_specimen(family, s, dx) for s in seeds[g0:g0 + group]]
        packed, off, body = _group_body(specs)
        rs = body.solve("adaptive", tol=tol, max_iterations=max_iterations, slice_iterations=2048, mass=mass,
                        progress=lambda d, r: log(f"    {d} iterations: {sum(x.converged for x in r)}/{len(r)} converged"))
        u = torch.from_numpy(body.u).cuda()
        F, s = features_and_scale(body, packed, off, hops)
        Xs.append(to_bf16(F))
        del F
        Ts.append((u / s[:, None]).to(torch.float32))
        Ws.append(torch.from_numpy((packed.fixed == 0).astype(np.float32)).cuda())
        its = [r.iterations for r in rs]
        stats.append({"cases": len(specs), "particles": packed.n, "converged": int(sum(r.converged for r in rs)),
                      "iterations_max": int(max(its)), "seconds": time.perf_counter() - t0})
        log(f"  data group {g0 // group}: {len(specs)} cases, {packed.n} particles, iterations "
            f"{min(its)}..{max(its)}, {stats[-1]['seconds']:.1f}s")
        del body
    return Xs, Ts, Ws, stats


def train(net, Xbf, T, Wd, epochs, batch_rows, lr0, lr_min, seed=0):
    """Adam over the device-resident data (BF16 features Xbf, FP32 targets T
    and weights Wd), shuffled once on the device, cosine learning rate."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    order = torch.randperm(Xbf.shape[0], device="cuda", generator=gen)
    Xbf, T, Wd = Xbf[order], T[order], Wd[order]
    del order
    n = Xbf.shape[0]
    nb = max(1, n // batch_rows)
    total = epochs * nb
    wsums = [max(float(Wd[bi * batch_rows:min(n, (bi + 1) * batch_rows)].sum()), 1.0) for bi in range(nb)]
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        ep_loss = 0.0
        for bi in range(nb):
            a, b = bi * batch_rows, min(n, (bi + 1) * batch_rows)
            lr = lr_min + 0.5 * (lr0 - lr_min) * (1 + np.cos(np.pi * step / total))
            loss = net.train_step(Xbf[a:b], T[a:b], Wd[a:b], wsums[bi], lr)
            ep_loss += loss
            step += 1
        hist.append(ep_loss / nb)
        log(f"  epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.3e} ({time.perf_counter() - t0:.1f}s)")
    return hist


def evaluate(net, seeds, dx, group, tol, max_iterations, mass=None, family="pressure", ref_tol=None):
    """Cold vs warm-started batched solves of the held-out cases of ``family``
    (the same DR mass for both: PDBody.solve's ``mass``).  Features and
    inference run on the device (predict_u_device, pdx_set_u_device); one
    untimed inference on the first group loads their GPU code first.  With
    ``ref_tol``, each group is also solved cold to ref_tol, and both solutions'
    distance to that one is reported (the error the tol criterion leaves)."""
    import torch
    from .warmstart import predict_u_device
    out = []
    for g0 in range(0, len(seeds), group):
        specs = [_specimen(family, s, dx) for s in seeds[g0:g0 + group]]
        packed, off, body = _group_body(specs)
        if g0 == 0:
            predict_u_device(net, body, packed, off)
        cold = np.atleast_1d(body.cold_residual())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = body.solve("adaptive", tol=tol, max_iterations=max_iterations, refs=cold, slice_iterations=2048,
                        progress=lambda d, r: None, mass=mass)
        torch.cuda.synchronize()
        t_cold = time.perf_counter() - t0
        u_cold = body.u
        body.u = packed.u0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        body.set_u_device(predict_u_device(net, body, packed, off))
        torch.cuda.synchronize()
        t_inf = time.perf_counter() - t0
        u_pred = body.u
        t_get = time.perf_counter() - t0 - t_inf  # the host copy of the prediction is not part of the warm start
        rw = body.solve("adaptive", tol=tol, max_iterations=max_iterations, refs=cold, slice_iterations=2048,
                        progress=lambda d, r: None, mass=mass)
        torch.cuda.synchronize()
        t_warm = time.perf_counter() - t0 - t_get
        u_warm = body.u
        err_pred = [float(np.linalg.norm(u_pred[a:b] - u_cold[a:b]) / np.linalg.norm(u_cold[a:b]))
                    for a, b in zip(off[:-1], off[1:])]
        err_final = [float(np.linalg.norm(u_warm[a:b] - u_cold[a:b]) / np.linalg.norm(u_cold[a:b]))
                     for a, b in zip(off[:-1], off[1:])]
        ic, iw = [r.iterations for r in rc], [r.iterations for r in rw]
        out.append({"cases": len(specs), "cold_s": t_cold, "warm_s": t_warm, "inference_s": t_inf,
                    "cold_iterations": ic, "warm_iterations": iw,
                    "all_converged": bool(all(r.converged for r in rc) and all(r.converged for r in rw)),
                    "prediction_rel_l2": err_pred, "final_rel_l2_vs_cold": err_final})
        if ref_tol is not None:
            body.u = packed.u0
            rr = body.solve("adaptive", tol=ref_tol, max_iterations=max_iterations, refs=cold,
                            slice_iterations=2048, progress=lambda d, r: None, mass=mass)
            u_ref = body.u
            dist = lambda u: [float(np.linalg.norm(u[a:b] - u_ref[a:b]) / np.linalg.norm(u_ref[a:b]))
                              for a, b in zip(off[:-1], off[1:])]
            out[-1].update({"ref_tol": ref_tol, "ref_converged": bool(all(r.converged for r in rr)),
                            "cold_rel_l2_vs_ref": dist(u_cold), "warm_rel_l2_vs_ref": dist(u_warm)})
        log(f"  held-out group {g0 // group}: cold {t_cold:.2f}s ({max(ic)} it max, {np.mean(ic):.0f} mean), "
            f"warm {t_warm:.2f}s incl. {t_inf:.2f}s inference ({max(iw)} it max, {np.mean(iw):.0f} mean), "
            f"prediction error {np.median(err_pred):.3f} median")
        del body
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=1024)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--dx", type=float, nargs="+", default=[0.8, 0.6],
                    help="training spacings; the cases are split evenly over them")
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iterations", type=int, default=60000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch-rows", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--seed", type=int, default=5000)
    ap.add_argument("--mass", default=None, choices=["gershgorin", "spectral"],
                    help="adaptive-DR mass of the held-out cold and warm solves")
    ap.add_argument("--data-mass", default=None, choices=["gershgorin", "spectral"],
                    help="adaptive-DR mass of the training-data solves")
    ap.add_argument("--hops", type=int, default=2, help="neighbour-mean hops of the input features (0..4)")
    ap.add_argument("--family", default="pressure", choices=FAMILIES,
                    help="pressure: Neo-Hookean bodies under a top pressure; j2: J2 bodies compressed between platens; "
                         "vertebra: the headline's shape with random cancellous fields and loads; "
                         "vertebra_mc: the headline's shape and load with random cancellous fields; "
                         "batch: random boxes, cylinders and truncated ellipsoids of 10-200 k particles")
    ap.add_argument("--ref-tol", type=float, default=None,
                    help="also solve each held-out group cold to this tolerance and report both solutions' distance")
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)

    import torch
    assert torch.cuda.is_available(), "training runs on the GPU"
    from .features import feature_width
    from .net import DIMS, WarmStartNet
    t_all = time.perf_counter()
    train_seeds = np.arange(args.seed, args.seed + args.train)
    held_seeds = np.arange(args.seed + 100000, args.seed + 100000 + args.heldout)
    parts = np.array_split(train_seeds, len(args.dx))
    net = WarmStartNet(dims=(feature_width(args.hops),) + DIMS[1:], seed=args.seed)
    Xs, Ts, Ws, gstats = [], [], [], []
    t_gen = time.perf_counter()
    for dx, seeds in zip(args.dx, parts):
        log(f"generating {len(seeds)} training cases at dx {dx}")
        X_, T_, W_, g_ = generate(seeds, dx, args.group, args.tol, args.max_iterations, net.to_bf16,
                                  mass=args.data_mass, hops=args.hops, family=args.family)
        Xs += X_
        Ts += T_
        Ws += W_
        gstats += g_
    t_gen = time.perf_counter() - t_gen
    Xbf, T, Wd = torch.cat(Xs), torch.cat(Ts), torch.cat(Ws)
    del Xs, Ts, Ws
    log(f"training set: {Xbf.shape[0]} particles, {int(Wd.sum())} free ({t_gen:.0f}s to generate)")
    t_tr = time.perf_counter()
    hist = train(net, Xbf, T, Wd, args.epochs, args.batch_rows, args.lr, args.lr_min)
    t_tr = time.perf_counter() - t_tr
    del Xbf, T, Wd
    if args.out:
        net.save(args.out, dx=np.array(args.dx), train_cases=args.train, epochs=args.epochs, final_loss=hist[-1],
                 family=args.family)
        log(f"weights -> {args.out}")
    log(f"evaluating on {args.heldout} held-out cases")
    ev = evaluate(net, held_seeds, args.dx[0], args.group, args.tol, args.max_iterations, mass=args.mass,
                  family=args.family, ref_tol=args.ref_tol)
    cold = sum(e["cold_s"] for e in ev)
    warm = sum(e["warm_s"] for e in ev)
    ic = [i for e in ev for i in e["cold_iterations"]]
    iw = [i for e in ev for i in e["warm_iterations"]]
    rep = {"train_cases": args.train, "heldout_cases": args.heldout, "dx": list(args.dx), "epochs": args.epochs,
           "hops": args.hops, "family": args.family,
           "train_rows": int(sum(g["particles"] for g in gstats)), "data_s": t_gen, "train_s": t_tr,
           "mass": args.mass, "loss_history": hist, "data_groups": gstats, "heldout": ev,
           "summary": {"cold_s": cold, "warm_s": warm, "speedup_wall": cold / max(warm, 1e-9),
                       "cold_iterations_mean": float(np.mean(ic)), "warm_iterations_mean": float(np.mean(iw)),
                       "speedup_iterations_mean": float(np.mean(ic) / max(np.mean(iw), 1e-9)),
                       "total_s": time.perf_counter() - t_all}}
    log(json.dumps(rep["summary"]))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(rep, fh)
    return rep


if __name__ == "__main__":
    main()
