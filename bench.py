"""Benchmark: body-steps/s of the per-step contact pipeline on one MI355X.

Default workload (cfg3, the configuration the headline target is quoted on): 262,144
spheres with log-normal radii (median 0.05, sigma 0.25), poured into a box with 5 static
planes and settled, 3D Coulomb mu = 0.4, CoR = 0, Verlet, dt = 0.005, FB residual
tolerance 1e-6 (max over contacts of ||phi||^2; max ||phi|| is reported beside it),
FP32 sweeps + FP64 residual/polish, stepped with the learned warm start (a 16-128-128-3
BF16 perceptron trained on the device from samples of the transported-cache solver).
From the same settled state the plain coloured Gauss-Seidel (cold start, as the
reference's 3D geometric map) and the transported impulse cache are measured as well
(learned and transported step by step on the same checkpointed states of one trajectory);
the sweep and body-step ratios go into config.warm_start_comparison.

Other workloads (--config): cfg1, 1,048,576 RSA circles (2D, mu 0.3, dt 0.005, 100
steps); cfg2, the pile with the plain coloured GS as the stepping mode; cfg4, an ensemble
of 512 independent piles (learned warm start, per-scene p50/p99 latency).

A step is one GeometricImpactFrictionMap step over all bodies (predictor, broad and
narrow phase, contact setup, warm start, colouring, coloured Gauss-Seidel to tolerance,
final update) with the state resident in HBM. --steps sets only the number of timed steps
(each ends with a stream synchronisation); everything else (settling, sample collection,
training, the comparison) is untimed and fixed by its own options. value = bodies * steps
/ wall time of the timed steps. roofline: the dominant kernel by device time, its
algorithmic bytes per launch (scisim_amd/roofline.py) over its mean launch duration from
HIP events on the library's stream; traffic is null unless --traffic gives HBM bytes per
launch measured by rocprofv3 --pmc in the same command. cpu_baseline: the FP64 oracle (the
project's restatement of the reference path: natural-order Gauss-Seidel from a cold
start, as the reference) on the state in the middle of the timed steps, on all host cores
of this process (parallel residual evaluation and per-contact set-up; the sweep is
sequential as in the reference) and on one thread, each on a bounded sample.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def make_scene(cfg, seed):
    from scisim_amd import scenes
    if cfg == "cfg1":
        s = scenes.ball2d_rsa(n=1048576, rmin=0.02, rmax=0.06, phi=0.55, seed=seed)
        prm = dict(dt=0.005, g=[0.0, -9.81], mu=0.3, CoR=0.0, tol=1e-6, max_iters=50000, eval_every=1,
                   integrator=0)
        desc = {"workload": "ball2d_rsa_1M", "bodies": 1048576, "dim": 2, "radii": "U[0.02,0.06]",
                "packing": "RSA phi=0.55", "planes": 4, "mu": 0.3, "CoR": 0.0, "dt": 0.005,
                "integrator": "symplectic_euler", "tol": 1e-6}
    elif cfg == "cfg4":
        s, body_off, plane_off, mus, counts = scenes.ensemble_piles(nscenes=512, seed=seed)
        s["ensemble"] = (body_off, plane_off, mus)
        prm = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.5, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1,
                   integrator=1)
        desc = {"workload": "spheres3d_ensemble_512", "scenes": 512, "bodies": int(body_off[-1]),
                "bodies_per_scene": "loguniform[500,20000] (min %d, max %d)" % (counts.min(), counts.max()),
                "dim": 3, "radii": "U[0.03,0.07]", "mu": "U[0.1,0.8] per scene",
                "packing": "poured piles (sequential deposition), settled", "planes_per_scene": 5, "CoR": 0.0,
                "dt": 0.005, "integrator": "verlet", "tol": 1e-6}
    else:
        s = scenes.spheres_pile(n=262144, rmed=0.05, sigma=0.25, seed=seed)
        prm = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, CoR=0.0, tol=1e-6, max_iters=50000, eval_every=1,
                   integrator=1)
        desc = {"workload": "spheres3d_pile_262k" + ("_learned_warm_start" if cfg == "cfg3" else "_plain_gs"),
                "bodies": 262144, "dim": 3, "radii": "lognormal(0.05,0.25)",
                "packing": "poured pile (sequential deposition), settled", "planes": 5, "mu": 0.4, "CoR": 0.0,
                "dt": 0.005, "integrator": "verlet", "tol": 1e-6}
    return s, prm, desc


def host_threads():
    """Host threads this process may use: OMP_NUM_THREADS if set (the GPU hosts set it to
    the job's CPU share), else the CPUs of its affinity mask."""
    try:
        env = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        env = 0
    return env if env > 0 else len(os.sched_getaffinity(0))


def collect_and_train(ctx, s, prm, precision, args, log):
    """Samples of `collect` transported-cache steps (a fraction of the contacts of each
    step, seeded), then the perceptron trained on them on the device."""
    import scisim_amd as scx
    from scisim_amd import warmstart
    rng = np.random.default_rng(args.seed)
    tp = scx.step_params(precision=precision, **dict(prm, warm_start=scx.WARM_TRANSPORT))
    Xs, Ts = [], []
    frac = 0.25 if s["r"].size <= 300000 else 0.05  # samples kept per step (host memory)
    t0 = time.perf_counter()
    for k in range(args.collect):
        st = ctx.step(tp)
        X, T = ctx.samples()
        keep = rng.random(X.shape[0]) < frac
        Xs.append(X[keep])
        Ts.append(T[keep])
        log("collect", k, "contacts", st["ncontacts"], "sweeps", st["sweeps"], "%.1fs" % (time.perf_counter() - t0))
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    ctx.set_warm_model(warmstart.init_params(seed=args.seed))
    t0 = time.perf_counter()
    loss = ctx.train_warm_model(X, T, epochs=args.epochs, batch=2048, lr=1e-3, seed=args.seed)
    train_s = time.perf_counter() - t0
    log("trained on", X.shape[0], "samples, loss", loss, "%.1fs" % train_s)
    return {"training_samples": int(X.shape[0]), "collect_steps": args.collect, "epochs": args.epochs,
            "train_s": train_s, "final_mse": loss}


def compare_warm_starts(ctx, s, prm, precision, args, log):
    """The transported cache and the learned warm start compared step by step on one
    trajectory, and the plain coloured Gauss-Seidel (cold start) from its first state.

    Two warm starts that both meet the tolerance leave slightly different states, and the
    pile's trajectories diverge within a few steps, so means over two independent runs of
    20 steps differ by about 10 % from that alone. Instead each of `compare_steps` states of
    one reference trajectory (advanced by transported-cache steps) is restored from a
    checkpoint (scx_save_checkpoint: bodies, impulse cache, model) and stepped once with
    each warm start: paired sweep counts of the same step. The context ends in the state
    it started from."""
    import scisim_amd as scx

    def params(mode):
        return scx.step_params(precision=precision, **dict(prm, warm_start=mode))

    img0 = ctx.save_checkpoint()
    img = img0
    out = {}
    paired = {"transport": [], "learned": []}
    secs = {"transport": 0.0, "learned": 0.0}
    ok = {"transport": 0, "learned": 0}
    for k in range(args.compare_steps):
        # learned first, then the transported-cache step that also advances the trajectory
        for name, mode in (("learned", scx.WARM_LEARNED), ("transport", scx.WARM_TRANSPORT)):
            ctx.load_checkpoint(img)
            t0 = time.perf_counter()
            st = ctx.step(params(mode))
            secs[name] += time.perf_counter() - t0
            paired[name].append(st["sweeps"])
            ok[name] += int(st["succeeded"])
        img = ctx.save_checkpoint()
        log("compare step", k, "sweeps learned", paired["learned"][-1], "transport", paired["transport"][-1])
    for name in ("transport", "learned"):
        if paired[name]:
            out[name] = {"steps": len(paired[name]), "mean_sweeps": float(np.mean(paired[name])),
                         "steps_at_tolerance": ok[name], "body_steps_per_s": s["r"].size * len(paired[name]) / secs[name]}
    if args.cold_steps > 0:
        ctx.load_checkpoint(img0)
        sw, okc = [], 0
        t0 = time.perf_counter()
        for k in range(args.cold_steps):
            st = ctx.step(params(scx.WARM_NONE))
            sw.append(st["sweeps"])
            okc += int(st["succeeded"])
            log("cold step", k, "sweeps", st["sweeps"], "%.1fs" % (time.perf_counter() - t0))
        out["cold"] = {"steps": args.cold_steps, "mean_sweeps": float(np.mean(sw)), "steps_at_tolerance": okc,
                       "body_steps_per_s": s["r"].size * args.cold_steps / (time.perf_counter() - t0)}
    ctx.load_checkpoint(img0)
    if paired["learned"]:
        r = np.asarray(paired["learned"], dtype=np.float64) / np.maximum(1.0, np.asarray(paired["transport"]))
        se = float(r.std(ddof=1) / np.sqrt(r.size)) if r.size > 1 else float("nan")
        out["paired_sweep_ratio_learned_over_transport"] = {"mean": float(r.mean()), "two_se": 2.0 * se,
                                                            "steps": int(r.size)}
    if "cold" in out and "learned" in out:
        out["sweeps_cold_over_learned"] = out["cold"]["mean_sweeps"] / max(1e-9, out["learned"]["mean_sweeps"])
        out["body_steps_learned_over_cold"] = (out["learned"]["body_steps_per_s"] /
                                               out["cold"]["body_steps_per_s"])
    if "transport" in out and "learned" in out:
        out["sweeps_learned_over_transport"] = out["learned"]["mean_sweeps"] / max(1e-9,
                                                                                  out["transport"]["mean_sweeps"])
        out["body_steps_learned_over_transport"] = (out["learned"]["body_steps_per_s"] /
                                                    out["transport"]["body_steps_per_s"])
    if "cold" in out and "transport" in out:
        out["sweeps_cold_over_transport"] = out["cold"]["mean_sweeps"] / max(1e-9, out["transport"]["mean_sweeps"])
    for name in ("transport", "learned", "cold"):
        if name in out:
            log(name, out[name])
    return out


def cpu_baseline(s, state, prm, budget_s, sweeps_hint, threads):
    """The FP64 oracle (natural-order Gauss-Seidel, cold start as the reference's maps,
    `threads` threads for its parallel parts) on the state in the middle of the timed
    steps: one step, its sweep limit cut so that it fits the budget. If that step reaches
    the tolerance the rate is measured; otherwise the step time is scaled linearly in the
    sweeps to `sweeps_hint` (the GPU's cold sweep count from the same kind of state: the
    natural and the coloured orders need about the same sweeps)."""
    import oracle
    oracle.set_threads(threads)
    q, v, w = state
    D = s["dim"]

    def one(k0, k1, p0, p1, mu, max_iters):
        sim = oracle.Sim(D, q[D * k0:D * k1], v[D * k0:D * k1], s["r"][k0:k1], s["m"][k0:k1], s["px"][p0:p1],
                         s["pn"][p0:p1], w=None if w is None else w[3 * k0:3 * k1],
                         inertia=None if s.get("inertia") is None else s["inertia"][k0:k1])
        t0 = time.perf_counter()
        st = sim.step(order=0, **dict(prm, mu=mu, warm_start=0, max_iters=max_iters))
        return time.perf_counter() - t0, st

    if "ensemble" in s:
        # one step of each scene alone, scene after scene, to the tolerance, within the budget
        body_off, plane_off, mus = s["ensemble"]
        t_all = done = k = ok = 0
        while k < len(mus) and t_all < budget_s:
            t, st = one(int(body_off[k]), int(body_off[k + 1]), int(plane_off[k]), int(plane_off[k + 1]),
                        float(mus[k]), prm["max_iters"])
            t_all += t
            done += int(body_off[k + 1] - body_off[k])
            ok += int(st["succeeded"])
            k += 1
        return {"value": done / t_all, "unit": "body-steps/s", "cores": threads, "kind": "port",
                "sample": "FP64 oracle, natural-order GS, one step of each of the first %d scenes alone from the "
                          "state in the middle of the timed steps (%d bodies, %d of them reached the tolerance), "
                          "%.1f s" % (k, done, ok, t_all)}
    n = s["r"].size
    P = s["px"].size // D
    # set-up plus one sweep, then sixteen: the time per sweep sets the limit of the step
    # that uses the rest of the budget
    t1, st1 = one(0, n, 0, P, prm["mu"], 1)
    t, st, m = t1, st1, 1
    if not st1["succeeded"] and t1 < budget_s:
        t16, st16 = one(0, n, 0, P, prm["mu"], 16)
        t, st, m = t16, st16, 16
        if not st16["succeeded"] and t1 + t16 < budget_s:
            per = max((t16 - t1) / 15.0, 1e-7)
            m = int(min(prm["max_iters"], max(16, (budget_s - t1 - t16 - t1) / per)))
            t, st = one(0, n, 0, P, prm["mu"], m)
    if st["succeeded"]:
        return {"value": n / t, "unit": "body-steps/s", "cores": threads, "kind": "port",
                "sample": "FP64 oracle on %d thread(s), natural-order GS from a cold start, one step from the state "
                          "in the middle of the timed steps, to the tolerance in %d sweeps, %.1f s"
                          % (threads, st["sweeps"], t)}
    per = max((t - t1) / max(m - 1, 1), 1e-9)
    est = (t1 - per) + per * max(sweeps_hint, m)
    return {"value": n / est, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": "FP64 oracle on %d thread(s), natural-order GS from a cold start, one step from the state in "
                      "the middle of the timed steps cut at %d sweeps (%.1f s; set-up %.2f s, %.2f ms per sweep), "
                      "scaled to %.0f sweeps (the GPU's cold-start sweeps on this workload): %.1f s per step"
                      % (threads, m, t, t1 - per, 1e3 * per, sweeps_hint, est)}


def read_traffic(path):
    """HBM bytes per launch of the dominant kernel, from a JSON file written in the same
    command by the rocprofv3 --pmc passes (keys: bytes_per_launch, and how it was made)."""
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: cfg1 100, the workload's step count; the piles 20)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cfg3", choices=["cfg1", "cfg2", "cfg3", "cfg4"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed settling steps of the piles under the transported cache (default: cfg2, cfg3 "
                         "300; cfg4 200)")
    ap.add_argument("--collect", type=int, default=None,
                    help="steps whose samples train the warm start (default: cfg3 40, cfg4 10)")
    ap.add_argument("--epochs", type=int, default=12, help="training passes over the samples")
    ap.add_argument("--compare-steps", type=int, default=20,
                    help="cfg3: steps of the transported-cache and learned warm starts in the comparison")
    ap.add_argument("--cold-steps", type=int, default=2,
                    help="cfg2/cfg3: steps of the plain coloured GS (cold start) in the comparison")
    ap.add_argument("--warm", default=None, choices=["none", "cache", "normal", "transport"],
                    help="impulse cache warm start of the timed steps (cfg1, cfg2; default none)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--max-iters", type=int, default=None, help="override the sweep limit (profiling)")
    ap.add_argument("--eval-every", type=int, default=None, help="override the residual period (profiling)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of each CPU baseline run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=None, metavar="JSON",
                    help="HBM bytes per launch of the dominant kernel measured by rocprofv3 --pmc in this command")
    ap.add_argument("--autotune", default="auto", choices=["auto", "on", "off"],
                    help="online autotuning of the persistent sweep launch width (scx_set_autotune); auto: on "
                         "for the piles and the ensemble, whose untimed settling steps train it, off for the "
                         "circle packing, whose active set grows through new launch classes every few steps")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 plain device copies of BYTES (the reference kernel that scales the HBM "
                         "byte counters of rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.config == "cfg1" else 20
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU only (requested --gpus %d); the library has no multi-GPU "
              "path" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch
        src = torch.empty(args.copy_calibration // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "copies": 5}))
        return

    import scisim_amd as scx
    from scisim_amd import roofline as rl

    t_start = time.perf_counter()

    def log(*a):
        if args.verbose:
            print("[%6.1fs]" % (time.perf_counter() - t_start), *a, file=sys.stderr, flush=True)

    s, prm, desc = make_scene(args.config, args.seed)
    dim = s["dim"]
    n = s["r"].size
    log("scene ready", desc["workload"], n)
    ctx = scx.Context(dim, n, max_planes=max(8, s["px"].size // dim), device=0)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s.get("inertia"))
    if "ensemble" in s:
        ctx.set_scenes(*s["ensemble"])
    if args.autotune == "auto":
        args.autotune = "off" if args.config == "cfg1" else "on"
    ctx.set_autotune(args.autotune == "on")
    desc["autotune_enabled"] = args.autotune == "on"
    precision = scx.FP32_POLISH if args.precision == "fp32" else scx.FP64
    it = 1
    pile = args.config in ("cfg2", "cfg3", "cfg4")
    if args.settle is None:
        # the single pile needs about 225 steps before its steps meet the tolerance within
        # the settling sweep limit; 300 leaves it settled (the ensemble's piles are smaller)
        args.settle = 200 if args.config == "cfg4" else 300
    if args.collect is None:
        args.collect = 10 if args.config == "cfg4" else 40
    if pile and args.settle > 0:
        # the deposited pile settles under the transported impulse cache (untimed)
        sp = scx.step_params(precision=precision, **dict(prm, max_iters=1000, warm_start=scx.WARM_TRANSPORT))
        t0 = time.perf_counter()
        for k in range(args.settle):
            st = ctx.step(sp)
            if k % 25 == 24:
                log("settle", k, "contacts", st["ncontacts"], "sweeps", st["sweeps"],
                    "%.1fs" % (time.perf_counter() - t0))
        it += args.settle
        desc["settle_steps"] = args.settle
    if args.max_iters is not None:
        prm["max_iters"] = args.max_iters
    if args.eval_every is not None:
        prm["eval_every"] = args.eval_every
    training = comparison = None
    warm = args.warm or "none"
    if args.config in ("cfg3", "cfg4"):
        training = collect_and_train(ctx, s, prm, precision, args, log)
        warm = "learned"
    if args.config in ("cfg2", "cfg3"):
        comparison = compare_warm_starts(ctx, s, prm, precision, args, log) if args.config == "cfg3" else None
        if args.config == "cfg2" and args.cold_steps > 0:
            desc["note"] = "plain coloured GS from a cold start in every timed step"
    prm["warm_start"] = {"none": scx.WARM_NONE, "cache": scx.WARM_CACHE, "normal": scx.WARM_NORMAL,
                         "transport": scx.WARM_TRANSPORT, "learned": scx.WARM_LEARNED}[warm]
    desc["warm_start"] = warm
    params = scx.step_params(precision=precision, iteration=it, **prm)
    for _ in range(args.warmup):
        ctx.step(params)
    log("warmup done")
    # timed region: every step synchronises at its end; per-phase device times from HIP events
    ctx.set_timing(True)
    phase_ms = {}
    pairs_bytes = gs_bytes = bp_bytes = 0.0
    contacts, sweeps, errors = [], [], []
    evals_full = evals_skipped = 0
    fp32 = precision == scx.FP32_POLISH
    step_ms = []
    converged = 0
    sample, sample_s = None, 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        if k == args.steps // 2:
            # the state the CPU baseline starts from: the middle of the timed steps (the copy
            # is taken out of the timed wall time)
            tc = time.perf_counter()
            sample = ctx.get_state()
            sample_s += time.perf_counter() - tc
        ts = time.perf_counter()
        st = ctx.step(params)
        step_ms.append(1000.0 * (time.perf_counter() - ts))
        tm = ctx.timing()
        for name, ms in tm:
            phase_ms[name] = phase_ms.get(name, 0.0) + ms
        C = st["ncontacts"]
        bp_bytes += rl.broadphase_bytes(dim, n, st["npairs"])
        for write_pass in (False, True):
            pairs_bytes += rl.pairs_launch_bytes(dim, n, st["ncells"], st["nentries"], st["npairs"], write_pass)
        s32 = st["sweeps_fp32"] if fp32 else 0
        s64 = st["sweeps"] - s32
        if "ensemble" in s:
            # a scene is frozen once it meets the tolerance: count each scene's contacts only
            # for the sweeps it took (per-scene stats read outside the timed wall time)
            tc = time.perf_counter()
            cnt_k, _, ssw_k = ctx.scene_stats()
            sample_s += time.perf_counter() - tc
            cnt_k = cnt_k.astype(np.float64)
            sw32 = np.minimum(ssw_k, s32).astype(np.float64)
            sw64 = np.clip(ssw_k.astype(np.int64) - s32, 0, s64).astype(np.float64)
            if fp32:
                gs_bytes += float(np.sum(cnt_k * (sw32 * rl.gs_sweep_bytes(dim, True) +
                                                  (sw32 + 1) * rl.gs_eval_bytes(dim, True))))
            gs_bytes += float(np.sum(cnt_k * (sw64 * rl.gs_sweep_bytes(dim, False) +
                                              (sw64 + 1) * rl.gs_eval_bytes(dim, False))))
        else:
            # evaluations as run: the FP32 phase replaces most of them by a check of the
            # contacts the last full one listed (a few hundred contacts, not counted)
            ev = ctx.sweep_counters()
            evals_full += ev["full"]
            evals_skipped += ev["skipped"]
            e32 = ev["full_fp32"] if fp32 else 0
            if fp32:
                gs_bytes += rl.gs_launch_bytes(dim, C, s32, e32, True)
            gs_bytes += rl.gs_launch_bytes(dim, C, s64, ev["full"] - e32, False)
        contacts.append(C)
        sweeps.append(st["sweeps"])
        errors.append(st["error"])
        converged += int(st["succeeded"])
        log("step", k, st, " ".join("%s %.3f" % (name, ms) for name, ms in tm if ms > 0.02))
    wall = time.perf_counter() - t0 - sample_s
    ctx.set_timing(False)
    ms_per_step = 1000.0 * wall / args.steps
    value = n * args.steps / wall
    # dominant kernel: the neighbour sweep (two launches per step) or the Gauss-Seidel
    # sweeps (one FP32 and one FP64 launch per step), by device time
    pairs_ms = phase_ms.get("pairs_count", 0.0) + phase_ms.get("pairs_write", 0.0)
    gs_ms = phase_ms.get("gs_fp32", 0.0) + phase_ms.get("gs_fp64", 0.0)
    if pairs_ms >= gs_ms:
        kernel, launches, kbytes, kms = "k_pairs (grid neighbour sweep + FP64 narrow phase)", 2, pairs_bytes, pairs_ms
    else:
        kernel, launches, kbytes, kms = "k_gs (coloured Gauss-Seidel, persistent)", 1 + int(fp32), gs_bytes, gs_ms
    traffic = None
    if args.traffic:
        traffic = read_traffic(args.traffic).get("bytes_per_launch")
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    # the broad phase as a whole: predictor, binning and both neighbour sweeps
    bp_ms = sum(phase_ms.get(k, 0.0) for k in ("predict", "grid", "pairs_count", "pairs_scan", "pairs_write"))
    bp_gbs = bp_bytes / (bp_ms * 1e-3) / 1e9 if bp_ms > 0 else 0.0
    broadphase = {"bound": "hbm", "achieved": bp_gbs, "peak": rl.HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": bp_gbs / rl.HBM_PEAK_GBS, "bytes_per_step": bp_bytes / args.steps,
                  "us_per_step": 1000.0 * bp_ms / args.steps,
                  "phases": "predict + grid + pairs_count + pairs_scan + pairs_write",
                  "bytes": "172 B/body + 8 B/pair (2D), 208 B/body + 8 B/pair (3D)"}
    tuned = ctx.autotune_table()
    if tuned:
        desc["autotune"] = [[e["phase"], e["contacts_half_octave"], e["exponent"], e["blocks"],
                             round(e["us_per_sweep"], 2), e["settled"]] for e in tuned]
    if "ensemble" in s:
        cnt, err, ssw = ctx.scene_stats()
        desc["step_latency_ms_p50"] = float(np.percentile(step_ms, 50))
        desc["step_latency_ms_p99"] = float(np.percentile(step_ms, 99))
        desc["scene_contacts_p50_p99"] = [float(np.percentile(cnt, 50)), float(np.percentile(cnt, 99))]
        desc["scene_residual_max"] = float(err.max())
        desc["scenes_at_tolerance_last_step"] = int((err < prm["tol"]).sum())
        # per-scene latency of the last step: its share of the sweeps (a scene is frozen once it
        # meets the tolerance) plus everything outside the sweeps
        gs_last = sum(ms for name, ms in tm if name.startswith("gs_"))
        other = step_ms[-1] - gs_last
        lat = other + gs_last * ssw / max(1, sweeps[-1])
        desc["scene_sweeps_p50_p99"] = [float(np.percentile(ssw, 50)), float(np.percentile(ssw, 99))]
        desc["scene_latency_ms_p50_p99_last_step"] = [float(np.percentile(lat, 50)), float(np.percentile(lat, 99))]
    emax = float(np.max(errors)) if errors else 0.0
    config = dict(desc, mean_contacts=float(np.mean(contacts)), mean_sweeps=float(np.mean(sweeps)),
                  max_iters=prm["max_iters"], steps_at_tolerance=converged,
                  residual_max_phi_squared=emax, residual_max_phi=float(np.sqrt(emax)),
                  phase_ms_per_step={k: v / args.steps for k, v in phase_ms.items()})
    if "ensemble" not in s:
        config["residual_evaluations_per_step"] = {"full": evals_full / args.steps,
                                                   "skipped": evals_skipped / args.steps}
    if training:
        config["warm_start_training"] = training
    if comparison:
        config["warm_start_comparison"] = comparison
    out = {
        "metric": "body-steps/sec at FB residual<=1e-6",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps + f64 residual/polish" if fp32 else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": rl.HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / rl.HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kbytes / (launches * args.steps),
                     "us_per_launch": 1000.0 * kms / (launches * args.steps)},
        "broadphase_roofline": broadphase,
    }
    ctx.close()
    if not args.no_cpu:
        # the GPU's cold-start sweeps on this workload scale a CPU step cut short of the
        # tolerance (cfg3: measured in the comparison; otherwise the timed steps' mean)
        hint = float(np.mean(sweeps))
        if comparison and "cold" in comparison:
            hint = comparison["cold"]["mean_sweeps"]
        cores = host_threads()
        base = cpu_baseline(s, sample, prm, args.cpu_budget, hint, cores)
        if cores > 1:
            base["one_thread"] = cpu_baseline(s, sample, prm, args.cpu_budget, hint, 1)
        out["cpu_baseline"] = base
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
