"""Benchmark: particle-steps/s of the plain 2D SPH step on one MI355X.

Workload: the 1,048,576-particle dam break (1024 x 1024 block in a 4 x 2 box,
dx = 0.25/128, h = 1.3 dx, Tait EOS, Monaghan viscosity, reflective walls),
FP32, synthetic seeded initial state. One step = forces + integration + cell
keys -> stable radix sort -> permutation + cell ranges -> density + EOS +
neighbor list (the fused pipeline of sph_step).

Prints one JSON line. `roofline` is for the kernel with the largest share of
the timed region, from hipEvents recorded on the context stream around every
launch; `stages` gives every kernel's share. `cpu_baseline` times the
reference-style OpenMP CPU step (the oracle's baseline build) on a bounded
sample of the same workload.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_particle(stage, k, occ, passes, variant, resort="full", movers_frac=0.0,
                           flags_in_force=False):
    """Algorithmic HBM bytes per particle of one launch of each stage (FP32
    state, u64 keys, u32 indices; neighbor and halo re-reads counted as cache
    hits). flags_in_force: the force kernel reports the re-sort's movers
    (reading the previous key; flag bits and the mover list). None for a stage
    without a model."""
    if variant == "split_block":
        # density: R r 8, W rho,pr 8; force: R r,v,rho,pr 24, W a 8, T r,v 16, key 8
        phys = {"approximate_density": 16.0, "force_computation": 56.0}
    elif variant == "fused_block":
        # R r,v 16; W rho,pr 8, a 8, T r,v 16, key 8
        phys = {"block_step": 56.0}
    else:
        # density: R r 8, W the packed (rho, pr) copy 8, count 4, list 4k;
        # force: R r,v 16, rho,pr 8, count 4, list 4k, T 16, key 8 (+ R
        # previous key 8, W one flag bit and the movers' list entries 12 f when
        # it reports the movers); the separate rho, pr and a arrays (8 + 8) are
        # written only in the last step of a sph_step call; neighbors come from
        # the packed copies (cache hits)
        phys = {"approximate_density": 20.0 + 4.0 * k,
                "force_computation": 52.0 + 4.0 * k
                + ((8.125 + 12.0 * movers_frac) if flags_in_force else 0.0)}
    if resort == "full":
        sort = {
            # R new keys (8) + previous sorted keys (8); clear previous ranges (8/occ. cell)
            "sort_histogram": 16.0 + 8.0 * occ,
            # first pass R key 8, W key+idx 12; later passes R 12, W 12 (mean per launch)
            "sort_pass": (20.0 + 24.0 * (passes - 1)) / passes,
            # R perm 4, sorted key 8, T r,v 16, pid 4; W r,v 16 (+ packed copy
            # 16), pid 4; ranges 8/occ. cell
            "neighbor_search": 68.0 + 8.0 * occ,
        }
    else:
        f = movers_frac
        sort = {
            # up to 4096 movers: R tile-first keys 8 and W tile prefixes and
            # window starts 8 per 256-particle tile; R list key+idx 12, W sorted
            # key+idx 12 per mover (+ R new and old keys 16, W flag bit and list
            # 12 f when no kernel before reported the movers)
            "movers": 16.0 / 256.0 + 24.0 * f
            + (0.0 if flags_in_force else 16.125 + 12.0 * f),
            # R old key 8, T r,v 16, pid 4; W r,v 16 (only the packed copy
            # inside sph_step), pid 4, key 8 (the permutation 4 only in the
            # last step of a call); clear the old ranges 8 occ
            "merge": 56.0 + 8.0 * occ,
            # ranges: R key 8, W 8 occ
            "neighbor_search": 8.0 + 8.0 * occ,
        }
    return {**phys, **sort}.get(stage)


STAGE_KERNEL = {"approximate_density": "k_density", "force_computation": "k_forces",
                "movers": "k_movers", "merge": "k_merge", "neighbor_search": "k_ranges"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at the 1M workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel),
    or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_dam_break_1m.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            for item in line.split()[1:]:
                name, _, val = item.partition("=")
                if name == kernel:
                    return int(val)
    return None


def measured_issue(kernel):
    """VALU and texture-address busy fractions of `kernel` at the 1M workload,
    from the committed rocprofv3 counter summary, or None: the step kernels are
    limited by instruction issue and the address path rather than by HBM."""
    path = os.path.join(ROOT, "profiles", "issue_bound_dam_break_1m.md")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        for key in ("valu_busy", "ta_busy"):
            if line.startswith(key + ":"):
                for item in line.split()[1:]:
                    name, _, val = item.partition("=")
                    if name == kernel:
                        out[key] = float(val)
    return out or None


def cpu_baseline(sph, scene, budget_s):
    from oracle import oracle as orc
    threads, t, _ = orc.baseline_run(scene.params, scene.rx, scene.ry, scene.vx, scene.vy, 2)
    per_step = t["main_loop"] / 2.0
    steps = int(max(2, min(60, budget_s / max(per_step, 1e-6))))
    log(f"cpu baseline: {per_step:.3f} s/step on {threads} threads -> sampling {steps} steps")
    threads, t, _ = orc.baseline_run(scene.params, scene.rx, scene.ry, scene.vx, scene.vy, steps)
    value = scene.n * steps / t["main_loop"]
    return {
        "value": value,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{scene.name}: {scene.n} particles x {steps} steps of the reference-style "
                  f"OpenMP mp-lists/Morton CPU step (serial sort/gather/ranges/scan/EOS/"
                  f"integrator), -O3 -ffast-math",
        "stage_seconds": {k: round(v, 4) for k, v in t.items()},
    }


def bench_batched(sph, steps, trial_steps=10):
    """The 256-scene batched workload: the library's autotuner picks cell size,
    tile and re-sort from its built-in candidates, then `steps` graph-replayed
    steps are timed (host clock) and once more with per-launch hipEvents."""
    b = sph.scenes.batched_dam_breaks()
    n = b.n
    log(f"batched: {len(b.counts)} scenes, n={n}")
    ctx = sph.Batch(b.params, b.counts)
    ctx.upload(b.rx, b.ry, b.vx, b.vy)
    a0 = time.perf_counter()
    best, times = ctx.autotune(trial_steps=trial_steps)
    tune_s = time.perf_counter() - a0
    log(f"batched autotune {tune_s:.2f} s -> {best}")
    ctx.prepare()
    ctx.set_graph(True)
    ctx.step(10)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    value = n * steps / wall
    ctx.set_timing(True)
    ctx.step(min(steps, 50))
    ctx.synchronize()
    stage = ctx.stage_times()
    ctx.set_timing(False)
    ev_steps = min(steps, 50)
    mfrac = ctx.movers() / n
    keys = ctx.sorted_keys()
    occ = float(np.unique(keys).size) / n
    ctx.density_pressure()
    k = float(ctx.ell_counts().mean())
    passes = (int(ctx.grid().key_bits) + 7) // 8
    resort = "incremental" if best[2] == sph.RESORT_INCREMENTAL else "full"
    stages = {}
    step_bytes = 0.0
    kern_ms = 0.0
    for name, (ms, launches) in stage.items():
        if launches == 0:
            continue
        per = ms / launches
        bpp = alg_bytes_per_particle(name, k, occ, passes, "particle", resort, mfrac)
        stages[name] = {"launches": launches, "us_per_launch": round(per * 1e3, 3)}
        kern_ms += ms
        if bpp is None:
            continue
        stages[name]["alg_bytes_per_launch"] = int(bpp * n)
        stages[name]["achieved_GBs"] = round(bpp * n / (per * 1e-3) / 1e9, 1)
        step_bytes += bpp * launches / ev_steps
    dom = max((s for s in stages if "achieved_GBs" in stages[s]),
              key=lambda s: stages[s]["us_per_launch"] * stages[s]["launches"])
    ctx.close()
    return {
        "workload": b.name, "scenes": int(len(b.counts)), "particles": n,
        "counts_min_max": [int(b.counts.min()), int(b.counts.max())],
        "value": value, "unit": "particle-steps/s", "steps": steps,
        "ms_per_step": wall * 1e3 / steps,
        "tuning": {"cell_factor": best[0], "tile": best[1], "resort": resort},
        "autotune_s": round(tune_s, 2),
        "autotune_ms_per_step": {f"{c[0]}/{c[1]}/{'inc' if c[2] else 'full'}": round(v, 4)
                                 for c, v in times.items()},
        "mean_neighbors": round(k, 3), "occupied_cells_per_particle": round(occ, 4),
        "radix_passes": passes,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": stages[dom]["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": stages[dom]["achieved_GBs"] / HBM_PEAK_GBS},
        "step_alg_bytes_per_particle": round(step_bytes, 1),
        "step_frac_of_hbm": step_bytes * value / (HBM_PEAK_GBS * 1e9),
        "kernel_ms_per_step": kern_ms / ev_steps,
        "stages": stages,
    }


def bench_surrogate(sph, scene, steps, train_iters, train_scenes=16, train_steps=1000,
                    check_every=10, threshold=0.02):
    """Surrogate in the loop on the 1M-particle dam break: a network trained on
    solver snapshots (`train_scenes` 65,536-particle dam breaks of `train_steps`
    steps, a snapshot every 10, the last scene held out; Adam with a gradient-norm
    limit) supplies the accelerations; plain SPH is evaluated every `check_every` steps
    and a window whose RMS relative acceleration error exceeds `threshold` is
    rolled back and recomputed with plain SPH. Compared with a plain run of the
    same steps from the same state (throughput and drift)."""
    sur = importlib.import_module("sph-simulations_amd.surrogate")
    n = scene.n
    log(f"surrogate: training ({train_iters} iterations)")
    trained, rep = sur.train_network(n_scenes=train_scenes, n_particles=65536, steps=train_steps,
                                     every=10, iters=train_iters,
                                     log=lambda m: log(f"train {m}"))
    theta = trained.params()
    trained.close()
    net = sur.Network(scene.params, n, layers=3, train=False)
    net.set_params(theta)
    out = {"training": rep, "steps": steps, "check_every": check_every, "threshold": threshold}
    plain = sph.Context(scene.params, n)
    surr = sph.Context(scene.params, n)
    for c in (plain, surr):
        c.upload(scene.rx, scene.ry, scene.vx, scene.vy)
        c.prepare()
    plain.set_graph(True)
    plain.step(2)
    plain.synchronize()
    # warm-up of the network path on a throwaway context
    warm = sph.Context(scene.params, n)
    warm.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    warm.prepare()
    sur.surrogate_run(warm, net, 2, check_every, threshold, 1)
    warm.close()
    # surrogate in the loop (window rollback)
    surr.synchronize()
    t0 = time.perf_counter()
    st = sur.surrogate_run(surr, net, steps, check_every, threshold, 1)
    surr.synchronize()
    t_s = time.perf_counter() - t0
    # plain run over the same steps from the same state (graph replay)
    plain2 = sph.Context(scene.params, n)
    plain2.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    plain2.prepare()
    plain2.set_graph(True)
    plain2.synchronize()
    t0 = time.perf_counter()
    plain2.step(steps)
    plain2.synchronize()
    t_p = time.perf_counter() - t0
    # network-only steps (every check accepted): the surrogate path's best case
    fast = sph.Context(scene.params, n)
    fast.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    fast.prepare()
    fast.synchronize()
    t0 = time.perf_counter()
    st_fast = sur.surrogate_run(fast, net, steps, check_every, 1e30, 0)
    fast.synchronize()
    t_f = time.perf_counter() - t0
    # drift of the surrogate trajectory against the plain one (matched ids)
    def state(c):
        d = c.download()
        c.prepare()
        f = c.fields()
        o = np.argsort(d["pid"])
        return {k: d[k][o].astype(np.float64) for k in ("rx", "ry")} | {
            "rho": f["rho"][o].astype(np.float64)}
    a, b, f_ = state(surr), state(plain2), state(fast)
    h = scene.params.h
    rho0 = scene.params.rho0

    def drift(x):
        dpos = np.hypot(x["rx"] - b["rx"], x["ry"] - b["ry"])
        drho = np.abs(x["rho"] - b["rho"]) / rho0
        return {"position_rms_over_h": float(np.sqrt(np.mean(dpos ** 2)) / h),
                "position_max_over_h": float(dpos.max() / h),
                "density_mean_rel": float(drho.mean()), "density_max_rel": float(drho.max())}

    out.update({
        "value": n * steps / t_s, "unit": "particle-steps/s", "ms_per_step": t_s * 1e3 / steps,
        "plain_value": n * steps / t_p, "plain_ms_per_step": t_p * 1e3 / steps,
        "vs_plain": (n * steps / t_s) / (n * steps / t_p),
        "fallback_rate": st["failed_checks"] / max(st["checks"], 1),
        "stats": st,
        "drift_vs_plain": drift(a),
        "network_only": {"value": n * steps / t_f, "ms_per_step": t_f * 1e3 / steps,
                         "vs_plain": t_p / t_f, "mean_error": st_fast["mean_error"],
                         "max_error": st_fast["max_error"], "drift_vs_plain": drift(f_)},
    })
    for c in (plain, surr, plain2, fast):
        c.close()
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--batched-steps", type=int, default=200,
                    help="steps of the 256-scene batched workload (0 disables)")
    ap.add_argument("--surrogate-steps", type=int, default=200,
                    help="steps of the surrogate-in-the-loop comparison (0 disables)")
    ap.add_argument("--train-iters", type=int, default=4000,
                    help="training iterations of the surrogate network")
    ap.add_argument("--train-scenes", type=int, default=16,
                    help="65,536-particle dam breaks generating the training snapshots")
    ap.add_argument("--train-steps", type=int, default=1000,
                    help="plain-SPH steps of each training scene")
    ap.add_argument("--variant", default="",
                    help="step kernel/re-sort to use without timing the others, e.g. "
                         "particle/incremental (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied device-to-device before the workload (a copy kernel of "
                         "known size for scaling HBM counters)")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU work for the cpu_baseline sample (0 disables)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); multi-GPU sharding of this path is described in DESIGN.md, not run.",
              file=sys.stderr)
        sys.exit(2)

    sph = importlib.import_module("sph-simulations_amd")
    if args.calibrate_copy > 0:
        nf = args.calibrate_copy * (1 << 20) // 4
        src, dst = sph.capi.DeviceBuffer(nf), sph.capi.DeviceBuffer(nf)
        rc = sph.lib().sph_device_copy(sph.capi.C.c_void_p(dst.ptr()),
                                       sph.capi.C.c_void_p(src.ptr()), nf * 4, 2)
        if rc != 0:
            raise sph.SphError(f"calibration copy -> {rc}")
        src.close()
        dst.close()
        log(f"calibration copy of {args.calibrate_copy} MiB done")
    scene = sph.scenes.by_name(args.workload)
    n = scene.n
    log(f"workload {scene.name}: n={n}")
    ctx = sph.Context(scene.params, n)
    ctx.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    ctx.prepare()
    ctx.set_graph(True)
    ctx.step(args.warmup)
    ctx.synchronize()
    log(f"warmup {args.warmup} steps done")

    # step-kernel and re-sort choice: time each combination on graph replay (all
    # give identical results)
    kernels = {"split_block": sph.STEP_SPLIT_BLOCK, "fused_block": sph.STEP_FUSED_BLOCK,
               "particle": sph.STEP_PARTICLE}
    resorts = {"full": sph.RESORT_FULL, "incremental": sph.RESORT_INCREMENTAL}
    tune = {}
    combos = [(k, r) for k in kernels for r in resorts]
    if args.variant:
        combos = [tuple(args.variant.split("/"))]
    for kname, rname in combos:
        ctx.set_step_kernel(kernels[kname])
        ctx.set_resort(resorts[rname])
        ctx.step(4)
        ctx.synchronize()
        a0 = time.perf_counter()
        ctx.step(20)
        ctx.synchronize()
        tune[f"{kname}/{rname}"] = (time.perf_counter() - a0) * 1e3 / 20
        log(f"variant {kname}/{rname}: {tune[f'{kname}/{rname}']:.4f} ms/step")
    best_combo = min(tune, key=tune.get)
    best, best_resort = best_combo.split("/")
    ctx.set_step_kernel(kernels[best])
    ctx.set_resort(resorts[best_resort])
    bstats = ctx.block_stats()
    log(f"block stats {bstats}")
    # movers per step (particles whose cell changed) over a short sample
    mv = []
    for _ in range(20):
        ctx.step(1)
        mv.append(ctx.movers())
    mover_stats = {"min": int(min(mv)), "mean": float(np.mean(mv)), "max": int(max(mv))}
    log(f"movers per step {mover_stats}")

    # timed region: exactly `steps` graph-replayed steps, synchronised on both sides
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = n * args.steps / wall
    log(f"timed {args.steps} steps: {wall * 1e3 / args.steps:.4f} ms/step")

    # the same number of steps again with every launch bracketed by hipEvents on
    # the context stream: per-kernel durations (the roofline) and their sum
    ctx.set_timing(True)
    ctx.synchronize()
    e0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    e1 = time.perf_counter()
    stage = ctx.stage_times()
    ctx.set_timing(False)
    ev_ms = (e1 - e0) * 1e3 / args.steps
    log(f"per-launch events pass: {ev_ms:.4f} ms/step")

    mfrac = ctx.movers() / n
    k = ctx.build_neighbor_lists(False) / n
    st, _ = ctx.cell_ranges()
    occ = float((st >= 0).sum()) / n
    passes = (int(ctx.grid().key_bits) + 7) // 8
    m = ctx.measure()

    stages = {}
    kern_total = 0.0
    for name, (ms, launches) in stage.items():
        if launches == 0:
            continue
        per_launch_ms = ms / launches
        entry = {"launches": launches, "ms_total": round(ms, 4),
                 "us_per_launch": round(per_launch_ms * 1e3, 3)}
        bpp = alg_bytes_per_particle(name, k, occ, passes, best, best_resort, mfrac,
                                     best == "particle" and best_resort == "incremental")
        if bpp is not None:
            b = bpp * n
            entry["alg_bytes_per_launch"] = int(b)
            entry["achieved_GBs"] = round(b / (per_launch_ms * 1e-3) / 1e9, 1)
        stages[name] = entry
        kern_total += ms
    dom_name = max((s for s in stages if "alg_bytes_per_launch" in stages[s]),
                   key=lambda s: stages[s]["ms_total"])
    dom = stages[dom_name]
    achieved = dom["achieved_GBs"]
    step_bytes = sum(alg_bytes_per_particle(s, k, occ, passes, best, best_resort, mfrac,
                                            best == "particle" and best_resort == "incremental")
                     * stages[s]["launches"] / args.steps
                     for s in stages if "alg_bytes_per_launch" in stages[s])
    out = {
        "metric": "particle-steps/s at 1M particles: surrogate vs plain SPH, \u22642% accel error",
        "value_path": "plain GPU SPH (the faster of the two paths within the 2% bound; the "
                      "surrogate-in-the-loop figures are under 'surrogate')",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded dam-break lattice with jitter)",
        "config": {"workload": scene.name, "particles": n, "box": list(scene.params.hi),
                   "h": scene.params.h, "dt": scene.params.dt, "physics": "wcsph",
                   "mean_neighbors": round(k, 3), "occupied_cells_per_particle": round(occ, 4),
                   "radix_passes": passes, "step_kernel": best, "resort": best_resort,
                   "movers_last_step": ctx.movers(), "movers_per_step": mover_stats,
                   "step_kernel_ms": {k2: round(v, 4) for k2, v in tune.items()},
                   "block_stats": bstats},
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNEL.get(dom_name, dom_name),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(STAGE_KERNEL.get(dom_name, dom_name)),
                     "issue": measured_issue(STAGE_KERNEL.get(dom_name, dom_name))},
        "stages": stages,
        "step": {"alg_bytes_per_particle": round(step_bytes, 1),
                 "step_frac_of_hbm": step_bytes * value / (HBM_PEAK_GBS * 1e9),
                 "kernel_ms_per_step": kern_total / args.steps,
                 "events_pass_ms_per_step": ev_ms},
        "diagnostics": {"kinetic_energy": m[0], "max_v2": m[1], "inside": int(m[2])},
    }
    ctx.close()
    if args.batched_steps > 0:
        out["batched"] = bench_batched(sph, args.batched_steps)
    if args.surrogate_steps > 0:
        out["surrogate"] = bench_surrogate(sph, scene, args.surrogate_steps, args.train_iters,
                                           args.train_scenes, args.train_steps)
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(sph, scene, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
