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


_HIP = None


def device_sync():
    """hipDeviceSynchronize (what torch.cuda.synchronize calls), through the
    HIP runtime the library links: waits for all work on the device, every
    stream, without the library's read of its device status word."""
    global _HIP
    if _HIP is None:
        import ctypes
        for name in ("libamdhip64.so.7", "libamdhip64.so"):
            try:
                _HIP = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("HIP runtime library not found")
    rc = _HIP.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize -> {rc}")


def timed_steps(ctx, steps):
    """Wall time of one `steps`-step call: the call alone between two
    device-wide synchronisations; the context's status word (set on the
    device when a bounded wait gave up) is checked after the clock stops,
    so a failed run still raises."""
    ctx.synchronize()
    device_sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    device_sync()
    wall = time.perf_counter() - t0
    ctx.synchronize()
    return wall


def alg_bytes_per_particle(stage, k, occ, passes, variant, resort="full", movers_frac=0.0,
                           flags_in_force=False, narrow=False, bucket_in_force=False):
    """Algorithmic HBM bytes per particle of one launch of each stage (FP32
    state, u64 keys, u32 indices; neighbor and halo re-reads counted as cache
    hits). flags_in_force: the force kernel reports the re-sort's movers
    (reading the previous key, writing one flag bit). bucket_in_force: it also
    puts each mover into its owner tile's bucket, and the merge forms the
    chunk counters (no bucketing launch). narrow: the step keeps its keys as
    u32 (kb = 4 bytes, else 8). None for a stage without a model."""
    kb = 4.0 if narrow else 8.0
    f = movers_frac
    if variant == "split_block":
        # density: R r 8, W rho,pr 8; force: R r,v,rho,pr 24, W a 8, T r,v 16, key 8
        phys = {"approximate_density": 16.0, "force_computation": 56.0}
    elif variant == "fused_block":
        # R r,v 16; W rho,pr 8, a 8, T r,v 16, key 8
        phys = {"block_step": 56.0}
    else:
        # density: R r 8, W the packed (rho, pr) copy 8, count 4, list 4k;
        # force: R r,v 16, rho,pr 8, count 4, list 4k, T 16, key kb (+ R
        # previous key kb and W one flag bit when it reports the movers); the
        # separate rho, pr and a arrays (8 + 8) are written only in the last
        # step of a sph_step call; neighbors come from the packed copies (cache
        # hits)
        phys = {"approximate_density": 20.0 + 4.0 * k,
                "force_computation": 44.0 + kb + 4.0 * k
                + ((kb + 0.125) if flags_in_force else 0.0)
                # per mover: R its new cell's range 8, pid 4; W its bucket
                # entry 32, its owner's counter 8
                + (52.0 * f if bucket_in_force else 0.0)}
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
        sort = {
            # k_bucket: R the flag bits 1/8, a packed 8-byte counter per
            # 256-particle tile; per mover R new key kb, T r,v 16, pid 4, its
            # new cell's range 8, W its bucket entry 32, its owner's counter 8
            # (+ R new and old keys 2 kb, W the flag bit when no kernel before
            # reported the movers)
            "movers": 0.125 + 8.0 / 256.0 + (68.0 + kb) * f
            + (0.0 if flags_in_force else 2.0 * kb + 0.125),
            # k_merge: R old key kb, T r,v 16, pid 4, flag bits 1/8, the
            # movers' bucket entries 32 f; W r,v 16 (only the packed copy
            # inside sph_step), pid 4, key kb (the permutation 4 only in the
            # last step of a call); clear the old ranges 8 occ
            "merge": 40.125 + 2.0 * kb + 32.0 * f + 8.0 * occ
            # (bucket_in_force: R the tile counters, W the chunk counters)
            + (16.0 / 256.0 if bucket_in_force else 0.0),
            # k_ranges: R key kb, W 8 occ, the tile-first key and the zeroed
            # counter 16 per 256-particle tile
            "neighbor_search": kb + 16.0 / 256.0 + 8.0 * occ,
        }
    return {**phys, **sort}.get(stage)


STAGE_KERNEL = {"approximate_density": "k_density", "force_computation": "k_forces",
                "movers": "k_bucket", "merge": "k_merge", "neighbor_search": "k_ranges"}


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


def cpu_info():
    """CPU model and the physical cores this process may run on (affinity)."""
    model, cores = "unknown", set()
    try:
        allowed = os.sched_getaffinity(0)
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if "processor" in cur and int(cur["processor"]) in allowed:
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                cur = {}
                continue
            k, _, v = line.partition(":")
            cur[k.strip()] = v.strip()
            if k.strip() == "model name" and model == "unknown":
                model = v.strip()
        if "processor" in cur and int(cur["processor"]) in allowed:
            cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
    except OSError:
        pass
    return model, len(cores) or (os.cpu_count() or 1)


def cpu_baseline(scene, budget_s, max_steps=200):
    """The reference-style OpenMP CPU step (oracle/cpu_baseline.cpp: compute_type
    mp-lists, and sorted) on a bounded sample of the same workload, in a fresh
    child process started with the reference's OpenMP recipe
    (blog/blog.md:126-127): OMP_SCHEDULE=static, OMP_PROC_BIND=true,
    OMP_NUM_THREADS = the physical cores this process may use (capped by an
    OMP_NUM_THREADS already set, the host's share for one GPU)."""
    import dataclasses
    import subprocess
    import tempfile
    model, phys = cpu_info()
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(phys, cap) if cap > 0 else phys
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_SCHEDULE="static",
               OMP_PROC_BIND="true")
    with tempfile.TemporaryDirectory() as tmp:
        npz = os.path.join(tmp, "state.npz")
        np.savez(npz, rx=scene.rx, ry=scene.ry, vx=scene.vx, vy=scene.vy)
        prm = os.path.join(tmp, "params.json")
        json.dump(dataclasses.asdict(scene.params), open(prm, "w"))
        log(f"cpu baseline: {threads} threads ({phys} physical cores visible, {model})")
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline_run", npz, prm,
                            str(budget_s), "mp-lists,sorted", str(max_steps)], cwd=ROOT,
                           env=env, capture_output=True, text=True,
                           timeout=max(120.0, 6 * budget_s))
    sys.stderr.write(r.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({r.returncode})")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    main = res["mp-lists"]
    return {
        "value": main["value"],
        "unit": "particle-steps/s",
        "cores": main["threads"],
        "kind": "port",
        "sample": f"{scene.name}: {scene.n} particles x {main['steps']} steps of the "
                  f"reference-style OpenMP mp-lists/Morton CPU step (serial sort/gather/"
                  f"ranges/scan/EOS/integrator), -O3 -ffast-math; the sorted (cell-stencil) "
                  f"variant under 'variants'",
        "cpu_model": model,
        "physical_cores_visible": phys,
        "omp_env": {"OMP_NUM_THREADS": threads, "OMP_SCHEDULE": "static",
                    "OMP_PROC_BIND": "true"},
        "variants": res,
        "stage_seconds": main["stage_seconds"],
    }


def stage_table(stage, steps, n, k, occ, passes, variant, resort, mfrac):
    """Per-stage launches, time and achieved algorithmic GB/s of an events pass
    of `steps` steps; returns (stages, algorithmic bytes per particle-step, kernel ms)."""
    flags_in_force = variant == "particle" and resort == "incremental"
    # steps of a multi-step call with keys below 2^32 keep them as u32
    narrow = flags_in_force and passes <= 4 and steps >= 2
    stages, step_bytes, kern_ms = {}, 0.0, 0.0
    for name, (ms, launches) in stage.items():
        if launches == 0:
            continue
        per = ms / launches
        entry = {"launches": launches, "ms_total": round(ms, 4),
                 "us_per_launch": round(per * 1e3, 3)}
        bpp = alg_bytes_per_particle(name, k, occ, passes, variant, resort, mfrac,
                                     flags_in_force, narrow, bucket_in_force=flags_in_force)
        if bpp is not None:
            entry["alg_bytes_per_launch"] = int(bpp * n)
            entry["achieved_GBs"] = round(bpp * n / (per * 1e-3) / 1e9, 1)
            step_bytes += bpp * launches / steps
        stages[name] = entry
        kern_ms += ms
    return stages, step_bytes, kern_ms / steps


def full_run_profile(sph, scene, total, chunk=100):
    """The workload's whole run (`total` steps from rest, graph-replayed), timed
    in chunks of `chunk` steps; after each chunk the movers of its last step and
    the way the re-sort ordered them ("bucket" or "sorted"). Returns the run's
    mean step time and the profile."""
    n = scene.n
    ctx = sph.Context(scene.params, n)
    ctx.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    ctx.prepare()
    ctx.set_graph(True)
    rows, t_sum, done = [], 0.0, 0
    while done < total:
        kk = min(chunk, total - done)
        dt = timed_steps(ctx, kk)
        t_sum += dt
        done += kk
        rows.append([done, round(dt * 1e3 / kk, 4), ctx.movers(), ctx.resort_path()])
    m = ctx.measure()
    ctx.close()
    paths = {}
    for r in rows:
        paths[r[3]] = paths.get(r[3], 0) + 1
    q = max(1, len(rows) // 10)
    out = {"steps": total, "value": n * total / t_sum, "ms_per_step": t_sum * 1e3 / total,
           "chunk_steps": chunk,
           "ms_per_step_first_tenth": float(np.mean([r[1] for r in rows[:q]])),
           "ms_per_step_last_tenth": float(np.mean([r[1] for r in rows[-q:]])),
           "movers_last_step_of_chunk": {"min": int(min(r[2] for r in rows)),
                                         "max": int(max(r[2] for r in rows)),
                                         "mean": float(np.mean([r[2] for r in rows]))},
           "mover_path_chunks": paths,
           "chunks": rows,
           "final": {"kinetic_energy": m[0], "max_v2": m[1], "inside": int(m[2])}}
    log(f"full run {total} steps: {out['ms_per_step']:.4f} ms/step "
        f"(first tenth {out['ms_per_step_first_tenth']:.4f}, last "
        f"{out['ms_per_step_last_tenth']:.4f}), movers up to {out['movers_last_step_of_chunk']['max']}")
    return out


def bench_periodic(sph, steps, ref_steps=20):
    """The 4,194,304-particle periodic square: `steps` graph-replayed fused steps
    timed (host clock), again with per-launch hipEvents (the fused pipeline's
    kernels), and `ref_steps` iterations of the reference's stage order through
    the stage API (integrate, find neighbors = keys + radix sort + permutation +
    cell ranges, CSR lists, density + EOS over the lists, forces over the
    lists) timed per stage under the reference's timer names
    (app/simulation_configuration.h:72-91)."""
    sc = sph.scenes.periodic_shear()
    n = sc.n
    log(f"periodic: n={n}")
    ctx = sph.Context(sc.params, n)
    ctx.upload(sc.rx, sc.ry, sc.vx, sc.vy)
    ctx.prepare()
    ctx.set_graph(True)
    ctx.step(10)
    ctx.step_graphs(steps)
    wall = timed_steps(ctx, steps)
    value = n * steps / wall
    ev = min(steps, 100)
    ctx.set_timing(True)
    ctx.step(ev)
    ctx.synchronize()
    stage = ctx.stage_times()
    ctx.set_timing(False)
    mfrac = ctx.movers() / n
    k = ctx.build_neighbor_lists(False) / n
    st, _ = ctx.cell_ranges()
    occ = float((st >= 0).sum()) / n
    passes = (int(ctx.grid().key_bits) + 7) // 8
    stages, step_bytes, kern_ms = stage_table(stage, ev, n, k, occ, passes, "particle",
                                              "incremental", mfrac)
    # the reference's stage order, each stage its own launches
    ctx.find_neighbors()
    ctx.build_neighbor_lists(False)
    ctx.density_pressure(sph.COMPUTE_LISTS)
    ctx.forces(sph.COMPUTE_LISTS)
    ctx.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(ref_steps):
        ctx.integrate()
        ctx.find_neighbors()
        ctx.build_neighbor_lists(False)
        ctx.density_pressure(sph.COMPUTE_LISTS)
        ctx.forces(sph.COMPUTE_LISTS)
    ctx.synchronize()
    ref_wall = time.perf_counter() - t0
    rs = ctx.stage_times()
    ctx.set_timing(False)
    ctx.close()
    per = lambda name: rs[name][0] / ref_steps
    reference_stages_ms = {
        "integration": per("integration"),
        "neighbor_search": per("neighbor_search") + per("sort_histogram") + per("sort_pass"),
        "morton_sort (part of neighbor_search)": per("sort_histogram") + per("sort_pass"),
        "cell_build (part of neighbor_search: keys, permutation, ranges)": per("neighbor_search"),
        "neighbor_list": per("neighbor_list"),
        "approximate_density (with calculate_pressure)": per("approximate_density"),
        "force_computation": per("force_computation"),
    }
    b_canon = 296.0
    return {
        "workload": sc.name, "particles": n, "steps": steps, "value": value,
        "unit": "particle-steps/s", "ms_per_step": wall * 1e3 / steps,
        "target_60pct_of_roofline": 1.62e10, "vs_target": value / 1.62e10,
        "roofline_b_canon": {"bytes_per_particle_step": b_canon,
                             "frac": value * b_canon / (HBM_PEAK_GBS * 1e9)},
        "step_alg_bytes_per_particle": round(step_bytes, 1),
        "step_frac_of_hbm": step_bytes * value / (HBM_PEAK_GBS * 1e9),
        "mean_neighbors": round(k, 3), "occupied_cells_per_particle": round(occ, 4),
        "movers_last_step": int(round(mfrac * n)),
        "fused_stages": stages, "fused_kernel_ms_per_step": kern_ms,
        "reference_order": {"steps": ref_steps, "ms_per_step": ref_wall * 1e3 / ref_steps,
                            "stage_ms_per_step": {kk: round(v, 4) for kk, v in
                                                  reference_stages_ms.items()}},
    }


def bench_small(sph, steps):
    """The 16,384-particle dam break (1,000 steps of the reference's CPU-sized
    case): one graph-replayed sph_step call of `steps` steps, then 100 steps
    with per-launch hipEvents. Launch-latency bound: reported as latency."""
    sc = sph.scenes.dam_break()
    n = sc.n
    ctx = sph.Context(sc.params, n)
    ctx.upload(sc.rx, sc.ry, sc.vx, sc.vy)
    ctx.prepare()
    ctx.set_graph(True)
    ctx.step(20)
    ctx.step_graphs(steps)
    wall = timed_steps(ctx, steps)
    ctx.set_timing(True)
    ctx.step(100)
    ctx.synchronize()
    stage = ctx.stage_times()
    ctx.set_timing(False)
    kern = {kk: round(v[0] / v[1] * 1e3, 3) for kk, v in stage.items() if v[1]}
    ctx.close()
    log(f"small: {wall * 1e3 / steps:.4f} ms/step")
    # the cell-block mappings (a workgroup per 8 x 8-cell Morton block, halo in
    # LDS) against one thread per particle at the batched scenes' sizes
    mappings = {}
    for side in (128, 256):
        sc2 = sph.scenes.dam_break(side, side)
        c2 = sph.Context(sc2.params, sc2.n)
        c2.upload(sc2.rx, sc2.ry, sc2.vx, sc2.vy)
        c2.prepare()
        c2.set_graph(True)
        row = {}
        for kname, kv in (("particle", sph.STEP_PARTICLE), ("split_block", sph.STEP_SPLIT_BLOCK),
                          ("fused_block", sph.STEP_FUSED_BLOCK)):
            c2.set_step_kernel(kv)
            c2.step(10)
            c2.step_graphs(200)
            row[kname] = round(timed_steps(c2, 200) * 1e6 / 200, 2)
        c2.close()
        mappings[f"{sc2.n}_particles_us_per_step"] = row
    log(f"mappings {mappings}")
    return {"workload": sc.name, "particles": n, "steps": steps,
            "value": n * steps / wall, "unit": "particle-steps/s",
            "us_per_step": wall * 1e6 / steps, "us_per_launch_events": kern,
            "bound": "launch latency (five dependent kernels per step, graph-replayed)",
            "cell_block_mappings": mappings}


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
    ctx.step_graphs(steps)
    wall = timed_steps(ctx, steps)
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
        # the batched force kernel reports the movers itself (incremental re-sort)
        bpp = alg_bytes_per_particle(name, k, occ, passes, "particle", resort, mfrac,
                                     flags_in_force=resort == "incremental",
                                     narrow=resort == "incremental" and passes <= 4)
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
                    check_every=10, threshold=0.02, fast_steps=500):
    """Surrogate in the loop on the 1M-particle dam break: a network trained on
    solver snapshots (`train_scenes` 65,536-particle dam breaks of `train_steps`
    steps, a snapshot every 10, the last scene held out; Adam with a gradient-norm
    limit) supplies the accelerations; plain SPH is evaluated every `check_every` steps
    and a window whose RMS relative acceleration error exceeds `threshold` is
    rolled back and recomputed with plain SPH. Compared with a plain run of the
    same steps from the same state (throughput and drift). The network-only run
    (every check accepted) takes `fast_steps` steps against a plain run of as
    many."""
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
    log(f"surrogate in the loop: {steps} steps, {t_s * 1e3 / steps:.3f} ms/step")
    # network-only steps (every check accepted): the surrogate path's best case,
    # against a plain run of as many steps
    fs = min(fast_steps, steps)
    fast = sph.Context(scene.params, n)
    fast.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    fast.prepare()
    fast.synchronize()
    t0 = time.perf_counter()
    st_fast = sur.surrogate_run(fast, net, fs, check_every, 1e30, 0)
    fast.synchronize()
    t_f = time.perf_counter() - t0
    plain3 = sph.Context(scene.params, n)
    plain3.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    plain3.prepare()
    plain3.set_graph(True)
    plain3.step(fs)
    plain3.synchronize()
    log(f"network only: {fs} steps, {t_f * 1e3 / fs:.3f} ms/step")

    # drift of the surrogate trajectory against the plain one (matched ids)
    def state(c):
        d = c.download()
        c.prepare()
        f = c.fields()
        o = np.argsort(d["pid"])
        return {k: d[k][o].astype(np.float64) for k in ("rx", "ry")} | {
            "rho": f["rho"][o].astype(np.float64)}
    a, b, f_, b3 = state(surr), state(plain2), state(fast), state(plain3)
    h = scene.params.h
    rho0 = scene.params.rho0

    def drift(x, b=b):
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
        "network_only": {"steps": fs, "value": n * fs / t_f, "ms_per_step": t_f * 1e3 / fs,
                         "vs_plain": (t_p / steps) / (t_f / fs),
                         "mean_error": st_fast["mean_error"],
                         "max_error": st_fast["max_error"], "drift_vs_plain": drift(f_, b3)},
    })
    # the network alone against the BF16 MFMA peak: FLOPs of its node GEMMs over
    # the time of whole forward passes (edge passes included)
    fwd = sph.Context(scene.params, n)
    fwd.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    fwd.prepare()
    L = sph.lib()
    if L.sph_nn_forward(net.handle, fwd.handle, None, None) != 0:
        raise sph.SphError("sph_nn_forward failed")
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        L.sph_nn_forward(net.handle, fwd.handle, None, None)
    fwd_ms = (time.perf_counter() - t0) * 1e3 / reps
    fwd.close()
    W, K0 = 128, 32
    flops = n * (2 * K0 * W + 2 * W * W + 3 * (4 * 2 * W * W + 2 * 2 * W * W) + 2 * 2 * W * W)
    tflops = flops / (fwd_ms * 1e-3) / 1e12
    out["roofline"] = {"bound": "mfma", "kernel": "network forward (node GEMMs, BF16 MFMA)",
                       "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                       "frac": tflops / 2500.0, "forward_ms": fwd_ms,
                       "gemm_flops_per_forward": flops}
    for c in (plain, surr, plain2, fast, plain3):
        c.close()
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--window-start", type=int, default=2500,
                    help="steps of the run from rest before the timed window (the collapse "
                         "is under way; movers per step grow with it)")
    ap.add_argument("--settle", type=int, default=200,
                    help="steps run in one call after the variant timings and the mover "
                         "sample, before the warm-up: the timed steps then start from the "
                         "state a long run is in (the first ~200 steps after short, "
                         "synchronised calls run ~1.5 % slower)")
    ap.add_argument("--full-steps", type=int, default=5000,
                    help="length of the whole-run profile from rest (0 disables)")
    ap.add_argument("--periodic-steps", type=int, default=200,
                    help="steps of the 4,194,304-particle periodic workload (0 disables)")
    ap.add_argument("--small-steps", type=int, default=1000,
                    help="steps of the 16,384-particle dam break (0 disables)")
    ap.add_argument("--batched-steps", type=int, default=500,
                    help="steps of the 256-scene batched workload (0 disables)")
    ap.add_argument("--surrogate-steps", type=int, default=5000,
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
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of CPU work per compute type for the cpu_baseline sample "
                         "(0 disables)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}): the step shards by Morton-key ranges with a two-cell halo "
              "exchange, a design that is not built here.",
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
    full = full_run_profile(sph, scene, args.full_steps) if args.full_steps > 0 else None

    ctx = sph.Context(scene.params, n)
    ctx.upload(scene.rx, scene.ry, scene.vx, scene.vy)
    ctx.prepare()
    if args.window_start > 0:
        # eager launches (no graph) for the untimed run up to the window: the
        # profiler's kernel tracing has crashed here on long graph replays
        # right after the set-up
        done = 0
        while done < args.window_start:
            kk = min(100, args.window_start - done)
            ctx.step(kk)
            ctx.synchronize()
            done += kk
        log(f"advanced {args.window_start} steps to the timed window")
    ctx.set_graph(True)

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
    # movers per step (particles whose cell changed) over a short sample
    mv, mp = [], {}
    for _ in range(20):
        ctx.step(1)
        mv.append(ctx.movers())
        pth = ctx.resort_path()
        mp[pth] = mp.get(pth, 0) + 1
    mover_stats = {"min": int(min(mv)), "mean": float(np.mean(mv)), "max": int(max(mv)),
                   "path_steps": mp}
    log(f"movers per step {mover_stats}")
    # the run continues without a break before the warm-up (the variant
    # timings and the mover sample are short calls with synchronisations
    # between them; the steps after such calls run slower until the device is
    # back in the state of a long run)
    if args.settle > 0:
        ctx.step(args.settle)
        ctx.synchronize()
        log(f"settle {args.settle} steps done")
    # warm-up steps right before the timed call, as one call (a step call that
    # follows another one reuses the packed state it left)
    ctx.step(args.warmup)
    ctx.synchronize()
    log(f"warmup {args.warmup} steps done")
    window_first = args.window_start + 24 * len(combos) + 20 + max(args.settle, 0) + args.warmup

    # timed region: exactly `steps` graph-replayed steps, a device-wide
    # synchronisation on both sides (the graphs that call replays are captured
    # beforehand; the status check follows the clock)
    ctx.step_graphs(args.steps)
    wall = timed_steps(ctx, args.steps)
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
    stages, step_bytes, kern_ms = stage_table(stage, args.steps, n, k, occ, passes, best,
                                              best_resort, mfrac)
    dom_name = max((s_ for s_ in stages if "alg_bytes_per_launch" in stages[s_]),
                   key=lambda s_: stages[s_]["ms_total"])
    dom = stages[dom_name]
    achieved = dom["achieved_GBs"]
    kernel = STAGE_KERNEL.get(dom_name, dom_name)
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
                   "window": {"first_step": window_first, "steps": args.steps,
                              "settle_steps": max(args.settle, 0),
                              "note": "timed late in the collapse, after the run from rest"},
                   "mean_neighbors": round(k, 3), "occupied_cells_per_particle": round(occ, 4),
                   "radix_passes": passes, "step_kernel": best, "resort": best_resort,
                   "movers_last_step": ctx.movers(), "movers_per_step": mover_stats,
                   "step_kernel_ms": {k2: round(v, 4) for k2, v in tune.items()},
                   "block_stats": bstats},
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel),
                     "traffic_source": "profiles/hbm_traffic_dam_break_1m.md (rocprofv3 "
                                       "FETCH_SIZE / WRITE_SIZE per dispatch, calibrated "
                                       "against a device copy; not measured in this run)",
                     "issue": measured_issue(kernel)},
        "stages": stages,
        "step": {"alg_bytes_per_particle": round(step_bytes, 1),
                 "step_frac_of_hbm": step_bytes * value / (HBM_PEAK_GBS * 1e9),
                 "b_canon_frac_of_hbm": 304.0 * value / (HBM_PEAK_GBS * 1e9),
                 "kernel_ms_per_step": kern_ms,
                 "events_pass_ms_per_step": ev_ms},
        "diagnostics": {"kinetic_energy": m[0], "max_v2": m[1], "inside": int(m[2])},
        "full_run": full,
    }
    ctx.close()
    if args.periodic_steps > 0:
        out["periodic"] = bench_periodic(sph, args.periodic_steps)
    if args.small_steps > 0:
        out["small"] = bench_small(sph, args.small_steps)
    if args.batched_steps > 0:
        out["batched"] = bench_batched(sph, args.batched_steps)
    if args.surrogate_steps > 0:
        out["surrogate"] = bench_surrogate(sph, scene, args.surrogate_steps, args.train_iters,
                                           args.train_scenes, args.train_steps)
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget)
        # the other plain-SPH configurations on the same host: the 16,384-particle
        # dam break (up to its 1,000 steps) and the 4M periodic square (bounded)
        cfgs = {}
        for sc, budget, cap in ((sph.scenes.dam_break(), args.cpu_budget / 2, 1000),
                                (sph.scenes.periodic_shear(), args.cpu_budget, 200)):
            r = cpu_baseline(sc, budget, cap)
            cfgs[sc.name] = {"particles": sc.n, "configured_steps": sc.steps,
                             "value": r["value"], "unit": "particle-steps/s",
                             "variants": {k: {kk: v[kk] for kk in ("steps", "value", "s_per_step",
                                                                   "stage_seconds")}
                                          for k, v in r["variants"].items()}}
        out["cpu_baseline"]["configs"] = cfgs
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
