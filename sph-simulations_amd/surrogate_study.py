"""Measured study of the surrogate's two targets: RMS relative acceleration error
<= 2 % against plain SPH, and more particle-steps/s than the plain GPU step.

    python sph-simulations_amd/surrogate_study.py [--quick]

Parts (one JSON line on stdout):
  hold      the RMS relative error of reusing the plain-SPH accelerations of
            step t at step t + K, K = 1, 2, 5, 10, on the 1,048,576-particle
            dam break late in its collapse: a lower bound for any scheme that
            evaluates an acceleration model once every K steps and holds it
            (a perfect model held for K steps has exactly this error);
  training  the network (3 message-passing layers, width 128) trained on
            solver snapshots with the configuration's constant Adam lr 1e-3
            and with the schedule train_surrogate.py uses (2e-3 decaying to
            1e-4, gradient norm limited to 1); held-out RMS relative errors;
  split     the scheduled network's error on held-out snapshots split into
            interior particles (at least 4 neighbors, the lattice count) and
            free-surface particles (fewer), with each class's share of the
            squared error;
  cost      one network evaluation at 1,048,576 particles against one plain
            step: time, FLOPs of the node GEMMs (BF16 MFMA), their achieved
            rate against the 2.5 PFLOP/s BF16 peak.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[study] {msg}", file=sys.stderr, flush=True)


def gemm_flops_per_node(layers=3, width=128, node_k=32):
    """Multiply-adds x 2 of the node linear layers of one forward pass:
    encoder (node_k -> W, W -> W), per layer A, B, D, F (W -> W) and E (2W -> W),
    readout A, B (W -> W)."""
    W = width
    enc = 2 * node_k * W + 2 * W * W
    per_layer = 4 * (2 * W * W) + 2 * (2 * W) * W
    ro = 2 * (2 * W * W)
    return enc + layers * per_layer + ro


def hold_study(sph, pre_steps, ks):
    s = sph.scenes.dam_break_1m()
    ctx = sph.Context(s.params, s.n)
    ctx.upload(s.rx, s.ry, s.vx, s.vy)
    ctx.prepare()
    done = 0
    while done < pre_steps:
        k = min(100, pre_steps - done)
        ctx.step(k)
        ctx.synchronize()
        done += k
    kmax = max(ks)
    acc = []
    for _ in range(kmax + 1):
        ctx.step(1)
        d = ctx.download()
        f = ctx.fields()  # the accelerations of the state this step started from
        a = np.empty((s.n, 2), np.float64)
        a[d["pid"], 0] = f["ax"]
        a[d["pid"], 1] = f["ay"]
        acc.append(a)
    ctx.close()
    out = {}
    for K in ks:
        errs = []
        for t0 in range(0, kmax + 1 - K):
            tgt = acc[t0 + K]
            diff = acc[t0] - tgt
            errs.append(math.sqrt(float(np.sum(diff * diff)) / float(np.sum(tgt * tgt))))
        out[str(K)] = float(np.mean(errs))
    return {"workload": "dam_break_1m", "from_step": pre_steps, "rms_rel_error_by_stride": out}


def split_errors(sur, tr, net, pairs):
    """Held-out error split into interior (>= 4 neighbors) and surface particles."""
    sums = {"interior": [0.0, 0.0, 0], "surface": [0.0, 0.0, 0]}
    total = [0.0, 0.0]
    for t, s in pairs:
        tr.load(t, s)
        f = tr.ctx.fields()
        cnt = tr.ctx.ell_counts()
        ax, ay = net.forward(tr.ctx)
        tx = f["ax"].astype(np.float64)
        ty = f["ay"].astype(np.float64)
        e2 = (ax - tx) ** 2 + (ay - ty) ** 2
        t2 = tx * tx + ty * ty
        inner = cnt >= 4
        for name, m in (("interior", inner), ("surface", ~inner)):
            sums[name][0] += float(np.sum(e2[m]))
            sums[name][1] += float(np.sum(t2[m]))
            sums[name][2] += int(np.sum(m))
        total[0] += float(np.sum(e2))
        total[1] += float(np.sum(t2))
    out = {"rms_rel_all": math.sqrt(total[0] / total[1])}
    for name, (e2, t2, cnt) in sums.items():
        out[name] = {"particles": cnt, "rms_rel": math.sqrt(e2 / t2) if t2 > 0 else None,
                     "share_of_squared_error": e2 / total[0] if total[0] > 0 else None,
                     "share_of_squared_target": t2 / total[1] if total[1] > 0 else None}
    return out


def train_and_split(sur, args, lr, lr_final, clip, split):
    scs = sur.training_scenes(n_scenes=args.scenes, n_particles=args.particles, seed=args.seed)
    data = sur.generate_dataset(scs, steps=args.steps, every=10)
    net = sur.Network(scs[0].params, args.particles, layers=3, train=True)
    net.set_clip(clip)
    tr = sur.Trainer(net, scs[0].params, data, seed=args.seed)
    held = args.scenes - 1
    pairs = [(t, held) for t in range(1, data.T, max(1, data.T // 12))]
    t0 = time.perf_counter()
    tr.train(args.iters, lr=lr, scenes=held, lr_final=lr_final,
             log_every=max(1, args.iters // 4), log=lambda m: log(f"lr {lr}: {m}"))
    train_s = time.perf_counter() - t0
    e_nn, e_hold = tr.evaluate(pairs)
    rep = {"lr": lr, "lr_final": lr_final, "clip": clip, "iterations": args.iters,
           "train_s": round(train_s, 2), "held_out_rms_rel": e_nn,
           "held_out_zero_order_hold_10_steps": e_hold}
    if split:
        rep["split"] = split_errors(sur, tr, net, pairs)
    theta = net.params()
    tr.close()
    data.close()
    net.close()
    return rep, theta, scs[0].params


def cost_study(sph, sur, theta, params_train, reps=10):
    s = sph.scenes.dam_break_1m()
    n = s.n
    ctx = sph.Context(s.params, n)
    ctx.upload(s.rx, s.ry, s.vx, s.vy)
    ctx.prepare()
    ctx.set_graph(True)
    ctx.step(20)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(200)
    ctx.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / 200
    ctx.prepare()
    net = sur.Network(s.params, n, layers=3, train=False)
    net.set_params(theta)
    L = sph.lib()
    import ctypes as C
    rc = L.sph_nn_forward(net.handle, ctx.handle, None, None)
    if rc != 0:
        raise sph.SphError(f"sph_nn_forward -> {rc}")
    t0 = time.perf_counter()
    for _ in range(reps):
        L.sph_nn_forward(net.handle, ctx.handle, None, None)
    fwd_ms = (time.perf_counter() - t0) * 1e3 / reps
    net.close()
    ctx.close()
    flops = gemm_flops_per_node() * n
    achieved = flops / (fwd_ms * 1e-3) / 1e12
    return {"particles": n, "plain_step_ms": plain_ms, "network_forward_ms": fwd_ms,
            "forward_vs_plain_step": fwd_ms / plain_ms,
            "gemm_flops_per_forward": flops,
            "gemm_tflops_over_whole_forward": achieved,
            "frac_of_bf16_peak": achieved / BF16_PEAK_TFLOPS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="fewer scenes and iterations")
    ap.add_argument("--scenes", type=int, default=16)
    ap.add_argument("--particles", type=int, default=65536)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--hold-from", type=int, default=2500)
    ap.add_argument("--seed", type=lambda v: int(v, 0), default=0x5EED0003)
    args = ap.parse_args()
    if args.quick:
        args.scenes, args.steps, args.iters = 4, 200, 300
    sph = importlib.import_module("sph-simulations_amd")
    sur = importlib.import_module("sph-simulations_amd.surrogate")
    out = {}
    log("hold study")
    out["hold"] = hold_study(sph, args.hold_from, (1, 2, 5, 10))
    log(f"hold errors {out['hold']['rms_rel_error_by_stride']}")
    log("training with constant lr 1e-3, no norm limit, then with the limit")
    out["training_constant_lr"], _, _ = train_and_split(sur, args, 1e-3, None, 0.0, False)
    out["training_constant_lr_clipped"], _, _ = train_and_split(sur, args, 1e-3, None, 1.0,
                                                                False)
    log("training with the schedule")
    rep, theta, prm = train_and_split(sur, args, 2e-3, 1e-4, 1.0, True)
    out["training_schedule"] = rep
    log("cost")
    out["cost"] = cost_study(sph, sur, theta, prm)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
