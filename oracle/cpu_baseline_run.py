"""Child process of bench.py's cpu_baseline leg -- MEASUREMENT INFRASTRUCTURE ONLY.

Started by bench.py with OMP_NUM_THREADS / OMP_SCHEDULE=static /
OMP_PROC_BIND=true already in its environment (the reference's recipe,
blog/blog.md:126-127), so that the OpenMP runtime of libcpu_baseline.so reads
them when it first starts. Reads the workload (npz: rx, ry, vx, vy; json: the
parameters), times the reference-style CPU step for each compute type on a
bounded sample and prints one JSON line.

usage: python -m oracle.cpu_baseline_run <state.npz> <params.json> <budget_s> [types] [max_steps]
"""
import json
import sys
from types import SimpleNamespace

import numpy as np

from oracle import oracle as orc


def main():
    state, prm_path, budget = sys.argv[1], sys.argv[2], float(sys.argv[3])
    types = sys.argv[4].split(",") if len(sys.argv) > 4 else ["mp-lists", "sorted"]
    max_steps = int(sys.argv[5]) if len(sys.argv) > 5 else 200
    z = np.load(state)
    d = json.load(open(prm_path))
    p = SimpleNamespace(**d)
    p.lo, p.hi = tuple(d["lo"]), tuple(d["hi"])
    n = int(z["rx"].size)
    out = {}
    for ct in types:
        threads, t, _ = orc.baseline_run(p, z["rx"], z["ry"], z["vx"], z["vy"], 2, ct)
        per_step = t["main_loop"] / 2.0
        steps = int(max(2, min(max_steps, budget / max(per_step, 1e-6))))
        threads, t, _ = orc.baseline_run(p, z["rx"], z["ry"], z["vx"], z["vy"], steps, ct)
        out[ct] = {"steps": steps, "threads": threads,
                   "value": n * steps / t["main_loop"],
                   "s_per_step": t["main_loop"] / steps,
                   "stage_seconds": {k: round(v, 4) for k, v in t.items()}}
        print(f"[cpu] {ct}: {steps} steps on {threads} threads, "
              f"{t['main_loop'] / steps:.3f} s/step", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
