"""Run-time diagnostics around the step: the reference driver's per-step line,
adaptive time stepping from the CFL and force criteria, and the reference's
plain-text position checkpoints.
"""
from __future__ import annotations

import numpy as np

from .capi import Context


def step_header(step: int, measure) -> str:
    """app/simulation_configuration.h:152-158: 'Step= <k>\\t N=<inside>\\t E_k=<KE>\\t
    |v|_max=<max |v|^2>' (the reference prints ComputeMaxVelocity, which is the
    largest squared speed) with C++ stream default formatting (6 significant
    digits)."""
    return (f"Step= {step}\t N={int(measure[2])}\t E_k={measure[0]:g}\t "
            f"|v|_max={measure[1]:g}\n")


def run_adaptive(ctx: Context, steps: int, cfl: float = 0.25, every: int = 10,
                 dt_max: float | None = None, log=None):
    """Advance `steps` steps, recomputing dt from sph_cfl every `every` steps
    (FP64 maxima of |v| and |a|; dt capped at dt_max if given). Needs a prepared
    context; returns the list of (step, dt, v_max, a_max)."""
    hist = []
    done = 0
    while done < steps:
        c = ctx.cfl(cfl)
        dt = c["dt"] if dt_max is None else min(c["dt"], dt_max)
        ctx.set_dt(dt)
        k = min(every, steps - done)
        ctx.step(k)
        hist.append((done, dt, c["v_max"], c["a_max"]))
        if log:
            log(step_header(done, ctx.measure()).rstrip())
        done += k
    return hist


def write_positions_txt(ctx: Context, prefix: str, step: int):
    """core/utils/CheckPointWriter.cpp:12-27: '<prefix><step>.txt' with one 'x y z'
    line per particle (z = 0 in 2D), particles in id order."""
    d = ctx.download()
    o = np.argsort(d["pid"], kind="stable")
    path = f"{prefix}{step}.txt"
    with open(path, "w") as f:
        for x, y in zip(d["rx"][o], d["ry"][o]):
            f.write(f"{x:g} {y:g} 0\n")
    return path
