"""Initial conditions: the reference's lattice and the synthetic workloads.

All randomness comes from one project-written generator (SplitMix64 -> uniform
doubles, Box-Muller -> Gaussians) evaluated on the host, so the GPU and the CPU
oracle start from bit-identical float32 arrays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .capi import (BC_PERIODIC, BC_WALLS, ORDER_MORTON, ORDER_ROWMAJOR, PHYS_REFERENCE,
                   PHYS_WCSPH, Params)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count outputs of SplitMix64 started at `seed` (vectorised, wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int) -> np.ndarray:
    """U[0,1) doubles from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed: int, count: int) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = np.maximum(u[:m], 1e-300)
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:count]


@dataclass
class Scene:
    name: str
    params: Params
    rx: np.ndarray
    ry: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    steps: int

    @property
    def n(self) -> int:
        return int(self.rx.size)


def lattice(nx: int, ny: int, dx: float, dy: float, x0: float = 0.0, y0: float = 0.0):
    """LatticeCreator::CreateLattice (core/container/Lattice.h:12-35): particle
    (i, j) at (dX*i, dY*j) evaluated in double and stored as float, index
    i + nx*j (x fastest)."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    X = np.tile(dx * i, ny) + x0
    Y = np.repeat(dy * j, nx) + y0
    return X, Y


def reference_example(order: int = ORDER_MORTON) -> Scene:
    """config/example.txt:1-14 with the reference physics macros
    (core/ConfigVariables.hpp:9-15): 96x48 box, lattice dx=dy=0.5 over 40% x 80%
    of it (78 x 78 = 6084 particles), dt=0.005, 50 steps, H=1, mass 1."""
    Lx, Ly, d = 96.0, 48.0, 0.5
    nx = math.ceil(0.4 * Lx / d) + 1
    ny = math.ceil(0.8 * Ly / d) + 1
    X, Y = lattice(nx, ny, d, d)
    p = Params(lo=(0.0, 0.0), hi=(Lx, Ly), h=1.0, dt=0.005, mass=1.0, rho0=25.0,
               gravity_y=-10.0, eos_k=30.0, mu=0.01, physics=PHYS_REFERENCE, boundary=BC_WALLS,
               order=order, ell_width=64)
    z = np.zeros(nx * ny, np.float32)
    return Scene("reference_example", p, X.astype(np.float32), Y.astype(np.float32), z, z.copy(), 50)


def wcsph_params(dx: float, lo, hi, boundary=BC_WALLS, gravity_y=-9.81, order=ORDER_MORTON,
                 ell_width=32) -> Params:
    """Weakly-compressible physics of the dam-break configuration: rest density
    1000, mass 1000 dx^2, Tait gamma 7 with c0 = 10 sqrt(2 g 0.25), h = 1.3 dx,
    Monaghan alpha 0.1, dt = 0.25 h / c0."""
    h = 1.3 * dx
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * 0.25)
    return Params(lo=tuple(lo), hi=tuple(hi), h=h, dt=0.25 * h / c0, mass=1000.0 * dx * dx,
                  rho0=1000.0, gravity_y=gravity_y, c0=c0, alpha=0.1, eps_visc=0.01, gamma=7,
                  physics=PHYS_WCSPH, boundary=boundary, order=order, ell_width=ell_width)


def _jitter(seed: int, n: int, amp: float):
    u = uniform(seed, 2 * n)
    return (2.0 * u[:n] - 1.0) * amp, (2.0 * u[n:] - 1.0) * amp


def dam_break(nx: int = 128, ny: int = 128, box=(1.0, 1.0), seed: int = 0x5EED0000,
              steps: int = 1000, name: str = "dam_break_16k", order=ORDER_MORTON) -> Scene:
    """Dam break: nx x ny lattice at dx = 0.25/128 from the origin, seeded
    uniform jitter of +-0.01 dx, at rest, reflective walls, gravity -9.81."""
    dx = 0.25 / 128
    X, Y = lattice(nx, ny, dx, dx)
    jx, jy = _jitter(seed, X.size, 0.01 * dx)
    p = wcsph_params(dx, (0.0, 0.0), box, order=order)
    z = np.zeros(X.size, np.float32)
    return Scene(name, p, (X + jx).astype(np.float32), (Y + jy).astype(np.float32), z,
                 z.copy(), steps)


def periodic_shear(n_side: int = 2048, seed: int = 0x5EED0001, steps: int = 200,
                   order=ORDER_MORTON) -> Scene:
    """Periodic unit square, n_side^2 lattice at dx = 1/n_side with +-0.05 dx
    jitter, u = 0.1 sin(2 pi y), v = 0.1 sin(2 pi x) plus N(0, 0.001^2) noise,
    no gravity."""
    dx = 1.0 / n_side
    X, Y = lattice(n_side, n_side, dx, dx)
    n = X.size
    jx, jy = _jitter(seed, n, 0.05 * dx)
    X = np.mod(X + jx, 1.0)
    Y = np.mod(Y + jy, 1.0)
    g = gaussian(seed ^ 0xA5A5A5A5, 2 * n) * 0.001
    u = 0.1 * np.sin(2 * math.pi * Y) + g[:n]
    v = 0.1 * np.sin(2 * math.pi * X) + g[n:]
    p = wcsph_params(dx, (0.0, 0.0), (1.0, 1.0), boundary=BC_PERIODIC, gravity_y=0.0,
                     order=order)
    X32 = X.astype(np.float32)
    Y32 = Y.astype(np.float32)
    X32[X32 >= 1.0] = 0.0
    Y32[Y32 >= 1.0] = 0.0
    return Scene("periodic_shear", p, X32, Y32, u.astype(np.float32), v.astype(np.float32), steps)


def dam_break_1m(seed: int = 0x5EED0004, steps: int = 5000) -> Scene:
    """1024 x 1024 dam-break block (1,048,576 particles) in a 4 x 2 box with the
    16k dam-break physics."""
    return dam_break(1024, 1024, box=(4.0, 2.0), seed=seed, steps=steps, name="dam_break_1m")


@dataclass
class BatchScene:
    """Independent scenes advanced together; arrays concatenated in scene order."""
    name: str
    params: list
    counts: np.ndarray
    rx: np.ndarray
    ry: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    steps: int

    @property
    def n(self) -> int:
        return int(self.rx.size)

    @property
    def first(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.counts)[:-1]]).astype(np.int64)

    def scene(self, s: int) -> Scene:
        a, b = int(self.first[s]), int(self.first[s] + self.counts[s])
        return Scene(f"{self.name}[{s}]", self.params[s], self.rx[a:b], self.ry[a:b],
                     self.vx[a:b], self.vy[a:b], self.steps)


def batched_dam_breaks(n_scenes: int = 256, seed: int = 0x5EED0002, n_lo: int = 2048,
                       n_hi: int = 65536, steps: int = 500, name: str = "batched_dam_breaks",
                       order=ORDER_MORTON) -> BatchScene:
    """Independent dam breaks in their own unit boxes. Scene s draws (from
    uniform(seed), three numbers per scene) a particle count log-uniform in
    [n_lo, n_hi], a block aspect ratio A = height/width uniform in [0.5, 2] and a
    block width W uniform in [0.2, 0.5]. Lattice nx = round(sqrt(n / A)),
    ny = round(n / nx) (at most 1/dx), dx = W / nx from the origin with +-0.01 dx
    jitter (seeded seed + 1 + s); otherwise the 16k dam break's physics: h =
    1.3 dx, mass 1000 dx^2, Tait gamma 7 with the same c0 = 10 sqrt(2 g 0.25)
    for every scene, dt = 0.25 h / c0, reflective walls, gravity -9.81."""
    u = uniform(seed, 3 * n_scenes)
    params, xs, ys = [], [], []
    counts = np.zeros(n_scenes, np.int64)
    for s in range(n_scenes):
        n = round(math.exp(math.log(n_lo) + u[3 * s] * (math.log(n_hi) - math.log(n_lo))))
        A = 0.5 + 1.5 * u[3 * s + 1]
        W = 0.2 + 0.3 * u[3 * s + 2]
        nx = max(1, round(math.sqrt(n / A)))
        dx = W / nx
        ny = max(1, min(round(n / nx), int(math.floor(1.0 / dx))))
        X, Y = lattice(nx, ny, dx, dx)
        jx, jy = _jitter(seed + 1 + s, X.size, 0.01 * dx)
        p = wcsph_params(dx, (0.0, 0.0), (1.0, 1.0), order=order)
        params.append(p)
        xs.append((X + jx).astype(np.float32))
        ys.append((Y + jy).astype(np.float32))
        counts[s] = X.size
    rx = np.concatenate(xs)
    ry = np.concatenate(ys)
    z = np.zeros(rx.size, np.float32)
    return BatchScene(name, params, counts, rx, ry, z, z.copy(), steps)


def by_name(name: str) -> Scene:
    return {
        "reference_example": reference_example,
        "dam_break_16k": dam_break,
        "periodic_shear": periodic_shear,
        "dam_break_1m": dam_break_1m,
    }[name]()


__all__ = ["Scene", "BatchScene", "batched_dam_breaks", "lattice", "reference_example", "dam_break", "periodic_shear",
           "dam_break_1m", "wcsph_params", "splitmix64", "uniform", "gaussian", "by_name",
           "ORDER_MORTON", "ORDER_ROWMAJOR"]
