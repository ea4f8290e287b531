"""Writes tests/golden/golden.json: known answers computed from first principles,
independently of the oracle's code.

- Morton codes by interleaving binary digit strings (x digit in even positions),
  the definition of MortonOrder::MortonCode2d (core/alg/orders/MortonOrder.h:117-126),
  including the known answers listed for it (0xFFFF,0) -> 0x55555555,
  (0,0xFFFF) -> 0xAAAAAAAA.
- Lattice densities: the poly6 sum over the lattice offsets inside the support,
  enumerated geometrically, for the reference example (dx = 0.5, h = 1, mass 1,
  3D constant 315/(64 pi h^9), core/Kernel.cpp:8-10) and for the 16k dam break
  lattice (dx = 0.25/128, h = 1.3 dx, mass 1000 dx^2, 2D constant 4/(pi h^8)).
- Symplectic-Euler free fall: y_n = y_0 + g dt^2 n (n+1) / 2.

Run: python tests/golden/make_golden.py
"""
import json
import math
import os
import random


def morton_str(x, y, bits=32):
    xs = format(x, f"0{bits}b")
    ys = format(y, f"0{bits}b")
    out = "".join(yb + xb for xb, yb in zip(xs, ys))  # MSB first: y then x per level
    return int(out, 2)


def lattice_sum(dx, h, term):
    tot = 0.0
    m = int(math.ceil(h / dx)) + 1
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            r2 = (i * dx) ** 2 + (j * dx) ** 2
            if r2 <= h * h:
                tot += term(r2)
    return tot


def main():
    rnd = random.Random(1234)
    kats = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (3, 5), (0xFFFF, 0), (0, 0xFFFF),
            (1023, 2047), (2047, 1023), (1575, 787), (0xFFFFFFFF, 0), (0, 0xFFFFFFFF)]
    kats += [(rnd.getrandbits(16), rnd.getrandbits(16)) for _ in range(40)]
    kats += [(rnd.getrandbits(32), rnd.getrandbits(32)) for _ in range(20)]
    morton = [[x, y, morton_str(x, y)] for x, y in kats]

    # reference example lattice (dx = 0.5, h = 1): 3D poly6 constant
    h = 1.0
    ref_sum = lattice_sum(0.5, h, lambda r2: (h * h - r2) ** 3)
    ref_rho = 315.0 / (64.0 * math.pi * h ** 9) * ref_sum
    ref_neighbors = sum(1 for i in range(-3, 4) for j in range(-3, 4)
                        if (i, j) != (0, 0) and (0.5 * i) ** 2 + (0.5 * j) ** 2 <= 1.0)

    dx = 0.25 / 128
    h = 1.3 * dx
    wc_sum = lattice_sum(dx, h, lambda r2: (h * h - r2) ** 3)
    wc_rho = 1000.0 * dx * dx * 4.0 / (math.pi * h ** 8) * wc_sum
    wc_neighbors = sum(1 for i in range(-3, 4) for j in range(-3, 4)
                       if (i, j) != (0, 0) and (dx * i) ** 2 + (dx * j) ** 2 <= h * h)

    g, dt, y0 = -9.81, 1e-3, 1.0
    free_fall = [y0 + g * dt * dt * n * (n + 1) / 2 for n in range(0, 21)]

    out = {
        "morton": morton,
        "reference_lattice": {"dx": 0.5, "h": 1.0, "kernel_sum": ref_sum, "rho": ref_rho,
                              "neighbors": ref_neighbors},
        "wcsph_lattice": {"dx": dx, "h": h, "kernel_sum_over_h6": wc_sum / h ** 6,
                          "rho": wc_rho, "neighbors": wc_neighbors},
        "free_fall": {"g": g, "dt": dt, "y0": y0, "y": free_fall},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
