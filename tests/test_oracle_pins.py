"""CPU tests that pin the two oracle pieces with no reference counterpart that the
plain-step parity rests on: Monaghan's artificial viscosity and the periodic
minimum image.

Each expected value is computed here from the published formula, by code that
shares nothing with the oracle, and each check is shown to reject the oracle
built with a planted mistake (`make -C oracle mutants`; ORC_MUTANT in
oracle/sph_oracle_impl.hpp):
  1 Monaghan gate flipped (viscosity on receding pairs),
  2 mean density without the 1/2,
  3 +alpha instead of -alpha,
  4 minimum image that wraps one side only,
  5 minimum image that wraps the wrong way.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


@pytest.fixture(scope="module")
def mutants(orc):
    r = subprocess.run(["make", "-j5", "-C", os.path.join(ROOT, "oracle"), "mutants"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return {k: orc.mutant_path(k) for k in range(1, 6)}


# ------------------------------------------------------------ Monaghan viscosity
def spiky_dwdr(h):
    """dW/dr of the 2D spiky kernel W = A (h - r)^3, with A fixed by integrating
    W over the disc to 1 (no use of the oracle's constants)."""
    from scipy.integrate import quad
    A = 1.0 / quad(lambda r: 2.0 * math.pi * r * (h - r) ** 3, 0.0, h)[0]
    return lambda r: -3.0 * A * (h - r) ** 2


def monaghan_pair(m, h, c0, alpha, eps, dwdr, r, v, rho, P):
    """Accelerations of a pair from the SPH momentum equation with Monaghan's
    artificial viscosity (J. J. Monaghan, Annu. Rev. Astron. Astrophys. 30
    (1992) 543, with beta = 0):
        dv_a/dt = -sum_b m_b (P_a/rho_a^2 + P_b/rho_b^2 + Pi_ab) grad_a W_ab
        Pi_ab = -alpha c mu_ab / rho_bar_ab  if v_ab . r_ab < 0, else 0
        mu_ab = h v_ab . r_ab / (r_ab^2 + eta^2),  eta^2 = eps h^2
        rho_bar_ab = (rho_a + rho_b) / 2
    Written with numpy vectors for particles a = 0, b = 1."""
    r_ab = r[0] - r[1]
    v_ab = v[0] - v[1]
    dist = float(np.linalg.norm(r_ab))
    vr = float(v_ab @ r_ab)
    if vr < 0:
        mu = h * vr / (dist ** 2 + eps * h * h)
        visc = -alpha * c0 * mu / (0.5 * (rho[0] + rho[1]))
    else:
        visc = 0.0
    grad_a = dwdr(dist) * r_ab / dist       # grad_0 W_01
    s = P[0] / rho[0] ** 2 + P[1] / rho[1] ** 2 + visc
    a0 = -m * s * grad_a
    a1 = -m * s * (-grad_a)                  # grad_1 W_10 = -grad_0 W_01
    return np.array([a0[0], a0[1], a1[0], a1[1]]), visc


def pair_case(sph, approaching):
    h = 0.1
    p = sph.scenes.wcsph_params(h / 1.3, (0.0, 0.0), (1.0, 1.0))
    p.h = h
    p.gravity_y = 0.0
    r = np.array([[0.40, 0.50], [0.43, 0.52]])
    v = np.array([[0.30, 0.10], [-0.20, 0.05]])
    if not approaching:
        v = -v
    rho = np.array([1005.0, 1012.0])
    B = p.rho0 * p.c0 ** 2 / p.gamma
    P = B * ((rho / p.rho0) ** p.gamma - 1.0)
    return p, r, v, rho, P


def oracle_pair(orc, p, r, v, rho, P):
    off = np.array([0, 1, 2], np.int64)
    idx = np.array([1, 0], np.int32)
    ax, ay = orc.forces(p, r[:, 0], r[:, 1], v[:, 0], v[:, 1], rho, P / rho ** 2, off, idx,
                        dtype=np.float64)
    return np.array([ax[0], ay[0], ax[1], ay[1]])


def monaghan_matches(sph, orc, approaching):
    p, r, v, rho, P = pair_case(sph, approaching)
    expect, visc = monaghan_pair(p.mass, p.h, p.c0, p.alpha, p.eps_visc, spiky_dwdr(p.h), r, v,
                                 rho, P)
    got = oracle_pair(orc, p, r, v, rho, P)
    return np.allclose(got, expect, rtol=1e-9, atol=1e-12 * np.abs(expect).max()), visc


@pytest.mark.parametrize("approaching", [True, False])
def test_monaghan_pair_closed_form(sph, orc, approaching):
    ok, visc = monaghan_matches(sph, orc, approaching)
    assert ok
    if approaching:
        # the viscous term is a sizeable share of the pressure term here, so
        # a wrong factor in it cannot hide in the tolerance
        p, r, v, rho, P = pair_case(sph, True)
        assert visc > 0.1 * (P[0] / rho[0] ** 2 + P[1] / rho[1] ** 2)
    else:
        assert visc == 0.0


@pytest.mark.parametrize("mutant", [1, 2, 3])
def test_monaghan_check_rejects_planted_mistakes(sph, orc, mutants, mutant):
    with orc.variant(mutants[mutant]):
        results = [monaghan_matches(sph, orc, a)[0] for a in (True, False)]
    assert not all(results), f"mutant {mutant} passed the closed-form check"


# ------------------------------------------------------------ periodic minimum image
def periodic_points(n=1500, seed=21):
    rng = np.random.default_rng(seed)
    rx = rng.random(n).astype(np.float32)
    ry = rng.random(n).astype(np.float32)
    # points on the seams and at the largest float below the box length
    top = np.nextafter(np.float32(1.0), np.float32(0.0))
    rx[:6] = [0.0, top, 0.0, top, 0.5, 0.01]
    ry[:6] = [0.0, top, top, 0.0, top, 0.0]
    return rx, ry


def image_neighbors(rx, ry, L, h2):
    """Brute force over the 3 x 3 explicit periodic images of every particle, in
    float64 (exact for float32 inputs in [0, L) with L = 1): row i = the j != i
    for which some image lies within h. Also returns each pair's smallest image
    distance^2 and the number of pairs only an image brings within h."""
    X = rx.astype(np.float64)
    Y = ry.astype(np.float64)
    rows, d2rows, seam_pairs = [], [], 0
    for i in range(X.size):
        best = np.full(X.size, np.inf)
        for sx in (-L, 0.0, L):
            for sy in (-L, 0.0, L):
                dx = X[i] - (X + sx)
                dy = Y[i] - (Y + sy)
                best = np.minimum(best, dx * dx + dy * dy)
        best[i] = np.inf
        direct = (X[i] - X) ** 2 + (Y[i] - Y) ** 2
        direct[i] = np.inf
        sel = np.nonzero(best <= h2)[0]
        seam_pairs += int(np.sum(direct[sel] > h2))
        rows.append(sel)
        d2rows.append(best)
    return rows, d2rows, seam_pairs


def sorted_state(orc, p, rx, ry):
    k = orc.keys(p, rx, ry)
    perm, sk = orc.stable_sort(k)
    st, en = orc.cell_ranges(sk, orc.grid(p).n_keys)
    return rx[perm], ry[perm], st, en


def periodic_rows_match(sph, orc, dtype):
    p = sph.scenes.wcsph_params(0.05 / 1.3, (0.0, 0.0), (1.0, 1.0), boundary=sph.BC_PERIODIC)
    rx, ry = periodic_points()
    srx, sry, st, en = sorted_state(orc, p, rx, ry)
    h2 = p.h * p.h
    rows, d2, seam_pairs = image_neighbors(srx, sry, 1.0, h2)
    cnt, off, idx = orc.neighbor_csr(p, srx, sry, st, en, dtype=dtype)
    for i, want in enumerate(rows):
        got = np.sort(idx[off[i]:off[i + 1]])
        if dtype == np.float64:
            if not np.array_equal(got, want):
                return False, seam_pairs
        else:
            # FP32 separations round differently from the exact ones: only pairs
            # within a relative 1e-5 of the support may differ
            diff = np.setxor1d(got, want)
            if diff.size and np.any(np.abs(d2[i][diff] - h2) > 1e-5 * h2):
                return False, seam_pairs
    return True, seam_pairs


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_periodic_lists_match_explicit_images(sph, orc, dtype):
    ok, seam_pairs = periodic_rows_match(sph, orc, dtype)
    assert ok
    assert seam_pairs > 100  # the check exercises neighbors across the seams


def periodic_lattice_ok(sph, orc):
    """64 x 64 lattice filling the periodic unit square (x = i / 64): every
    particle, the seam rows included, has exactly 4 neighbors at h = 1.3 dx and
    the lattice density 958.50 (tests/golden/make_golden.py, closed form)."""
    n = 64
    dx = 1.0 / n
    p = sph.scenes.wcsph_params(dx, (0.0, 0.0), (1.0, 1.0), boundary=sph.BC_PERIODIC)
    X, Y = sph.scenes.lattice(n, n, dx, dx)
    srx, sry, st, en = sorted_state(orc, p, X.astype(np.float32), Y.astype(np.float32))
    cnt, off, idx = orc.neighbor_csr(p, srx, sry, st, en)
    if not np.all(cnt == 4):
        return False
    rho, _, _ = orc.density(p, srx.astype(np.float64), sry.astype(np.float64), off, idx,
                            dtype=np.float64)
    return bool(np.allclose(rho, GOLDEN["wcsph_lattice"]["rho"], rtol=1e-5))


def test_periodic_lattice_closed_form(sph, orc):
    assert periodic_lattice_ok(sph, orc)


@pytest.mark.parametrize("mutant", [4, 5])
def test_periodic_checks_reject_planted_mistakes(sph, orc, mutants, mutant):
    with orc.variant(mutants[mutant]):
        assert not periodic_rows_match(sph, orc, np.float64)[0]
        assert not periodic_lattice_ok(sph, orc)


def test_unmutated_build_passes_the_same_checks(sph, orc):
    # the same helpers on the checker itself (guards the helpers, not the oracle)
    assert monaghan_matches(sph, orc, True)[0] and monaghan_matches(sph, orc, False)[0]
    assert periodic_lattice_ok(sph, orc)
