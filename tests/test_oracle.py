"""CPU tests that pin the oracle (oracle/) before it is trusted as the checker.

Every check compares the oracle with something it does not share code with:
known answers computed from first principles (tests/golden/golden.json),
textbook routines (numpy's stable argsort), all-pairs brute force, lattice
closed forms, numerical integration of the kernels, conservation laws and the
exact symplectic-Euler free-fall trajectory.
"""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


# ------------------------------------------------------------------ Morton
def test_morton_known_answers(orc):
    for x, y, key in GOLDEN["morton"]:
        assert orc.morton(x, y) == key, (x, y)
        assert orc.morton_magic(x, y) == key, (x, y)
        assert orc.morton_decode(key) == (x, y)


def test_morton_reference_kats(orc):
    # MortonOrder.h:117-126 semantics: x -> even bits, y -> odd bits
    assert [orc.morton(*c) for c in [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0)]] == [0, 1, 2, 3, 4]
    assert orc.morton(0xFFFF, 0) == 0x55555555
    assert orc.morton(0, 0xFFFF) == 0xAAAAAAAA


# ------------------------------------------------------------------ grid
def test_grid_reference_example(sph, orc):
    s = sph.scenes.reference_example()
    g = orc.grid(s.params)
    assert (g.nx, g.ny, g.nxp, g.nyp, g.n_keys) == (96, 48, 128, 64, 8192)
    assert s.n == 6084  # 78 x 78 (Lattice.h:13-17 with example.txt)


def test_grid_dam_break_1m(sph, orc):
    p = sph.scenes.wcsph_params(0.25 / 128, (0, 0), (4, 2))
    g = orc.grid(p)
    assert (g.nx, g.ny, g.nxp, g.nyp) == (1576, 788, 2048, 1024)
    assert g.n_keys == 2048 * 1024 and g.key_bits == 21


def test_grid_periodic(sph, orc):
    p = sph.scenes.wcsph_params(1 / 2048, (0, 0), (1, 1), boundary=sph.BC_PERIODIC)
    g = orc.grid(p)
    assert g.nx == g.ny == 1575  # floor(1 / (1.3/2048))
    assert g.cs_x >= np.float32(p.h)


def test_grid_rejects_small_cells(sph, orc):
    p = sph.scenes.wcsph_params(0.01, (0, 0), (1, 1))
    p.cell_size = 0.5 * p.h
    with pytest.raises(ValueError):
        orc.grid(p)


def test_cell_coord_edges(orc):
    cs, n = np.float32(0.25), 8
    assert orc.cell_coord(0.0, 0.0, cs, n) == 0
    assert orc.cell_coord(0.25, 0.0, cs, n) == 1      # x = k cs exactly -> k
    assert orc.cell_coord(0.2499999, 0.0, cs, n) == 0
    assert orc.cell_coord(2.0, 0.0, cs, n) == 7       # x = L -> clamp n-1
    assert orc.cell_coord(5.0, 0.0, cs, n) == 7
    assert orc.cell_coord(-0.01, 0.0, cs, n) == 0     # below the box -> 0
    assert orc.cell_coord(float("nan"), 0.0, cs, n) == 0
    assert orc.cell_coord(float("inf"), 0.0, cs, n) == 7


# ------------------------------------------------------------------ sort, ranges
@pytest.mark.parametrize("n,kmax", [(1, 4), (17, 3), (1000, 50), (20000, 5000), (5000, 1)])
def test_stable_sort_matches_numpy(orc, n, kmax):
    rng = np.random.default_rng(n)
    k = rng.integers(0, kmax, n).astype(np.uint64)
    perm, sk = orc.stable_sort(k)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32))
    assert np.array_equal(sk, k[ref])
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))


def test_cell_ranges_closed_form(orc):
    rng = np.random.default_rng(7)
    k = np.sort(rng.integers(0, 300, 2000)).astype(np.uint64)
    s, e = orc.cell_ranges(k, 512)
    for c in range(512):
        idx = np.nonzero(k == c)[0]
        if idx.size == 0:
            assert s[c] == -1 and e[c] == -1
        else:
            assert s[c] == idx[0] and e[c] == idx[-1] + 1
    occ = s >= 0
    assert np.sum(e[occ] - s[occ]) == k.size


# ------------------------------------------------------------------ neighbor lists
def _sorted_state(orc, p, rx, ry):
    k = orc.keys(p, rx, ry)
    perm, sk = orc.stable_sort(k)
    g = orc.grid(p)
    s, e = orc.cell_ranges(sk, g.n_keys)
    return rx[perm], ry[perm], s, e


def _rows(off, idx):
    return [sorted(idx[off[i]:off[i + 1]].tolist()) for i in range(off.size - 1)]


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("cell_factor", [0.0, 1.7])
def test_csr_matches_brute_force(sph, orc, boundary, cell_factor):
    rng = np.random.default_rng(11 + boundary)
    p = sph.scenes.wcsph_params(0.05, (0.0, 0.0), (1.0, 1.0), boundary=boundary)
    if cell_factor:
        p.cell_size = cell_factor * p.h
    n = 1500
    rx = rng.random(n).astype(np.float32)
    ry = rng.random(n).astype(np.float32)
    # clusters and exact duplicates
    rx[:50] = rx[50:100]
    ry[:50] = ry[50:100]
    rx[100:120] = np.float32(0.5)
    ry[100:120] = np.float32(0.5)
    srx, sry, s, e = _sorted_state(orc, p, rx, ry)
    cnt, off, idx = orc.neighbor_csr(p, srx, sry, s, e)
    bcnt, boff, bidx = orc.brute_neighbors(p, srx, sry)
    assert np.array_equal(cnt, bcnt)
    assert _rows(off, idx) == _rows(boff, bidx)
    # self-inclusive variant: exactly one more entry, i itself
    cnt2, off2, idx2 = orc.neighbor_csr(p, srx, sry, s, e, include_self=True)
    assert np.array_equal(cnt2, cnt + 1)
    for i in range(0, n, 97):
        assert sorted(idx2[off2[i]:off2[i + 1]].tolist()) == sorted(
            idx[off[i]:off[i + 1]].tolist() + [i])


def test_csr_stencil_order(sph, orc):
    # rows follow the reference's stencil order: x offset outer, y inner, index ascending
    s = sph.scenes.reference_example()
    srx, sry, st, en = _sorted_state(orc, s.params, s.rx, s.ry)
    cnt, off, idx = orc.neighbor_csr(s.params, srx, sry, st, en)
    g = orc.grid(s.params)
    for i in range(0, s.n, 331):
        cx = int(math.floor(srx[i])); cy = int(math.floor(sry[i]))
        expect = []
        for ox in (-1, 0, 1):
            for oy in (-1, 0, 1):
                x, y = cx + ox, cy + oy
                if not (0 <= x < g.nx and 0 <= y < g.ny):
                    continue
                k = orc.morton(x, y)
                if st[k] < 0:
                    continue
                for j in range(st[k], en[k]):
                    if j == i:
                        continue
                    d = (srx[i] - srx[j]) ** 2 + (sry[i] - sry[j]) ** 2
                    if np.float32(d) <= np.float32(1.0):
                        expect.append(j)
        assert idx[off[i]:off[i + 1]].tolist() == expect


def test_lattice_neighbor_counts(sph, orc):
    s = sph.scenes.reference_example()
    srx, sry, st, en = _sorted_state(orc, s.params, s.rx, s.ry)
    cnt, _, _ = orc.neighbor_csr(s.params, srx, sry, st, en)
    interior = (srx > 2) & (srx < 36) & (sry > 2) & (sry < 36)
    assert np.all(cnt[interior] == GOLDEN["reference_lattice"]["neighbors"])  # 12
    d = sph.scenes.dam_break()
    rx, ry = sph.scenes.lattice(128, 128, 0.25 / 128, 0.25 / 128)
    srx, sry, st, en = _sorted_state(orc, d.params, rx.astype(np.float32), ry.astype(np.float32))
    cnt, _, _ = orc.neighbor_csr(d.params, srx, sry, st, en)
    interior = (srx > 0.01) & (srx < 0.24) & (sry > 0.01) & (sry < 0.24)
    assert np.all(cnt[interior] == GOLDEN["wcsph_lattice"]["neighbors"])  # 4


# ------------------------------------------------------------------ density
def _prepare(orc, p, rx, ry, dtype=np.float32):
    srx, sry, st, en = _sorted_state(orc, p, rx, ry)
    cnt, off, idx = orc.neighbor_csr(p, srx, sry, st, en)
    return srx, sry, off, idx


def test_density_reference_lattice(sph, orc):
    s = sph.scenes.reference_example()
    srx, sry, off, idx = _prepare(orc, s.params, s.rx, s.ry)
    rho, pres, pr = orc.density(s.params, srx, sry, off, idx, dtype=np.float64)
    interior = (srx > 2) & (srx < 36) & (sry > 2) & (sry < 36)
    assert np.allclose(rho[interior], GOLDEN["reference_lattice"]["rho"], rtol=1e-12)
    assert np.allclose(pres, 30.0 * (rho - 25.0), rtol=1e-12)  # ConfigVariables.hpp:11-13
    assert np.allclose(pr, pres / rho ** 2, rtol=1e-12)


def test_density_wcsph_lattice(sph, orc):
    d = sph.scenes.dam_break()
    rx, ry = sph.scenes.lattice(128, 128, 0.25 / 128, 0.25 / 128)
    srx, sry, off, idx = _prepare(orc, d.params, rx.astype(np.float32), ry.astype(np.float32))
    rho, pres, _ = orc.density(d.params, srx.astype(np.float64), sry.astype(np.float64), off, idx,
                               dtype=np.float64)
    interior = (srx > 0.01) & (srx < 0.24) & (sry > 0.01) & (sry < 0.24)
    assert np.allclose(rho[interior], GOLDEN["wcsph_lattice"]["rho"], rtol=1e-5)
    B = 1000.0 * d.params.c0 ** 2 / 7.0
    assert np.allclose(pres, B * ((rho / 1000.0) ** 7 - 1.0), rtol=1e-9, atol=1e-6)


def test_density_normalisation_2d(sph, orc):
    # a fine lattice (h = 4 dx) must reproduce rho0 within the quadrature error:
    # pins the 2D poly6 normalisation 4 / (pi h^8)
    dx = 0.01
    p = sph.scenes.wcsph_params(dx, (0, 0), (1, 1))
    p.h = 4 * dx
    rx, ry = sph.scenes.lattice(100, 100, dx, dx)
    srx, sry, off, idx = _prepare(orc, p, rx.astype(np.float32), ry.astype(np.float32))
    rho, _, _ = orc.density(p, srx.astype(np.float64), sry.astype(np.float64), off, idx,
                            dtype=np.float64)
    interior = (srx > 0.1) & (srx < 0.9) & (sry > 0.1) & (sry < 0.9)
    assert abs(np.mean(rho[interior]) / 1000.0 - 1.0) < 5e-3


def test_spiky_gradient_matches_normalised_kernel(sph, orc):
    # two particles: the oracle's pressure acceleration must equal
    # -m (P_i/rho_i^2 + P_j/rho_j^2) dW/dr with W the spiky kernel normalised by
    # numerical integration in 2D (no reference to the oracle's constants)
    from scipy.integrate import quad
    h = 0.1
    A = 1.0 / quad(lambda r: 2 * math.pi * r * (h - r) ** 3, 0, h)[0]
    dWdr = lambda r: -3 * A * (h - r) ** 2
    p = sph.scenes.wcsph_params(h / 1.3, (0, 0), (1, 1))
    p.gravity_y = 0.0
    p.h = h
    r = 0.037
    rx = np.array([0.5, 0.5 + r])
    ry = np.array([0.5, 0.5])
    off = np.array([0, 1, 2], np.int64)
    idx = np.array([1, 0], np.int32)
    rho = np.array([1050.0, 1050.0])
    P = 1000.0 * p.c0 ** 2 / 7 * ((1.05) ** 7 - 1)
    pr = np.array([P / 1050.0 ** 2] * 2)
    v = np.zeros(2)
    ax, ay = orc.forces(p, rx, ry, v, v, rho, pr, off, idx, dtype=np.float64)
    # a_0 = -m (P_0 + P_1) grad_0 W, grad_0 W = dW/dr (x_0 - x_1)/r = -dW/dr (> 0):
    # positive pressure pushes particle 0 towards -x
    expect = -p.mass * 2 * pr[0] * (-dWdr(r))
    assert expect < 0
    assert ax[0] == pytest.approx(expect, rel=1e-6)
    assert ax[1] == pytest.approx(-expect, rel=1e-6)  # antisymmetric
    assert ay[0] == 0 and ay[1] == 0


def test_reference_force_two_particles(sph, orc):
    # reference physics, two particles along x with unequal densities: pressure
    # term 1.5 cp m (h-r)^2/r (P_i+P_j) d with cp = 15/(pi h^6) and viscosity
    # 45 mu/(pi h^6)(h-r)/rho_j (v_j - v_i)  (GenericForceComputation.tpp:229-265)
    s = sph.scenes.reference_example()
    p = s.params
    p.gravity_y = 0.0
    r = 0.6
    rx = np.array([10.0, 10.0 + r]); ry = np.array([5.0, 5.0])
    vx = np.array([0.0, 0.3]); vy = np.array([0.1, 0.0])
    rho = np.array([4.0, 6.0])
    pres = 30.0 * (rho - 25.0)
    pr = pres / rho ** 2
    off = np.array([0, 1, 2], np.int64); idx = np.array([1, 0], np.int32)
    ax, ay = orc.forces(p, rx, ry, vx, vy, rho, pr, off, idx, dtype=np.float64)
    grad = -3 * 15 / math.pi * (1 - r) ** 2   # d/dr of 15/pi (h-r)^3, h = 1
    # reference: fp += grad/r * d * (-1/2)(P_i/rho_i^2 + P_j/rho_j^2) m, d = r_i - r_j
    fp0 = grad / r * (-r) * (-0.5) * (pr[0] + pr[1])
    visc = lambda rho_j: 0.01 / (rho_j * math.pi) * 45 * (1 - r)
    assert ax[0] == pytest.approx(fp0 + visc(rho[1]) * (vx[1] - vx[0]), rel=1e-9)
    assert ay[0] == pytest.approx(visc(rho[1]) * (vy[1] - vy[0]), rel=1e-9)
    assert ax[1] == pytest.approx(-fp0 + visc(rho[0]) * (vx[0] - vx[1]), rel=1e-9)


def test_momentum_conservation_wcsph(sph, orc):
    # pressure + Monaghan viscosity are pairwise antisymmetric: sum m a = N m g
    d = sph.scenes.dam_break(32, 32, seed=3)
    rng = np.random.default_rng(5)
    vx = rng.normal(0, 0.5, d.n); vy = rng.normal(0, 0.5, d.n)
    srx, sry, st, en = _sorted_state(orc, d.params, d.rx, d.ry)
    _, off, idx = orc.neighbor_csr(d.params, srx, sry, st, en)
    X, Y = srx.astype(np.float64), sry.astype(np.float64)
    rho, _, pr = orc.density(d.params, X, Y, off, idx, dtype=np.float64)
    ax, ay = orc.forces(d.params, X, Y, vx, vy, rho, pr, off, idx, dtype=np.float64)
    scale = np.abs(ax).sum() + np.abs(ay).sum()
    assert abs(ax.sum()) < 1e-10 * scale
    assert abs(ay.sum() - d.n * d.params.gravity_y) < 1e-10 * scale


def test_interior_lattice_force_vanishes(sph, orc):
    d = sph.scenes.dam_break(64, 64)
    rx, ry = sph.scenes.lattice(64, 64, 0.25 / 128, 0.25 / 128, 0.3, 0.3)
    srx, sry, st, en = _sorted_state(orc, d.params, rx.astype(np.float32), ry.astype(np.float32))
    _, off, idx = orc.neighbor_csr(d.params, srx, sry, st, en)
    X, Y = rx.astype(np.float32).astype(np.float64), ry.astype(np.float32).astype(np.float64)
    X, Y = srx.astype(np.float64), sry.astype(np.float64)
    rho, _, pr = orc.density(d.params, X, Y, off, idx, dtype=np.float64)
    z = np.zeros(X.size)
    ax, ay = orc.forces(d.params, X, Y, z, z, rho, pr, off, idx, dtype=np.float64)
    dxl = 0.25 / 128
    interior = (X > 0.3 + 3 * dxl) & (X < 0.3 + 60 * dxl) & (Y > 0.3 + 3 * dxl) & (Y < 0.3 + 60 * dxl)
    edge_scale = np.abs(ax).max()
    assert np.abs(ax[interior]).max() < 1e-6 * edge_scale
    assert np.abs(ay[interior] - d.params.gravity_y).max() < 1e-6 * edge_scale


# ------------------------------------------------------------------ integrator
def test_free_fall_symplectic_euler(sph, orc):
    ff = GOLDEN["free_fall"]
    p = sph.scenes.wcsph_params(0.01, (0, -100), (1, 100))
    p.dt = ff["dt"]
    x = np.array([0.5]); y = np.array([ff["y0"]]); vx = np.zeros(1); vy = np.zeros(1)
    a = np.array([0.0]); g = np.array([ff["g"]])
    for n in range(1, len(ff["y"])):
        x, y, vx, vy = orc.integrate(p, x, y, vx, vy, a, g, dtype=np.float64)
        assert y[0] == pytest.approx(ff["y"][n], rel=1e-13, abs=1e-15)


def test_wall_reflection_and_periodic_wrap(sph, orc):
    p = sph.scenes.wcsph_params(0.01, (0, 0), (1, 1))
    p.dt = 0.1
    x, y, vx, vy = orc.integrate(p, [0.95, 0.05], [0.5, 0.02], [1.0, -1.0], [0.0, -0.5],
                                 [0, 0], [0, 0], dtype=np.float64)
    assert x[0] == 1.0 and vx[0] == -1.0   # Integrator.cpp:60-64
    assert x[1] == 0.0 and vx[1] == 1.0    # Integrator.cpp:66-70
    assert y[1] == 0.0 and vy[1] == 0.5
    p.boundary = 1
    x, y, vx, vy = orc.integrate(p, [0.95, 0.05], [0.5, 0.5], [1.0, -1.0], [0, 0], [0, 0],
                                 [0, 0], dtype=np.float64)
    assert x[0] == pytest.approx(0.05) and vx[0] == 1.0
    assert x[1] == pytest.approx(0.95) and vx[1] == -1.0


# ------------------------------------------------------------------ diagnostics
def test_measure_matches_fsum(sph, orc):
    rng = np.random.default_rng(3)
    d = sph.scenes.dam_break(64, 64)
    vx = rng.normal(0, 1, d.n).astype(np.float32)
    vy = rng.normal(0, 1, d.n).astype(np.float32)
    rx = d.rx.copy(); rx[:3] = [-1.0, np.nan, 2.0]
    out = orc.measure(d.params, rx, d.ry, vx, vy)
    v2 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2
    m = d.params.mass
    assert out[0] == pytest.approx(math.fsum(0.5 * m * v2), rel=1e-14)
    assert out[1] == v2.max()
    inside = (np.isfinite(rx) & np.isfinite(d.ry) & (rx >= 0) & (rx <= 1) & (d.ry >= 0)
              & (d.ry <= 1))  # Domain::IsInside (core/Domain.h:27-37)
    assert out[2] == inside.sum() and inside.sum() < d.n - 3
    assert out[3] == pytest.approx(math.fsum(m * 9.81 * d.ry.astype(np.float64)), rel=1e-13)


# ------------------------------------------------------------------ delta encoder
def test_delta_round_trip(orc):
    lst = np.array([3, 4, 9, 10, 200, 201, 5000], np.int32)
    d = orc.delta_encode(lst)
    assert d.tolist() == [3, 0, 4, 0, 189, 0, 4798]   # Encoder.h:36-44
    assert np.array_equal(orc.delta_decode(d), lst)
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = np.unique(rng.integers(0, 10 ** 6, rng.integers(1, 50))).astype(np.int32)
        assert np.array_equal(orc.delta_decode(orc.delta_encode(a)), a)


# ------------------------------------------------------------------ FP32 vs FP64
def test_fp32_restatement_tracks_fp64(sph, orc):
    d = sph.scenes.dam_break(48, 48)
    a32 = orc.run(d.params, d.rx, d.ry, d.vx, d.vy, 5)
    a64 = orc.run(d.params, d.rx, d.ry, d.vx, d.vy, 5, dtype=np.float64)
    o32 = np.argsort(a32["pid"]); o64 = np.argsort(a64["pid"])
    assert np.max(np.abs(a32["rho"][o32] - a64["rho"][o64])) < 1e-4 * 1000
    amag = np.sqrt(a64["ax"] ** 2 + a64["ay"] ** 2)
    err = np.hypot(a32["ax"][o32] - a64["ax"][o64], a32["ay"][o32] - a64["ay"][o64])
    assert err.max() < 1e-3 * amag.max()


def test_batched_scene_generator(sph):
    """Seeded, deterministic, counts within the drawn range, blocks inside the
    unit box, per-scene h, mass and dt from each scene's own lattice spacing and
    the 16k dam break's sound speed."""
    b1 = sph.scenes.batched_dam_breaks(n_scenes=16)
    b2 = sph.scenes.batched_dam_breaks(n_scenes=16)
    assert np.array_equal(b1.rx, b2.rx) and np.array_equal(b1.counts, b2.counts)
    assert b1.n == int(b1.counts.sum()) and b1.rx.dtype == np.float32
    assert np.all(b1.counts >= 1900) and np.all(b1.counts <= 66000)
    for s in range(16):
        sc = b1.scene(s)
        p = sc.params
        dx = p.h / 1.3
        assert abs(p.mass - 1000.0 * dx * dx) < 1e-12
        assert abs(p.dt - 0.25 * p.h / p.c0) < 1e-15
        assert p.c0 == 10.0 * math.sqrt(2.0 * 9.81 * 0.25)  # the 16k dam break's sound speed
        assert sc.ry.max() < 1.0 and sc.rx.max() < 0.51
    assert len({p.h for p in b1.params}) == 16


def test_batch_default_candidates(sph):
    c = sph.Batch.default_candidates()
    assert len(c) == 24 and len(set(c)) == 24
    assert {t for _, t, _ in c} == {64, 128, 256, 512}
