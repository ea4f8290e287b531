"""GPU parity: every stage of the HIP path against the CPU oracle, through the C ABI.

Integer outputs (keys, permutation, cell ranges, neighbor lists) must be
bit-exact. Float outputs are compared with the FP32 oracle (same operation
order; stated tolerance 4 ulp-scale: rtol 1e-6) and with the FP64 oracle
(stated tolerances per quantity below).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# FP64-oracle tolerances of one evaluation (FP32 on the GPU)
RHO_RTOL = 2e-6          # density: relative, per particle
ACC_RTOL = 2e-4          # acceleration: |a - a64| <= ACC_RTOL * max|a64|


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def scenes(sph):
    rng = np.random.default_rng(42)
    out = {}
    out["reference_example"] = sph.scenes.reference_example()
    out["reference_rowmajor"] = sph.scenes.reference_example(order=sph.ORDER_ROWMAJOR)
    out["dam_break_16k"] = sph.scenes.dam_break()
    s = sph.scenes.periodic_shear(n_side=96, seed=9)
    out["periodic_9k"] = s        # 73 x 73 cells (odd: stencil rows of three loads)
    out["periodic_even"] = sph.scenes.periodic_shear(n_side=100, seed=10)  # 76 x 76 cells
    # random gas with clusters / duplicates
    d = sph.scenes.dam_break(80, 80, seed=5, name="random")
    d.rx = rng.random(d.n).astype(np.float32) * 0.3
    d.ry = rng.random(d.n).astype(np.float32) * 0.3
    d.rx[:100] = d.rx[100:200]
    d.ry[:100] = d.ry[100:200]
    d.vx = rng.normal(0, 0.3, d.n).astype(np.float32)
    d.vy = rng.normal(0, 0.3, d.n).astype(np.float32)
    out["random_clusters"] = d
    # more particles in one Morton block's halo box than the LDS stage holds:
    # exercises the global-memory path of the block kernel
    e = sph.scenes.dam_break(60, 60, seed=6, name="dense")
    h = e.params.h
    e.rx = (0.3 + rng.random(e.n) * 4 * h).astype(np.float32)
    e.ry = (0.3 + rng.random(e.n) * 4 * h).astype(np.float32)
    e.rx[:1000] = (0.5 + rng.random(1000) * 30 * h).astype(np.float32)
    out["dense_cluster"] = e
    return out


@pytest.fixture(scope="module")
def all_scenes(sph):
    return scenes(sph)


def oracle_sorted(orc, p, rx, ry, vx, vy):
    k = orc.keys(p, rx, ry)
    perm, sk = orc.stable_sort(k)
    g = orc.grid(p)
    st, en = orc.cell_ranges(sk, g.n_keys)
    return dict(keys=k, perm=perm, sk=sk, start=st, end=en, rx=rx[perm], ry=ry[perm],
                vx=vx[perm], vy=vy[perm])


NAMES = ["reference_example", "reference_rowmajor", "dam_break_16k", "periodic_9k",
         "random_clusters"]


@pytest.mark.parametrize("name", NAMES)
def test_find_neighbors_bit_exact(sph, orc, all_scenes, name):
    s = all_scenes[name]
    note(f"[gpu] find_neighbors {name} n={s.n}")
    o = oracle_sorted(orc, s.params, s.rx, s.ry, s.vx, s.vy)
    with sph.Context(s.params, s.n) as ctx:
        g = ctx.grid()
        og = orc.grid(s.params)
        assert (g.nx, g.ny, g.nxp, g.nyp, g.n_keys, g.key_bits) == (
            og.nx, og.ny, og.nxp, og.nyp, og.n_keys, og.key_bits)
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.find_neighbors()
        assert np.array_equal(ctx.sorted_keys(), o["sk"])
        assert np.array_equal(ctx.permutation(), o["perm"])
        st, en = ctx.cell_ranges()
        assert np.array_equal(st, o["start"]) and np.array_equal(en, o["end"])
        d = ctx.download()
        assert np.array_equal(d["rx"], o["rx"]) and np.array_equal(d["vy"], o["vy"])
        assert np.array_equal(d["pid"], o["perm"])
        # a second sort of the already sorted state is the identity and leaves
        # no stale cell ranges behind
        ctx.find_neighbors()
        assert np.array_equal(ctx.permutation(), np.arange(s.n, dtype=np.uint32))
        st2, en2 = ctx.cell_ranges()
        assert np.array_equal(st2, o["start"]) and np.array_equal(en2, o["end"])


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("include_self", [False, True])
def test_csr_lists_bit_exact(sph, orc, all_scenes, name, include_self):
    s = all_scenes[name]
    note(f"[gpu] csr {name} self={include_self}")
    o = oracle_sorted(orc, s.params, s.rx, s.ry, s.vx, s.vy)
    cnt, off, idx = orc.neighbor_csr(s.params, o["rx"], o["ry"], o["start"], o["end"],
                                     include_self=include_self)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.find_neighbors()
        tot = ctx.build_neighbor_lists(include_self)
        assert tot == idx.size
        goff, gidx = ctx.neighbor_lists(tot)
        assert np.array_equal(goff, off)
        assert np.array_equal(gidx, idx)   # same stencil order, element by element


@pytest.mark.parametrize("name", NAMES)
def test_density_forces_stages(sph, orc, all_scenes, name):
    s = all_scenes[name]
    note(f"[gpu] density/forces {name}")
    o = oracle_sorted(orc, s.params, s.rx, s.ry, s.vx, s.vy)
    cnt, off, idx = orc.neighbor_csr(s.params, o["rx"], o["ry"], o["start"], o["end"])
    rho, pres, pr = orc.density(s.params, o["rx"], o["ry"], off, idx)
    ax, ay = orc.forces(s.params, o["rx"], o["ry"], o["vx"], o["vy"], rho, pr, off, idx)
    X = [o[k].astype(np.float64) for k in ("rx", "ry", "vx", "vy")]
    rho64, _, pr64 = orc.density(s.params, X[0], X[1], off, idx, dtype=np.float64)
    ax64, ay64 = orc.forces(s.params, *X, rho64, pr64, off, idx, dtype=np.float64)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.find_neighbors()
        results = {}
        for mode in ("sorted", "lists", "fused"):
            if mode == "sorted":
                ctx.density_pressure(sph.COMPUTE_SORTED)
                ctx.forces(sph.COMPUTE_SORTED)
            elif mode == "lists":
                ctx.build_neighbor_lists(False)
                ctx.density_pressure(sph.COMPUTE_LISTS)
                ctx.forces(sph.COMPUTE_LISTS)
            else:
                ctx.prepare()   # re-sorts an already sorted state: identity
                ctx.forces(sph.COMPUTE_LISTS)
            results[mode] = ctx.fields()
        for mode, f in results.items():
            # FP32 oracle: same operation order
            assert np.allclose(f["rho"], rho, rtol=1e-6, atol=0), mode
            assert np.allclose(f["p"], pres, rtol=1e-5, atol=1e-6 * np.abs(pres).max()), mode
            amax = max(np.abs(ax).max(), np.abs(ay).max())
            assert np.abs(f["ax"] - ax).max() <= 1e-5 * amax, mode
            assert np.abs(f["ay"] - ay).max() <= 1e-5 * amax, mode
            # FP64 oracle
            assert np.max(np.abs(f["rho"] - rho64) / np.abs(rho64)) <= RHO_RTOL, mode
            a64 = max(np.abs(ax64).max(), np.abs(ay64).max())
            assert np.abs(f["ax"] - ax64).max() <= ACC_RTOL * a64, mode
            assert np.abs(f["ay"] - ay64).max() <= ACC_RTOL * a64, mode
        exact = {m: (np.array_equal(f["rho"], rho), np.array_equal(f["ax"], ax)) for m, f in
                 results.items()}
        note(f"[gpu] bit-identical to FP32 oracle (rho, ax): {exact}")


@pytest.mark.parametrize("name", ["dam_break_16k", "periodic_9k", "periodic_even",
                                  "reference_example"])
def test_ell_lists_bit_exact(sph, orc, all_scenes, name):
    s = all_scenes[name]
    o = oracle_sorted(orc, s.params, s.rx, s.ry, s.vx, s.vy)
    cnt, off, idx = orc.neighbor_csr(s.params, o["rx"], o["ry"], o["start"], o["end"])
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        gcnt, slots = ctx.ell()
        assert np.array_equal(gcnt, cnt)
        w = slots.shape[1]
        for i in range(s.n):
            k = min(cnt[i], w)
            assert np.array_equal(slots[i, :k], idx[off[i]:off[i] + k])
        assert ctx.ell_overflow() == int(np.sum(cnt > w))


def corner_pairs(sph, orc, periodic):
    """Pairs straddling the support radius across stencil corners: one particle
    of each pair on a cell corner (the nearest float inside the diagonal cell),
    the other inside the cell at distance h (1 + e) from it, |e| <= 2e-5, for
    the four corners of cells spread over the box (and across the seam when
    periodic): the diagonal stencil cells at the edge of the support, where a
    walk that prunes cells or rounds the cell assignment differently would
    lose or gain neighbors."""
    base = (sph.scenes.periodic_shear(n_side=96, seed=3) if periodic
            else sph.scenes.dam_break(40, 40, seed=3, name="corner_pairs"))
    p = base.params
    g = orc.grid(p)
    lo = np.array([g.lo_x, g.lo_y], np.float64)
    cs = np.array([g.cs_x, g.cs_y], np.float64)
    nc = np.array([g.nx, g.ny])
    length = np.array([g.len_x, g.len_y], np.float64)
    rng = np.random.default_rng(11)
    first = 0 if periodic else 2
    cand = [(cx, cy) for cx in range(first, g.nx - first, 3) for cy in range(first, g.ny - first, 3)]
    picks = [cand[k] for k in rng.choice(len(cand), 40, replace=False)]
    if periodic:
        picks += [(0, 0), (g.nx - 1, g.ny - 1), (0, g.ny // 2)]
    pts = []
    for cx, cy in picks:
        for sx in (-1, 1):
            for sy in (-1, 1):
                c = np.array([cx, cy])
                s = np.array([sx, sy])
                corner = lo + (c + (s > 0)) * cs
                want = c + s
                if periodic:
                    want = np.mod(want, nc)
                elif np.any(want < 0) or np.any(want >= nc):
                    continue
                # the corner particle: the float32 nearest the corner that is
                # classified into the diagonal cell (wrapped across the seam)
                j = corner.astype(np.float32)
                ok = True
                for ax in (0, 1):
                    lo_a, hi_a = np.float32(lo[ax]), np.float32(lo[ax] + length[ax])
                    step = np.float32(-np.inf if s[ax] < 0 else np.inf)
                    for _ in range(16):
                        if periodic and j[ax] < lo_a:
                            j[ax] = np.nextafter(hi_a, np.float32(-np.inf))
                        elif periodic and j[ax] >= hi_a:
                            j[ax] = lo_a
                        if orc.cell_coord(float(j[ax]), float(lo_a), float(np.float32(cs[ax])),
                                          int(nc[ax])) == want[ax]:
                            break
                        j[ax] = np.nextafter(j[ax], step)
                    else:
                        ok = False
                if not ok:
                    continue
                for th in np.linspace(0.15, 1.42, 6):
                    for e in np.linspace(-2e-5, 2e-5, 9):
                        r = p.h * (1.0 + e)
                        i = corner - s * r * np.array([np.cos(th), np.sin(th)])
                        if periodic:
                            i = lo + np.mod(i - lo, length)
                        i32 = i.astype(np.float32)
                        if (orc.cell_coord(float(i32[0]), float(g.lo_x), float(g.cs_x), g.nx) != cx or
                                orc.cell_coord(float(i32[1]), float(g.lo_y), float(g.cs_y), g.ny) != cy):
                            continue
                        pts.append(i32)
                        pts.append(j.copy())
    pts = np.array(pts, np.float32)
    n = len(pts)
    z = np.zeros(n, np.float32)
    return p, pts[:, 0].copy(), pts[:, 1].copy(), z, z.copy()


@pytest.mark.parametrize("periodic", [False, True])
def test_corner_pairs_lists_and_density(sph, orc, periodic):
    """Neighbor lists and densities bit-exact for pairs across stencil corners
    at h (1 + e): accepted and rejected exactly as the oracle's 3x3 walk
    does."""
    p, rx, ry, vx, vy = corner_pairs(sph, orc, periodic)
    note(f"[gpu] corner pairs periodic={periodic}: n={rx.size}")
    o = oracle_sorted(orc, p, rx, ry, vx, vy)
    cnt, off, idx = orc.neighbor_csr(p, o["rx"], o["ry"], o["start"], o["end"])
    rho, _, _ = orc.density(p, o["rx"], o["ry"], off, idx)
    # the set exercises both sides of the threshold: some pairs whose FP32
    # squared distance is within 1e-4 of h^2 are accepted, some rejected
    x, y = o["rx"], o["ry"]
    h2 = np.float32(p.h) * np.float32(p.h)
    k0 = np.arange(0, rx.size, 2)
    a = np.argsort(o["perm"])  # sorted position of each input particle
    dx = (x[a[k0]] - x[a[k0 + 1]]).astype(np.float64)
    dy = (y[a[k0]] - y[a[k0 + 1]]).astype(np.float64)
    if periodic:
        dx -= np.round(dx / (p.hi[0] - p.lo[0])) * (p.hi[0] - p.lo[0])
        dy -= np.round(dy / (p.hi[1] - p.lo[1])) * (p.hi[1] - p.lo[1])
    rr = (dx * dx + dy * dy) / float(h2)
    near = np.abs(rr - 1.0) < 1e-4
    assert np.sum(near & (rr <= 1.0)) > 20 and np.sum(near & (rr > 1.0)) > 20
    with sph.Context(p, rx.size) as ctx:
        ctx.upload(rx, ry, vx, vy)
        ctx.prepare()
        gcnt, slots = ctx.ell()
        f = ctx.fields()
    assert np.array_equal(gcnt, cnt)
    w = slots.shape[1]
    for i in range(rx.size):
        k = min(cnt[i], w)
        assert np.array_equal(slots[i, :k], idx[off[i]:off[i] + k]), i
    assert np.array_equal(f["rho"], rho)


def test_ell_overflow_fallback(sph, orc):
    # a narrow list forces the stencil fallback; forces must not change
    s = sph.scenes.reference_example()
    s.params.ell_width = 4
    o = oracle_sorted(orc, s.params, s.rx, s.ry, s.vx, s.vy)
    cnt, off, idx = orc.neighbor_csr(s.params, o["rx"], o["ry"], o["start"], o["end"])
    rho, pres, pr = orc.density(s.params, o["rx"], o["ry"], off, idx)
    ax, ay = orc.forces(s.params, o["rx"], o["ry"], o["vx"], o["vy"], rho, pr, off, idx)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        assert ctx.ell_overflow() == int(np.sum(cnt > 4)) > 0
        ctx.forces(sph.COMPUTE_LISTS)
        f = ctx.fields()
        amax = np.abs(ay).max()
        assert np.abs(f["ax"] - ax).max() <= 1e-5 * amax
        assert np.abs(f["ay"] - ay).max() <= 1e-5 * amax


@pytest.mark.parametrize("resort", [0, 1])
@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("name,steps", [("reference_example", 50), ("reference_rowmajor", 20),
                                        ("dam_break_16k", 40), ("periodic_9k", 40),
                                        ("periodic_even", 30), ("random_clusters", 20),
                                        ("dense_cluster", 3)])
def test_multistep_run_bit_exact_fp32(sph, orc, all_scenes, name, steps, variant, resort):
    """The whole pipeline, `steps` iterations, with each step-kernel variant
    (split blocks, fused blocks, per particle) and each re-sort (full radix,
    incremental): identical sort permutations every step (same particle ids in
    the same order) and bit-identical FP32 state."""
    s = all_scenes[name]
    note(f"[gpu] run {name} steps={steps} step_kernel={variant} resort={resort}")
    o32 = orc.run(s.params, s.rx, s.ry, s.vx, s.vy, steps)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.set_step_kernel(variant)
        ctx.set_resort(resort)
        ctx.run(steps)
        d = ctx.download()
        f = ctx.fields()
    assert np.array_equal(d["pid"], o32["pid"])
    for k in ("rx", "ry", "vx", "vy"):
        assert np.array_equal(d[k], o32[k]), k
    assert np.array_equal(f["rho"], o32["rho"])
    assert np.array_equal(f["ax"], o32["ax"]) and np.array_equal(f["ay"], o32["ay"])


# FP64 comparison horizon per scene (the random-cluster gas is chaotic: nearly
# coincident particles), and stated tolerances after that many steps
FP64_CASES = [("reference_example", 10), ("dam_break_16k", 40), ("periodic_9k", 40),
              ("random_clusters", 2)]
FP64_POS_TOL_H = 1e-3   # max |r - r64| / h
FP64_RHO_RTOL = 1e-3    # max |rho - rho64| / rho64
FP64_ACC_TOL = 1e-2     # max |a - a64| / max |a64|


@pytest.mark.parametrize("name,steps", FP64_CASES)
def test_multistep_run_vs_fp64(sph, orc, all_scenes, name, steps):
    s = all_scenes[name]
    o64 = orc.run(s.params, s.rx, s.ry, s.vx, s.vy, steps, dtype=np.float64)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.run(steps)
        d = ctx.download()
        f = ctx.fields()
    h = s.params.h
    i32 = np.argsort(d["pid"])
    i64 = np.argsort(o64["pid"])
    perr = np.hypot(d["rx"][i32] - o64["rx"][i64], d["ry"][i32] - o64["ry"][i64]).max() / h
    rerr = (np.abs(f["rho"][i32] - o64["rho"][i64]) / o64["rho"][i64]).max()
    a64 = np.hypot(o64["ax"], o64["ay"]).max()
    aerr = np.hypot(f["ax"][i32] - o64["ax"][i64], f["ay"][i32] - o64["ay"][i64]).max() / a64
    note(f"[gpu] {name} after {steps} steps vs FP64: position {perr:.3e} h, density "
         f"{rerr:.3e} rel, acceleration {aerr:.3e} of max")
    assert perr <= FP64_POS_TOL_H
    assert rerr <= FP64_RHO_RTOL
    assert aerr <= FP64_ACC_TOL


@pytest.mark.parametrize("name", ["dam_break_16k", "random_clusters", "periodic_9k"])
def test_incremental_resort_movers_and_permutation(sph, orc, all_scenes, name):
    """One step with the incremental re-sort: the mover count equals the number of
    particles whose cell key changed, and the permutation equals the oracle's
    stable sort of the integrated keys."""
    s = all_scenes[name]
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        before = ctx.download()
        k_before = ctx.sorted_keys()
        ctx.step(1)
        after = ctx.download()
        perm = ctx.permutation()
        m = ctx.movers()
    # integrated positions in the previous order: after[...] through perm
    x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
    x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
    k_new = orc.keys(s.params, x_prev, y_prev)
    assert m == int(np.sum(k_new != k_before))
    operm, _ = orc.stable_sort(k_new)
    assert np.array_equal(perm, operm)
    assert np.array_equal(after["pid"], before["pid"][perm])
    note(f"[gpu] {name}: {m} movers of {s.n}")


@pytest.mark.parametrize("name", ["dam_break_16k", "random_clusters", "periodic_9k",
                                  "packed_cluster"])
def test_incremental_resort_cell_ranges(sph, orc, all_scenes, name):
    """Cell ranges after incremental re-sorts, read straight from the table the
    step leaves (no sort in between), against the oracle's ranges of the sorted
    keys: every occupied cell's (start, end) and -1 for every other cell.
    packed_cluster puts 2,600 particles in a 2h x 2h square, so one cell's run
    spans several 256-particle merge tiles."""
    if name == "packed_cluster":
        s = sph.scenes.dam_break(60, 60, seed=6, name="packed")
        rng = np.random.default_rng(8)
        h = s.params.h
        s.rx = (0.3 + rng.random(s.n) * 2 * h).astype(np.float32)
        s.ry = (0.3 + rng.random(s.n) * 2 * h).astype(np.float32)
        s.rx[:1000] = (0.5 + rng.random(1000) * 30 * h).astype(np.float32)
    else:
        s = all_scenes[name]
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        n_keys = int(ctx.grid().n_keys)
        for step in range(3):
            ctx.step(1)
            sk = ctx.sorted_keys()
            st, en = ctx.cell_ranges()
            ost, oen = orc.cell_ranges(sk, n_keys)
            note(f"[gpu] {name} step {step}: {ctx.movers()} movers, "
                 f"{int(np.sum(ost >= 0))} occupied cells, largest {int((oen - ost).max())}")
            assert np.all(sk[1:] >= sk[:-1])
            assert np.array_equal(st, ost) and np.array_equal(en, oen)
            if name == "packed_cluster" and step == 0:
                assert int((oen - ost).max()) > 256


def test_block_stats_counts_capacity_fallbacks(sph, all_scenes):
    s = all_scenes["dense_cluster"]
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        st = ctx.block_stats()
        note(f"[gpu] dense_cluster block stats {st}")
        assert st["density"]["blocks"] > 0 and st["density"]["over_capacity"] > 0
        d = sph.scenes.dam_break()
    with sph.Context(d.params, d.n) as ctx:
        ctx.upload(d.rx, d.ry, d.vx, d.vy)
        ctx.prepare()
        st = ctx.block_stats()
        assert st["density"]["over_capacity"] == 0 and st["force"]["over_capacity"] == 0
        assert st["density"]["staged"] >= d.n


def test_fused_step_equals_stage_path(sph, all_scenes):
    s = all_scenes["dam_break_16k"]
    with sph.Context(s.params, s.n) as a, sph.Context(s.params, s.n) as b:
        a.upload(s.rx, s.ry, s.vx, s.vy)
        b.upload(s.rx, s.ry, s.vx, s.vy)
        a.run(6)
        # stage by stage, the reference's loop order
        b.find_neighbors(); b.build_neighbor_lists(False)
        b.density_pressure(sph.COMPUTE_LISTS); b.forces(sph.COMPUTE_LISTS)
        for _ in range(5):
            b.integrate(); b.find_neighbors(); b.build_neighbor_lists(False)
            b.density_pressure(sph.COMPUTE_LISTS); b.forces(sph.COMPUTE_LISTS)
        da, db = a.download(), b.download()
        fa, fb = a.fields(), b.fields()
        for k in ("rx", "ry", "vx", "vy", "pid"):
            assert np.array_equal(da[k], db[k]), k
        for k in ("ax", "ay", "rho"):
            assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("pairs", [1, 4])
def test_graph_replay_equals_direct(sph, all_scenes, pairs):
    """Graphs of 2 * pairs iterations: several replays with eager work in
    between, then two replays and three direct iterations in one call, equal
    direct steps, state and fields."""
    s = all_scenes["periodic_9k"]
    per = 2 * pairs
    with sph.Context(s.params, s.n) as a, sph.Context(s.params, s.n) as b:
        for c in (a, b):
            c.upload(s.rx, s.ry, s.vx, s.vy)
            c.prepare()
        b.set_graph(True, pairs)
        for _ in range(3):   # several replays with eager work in between
            a.step(per)
            b.step(per)
            a.measure(); b.measure()
            da, db = a.download(), b.download()
            for k in ("rx", "ry", "vx", "vy", "pid"):
                assert np.array_equal(da[k], db[k]), k
        b.step(2 * per + 3)
        a.step(2 * per + 3)
        da, db = a.download(), b.download()
        for k in ("rx", "ry", "vx", "vy", "pid"):
            assert np.array_equal(da[k], db[k]), k
        fa, fb = a.fields(), b.fields()
        for k in ("ax", "ay", "rho"):
            assert np.array_equal(fa[k], fb[k]), k


def test_graph_cache_parities_and_remainders(sph, all_scenes):
    """Graphs are kept per buffer assignment (which alternates every step) and
    per length: calls that start after odd and even step counts, whole graphs
    followed by a shorter graph for the even rest and an odd last step, and
    graphs captured ahead by step_graphs (which runs nothing) all equal direct
    steps, state and fields."""
    s = all_scenes["dam_break_16k"]
    with sph.Context(s.params, s.n) as a, sph.Context(s.params, s.n) as b:
        for c in (a, b):
            c.upload(s.rx, s.ry, s.vx, s.vy)
            c.prepare()
        b.set_graph(True, 4)
        for k, ahead in ((3, False), (12, True), (1, False), (20, True), (5, False),
                         (20, False), (14, True), (9, True)):
            note(f"[gpu] graph cache: call of {k} steps, captured ahead={ahead}")
            if ahead:
                before = b.download()
                b.step_graphs(k)
                after = b.download()
                for q in ("rx", "vx", "pid"):
                    assert np.array_equal(before[q], after[q]), ("step_graphs ran steps", q)
            a.step(k)
            b.step(k)
            da, db = a.download(), b.download()
            for q in ("rx", "ry", "vx", "vy", "pid"):
                assert np.array_equal(da[q], db[q]), (k, q)
            fa, fb = a.fields(), b.fields()
            for q in ("ax", "ay", "rho"):
                assert np.array_equal(fa[q], fb[q]), (k, q)


def test_step_calls_repack_after_changes(sph, all_scenes):
    """A step call right after another one reuses the packed copy of the state
    the first one left; any other call that may change the state, and any
    write through a device view, makes the next step call pack afresh. Each
    case equals the same sequence with a state-preserving call (measure) that
    forces the pack, state and fields."""
    import ctypes as C
    s = all_scenes["dam_break_16k"]

    def run(seq):
        with sph.Context(s.params, s.n) as ctx:
            ctx.upload(s.rx, s.ry, s.vx, s.vy)
            ctx.prepare()
            ctx.set_graph(True, 1)
            view = None
            for op in seq:
                if op == "view":
                    view = ctx.device_view()
                elif op == "measure":
                    ctx.measure()
                elif op == "kick":
                    ctx.update_velocities()
                elif op == "write_vx":
                    # halve vx through the view's pointer, no call on the context
                    ctx.synchronize()
                    buf = np.empty(s.n, np.float32)
                    assert sph.lib().sph_device_copy(buf.ctypes.data_as(C.c_void_p),
                                                     C.c_void_p(view.vx), s.n * 4, 1) == 0
                    buf *= np.float32(0.5)
                    assert sph.lib().sph_device_copy(C.c_void_p(view.vx),
                                                     buf.ctypes.data_as(C.c_void_p),
                                                     s.n * 4, 0) == 0
                else:
                    ctx.step(op)
            if view is not None:
                assert ctx.device_view().vx == view.vx   # the pointer written through
            return ctx.download(), ctx.fields()

    cases = [([4, 4], [4, "measure", 4], [8]),
             ([4, "kick", 4], [4, "kick", "measure", 4], None),
             (["view", 4, "write_vx", 4], ["view", 4, "write_vx", "measure", 4], None),
             ([3, 5, 2], [3, "measure", 5, "measure", 2], [10])]
    for seq, ref, direct in cases:
        note(f"[gpu] step calls {seq} vs {ref}")
        (d, f), (d0, f0) = run(seq), run(ref)
        outs = [(d0, f0)] + ([run(direct)] if direct else [])
        for dr, fr in outs:
            for k in ("pid", "rx", "ry", "vx", "vy"):
                assert np.array_equal(d[k], dr[k]), (seq, k)
            for k in ("ax", "ay", "rho"):
                assert np.array_equal(f[k], fr[k]), (seq, k)
    # and the write through the view did change the outcome
    (d1, _), (d2, _) = run(["view", 4, "write_vx", 4]), run(["view", 4, 4])
    assert not np.array_equal(d1["vx"], d2["vx"])


@pytest.mark.parametrize("name", ["dam_break_16k", "random_clusters"])
def test_measure_fp64(sph, orc, all_scenes, name):
    s = all_scenes[name]
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        m = ctx.measure()
    o = orc.measure(s.params, s.rx, s.ry, s.vx, s.vy)
    assert m[0] == pytest.approx(o[0], rel=1e-12, abs=1e-300)
    assert m[1] == o[1]
    assert m[2] == o[2]
    assert m[3] == pytest.approx(o[3], rel=1e-12)
    assert m[4] == pytest.approx(o[4], rel=1e-15)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 2047, 2048, 2049, 4097])
def test_sort_edge_sizes(sph, orc, n):
    # ragged tails of the 2048-key tile and the 64-lane wave; one particle
    rng = np.random.default_rng(n)
    p = sph.scenes.wcsph_params(0.02, (0, 0), (1, 1))
    rx = rng.random(n).astype(np.float32)
    ry = rng.random(n).astype(np.float32)
    rx[: n // 3] = rx[0]
    ry[: n // 3] = ry[0]
    o = oracle_sorted(orc, p, rx, ry, rx, ry)
    with sph.Context(p, 5000) as ctx:
        ctx.upload(rx, ry, rx, ry)
        ctx.find_neighbors()
        assert np.array_equal(ctx.permutation(), o["perm"])
        st, en = ctx.cell_ranges()
        assert np.array_equal(st, o["start"]) and np.array_equal(en, o["end"])


def test_reupload_smaller_clears_old_ranges(sph, orc):
    p = sph.scenes.wcsph_params(0.02, (0, 0), (1, 1))
    rng = np.random.default_rng(1)
    with sph.Context(p, 4000) as ctx:
        for n in (4000, 37, 1500):
            rx = rng.random(n).astype(np.float32)
            ry = rng.random(n).astype(np.float32)
            o = oracle_sorted(orc, p, rx, ry, rx, ry)
            ctx.upload(rx, ry, rx, ry)
            ctx.find_neighbors()
            st, en = ctx.cell_ranges()
            assert np.array_equal(st, o["start"]) and np.array_equal(en, o["end"])


def test_capacity_and_state_errors(sph):
    p = sph.scenes.wcsph_params(0.02, (0, 0), (1, 1))
    with sph.Context(p, 10) as ctx:
        with pytest.raises(sph.SphError, match="capacity"):
            ctx.upload(*(np.zeros(11, np.float32),) * 4)
        ctx.upload(*(np.full(10, 0.5, np.float32),) * 4)
        with pytest.raises(sph.SphError, match="find_neighbors"):
            ctx.density_pressure()
        with pytest.raises(sph.SphError, match="prepare"):
            ctx.step(1)


def test_incremental_resort_buckets_many_chunks(sph, orc):
    """About 2,700 movers among 5625 re-sort tiles (1.44 M particles, 88
    counter chunks of 64 tiles, so the merge's prefixes sum chunk counters
    over more than one wave): every mover goes into its owner tile's bucket;
    the permutation is still the oracle's stable sort and the mover count the
    number of changed keys."""
    s = sph.scenes.dam_break(1200, 1200, box=(4.0, 4.0), seed=5)
    rng = np.random.default_rng(11)
    pick = rng.choice(s.n, 1000, replace=False)  # plus about 1,700 slow movers
    s.vx[pick] = rng.normal(0, 400.0, pick.size).astype(np.float32)  # several cells per step
    s.vy[pick] = rng.normal(0, 400.0, pick.size).astype(np.float32)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        k_before = ctx.sorted_keys()
        ctx.step(1)
        after = ctx.download()
        perm = ctx.permutation()
        m = ctx.movers()
        path = ctx.resort_path()
        sk = ctx.sorted_keys()
        st, en = ctx.cell_ranges()
        n_keys = int(ctx.grid().n_keys)
    note(f"[gpu] bucketed movers, {s.n} particles: {m} ({path})")
    assert 0 < m <= 4096 and path == "bucket"
    ost, oen = orc.cell_ranges(sk, n_keys)
    assert np.array_equal(st, ost) and np.array_equal(en, oen)
    x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
    x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
    k_new = orc.keys(s.params, x_prev, y_prev)
    assert m == int(np.sum(k_new != k_before))
    operm, _ = orc.stable_sort(k_new)
    assert np.array_equal(perm, operm)


def test_incremental_resort_bucket_rounds(sph, orc):
    """One bucketing workgroup's 16,384 particles with about one in seven of
    them moving one cell: more movers than its LDS list holds (listed in rounds
    of 2048) and windows of tens of movers per tile, within the 64 a bucket
    holds (the merge loads the entries past its first 8 in a second round);
    still the bucket path, and the oracle's permutation and cell ranges."""
    s = sph.scenes.dam_break(128, 128, seed=21)
    with sph.Context(s.params, s.n) as ctx:
        cs = float(ctx.grid().cs_x)
        rng = np.random.default_rng(23)
        # displacements of about 0.095 cell per axis and step
        s.vx = (0.095 * cs / s.params.dt * rng.normal(0, 1, s.n)).astype(np.float32)
        s.vy = (0.095 * cs / s.params.dt * rng.normal(0, 1, s.n)).astype(np.float32)
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        for step in range(2):
            k_before = ctx.sorted_keys()
            ctx.step(1)
            after = ctx.download()
            perm = ctx.permutation()
            m = ctx.movers()
            path = ctx.resort_path()
            note(f"[gpu] bucket rounds, step {step}: {m} of {s.n} ({path})")
            assert m > 2048 and path == "bucket"
            x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
            x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
            k_new = orc.keys(s.params, x_prev, y_prev)
            assert m == int(np.sum(k_new != k_before))
            operm, sk = orc.stable_sort(k_new)
            assert np.array_equal(perm, operm)
            st, en = ctx.cell_ranges()
            ost, oen = orc.cell_ranges(sk, int(ctx.grid().n_keys))
            assert np.array_equal(st, ost) and np.array_equal(en, oen)


@pytest.mark.parametrize("name", ["dam_break_16k", "periodic_9k", "random_clusters"])
def test_step_call_splits_equivalent(sph, all_scenes, name):
    """Six steps as one sph_step call or split over several calls give the same
    state and the same fields: single-step calls have the merge write the
    separate state arrays, longer calls keep only the packed copies until they
    return, and the accelerations and separate density arrays come from a
    call's last iteration only (the first of two graph-captured iterations
    skips them)."""
    s = all_scenes[name]
    outs = []
    for split in ([6], [1] * 6, [2, 3, 1], [1, 5], [3, 3]):
        with sph.Context(s.params, s.n) as ctx:
            ctx.upload(s.rx, s.ry, s.vx, s.vy)
            ctx.prepare()
            for k in split:
                ctx.step(k)
            outs.append((split, ctx.download(), ctx.fields()))
    _, d0, f0 = outs[0]
    for split, d, f in outs[1:]:
        note(f"[gpu] step split {split} vs [6] on {name}")
        for k in ("pid", "rx", "ry", "vx", "vy"):
            assert np.array_equal(d[k], d0[k]), (split, k)
        for k in ("ax", "ay", "rho"):
            assert np.array_equal(f[k], f0[k]), (split, k)


MOVER_PATTERNS = {
    # name: (particles sent, kind, mover bounds of the first step, its path)
    # (a sent particle's neighbors feel its approach through the artificial
    # viscosity and move too: about three movers per particle sent)
    "spread": (18000, "spread", (16384, 32768), "bucket"),
    "spread_cells": (9000, "spread", (4096, 16384), "bucket"),
    "converge": (2500, "converge", (4096, 16384), "sorted"),
    "converge_grid": (6000, "converge", (16384, 32768), "sorted"),
    "empty_cells": (4500, "empty", (4096, 16384), "sorted"),
}


@pytest.mark.parametrize("pattern", list(MOVER_PATTERNS))
def test_incremental_resort_mid_movers(sph, orc, pattern):
    """Thousands of movers over 5625 tiles of 256. "spread": fast movers over
    many tiles, each tile's owned movers within its bucket; "converge":
    particles sent into one spot, one tile owning thousands of movers, so its
    bucket overflows and the step takes the sorted path (ordered compaction
    and a stable radix sort in one workgroup); "empty_cells": particles sent
    into empty space, many of them owned by one tile (sorted path; the
    owner search in the sampled tile-first table runs for the "spread" movers
    that land in empty cells). Two steps in
    a row; the permutation is the oracle's stable sort, the count exact and
    the cell ranges the oracle's."""
    count, kind, (lo, hi), want = MOVER_PATTERNS[pattern]
    s = sph.scenes.dam_break(1200, 1200, box=(4.0, 4.0), seed=5)
    rng = np.random.default_rng(17)
    pick = rng.choice(s.n, count, replace=False)
    dt = s.params.dt
    if kind == "spread":
        s.vx[pick] = rng.normal(0, 400.0, pick.size).astype(np.float32)
        s.vy[pick] = rng.normal(0, 400.0, pick.size).astype(np.float32)
    elif kind == "converge":
        s.vx[pick] = ((2.0 - s.rx[pick]) / dt).astype(np.float32)
        s.vy[pick] = ((1.0 - s.ry[pick]) / dt).astype(np.float32)
    else:
        s.vx[pick] = ((3.0 + 0.5 * rng.random(pick.size) - s.rx[pick]) / dt).astype(np.float32)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        for step in range(2):
            k_before = ctx.sorted_keys()
            ctx.step(1)
            after = ctx.download()
            perm = ctx.permutation()
            m = ctx.movers()
            path = ctx.resort_path()
            sk = ctx.sorted_keys()
            st, en = ctx.cell_ranges()
            note(f"[gpu] mid movers ({pattern}), step {step}: {m} of {s.n} ({path})")
            if step == 0:
                assert lo < m <= hi
                assert path == want
            x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
            x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
            k_new = orc.keys(s.params, x_prev, y_prev)
            assert m == int(np.sum(k_new != k_before))
            operm, osk = orc.stable_sort(k_new)
            assert np.array_equal(perm, operm)
            assert np.array_equal(sk, osk)
            ost, oen = orc.cell_ranges(osk, int(ctx.grid().n_keys))
            assert np.array_equal(st, ost) and np.array_equal(en, oen)


@pytest.mark.parametrize("steps", [1, 2])
def test_incremental_resort_empty_cell_owners(sph, orc, steps):
    """Movers into cells that held no particle, few enough that every bucket
    holds its owner's movers (bucket path): their owner tiles come from the
    search of the tile-first keys, by the whole wave, or by one lane in the
    last, partial wave (10,000 particles: its last wave has 16 lanes, and the
    last 16 particles in sorted order are among the movers). One- and two-step
    calls (64- and 32-bit keys); the permutation is the oracle's stable sort."""
    s = sph.scenes.dam_break(100, 100, box=(1.0, 1.0), seed=21)
    assert s.n % 64 == 16
    rng = np.random.default_rng(23)
    dt, h = s.params.dt, s.params.h
    k0 = orc.keys(s.params, s.rx, s.ry)
    order, _ = orc.stable_sort(k0)
    last = order[-16:]
    pick = np.unique(np.concatenate([last, rng.choice(s.n, 60, replace=False)]))
    # targets in empty cells of the two Morton quadrants of 64 x 64 cells that
    # the 77 x 77-cell block only partly fills (cells 80-125 by 2-61 and 2-61
    # by 80-125), so their keys interleave with the block's and their owners
    # spread over many tiles (at most 40 per bucket on this scene)
    side = rng.random(pick.size) < 0.5
    a = (80 + 46 * rng.random(pick.size)) * h
    b = (2 + 60 * rng.random(pick.size)) * h
    tx = np.where(side, a, b)
    ty = np.where(side, b, a)
    s.vx[pick] = ((tx - s.rx[pick]) / dt).astype(np.float32)
    s.vy[pick] = ((ty - s.ry[pick]) / dt).astype(np.float32)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        k_before = ctx.sorted_keys()
        pid_before = ctx.download()["pid"]
        ctx.step(steps)
        after = ctx.download()
        perm = ctx.permutation()
        m = ctx.movers()
        path = ctx.resort_path()
        note(f"[gpu] empty-cell owners, {steps}-step call: {m} movers ({path})")
        assert path == "bucket"
        # the last step's movers, from the state it started from
        x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
        x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
        k_new = orc.keys(s.params, x_prev, y_prev)
        if steps == 1:
            assert m == int(np.sum(k_new != k_before))
            assert m >= pick.size
            # the last wave's particles were all sent away
            moved_ids = set(pid_before[np.nonzero(k_new != k_before)[0]].tolist())
            assert set(last.tolist()) <= moved_ids
        operm, osk = orc.stable_sort(k_new)
        assert np.array_equal(perm, operm)
        assert np.array_equal(ctx.sorted_keys(), osk)
        st, en = ctx.cell_ranges()
        ost, oen = orc.cell_ranges(osk, int(ctx.grid().n_keys))
        assert np.array_equal(st, ost) and np.array_equal(en, oen)


@pytest.mark.parametrize("graph", [False, True])
def test_sorted_path_inside_multistep_call(sph, orc, graph):
    """A bucket overflows inside a multi-step call (32-bit keys, the merge
    orders the movers in its first workgroup while the others wait): 2,000
    particles of a 90,000-particle dam break sent into one spot, three
    reference iterations (prepare + a two-step call + the final density and
    forces), against the FP32 oracle: the same order and bit-identical state."""
    s = sph.scenes.dam_break(300, 300, seed=31)
    rng = np.random.default_rng(37)
    pick = rng.choice(s.n, 2000, replace=False)
    dt = s.params.dt
    s.vx[pick] = ((0.3 - s.rx[pick]) / dt).astype(np.float32)
    s.vy[pick] = ((0.3 - s.ry[pick]) / dt).astype(np.float32)
    o = orc.run(s.params, s.rx, s.ry, s.vx, s.vy, 3)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.set_graph(graph, 1)
        ctx.run(3)
        path = ctx.resort_path()
        d = ctx.download()
    note(f"[gpu] sorted path in a two-step call (graph={graph}): last step {path}")
    assert np.array_equal(d["pid"], o["pid"])
    for k in ("rx", "ry", "vx", "vy"):
        assert np.array_equal(d[k], o[k]), k


@pytest.mark.parametrize("side,sigma", [(100, 40.0), (512, 400.0)])
def test_incremental_resort_many_movers(sph, orc, side, sigma):
    """Most particles moving 0.5 to 5 cells per step, two steps in a row: 10 K
    particles (one bucketing workgroup) and 262 K particles (16 of them, more
    than 65,536 movers); tiles own more movers than a bucket holds, so both
    take the sorted path; the permutation is still the oracle's stable sort
    and the mover count exact."""
    s = sph.scenes.dam_break(side, side, seed=12)
    rng = np.random.default_rng(3)
    s.vx = rng.normal(0, sigma, s.n).astype(np.float32)  # ~0.5 to 5 cells per step
    s.vy = rng.normal(0, sigma, s.n).astype(np.float32)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        for step in range(2):
            k_before = ctx.sorted_keys()
            ctx.step(1)
            after = ctx.download()
            perm = ctx.permutation()
            m = ctx.movers()
            path = ctx.resort_path()
            note(f"[gpu] many movers, step {step}: {m} of {s.n} ({path})")
            assert m > 4096 and path == "sorted"
            x_prev = np.empty_like(after["rx"]); y_prev = np.empty_like(after["ry"])
            x_prev[perm] = after["rx"]; y_prev[perm] = after["ry"]
            k_new = orc.keys(s.params, x_prev, y_prev)
            assert m == int(np.sum(k_new != k_before))
            operm, sk = orc.stable_sort(k_new)
            assert np.array_equal(perm, operm)
            st, en = ctx.cell_ranges()
            ost, oen = orc.cell_ranges(sk, int(ctx.grid().n_keys))
            assert np.array_equal(st, ost) and np.array_equal(en, oen)
