"""GPU parity of the batched step: many scenes in one context, one launch per
stage. Each scene's trajectory must be bit-identical to the FP32 oracle run of
that scene alone with cell size = cell_factor * h (same sort permutation, same
FP32 state), for every tuning the autotuner can choose."""
import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_batch(sph):
    # uneven counts (a few hundred to a few thousand), own dx / h / dt each
    return sph.scenes.batched_dam_breaks(n_scenes=6, seed=0x5EED0102, n_lo=300, n_hi=3000,
                                         steps=20)


def scene_params(p, factor):
    return dataclasses.replace(p, cell_size=factor * p.h)


def check_scenes_vs_oracle(orc, b, d, f, factor, steps):
    for s in range(len(b.counts)):
        sc = b.scene(s)
        a0, a1 = int(b.first[s]), int(b.first[s] + b.counts[s])
        o = orc.run(scene_params(sc.params, factor), sc.rx, sc.ry, sc.vx, sc.vy, steps)
        assert np.array_equal(d["pid"][a0:a1] - a0, o["pid"]), f"scene {s}: permutation"
        for k in ("rx", "ry", "vx", "vy"):
            assert np.array_equal(d[k][a0:a1], o[k]), f"scene {s}: {k}"
        assert np.array_equal(f["rho"][a0:a1], o["rho"]), f"scene {s}: rho"
        assert np.array_equal(f["ax"][a0:a1], o["ax"]), f"scene {s}: ax"
        assert np.array_equal(f["ay"][a0:a1], o["ay"]), f"scene {s}: ay"


@pytest.mark.parametrize("tuning", [(1.0, 256, 1), (1.0, 64, 0), (1.25, 512, 1),
                                    (1.5, 128, 0)])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_bit_exact_per_scene(sph, orc, small_batch, tuning, graph):
    b = small_batch
    steps = 20
    note(f"[gpu] batch {len(b.counts)} scenes n={b.n} tuning={tuning} graph={graph}")
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.set_tuning(*tuning)
        ctx.set_graph(graph)
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        ctx.run(steps)
        d = ctx.download()
        f = ctx.fields()
        assert ctx.tuning() == tuning
    check_scenes_vs_oracle(orc, b, d, f, tuning[0], steps)


def test_batch_segments_keys_and_ranges(sph, orc, small_batch):
    """Sorted keys are non-decreasing, every scene keeps its segment, and the key
    of each particle is the scene base plus the scene-local cell key."""
    b = small_batch
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.set_tuning(1.25, 128, 1)
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        ctx.prepare()
        ctx.step(3)
        keys = ctx.sorted_keys()
        d = ctx.download()
        st, en = ctx.cell_ranges()
        infos = [ctx.scene_info(s) for s in range(len(b.counts))]
    assert np.all(keys[1:] >= keys[:-1])
    base = 0
    for s, (first, count, g) in enumerate(infos):
        assert first == b.first[s] and count == b.counts[s]
        seg = slice(first, first + count)
        assert np.all((d["pid"][seg] >= first) & (d["pid"][seg] < first + count))
        p = scene_params(b.params[s], 1.25)
        local = orc.keys(p, d["rx"][seg], d["ry"][seg])
        assert np.array_equal(keys[seg] - np.uint64(base), local)
        # ranges of this scene's table: closed form from its sorted keys
        os_, oe = orc.cell_ranges(local, int(g.n_keys))
        got_s = st[base:base + int(g.n_keys)].copy()
        got_e = en[base:base + int(g.n_keys)].copy()
        occ = got_s >= 0
        got_s[occ] -= first
        got_e[occ] -= first
        assert np.array_equal(got_s, os_) and np.array_equal(got_e, oe)
        base += (int(g.n_keys) + 1) & ~1  # tables start at even entries


def test_batch_autotune_restores_state_and_applies_best(sph, orc, small_batch):
    b = small_batch
    cands = [(1.0, 256, 1), (1.25, 64, 1), (1.0, 512, 0), (1.5, 128, 1)]
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        before = ctx.download()
        best, times = ctx.autotune(cands, trial_steps=4)
        note(f"[gpu] autotune best {best}: {times}")
        assert best in cands and ctx.tuning() == best
        assert all(t > 0 for t in times.values())
        assert times[best] == min(times.values())
        after = ctx.download()
        for k in ("rx", "ry", "vx", "vy", "pid"):
            assert np.array_equal(after[k], before[k]), k
        ctx.run(10)
        d = ctx.download()
        f = ctx.fields()
    check_scenes_vs_oracle(orc, b, d, f, best[0], 10)


def test_batch_autotune_skips_inapplicable_candidates(sph, small_batch):
    b = small_batch
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        best, times = ctx.autotune([(0.5, 256, 1), (1.0, 100, 1), (1.0, 128, 1)],
                                   trial_steps=2)
        assert best == (1.0, 128, 1)
        assert times[(0.5, 256, 1)] == -1.0 and times[(1.0, 100, 1)] == -1.0


def test_batch_measure_per_scene(sph, orc, small_batch):
    b = small_batch
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        ctx.prepare()
        ctx.step(5)
        m = ctx.measure_scenes()
        d = ctx.download()
    for s in range(len(b.counts)):
        a0, a1 = int(b.first[s]), int(b.first[s] + b.counts[s])
        o = orc.measure(b.params[s], d["rx"][a0:a1], d["ry"][a0:a1], d["vx"][a0:a1],
                        d["vy"][a0:a1])
        assert m[s, 2] == o[2] and m[s, 1] == o[1]
        np.testing.assert_allclose(m[s], o, rtol=1e-12, atol=0)


def test_batch_periodic_scenes_bit_exact(sph, orc):
    """A batch of periodic scenes with odd and even column counts (46, 49 and
    76 cells a side): every scene bit-identical to its own oracle run, and
    the scenes' cell tables at their documented key bases."""
    parts = [sph.scenes.periodic_shear(n_side=ns, seed=0x5EED0200 + ns) for ns in (60, 64, 100)]
    cat = lambda k: np.concatenate([getattr(sc, k) for sc in parts])
    b = sph.scenes.BatchScene("periodic_batch", [sc.params for sc in parts],
                              np.array([sc.n for sc in parts], np.int64), cat("rx"), cat("ry"),
                              cat("vx"), cat("vy"), 15)
    assert sorted(orc.grid(sc.params).nx % 2 for sc in parts) == [0, 0, 1]
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.set_tuning(1.0, 256, 1)
        ctx.set_graph(True)
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        ctx.run(b.steps)
        d = ctx.download()
        f = ctx.fields()
        keys = ctx.sorted_keys()
        infos = [ctx.scene_info(s) for s in range(len(b.counts))]
    check_scenes_vs_oracle(orc, b, d, f, 1.0, b.steps)
    base = 0
    for s, (first, count, g) in enumerate(infos):
        seg = slice(first, first + count)
        local = orc.keys(scene_params(b.params[s], 1.0), d["rx"][seg], d["ry"][seg])
        assert np.array_equal(keys[seg] - np.uint64(base), local)
        base += (int(g.n_keys) + 1) & ~1


def test_batch_errors(sph, small_batch):
    b = small_batch
    with sph.Batch(b.params, b.counts) as ctx:
        with pytest.raises(sph.SphError):
            ctx.upload(b.rx[:-1], b.ry[:-1], b.vx[:-1], b.vy[:-1])  # not every particle
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        with pytest.raises(sph.SphError):
            ctx.step(1)  # not prepared
        ctx.prepare()
        for call in (ctx.integrate, ctx.measure, ctx.update_velocities,
                     lambda: ctx.build_neighbor_lists(False), lambda: ctx.set_dt(1e-5)):
            with pytest.raises(sph.SphError):
                call()
        for bad in ((0.9, 256, 1), (1.0, 96, 1), (1.0, 1024, 1), (1.0, 256, 7)):
            with pytest.raises(sph.SphError):
                ctx.set_tuning(*bad)
    p2 = [dataclasses.replace(b.params[0], boundary=sph.BC_PERIODIC)] + list(b.params[1:])
    with pytest.raises(sph.SphError):
        sph.Batch(p2, b.counts)
    with pytest.raises(sph.SphError):
        sph.Batch(b.params, np.zeros(len(b.counts), np.int64))


def test_batch_256_scenes_full_size(sph, orc):
    """The 256-scene workload (5.1 M particles): a few graph-replayed steps keep
    every scene in its segment with finite state; the smallest scene matches its
    own oracle run bit for bit."""
    b = sph.scenes.batched_dam_breaks()
    steps = 7  # run(): prepare, then six steps, one replay of a six-step graph
    note(f"[gpu] batch 256 scenes n={b.n}")
    with sph.Batch(b.params, b.counts) as ctx:
        ctx.set_graph(True, 3)
        ctx.upload(b.rx, b.ry, b.vx, b.vy)
        ctx.run(steps)
        d = ctx.download()
        f = ctx.fields()
        keys = ctx.sorted_keys()
        m = ctx.measure_scenes()
    assert np.all(keys[1:] >= keys[:-1])
    for k in ("rx", "ry", "vx", "vy"):
        assert np.all(np.isfinite(d[k]))
    assert np.all(m[:, 2] == b.counts)  # every particle inside its own box
    s = int(np.argmin(b.counts))
    a0, a1 = int(b.first[s]), int(b.first[s] + b.counts[s])
    sc = b.scene(s)
    o = orc.run(sc.params, sc.rx, sc.ry, sc.vx, sc.vy, steps)
    assert np.array_equal(d["pid"][a0:a1] - a0, o["pid"])
    for k in ("rx", "ry", "vx", "vy"):
        assert np.array_equal(d[k][a0:a1], o[k])
    assert np.array_equal(f["ax"][a0:a1], o["ax"])
