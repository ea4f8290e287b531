"""After sph_step, every array of the context is indexed by the current sorted
order: the last iteration's density, pressure and accelerations (computed
before that iteration's re-sort) follow their particles, so slot i of the
fields and slot i of the state describe the same particle. Checked against the
oracle's last iteration, particle by particle (matched by id), bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def scene(sph):
    s = sph.scenes.dam_break(48, 48, seed=3)
    rng = np.random.default_rng(4)
    # fast enough that many particles change cell every step
    s.vx = rng.normal(0, 2.0, s.n).astype(np.float32)
    s.vy = rng.normal(0, 2.0, s.n).astype(np.float32)
    return s


CASES = [("particle", "incremental", 1, False), ("particle", "incremental", 5, False),
         ("particle", "incremental", 4, True), ("particle", "full", 3, False),
         ("split_block", "incremental", 2, True), ("fused_block", "full", 3, False)]


@pytest.mark.parametrize("kernel,resort,steps,graph", CASES)
def test_step_fields_follow_particles(sph, orc, kernel, resort, steps, graph):
    s = scene(sph)
    p = s.params
    note(f"[gpu] fields order {kernel}/{resort} steps={steps} graph={graph} n={s.n}")
    kernels = {"particle": sph.STEP_PARTICLE, "split_block": sph.STEP_SPLIT_BLOCK,
               "fused_block": sph.STEP_FUSED_BLOCK}
    with sph.Context(p, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.set_step_kernel(kernels[kernel])
        ctx.set_resort(sph.RESORT_INCREMENTAL if resort == "incremental" else sph.RESORT_FULL)
        ctx.set_graph(graph)
        ctx.prepare()
        ctx.step(steps)
        movers = ctx.movers() if resort == "incremental" else -1
        d = ctx.download()
        f = ctx.fields()
        v = ctx.device_view()
        assert v.n == s.n
    # oracle: `steps` iterations end with the fields of the state the last
    # sph_step iteration started from; one more gives the state it ended with
    last = orc.run(p, s.rx, s.ry, s.vx, s.vy, steps)
    now = orc.run(p, s.rx, s.ry, s.vx, s.vy, steps + 1)
    assert np.array_equal(d["pid"], now["pid"])
    assert np.array_equal(d["rx"], now["rx"]) and np.array_equal(d["vy"], now["vy"])
    # the oracle's pressure of its last iteration (same lists, same density)
    k = orc.keys(p, last["rx"], last["ry"])
    perm, sk = orc.stable_sort(k)
    assert np.array_equal(perm, np.arange(s.n, dtype=np.uint32))  # already sorted
    st, en = orc.cell_ranges(sk, orc.grid(p).n_keys)
    _, off, idx = orc.neighbor_csr(p, last["rx"], last["ry"], st, en)
    rho, pres, _ = orc.density(p, last["rx"], last["ry"], off, idx)
    assert np.array_equal(rho, last["rho"])
    # match particles by id: slot i of the context holds particle d["pid"][i]
    where = np.empty(s.n, np.int64)
    where[last["pid"]] = np.arange(s.n)
    src = where[d["pid"]]
    assert np.array_equal(f["rho"], last["rho"][src])
    assert np.array_equal(f["p"], pres[src])
    assert np.array_equal(f["ax"], last["ax"][src])
    assert np.array_equal(f["ay"], last["ay"][src])
    note(f"[gpu] fields follow particles (movers in the last re-sort: {movers})")
    assert movers != 0
