"""GPU tests of the FP64 time-step criteria, adaptive stepping, the per-step
diagnostics line and binary snapshots (save -> load -> resume)."""
import dataclasses
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_cfl_matches_host_fp64(sph):
    s = sph.scenes.dam_break(64, 64, seed=3)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        ctx.step(20)
        got = ctx.cfl(0.25)
        d = ctx.download()
        f = ctx.fields()
    v2 = d["vx"].astype(np.float64) ** 2 + d["vy"].astype(np.float64) ** 2
    a2 = f["ax"].astype(np.float64) ** 2 + f["ay"].astype(np.float64) ** 2
    vmax, amax = math.sqrt(v2.max()), math.sqrt(a2.max())
    assert got["v_max"] == vmax and got["a_max"] == amax
    p = s.params
    dt = min(0.25 * p.h / (p.c0 + vmax), 0.25 * math.sqrt(p.h / amax))
    assert got["dt"] == dt
    note(f"[gpu] cfl dt {got['dt']:.4e} (fixed {p.dt:.4e}), |v|max {vmax:.3e}, |a|max {amax:.3e}")


def test_adaptive_run_is_stable(sph):
    s = sph.scenes.dam_break(64, 64, seed=4)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        lines = []
        hist = sph.diagnostics.run_adaptive(ctx, 60, cfl=0.25, every=10, log=lines.append)
        d = ctx.download()
        m = ctx.measure()
    assert len(hist) == 6 and all(h[1] > 0 for h in hist)
    assert hist[0][1] == pytest.approx(s.params.dt, rel=1e-12)  # at rest: dt = 0.25 h / c0
    for k in ("rx", "ry", "vx", "vy"):
        assert np.all(np.isfinite(d[k]))
    assert int(m[2]) == s.n
    assert lines[0].startswith("Step= 0\t N=") and "\t E_k=" in lines[0]


def test_snapshot_resume_is_bit_exact(sph, tmp_path):
    s = sph.scenes.dam_break(48, 48, seed=5)
    path = str(tmp_path / "state.sphsnap")
    with sph.Context(s.params, s.n) as a:
        a.upload(s.rx, s.ry, s.vx, s.vy)
        a.prepare()
        a.step(10)
        a.snapshot_save(path, 10)
        a.step(15)
        ref = a.download()
    p, n, step = sph.snapshot_info(path)
    assert (n, step) == (s.n, 10)
    assert dataclasses.asdict(p) == dataclasses.asdict(dataclasses.replace(s.params))
    with sph.Context(p, n) as b:
        assert b.snapshot_load(path) == 10
        b.prepare()
        b.step(15)
        got = b.download()
    for k in ("pid", "rx", "ry", "vx", "vy"):
        assert np.array_equal(got[k], ref[k]), k


def test_snapshot_load_rejects_bad_files(sph, tmp_path):
    s = sph.scenes.dam_break(32, 32, seed=6)
    path = str(tmp_path / "a.sphsnap")
    with sph.Context(s.params, s.n) as a:
        a.upload(s.rx, s.ry, s.vx, s.vy)
        a.prepare()
        a.snapshot_save(path, 1)
        # corrupted payload
        raw = bytearray(open(path, "rb").read())
        raw[200] ^= 0xFF
        bad = str(tmp_path / "bad.sphsnap")
        open(bad, "wb").write(bytes(raw))
        with pytest.raises(sph.SphError):
            a.snapshot_load(bad)
        # truncated
        short = str(tmp_path / "short.sphsnap")
        open(short, "wb").write(bytes(raw[:300]))
        with pytest.raises(sph.SphError):
            a.snapshot_load(short)
    other = dataclasses.replace(s.params, h=s.params.h * 1.1)
    with sph.Context(other, s.n) as b:
        with pytest.raises(sph.SphError):
            b.snapshot_load(path)
    with sph.Context(s.params, s.n - 1) as c:
        with pytest.raises(sph.SphError):
            c.snapshot_load(path)
    # a context with another dt takes the file's dt
    with sph.Context(dataclasses.replace(s.params, dt=s.params.dt * 0.5), s.n) as d:
        assert d.snapshot_load(path) == 1


def test_positions_text_checkpoint(sph, tmp_path):
    s = sph.scenes.dam_break(16, 16, seed=7)
    with sph.Context(s.params, s.n) as ctx:
        ctx.upload(s.rx, s.ry, s.vx, s.vy)
        ctx.prepare()
        ctx.step(3)
        path = sph.diagnostics.write_positions_txt(ctx, str(tmp_path / "positions_step_"), 3)
        d = ctx.download()
    rows = np.loadtxt(path)
    o = np.argsort(d["pid"])
    assert rows.shape == (s.n, 3) and np.all(rows[:, 2] == 0)
    np.testing.assert_allclose(rows[:, 0], d["rx"][o], rtol=1e-5, atol=1e-7)
