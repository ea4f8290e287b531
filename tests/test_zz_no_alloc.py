"""The allocation-free measurement (sph_measure_noalloc) against sph_measure:
the same FP64 values bit for bit, on two calls in a row (the second with fewer
particles, so stale partial sums of the first would show), called through the
C ABI on device arrays the test owns."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPH_ERR_CAPACITY = -3


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def moving_scene(sph, side, seed):
    s = sph.scenes.dam_break(side, side, seed=seed)
    rng = np.random.default_rng(seed)
    s.vx = rng.normal(0, 1.0, s.n).astype(np.float32)
    s.vy = rng.normal(0, 1.0, s.n).astype(np.float32)
    s.rx[:3] = [-0.5, 2.0, 0.5]   # two particles outside the box until the first step
    return s


def test_measure_noalloc_matches_measure(sph):
    L = sph.lib()
    out = sph.capi.DeviceBuffer(10)   # 5 doubles, owned by the caller
    try:
        for side, seed in ((160, 1), (70, 2)):   # 25,600 then 4,900 particles
            s = moving_scene(sph, side, seed)
            note(f"[gpu] sph_measure_noalloc n={s.n}")
            with sph.Context(s.params, s.n) as ctx:
                ctx.upload(s.rx, s.ry, s.vx, s.vy)
                ctx.prepare()
                ctx.step(3)
                ref = ctx.measure()
                v = ctx.device_view()
                prm = s.params.to_c()
                rc = L.sph_measure_noalloc(C.byref(prm), s.n, v.rx, v.ry, v.vx, v.vy,
                                           C.c_void_p(out.ptr()), v.stream)
                assert rc == 0, L.sph_last_error(None)
                ctx.synchronize()
                got = out.to_host(0, 10).view(np.float64)
            assert np.array_equal(got, ref), (got, ref)
            assert got[2] == s.n and got[0] > 0  # walls clamped the two outside particles
        # empty input: zeros, no launch failure
        prm = moving_scene(sph, 8, 3).params.to_c()
        rc = L.sph_measure_noalloc(C.byref(prm), 0, None, None, None, None,
                                   C.c_void_p(out.ptr()), None)
        assert rc == 0
        got = out.to_host(0, 10).view(np.float64)
        assert np.array_equal(got, np.zeros(5))
        # beyond the static scratch: refused before any launch
        rc = L.sph_measure_noalloc(C.byref(prm), 8388609, None, None, None, None,
                                   C.c_void_p(out.ptr()), None)
        assert rc == SPH_ERR_CAPACITY
        assert b"8388608" in L.sph_last_error(None)
    finally:
        out.close()
