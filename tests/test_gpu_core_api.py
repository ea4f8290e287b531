"""The reference's driver loop written against the C++ drop-in API
(tests/cpp/core_api_example.cpp over include/sph/core.hpp) reproduces the oracle
on the reference's own configuration (config/example.txt)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "core_api_example")


def read_output(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        total = int(np.frombuffer(f.read(8), np.int64)[0])
        rt = int(np.frombuffer(f.read(4), np.int32)[0])
        arrs = {k: np.frombuffer(f.read(4 * n), np.float32)
                for k in ("rx", "ry", "vx", "vy", "fx", "fy", "rho", "p")}
        arrs["off"] = np.frombuffer(f.read(4 * (n + 1)), np.int32)
        arrs["idx"] = np.frombuffer(f.read(4 * total), np.int32)
        arrs["cell_id"] = np.frombuffer(f.read(4 * n), np.uint32)
    return n, rt, arrs


def test_core_api_driver_matches_oracle(sph, orc, tmp_path):
    assert os.path.exists(EXE), "build() compiles tests/cpp first"
    steps = 6
    out = tmp_path / "core.bin"
    print(f"[gpu] core_api_example {steps} steps", flush=True)
    r = subprocess.run([EXE, str(out), str(steps)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    # StepHeader reproduces the printf line of the last step
    hdr = [l for l in lines if l.startswith(f"Step= {steps - 1}\t")]
    assert len(hdr) == 2 and hdr[0].split("\t")[1] == hdr[1].split("\t")[1]
    cfl = [l for l in lines if l.startswith("cfl dt=")]
    assert len(cfl) == 1 and float(cfl[0].split("=")[1]) > 0
    n, rt, g = read_output(out)
    assert rt == 1  # Codifier round trip
    s = sph.scenes.reference_example()
    assert n == s.n
    o = orc.run(s.params, s.rx, s.ry, s.vx, s.vy, steps)
    assert np.array_equal(g["rx"], o["rx"]) and np.array_equal(g["ry"], o["ry"])
    assert np.array_equal(g["vx"], o["vx"]) and np.array_equal(g["vy"], o["vy"])
    assert np.allclose(g["rho"], o["rho"], rtol=1e-6)
    amax = np.abs(o["ay"]).max()
    assert np.abs(g["fy"] - o["ay"]).max() <= 1e-5 * amax
    # self-inclusive CSR lists (the CPU convention of MultipassNeighborLists)
    k = orc.keys(s.params, g["rx"], g["ry"])
    assert np.array_equal(g["cell_id"], k.astype(np.uint32))
    perm, sk = orc.stable_sort(k)
    assert np.array_equal(perm, np.arange(n, dtype=np.uint32))
    st, en = orc.cell_ranges(sk, int(orc.grid(s.params).n_keys))
    cnt, off, idx = orc.neighbor_csr(s.params, g["rx"], g["ry"], st, en, include_self=True)
    assert np.array_equal(g["off"], off.astype(np.int32))
    assert np.array_equal(g["idx"], idx)
