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
        flags = np.frombuffer(f.read(12), np.int32)
        arrs["kat_ok"], arrs["rows_same"], arrs["rejected"] = (int(x) for x in flags)
        etotal = int(np.frombuffer(f.read(8), np.int64)[0])
        arrs["eoff"] = np.frombuffer(f.read(4 * (n + 1)), np.int32)
        arrs["eidx"] = np.frombuffer(f.read(4 * etotal), np.int32)
        for k in ("rho_e", "fx_e", "fy_e"):
            arrs[k] = np.frombuffer(f.read(4 * n), np.float32)
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
    assert g["kat_ok"] == 1  # Codifier [3,4,9,10,200,201,5000] -> [3,0,4,0,189,0,4798]
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


def test_list_force_classes_evaluate_edited_lists(sph, orc, tmp_path):
    """ListsForceComputation / MPListsForceComputation evaluate the lists they
    hold: the example edits the last neighbor lists (every third row loses its
    last non-self entry, every fifth is reversed) and both classes must give the
    oracle's list-based density and accelerations for those lists
    (ListsForceComputation.cpp:3-145, MPListsForceComputation.cpp:3-147)."""
    assert os.path.exists(EXE), "build() compiles tests/cpp first"
    out = tmp_path / "core.bin"
    print("[gpu] core_api_example edited lists", flush=True)
    r = subprocess.run([EXE, str(out), "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    n, _, g = read_output(out)
    assert g["rows_same"] == 1      # per-particle rows and CSR give identical results
    assert g["rejected"] == 1       # a row without its own particle is refused
    s = sph.scenes.reference_example()
    p = s.params
    eoff, eidx = g["eoff"].astype(np.int64), g["eidx"]
    rows = [eidx[eoff[i]:eoff[i + 1]] for i in range(n)]
    assert all(np.sum(r_ == i) == 1 for i, r_ in enumerate(rows))
    # the oracle takes self-excluded lists and adds the self term itself
    kept = [r_[r_ != i] for i, r_ in enumerate(rows)]
    off = np.concatenate([[0], np.cumsum([k.size for k in kept])]).astype(np.int64)
    idx = np.concatenate(kept).astype(np.int32)
    rx, ry, vx, vy = g["rx"], g["ry"], g["vx"], g["vy"]
    rho, pres, pr = orc.density(p, rx, ry, off, idx)
    ax, ay = orc.forces(p, rx, ry, vx, vy, rho, pr, off, idx)
    assert np.allclose(g["rho_e"], rho, rtol=1e-6, atol=0)
    amax = max(np.abs(ax).max(), np.abs(ay).max())
    assert np.abs(g["fx_e"] - ax).max() <= 1e-5 * amax
    assert np.abs(g["fy_e"] - ay).max() <= 1e-5 * amax
    # and the edit is visible: the stencil densities differ where rows lost a neighbor
    lost = np.array([i % 3 == 0 and rows[i].size < g["off"][i + 1] - g["off"][i]
                     for i in range(n)])
    assert lost.sum() > 100
    # (a removed neighbor at exactly r = h contributes nothing, so not every
    # shortened row changes)
    assert np.sum(g["rho_e"][lost] < g["rho"][lost] * (1 - 1e-4)) > 200
    assert np.all(g["rho_e"][lost] <= g["rho"][lost] * (1 + 1e-6))
    print(f"[gpu] edited lists bit-identical to the oracle: rho "
          f"{np.array_equal(g['rho_e'], rho)}, ax {np.array_equal(g['fx_e'], ax)}", flush=True)
