"""CPU checks of the product library's boundary: it loads without a GPU, it
exports exactly the entry points include/sph_amd.h declares, and the ctypes
structs match the C layout. No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sph_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sph):
    lib = ctypes.CDLL(sph.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(sph.capi.EXPORTED) == names


def test_binding_loads(sph):
    L = sph.lib()
    assert L.sph_last_error(None) is not None


def test_struct_sizes(sph, orc):
    # lo[2], hi[2] and 11 more doubles, then 6 int32, on both sides of the ABI
    assert ctypes.sizeof(sph.capi.SphParams) == 15 * 8 + 6 * 4
    assert ctypes.sizeof(orc.OrcParams) == ctypes.sizeof(sph.capi.SphParams)
    assert ctypes.sizeof(sph.capi.GridInfo) == ctypes.sizeof(orc.OrcGrid)


def test_create_rejects_bad_params_without_touching_gpu(sph):
    p = sph.Params(h=1.0, physics=7)
    with pytest.raises(sph.SphError, match="unknown physics"):
        sph.Context(p, 100)
    p = sph.Params(h=1.0, cell_size=0.5)
    with pytest.raises(sph.SphError, match="cell_size"):
        sph.Context(p, 100)
    with pytest.raises(sph.SphError, match="n_max"):
        sph.Context(sph.Params(h=1.0), 0)


def test_snapshot_header_format_on_cpu(sph, tmp_path):
    """The documented file layout, written by hand, is read back by
    sph_snapshot_info (no GPU needed); other files are rejected."""
    import struct
    p = sph.scenes.dam_break(8, 8).params
    c = p.to_c()
    params_bytes = bytes(memoryview(c))
    assert len(params_bytes) == 144
    hdr = b"SPHSNAP1" + struct.pack("<IIqq", 1, 192, 64, 77) + params_bytes + \
        struct.pack("<IIQ", 1, 0, 0)
    assert len(hdr) == 192
    path = str(tmp_path / "h.sphsnap")
    open(path, "wb").write(hdr)
    q, n, step = sph.snapshot_info(path)
    assert (n, step) == (64, 77) and q.h == p.h and q.dt == p.dt and q.gamma == p.gamma
    open(path, "wb").write(b"NOTASNAP" + hdr[8:])
    with pytest.raises(sph.SphError):
        sph.snapshot_info(path)
    open(path, "wb").write(hdr[:100])
    with pytest.raises(sph.SphError):
        sph.snapshot_info(path)


def test_step_header_line(sph):
    line = sph.diagnostics.step_header(12, [0.125, 2.5, 4096, 0.0, 1.0])
    assert line == "Step= 12\t N=4096\t E_k=0.125\t |v|_max=2.5\n"
