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
