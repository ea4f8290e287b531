"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
BASELINE_LIB = os.path.join(_HERE, "_build", "libcpu_baseline.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("lo", C.c_double * 2), ("hi", C.c_double * 2), ("h", C.c_double),
        ("cell_size", C.c_double), ("dt", C.c_double), ("mass", C.c_double),
        ("rho0", C.c_double), ("gravity_y", C.c_double), ("eos_k", C.c_double),
        ("mu", C.c_double), ("c0", C.c_double), ("alpha", C.c_double),
        ("eps_visc", C.c_double), ("gamma", C.c_int32), ("physics", C.c_int32),
        ("boundary", C.c_int32), ("order", C.c_int32), ("ell_width", C.c_int32),
        ("reserved", C.c_int32),
    ]


class OrcGrid(C.Structure):
    _fields_ = [
        ("cs_x", C.c_float), ("cs_y", C.c_float), ("lo_x", C.c_float), ("lo_y", C.c_float),
        ("len_x", C.c_float), ("len_y", C.c_float), ("nx", C.c_int32), ("ny", C.c_int32),
        ("nxp", C.c_int32), ("nyp", C.c_int32), ("n_keys", C.c_int64),
        ("key_bits", C.c_int32), ("order", C.c_int32),
    ]


CONST_NAMES = ("h", "h2", "tol", "mc_rho", "c_pres", "c_visc", "gx", "gy", "eos_k", "rho0",
               "inv_rho0", "tait_b", "neg_alpha_c0", "eps_h2", "dt", "lo_x", "lo_y", "hi_x",
               "hi_y", "len_x", "len_y")


class OrcConsts(C.Structure):
    _fields_ = [(n, C.c_float) for n in CONST_NAMES] + [
        ("physics", C.c_int32), ("boundary", C.c_int32), ("gamma", C.c_int32),
        ("pad", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB)
    return _lib


def mutant_path(k: int) -> str:
    """The checker built with planted mistake k (`make -C oracle mutants`)."""
    return os.path.join(_HERE, "_build", f"liboracle_m{k}.so")


@contextlib.contextmanager
def variant(path: str):
    """Route every call of this module through another build of the checker
    (a mutant) inside the `with` block."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def _load(path: str):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    vp, i32, i64, u32, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
    P = C.POINTER
    sig = {
        "orc_make_grid": (i32, [P(OrcParams), P(OrcGrid)]),
        "orc_make_consts": (i32, [P(OrcParams), P(OrcConsts)]),
        "orc_morton2d": (u64, [u32, u32]),
        "orc_morton2d_magic": (u64, [u32, u32]),
        "orc_morton2d_decode": (None, [u64, P(u32), P(u32)]),
        "orc_cell_coord": (i32, [C.c_float, C.c_float, C.c_float, i32]),
        "orc_keys_f32": (None, [P(OrcParams), i64, vp, vp, vp]),
        "orc_stable_sort": (None, [i64, vp, vp, vp]),
        "orc_cell_ranges": (None, [i64, vp, i64, vp, vp]),
        "orc_neighbor_csr_f32": (i64, [P(OrcParams), i64, vp, vp, vp, vp, i32, vp, vp, vp]),
        "orc_neighbor_csr_f64": (i64, [P(OrcParams), i64, vp, vp, vp, vp, i32, vp, vp, vp]),
        "orc_brute_neighbors_f32": (i64, [P(OrcParams), i64, vp, vp, vp, vp, vp]),
        "orc_density_f32": (None, [P(OrcParams), i64, vp, vp, vp, vp, vp, vp, vp]),
        "orc_density_f64": (None, [P(OrcParams), i64, vp, vp, vp, vp, vp, vp, vp]),
        "orc_forces_f32": (None, [P(OrcParams), i64] + [vp] * 10),
        "orc_forces_f64": (None, [P(OrcParams), i64] + [vp] * 10),
        "orc_integrate_f32": (None, [P(OrcParams), i64] + [vp] * 6),
        "orc_integrate_f64": (None, [P(OrcParams), i64] + [vp] * 6),
        "orc_measure_f32": (None, [P(OrcParams), i64, vp, vp, vp, vp, vp]),
        "orc_run_f32": (i32, [P(OrcParams), i64] + [vp] * 8 + [i32]),
        "orc_run_f64": (i32, [P(OrcParams), i64] + [vp] * 8 + [i32]),
        "orc_delta_encode": (None, [i32, vp, vp]),
        "orc_delta_decode": (None, [i32, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(p) -> OrcParams:
    """OrcParams from the product's Params dataclass (same field names)."""
    o = OrcParams()
    o.lo[0], o.lo[1] = p.lo
    o.hi[0], o.hi[1] = p.hi
    for name in ("h", "cell_size", "dt", "mass", "rho0", "gravity_y", "eos_k", "mu", "c0",
                 "alpha", "eps_visc", "gamma", "physics", "boundary", "order", "ell_width"):
        setattr(o, name, getattr(p, name))
    return o


def grid(p) -> OrcGrid:
    g = OrcGrid()
    rc = lib().orc_make_grid(C.byref(params(p)), C.byref(g))
    if rc != 0:
        raise ValueError(f"orc_make_grid -> {rc}")
    return g


def consts(p) -> dict:
    c = OrcConsts()
    rc = lib().orc_make_consts(C.byref(params(p)), C.byref(c))
    if rc != 0:
        raise ValueError(f"orc_make_consts -> {rc}")
    return {n: getattr(c, n) for n in CONST_NAMES}


def morton(cx: int, cy: int) -> int:
    return int(lib().orc_morton2d(cx, cy))


def morton_magic(cx: int, cy: int) -> int:
    return int(lib().orc_morton2d_magic(cx, cy))


def morton_decode(key: int):
    x, y = C.c_uint32(), C.c_uint32()
    lib().orc_morton2d_decode(key, C.byref(x), C.byref(y))
    return x.value, y.value


def cell_coord(x: float, lo: float, cs: float, n: int) -> int:
    return int(lib().orc_cell_coord(x, lo, cs, n))


def keys(p, rx, ry):
    rx = np.ascontiguousarray(rx, np.float32)
    ry = np.ascontiguousarray(ry, np.float32)
    out = np.empty(rx.size, np.uint64)
    lib().orc_keys_f32(C.byref(params(p)), rx.size, _p(rx), _p(ry), _p(out))
    return out


def stable_sort(k):
    k = np.ascontiguousarray(k, np.uint64)
    perm = np.empty(k.size, np.uint32)
    sk = np.empty(k.size, np.uint64)
    lib().orc_stable_sort(k.size, _p(k), _p(perm), _p(sk))
    return perm, sk


def cell_ranges(sorted_keys, n_keys):
    sk = np.ascontiguousarray(sorted_keys, np.uint64)
    s = np.empty(n_keys, np.int32)
    e = np.empty(n_keys, np.int32)
    lib().orc_cell_ranges(sk.size, _p(sk), n_keys, _p(s), _p(e))
    return s, e


def neighbor_csr(p, rx, ry, start, end, include_self=False, dtype=np.float32):
    f = lib().orc_neighbor_csr_f32 if dtype == np.float32 else lib().orc_neighbor_csr_f64
    rx = np.ascontiguousarray(rx, dtype)
    ry = np.ascontiguousarray(ry, dtype)
    n = rx.size
    cnt = np.empty(n, np.int32)
    off = np.empty(n + 1, np.int64)
    pp = params(p)
    tot = f(C.byref(pp), n, _p(rx), _p(ry), _p(start), _p(end), int(include_self), _p(cnt),
            _p(off), None)
    if tot < 0:
        raise ValueError("orc_neighbor_csr failed")
    idx = np.empty(max(tot, 1), np.int32)
    f(C.byref(pp), n, _p(rx), _p(ry), _p(start), _p(end), int(include_self), _p(cnt), _p(off),
      _p(idx))
    return cnt, off, idx[:tot]


def brute_neighbors(p, rx, ry):
    rx = np.ascontiguousarray(rx, np.float32)
    ry = np.ascontiguousarray(ry, np.float32)
    n = rx.size
    cnt = np.empty(n, np.int32)
    off = np.empty(n + 1, np.int64)
    pp = params(p)
    tot = lib().orc_brute_neighbors_f32(C.byref(pp), n, _p(rx), _p(ry), _p(cnt), _p(off), None)
    idx = np.empty(max(tot, 1), np.int32)
    lib().orc_brute_neighbors_f32(C.byref(pp), n, _p(rx), _p(ry), _p(cnt), _p(off), _p(idx))
    return cnt, off, idx[:tot]


def density(p, rx, ry, off, idx, dtype=np.float32):
    f = lib().orc_density_f32 if dtype == np.float32 else lib().orc_density_f64
    rx = np.ascontiguousarray(rx, dtype)
    ry = np.ascontiguousarray(ry, dtype)
    n = rx.size
    rho, pres, pr = (np.empty(n, dtype) for _ in range(3))
    f(C.byref(params(p)), n, _p(rx), _p(ry), _p(np.ascontiguousarray(off, np.int64)),
      _p(np.ascontiguousarray(idx, np.int32)), _p(rho), _p(pres), _p(pr))
    return rho, pres, pr


def forces(p, rx, ry, vx, vy, rho, pr, off, idx, dtype=np.float32):
    f = lib().orc_forces_f32 if dtype == np.float32 else lib().orc_forces_f64
    arrs = [np.ascontiguousarray(a, dtype) for a in (rx, ry, vx, vy, rho, pr)]
    n = arrs[0].size
    ax = np.empty(n, dtype)
    ay = np.empty(n, dtype)
    f(C.byref(params(p)), n, *[_p(a) for a in arrs], _p(np.ascontiguousarray(off, np.int64)),
      _p(np.ascontiguousarray(idx, np.int32)), _p(ax), _p(ay))
    return ax, ay


def integrate(p, rx, ry, vx, vy, ax, ay, dtype=np.float32):
    f = lib().orc_integrate_f32 if dtype == np.float32 else lib().orc_integrate_f64
    rx, ry, vx, vy = (np.array(a, dtype=dtype, copy=True) for a in (rx, ry, vx, vy))
    ax = np.ascontiguousarray(ax, dtype)
    ay = np.ascontiguousarray(ay, dtype)
    f(C.byref(params(p)), rx.size, _p(rx), _p(ry), _p(vx), _p(vy), _p(ax), _p(ay))
    return rx, ry, vx, vy


def measure(p, rx, ry, vx, vy):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (rx, ry, vx, vy)]
    out = np.zeros(5, np.float64)
    lib().orc_measure_f32(C.byref(params(p)), arrs[0].size, *[_p(a) for a in arrs], _p(out))
    return out


def run(p, rx, ry, vx, vy, steps, dtype=np.float32, pid=None):
    """`steps` reference iterations; returns the final sorted state dict."""
    f = lib().orc_run_f32 if dtype == np.float32 else lib().orc_run_f64
    rx, ry, vx, vy = (np.array(a, dtype=dtype, copy=True) for a in (rx, ry, vx, vy))
    n = rx.size
    ax = np.zeros(n, dtype)
    ay = np.zeros(n, dtype)
    rho = np.zeros(n, dtype)
    pid = np.arange(n, dtype=np.uint32) if pid is None else np.array(pid, np.uint32)
    rc = f(C.byref(params(p)), n, _p(rx), _p(ry), _p(vx), _p(vy), _p(ax), _p(ay), _p(rho),
           _p(pid), int(steps))
    if rc != 0:
        raise ValueError(f"orc_run -> {rc}")
    return dict(rx=rx, ry=ry, vx=vx, vy=vy, ax=ax, ay=ay, rho=rho, pid=pid)


def delta_encode(lst):
    a = np.ascontiguousarray(lst, np.int32)
    out = np.empty_like(a)
    lib().orc_delta_encode(a.size, _p(a), _p(out))
    return out


def delta_decode(d):
    a = np.ascontiguousarray(d, np.int32)
    out = np.empty_like(a)
    lib().orc_delta_decode(a.size, _p(a), _p(out))
    return out


def baseline_run(p, rx, ry, vx, vy, steps, compute_type="mp-lists"):
    """Timed CPU baseline (reference-style OpenMP step; compute_type "mp-lists"
    or "sorted", the reference's names)."""
    L = C.CDLL(BASELINE_LIB)
    f = L.orc_baseline_run2
    f.restype = C.c_int32
    f.argtypes = [C.POINTER(OrcParams), C.c_int64] + [C.c_void_p] * 4 + [
        C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    ct = {"mp-lists": 0, "sorted": 1}[compute_type]
    arrs = [np.ascontiguousarray(a, np.float32) for a in (rx, ry, vx, vy)]
    secs = np.zeros(7, np.float64)
    cs = np.zeros(1, np.float64)
    threads = f(C.byref(params(p)), arrs[0].size, *[_p(a) for a in arrs], int(steps), ct,
                _p(secs), _p(cs))
    names = ("integration", "neighbor_search", "neighbor_list", "approximate_density",
             "calculate_pressure", "force_computation", "main_loop")
    return threads, dict(zip(names, secs.tolist())), float(cs[0])
