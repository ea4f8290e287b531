"""ctypes binding of libsph_amd.so (the C ABI declared in include/sph_amd.h).

The library is the only compute path: importing this module on a machine where
the library is missing raises immediately, and every call checks the returned
status and raises SphError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsph_amd.so")

SPH_OK = 0
ORDER_MORTON, ORDER_ROWMAJOR = 0, 1
PHYS_REFERENCE, PHYS_WCSPH = 0, 1
BC_WALLS, BC_PERIODIC = 0, 1
COMPUTE_SORTED, COMPUTE_LISTS = 0, 1
STEP_SPLIT_BLOCK, STEP_FUSED_BLOCK, STEP_PARTICLE = 0, 1, 2
RESORT_FULL, RESORT_INCREMENTAL = 0, 1
STAGES = ("integration", "neighbor_search", "neighbor_list", "approximate_density",
          "force_computation", "sort_histogram", "sort_pass", "block_step", "measure",
          "movers", "merge")


class SphError(RuntimeError):
    pass


class SphParams(C.Structure):
    _fields_ = [
        ("lo", C.c_double * 2), ("hi", C.c_double * 2), ("h", C.c_double),
        ("cell_size", C.c_double), ("dt", C.c_double), ("mass", C.c_double),
        ("rho0", C.c_double), ("gravity_y", C.c_double), ("eos_k", C.c_double),
        ("mu", C.c_double), ("c0", C.c_double), ("alpha", C.c_double),
        ("eps_visc", C.c_double), ("gamma", C.c_int32), ("physics", C.c_int32),
        ("boundary", C.c_int32), ("order", C.c_int32), ("ell_width", C.c_int32),
        ("reserved", C.c_int32),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("cs_x", C.c_float), ("cs_y", C.c_float), ("lo_x", C.c_float), ("lo_y", C.c_float),
        ("len_x", C.c_float), ("len_y", C.c_float), ("nx", C.c_int32), ("ny", C.c_int32),
        ("nxp", C.c_int32), ("nyp", C.c_int32), ("n_keys", C.c_int64),
        ("key_bits", C.c_int32), ("order", C.c_int32),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("cell_factor", C.c_double), ("tile", C.c_int32), ("resort", C.c_int32)]

    def as_tuple(self):
        return (float(self.cell_factor), int(self.tile), int(self.resort))


class NnConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("layers", C.c_int32), ("train", C.c_int32),
                ("periodic", C.c_int32), ("max_nodes", C.c_int64), ("h", C.c_double),
                ("rho0", C.c_double), ("c0", C.c_double), ("v_scale", C.c_double),
                ("a_scale", C.c_double), ("gravity_x", C.c_double), ("gravity_y", C.c_double),
                ("len_x", C.c_double), ("len_y", C.c_double)]


class SurrogateStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("checks", C.c_int64), ("failed_checks", C.c_int64),
                ("nn_steps", C.c_int64), ("plain_steps", C.c_int64),
                ("last_error", C.c_double), ("max_error", C.c_double),
                ("mean_error", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceView(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("rx", C.c_void_p), ("ry", C.c_void_p), ("vx", C.c_void_p),
        ("vy", C.c_void_p), ("ax", C.c_void_p), ("ay", C.c_void_p), ("rho", C.c_void_p),
        ("pr", C.c_void_p), ("pid", C.c_void_p), ("ell_count", C.c_void_p),
        ("ell_index", C.c_void_p), ("ell_width", C.c_int32), ("pad", C.c_int32),
        ("cell_range", C.c_void_p), ("stream", C.c_void_p),
    ]


@dataclass
class Params:
    """Python mirror of sph_params; fills both the product and oracle structs."""
    lo: tuple = (0.0, 0.0)
    hi: tuple = (1.0, 1.0)
    h: float = 1.0
    cell_size: float = 0.0
    dt: float = 0.005
    mass: float = 1.0
    rho0: float = 25.0
    gravity_y: float = -10.0
    eos_k: float = 30.0
    mu: float = 0.01
    c0: float = 0.0
    alpha: float = 0.0
    eps_visc: float = 0.01
    gamma: int = 7
    physics: int = PHYS_REFERENCE
    boundary: int = BC_WALLS
    order: int = ORDER_MORTON
    ell_width: int = 0
    extra: dict = field(default_factory=dict)

    def fill(self, s):
        s.lo[0], s.lo[1] = self.lo
        s.hi[0], s.hi[1] = self.hi
        for name in ("h", "cell_size", "dt", "mass", "rho0", "gravity_y", "eos_k", "mu", "c0",
                     "alpha", "eps_visc", "gamma", "physics", "boundary", "order",
                     "ell_width"):
            setattr(s, name, getattr(self, name))
        s.reserved = 0
        return s

    def to_c(self) -> SphParams:
        return self.fill(SphParams())


_lib = None


def lib():
    """Load libsph_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SphError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "sph_create": (i32, [P(SphParams), i64, C.c_int, P(vp)]),
        "sph_destroy": (i32, [vp]),
        "sph_last_error": (C.c_char_p, [vp]),
        "sph_get_grid": (i32, [vp, P(GridInfo)]),
        "sph_get_constants": (i32, [vp, P(C.c_float), i32]),
        "sph_upload": (i32, [vp, i64, vp, vp, vp, vp, vp]),
        "sph_download": (i32, [vp, vp, vp, vp, vp, vp]),
        "sph_download_fields": (i32, [vp, vp, vp, vp, vp]),
        "sph_size": (i64, [vp]),
        "sph_find_neighbors": (i32, [vp]),
        "sph_build_neighbor_lists": (i32, [vp, i32, P(i64)]),
        "sph_set_neighbor_lists": (i32, [vp, vp, vp, i32]),
        "sph_density_pressure": (i32, [vp, i32]),
        "sph_forces": (i32, [vp, i32]),
        "sph_integrate": (i32, [vp]),
        "sph_update_velocities": (i32, [vp]),
        "sph_update_positions": (i32, [vp]),
        "sph_set_dt": (i32, [vp, d]),
        "sph_measure": (i32, [vp, P(d)]),
        "sph_prepare": (i32, [vp]),
        "sph_step": (i32, [vp, i32]),
        "sph_set_graph": (i32, [vp, i32]),
        "sph_step_graphs": (i32, [vp, i32]),
        "sph_ell_overflow": (i32, [vp, P(i64)]),
        "sph_set_step_kernel": (i32, [vp, i32]),
        "sph_set_resort": (i32, [vp, i32]),
        "sph_movers": (i32, [vp, P(i64)]),
        "sph_resort_path": (i32, [vp, P(i32)]),
        "sph_block_stats": (i32, [vp, P(i64)]),
        "sph_get_sorted_keys": (i32, [vp, vp]),
        "sph_get_permutation": (i32, [vp, vp]),
        "sph_get_cell_ranges": (i32, [vp, vp, vp]),
        "sph_get_neighbor_lists": (i32, [vp, vp, vp]),
        "sph_get_ell": (i32, [vp, vp, vp]),
        "sph_get_device_view": (i32, [vp, P(DeviceView)]),
        "sph_set_timing": (i32, [vp, i32]),
        "sph_stage_times": (i32, [vp, P(d), P(i64), i32]),
        "sph_synchronize": (i32, [vp]),
        "sph_batch_create": (i32, [P(SphParams), i32, P(i64), C.c_int, P(vp)]),
        "sph_batch_scenes": (i32, [vp]),
        "sph_batch_scene_info": (i32, [vp, i32, P(i64), P(i64), P(GridInfo)]),
        "sph_batch_set_tuning": (i32, [vp, P(BatchTuning)]),
        "sph_batch_get_tuning": (i32, [vp, P(BatchTuning)]),
        "sph_batch_default_candidates": (i32, [P(BatchTuning), i32]),
        "sph_batch_autotune": (i32, [vp, P(BatchTuning), i32, i32, P(d), P(i32)]),
        "sph_batch_measure": (i32, [vp, P(d)]),
        "sph_export_state": (i32, [vp, vp]),
        "sph_cfl": (i32, [vp, d, P(d)]),
        "sph_snapshot_info": (i32, [C.c_char_p, P(SphParams), P(i64), P(i64)]),
        "sph_snapshot_save": (i32, [vp, C.c_char_p, i64]),
        "sph_snapshot_load": (i32, [vp, C.c_char_p, P(i64)]),
        "sph_device_malloc": (i32, [C.c_int, i64, P(vp)]),
        "sph_device_free": (i32, [vp]),
        "sph_device_copy": (i32, [vp, vp, i64, i32]),
        "sph_upload_device": (i32, [vp, i64, vp, vp, vp, vp]),
        "sph_nn_create": (i32, [P(NnConfig), C.c_int, P(vp)]),
        "sph_nn_destroy": (i32, [vp]),
        "sph_nn_last_error": (C.c_char_p, [vp]),
        "sph_nn_num_params": (i64, [vp]),
        "sph_nn_set_params": (i32, [vp, vp, i64]),
        "sph_nn_get_params": (i32, [vp, vp, i64]),
        "sph_nn_get_grad": (i32, [vp, vp, i64]),
        "sph_nn_reset_optimizer": (i32, [vp]),
        "sph_nn_set_clip": (i32, [vp, d]),
        "sph_nn_forward": (i32, [vp, vp, vp, vp]),
        "sph_nn_loss_grad": (i32, [vp, vp, P(d)]),
        "sph_nn_adam": (i32, [vp, vp, d, d, d, d]),
        "sph_nn_train_step": (i32, [vp, vp, d, P(d)]),
        "sph_surrogate_run": (i32, [vp, vp, i32, i32, d, i32, P(SurrogateStats)]),
        "sph_measure_noalloc": (i32, [P(SphParams), i64, vp, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


EXPORTED = ("sph_create", "sph_destroy", "sph_last_error", "sph_get_grid", "sph_get_constants",
            "sph_upload", "sph_download", "sph_download_fields", "sph_size",
            "sph_find_neighbors", "sph_build_neighbor_lists", "sph_set_neighbor_lists",
            "sph_density_pressure",
            "sph_forces", "sph_integrate", "sph_update_velocities", "sph_update_positions",
            "sph_set_dt", "sph_measure", "sph_prepare", "sph_step",
            "sph_set_graph", "sph_step_graphs", "sph_ell_overflow", "sph_set_step_kernel",
            "sph_set_resort", "sph_movers", "sph_resort_path", "sph_block_stats",
            "sph_get_sorted_keys", "sph_get_permutation",
            "sph_get_cell_ranges", "sph_get_neighbor_lists", "sph_get_ell",
            "sph_get_device_view", "sph_set_timing", "sph_stage_times", "sph_synchronize",
            "sph_batch_create", "sph_batch_scenes", "sph_batch_scene_info",
            "sph_batch_set_tuning", "sph_batch_get_tuning", "sph_batch_default_candidates",
            "sph_batch_autotune", "sph_batch_measure", "sph_export_state", "sph_cfl",
            "sph_snapshot_info", "sph_snapshot_save", "sph_snapshot_load",
            "sph_device_malloc", "sph_device_free", "sph_device_copy",
            "sph_upload_device", "sph_nn_create", "sph_nn_destroy", "sph_nn_last_error",
            "sph_nn_num_params", "sph_nn_set_params", "sph_nn_get_params", "sph_nn_get_grad",
            "sph_nn_reset_optimizer", "sph_nn_set_clip", "sph_nn_forward", "sph_nn_loss_grad", "sph_nn_adam",
            "sph_nn_train_step", "sph_surrogate_run", "sph_measure_noalloc")


class DeviceBuffer:
    """Raw device memory owned by Python (sph_device_malloc); float32 views are
    addressed by element offset."""

    def __init__(self, n_floats: int, device: int = 0):
        self._L = lib()
        p = C.c_void_p()
        rc = self._L.sph_device_malloc(int(device), int(n_floats) * 4, C.byref(p))
        if rc != SPH_OK:
            raise SphError(f"sph_device_malloc({n_floats * 4} bytes) -> {rc}")
        self._p = p.value
        self.n = int(n_floats)

    def ptr(self, offset: int = 0) -> int:
        if not 0 <= offset <= self.n:
            raise IndexError("offset outside the buffer")
        return self._p + 4 * int(offset)

    def to_host(self, offset=0, count=None) -> np.ndarray:
        count = self.n - offset if count is None else int(count)
        if offset + count > self.n:
            raise IndexError("range outside the buffer")
        out = np.empty(count, np.float32)
        rc = self._L.sph_device_copy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr(offset)),
                                     count * 4, 1)
        if rc != SPH_OK:
            raise SphError(f"sph_device_copy -> {rc}")
        return out

    def close(self):
        if getattr(self, "_p", None):
            self._L.sph_device_free(C.c_void_p(self._p))
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Context:
    """One particle system on one GPU (owns an sph_ctx)."""

    def __init__(self, params: Params, n_max: int, device: int = 0):
        self._L = lib()
        self.params = params
        h = C.c_void_p()
        rc = self._L.sph_create(C.byref(params.to_c()), int(n_max), int(device), C.byref(h))
        if rc != SPH_OK:
            raise SphError(f"sph_create -> {rc}: {self._L.sph_last_error(None).decode()}")
        self._h = h
        self.n_max = int(n_max)

    # -- plumbing
    def _chk(self, rc, what):
        if rc != SPH_OK:
            msg = self._L.sph_last_error(self._h).decode()
            raise SphError(f"{what} -> {rc}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def n(self) -> int:
        return int(self._L.sph_size(self._h))

    def grid(self) -> GridInfo:
        g = GridInfo()
        self._chk(self._L.sph_get_grid(self._h, C.byref(g)), "sph_get_grid")
        return g

    def constants(self) -> np.ndarray:
        out = np.zeros(32, np.float32)
        m = self._L.sph_get_constants(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), 32)
        if m < 0:
            self._chk(m, "sph_get_constants")
        return out[:m]

    # -- data
    def upload(self, rx, ry, vx, vy, pid=None):
        rx, ry, vx, vy = _f32(rx), _f32(ry), _f32(vx), _f32(vy)
        p = None if pid is None else np.ascontiguousarray(pid, dtype=np.uint32)
        self._chk(self._L.sph_upload(self._h, rx.size, _ptr(rx), _ptr(ry), _ptr(vx), _ptr(vy),
                                     _ptr(p)), "sph_upload")

    def download(self):
        n = self.n
        rx, ry, vx, vy = (np.empty(n, np.float32) for _ in range(4))
        pid = np.empty(n, np.uint32)
        self._chk(self._L.sph_download(self._h, _ptr(rx), _ptr(ry), _ptr(vx), _ptr(vy),
                                       _ptr(pid)), "sph_download")
        return dict(rx=rx, ry=ry, vx=vx, vy=vy, pid=pid)

    def fields(self):
        n = self.n
        ax, ay, rho, p = (np.empty(n, np.float32) for _ in range(4))
        self._chk(self._L.sph_download_fields(self._h, _ptr(ax), _ptr(ay), _ptr(rho), _ptr(p)),
                  "sph_download_fields")
        return dict(ax=ax, ay=ay, rho=rho, p=p)

    # -- stages
    def find_neighbors(self):
        self._chk(self._L.sph_find_neighbors(self._h), "sph_find_neighbors")

    def build_neighbor_lists(self, include_self=False) -> int:
        tot = C.c_int64()
        self._chk(self._L.sph_build_neighbor_lists(self._h, int(bool(include_self)),
                                                   C.byref(tot)), "sph_build_neighbor_lists")
        return int(tot.value)

    def set_neighbor_lists(self, offsets, indices, include_self=False):
        """Caller-built CSR lists (offsets int64 [n + 1], indices int32) over the
        current sorted particles, evaluated by COMPUTE_LISTS as given."""
        off = np.ascontiguousarray(offsets, np.int64)
        idx = np.ascontiguousarray(indices, np.int32)
        if idx.size == 0:
            idx = np.zeros(1, np.int32)
        self._chk(self._L.sph_set_neighbor_lists(self._h, _ptr(off), _ptr(idx),
                                                 int(bool(include_self))),
                  "sph_set_neighbor_lists")

    def density_pressure(self, compute=COMPUTE_SORTED):
        self._chk(self._L.sph_density_pressure(self._h, int(compute)), "sph_density_pressure")

    def forces(self, compute=COMPUTE_SORTED):
        self._chk(self._L.sph_forces(self._h, int(compute)), "sph_forces")

    def integrate(self):
        self._chk(self._L.sph_integrate(self._h), "sph_integrate")

    def set_dt(self, dt: float):
        self._chk(self._L.sph_set_dt(self._h, float(dt)), "sph_set_dt")

    def update_velocities(self):
        self._chk(self._L.sph_update_velocities(self._h), "sph_update_velocities")

    def update_positions(self):
        self._chk(self._L.sph_update_positions(self._h), "sph_update_positions")

    def measure(self) -> np.ndarray:
        out = np.zeros(5, np.float64)
        self._chk(self._L.sph_measure(self._h, out.ctypes.data_as(C.POINTER(C.c_double))),
                  "sph_measure")
        return out

    def prepare(self):
        self._chk(self._L.sph_prepare(self._h), "sph_prepare")

    def step(self, n_steps=1):
        self._chk(self._L.sph_step(self._h, int(n_steps)), "sph_step")

    def set_graph(self, enable=True, pairs: int = 16):
        """Replay sph_step's iterations from a hipGraph of 2 * pairs iterations."""
        self._chk(self._L.sph_set_graph(self._h, int(pairs) if enable else 0), "sph_set_graph")

    def step_graphs(self, n_steps: int):
        """Capture (without running) the graphs that step(n_steps) replays from
        the current state."""
        self._chk(self._L.sph_step_graphs(self._h, int(n_steps)), "sph_step_graphs")

    def set_step_kernel(self, variant: int):
        self._chk(self._L.sph_set_step_kernel(self._h, int(variant)), "sph_set_step_kernel")

    def set_resort(self, mode: int):
        self._chk(self._L.sph_set_resort(self._h, int(mode)), "sph_set_resort")

    def movers(self) -> int:
        c = C.c_int64()
        self._chk(self._L.sph_movers(self._h, C.byref(c)), "sph_movers")
        return int(c.value)

    def resort_path(self) -> str:
        """'bucket' or 'sorted': how the last incremental re-sort ordered its movers."""
        c = C.c_int32()
        self._chk(self._L.sph_resort_path(self._h, C.byref(c)), "sph_resort_path")
        return "bucket" if c.value else "sorted"

    def block_stats(self):
        out = np.zeros(6, np.int64)
        self._chk(self._L.sph_block_stats(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))),
                  "sph_block_stats")
        keys = ("blocks", "over_capacity", "staged")
        return {"density": dict(zip(keys, out[:3].tolist())),
                "force": dict(zip(keys, out[3:].tolist()))}

    def ell_overflow(self) -> int:
        c = C.c_int64()
        self._chk(self._L.sph_ell_overflow(self._h, C.byref(c)), "sph_ell_overflow")
        return int(c.value)

    def run(self, steps: int):
        """Reference-equivalent run of `steps` iterations (pre-loop + steps-1 loops):
        the last iteration's density, pressure and forces are evaluated on the
        final state."""
        self.prepare()
        if steps > 1:
            self.step(steps - 1)
        self.density_pressure(COMPUTE_SORTED)
        self.forces(COMPUTE_SORTED)

    # -- inspection
    def sorted_keys(self):
        out = np.empty(self.n, np.uint64)
        self._chk(self._L.sph_get_sorted_keys(self._h, _ptr(out)), "sph_get_sorted_keys")
        return out

    def permutation(self):
        out = np.empty(self.n, np.uint32)
        self._chk(self._L.sph_get_permutation(self._h, _ptr(out)), "sph_get_permutation")
        return out

    def cell_ranges(self):
        nk = int(self.grid().n_keys)
        s = np.empty(nk, np.int32)
        e = np.empty(nk, np.int32)
        self._chk(self._L.sph_get_cell_ranges(self._h, _ptr(s), _ptr(e)), "sph_get_cell_ranges")
        return s, e

    def neighbor_lists(self, total: int):
        off = np.empty(self.n + 1, np.int64)
        idx = np.empty(max(total, 1), np.int32)
        self._chk(self._L.sph_get_neighbor_lists(self._h, _ptr(off), _ptr(idx)),
                  "sph_get_neighbor_lists")
        return off, idx[:total]

    def ell(self):
        w = self.params.ell_width or 32
        cnt = np.empty(self.n, np.int32)
        slots = np.empty((self.n, w), np.int32)
        self._chk(self._L.sph_get_ell(self._h, _ptr(cnt), _ptr(slots)), "sph_get_ell")
        return cnt, slots

    def ell_counts(self):
        cnt = np.empty(self.n, np.int32)
        self._chk(self._L.sph_get_ell(self._h, _ptr(cnt), None), "sph_get_ell")
        return cnt

    def cfl(self, cfl=0.25) -> dict:
        """sph_cfl: dt = min(cfl h / (c0 + max|v|), 0.25 sqrt(h / max|a|)) (FP64 maxima)."""
        out = np.zeros(3, np.float64)
        self._chk(self._L.sph_cfl(self._h, float(cfl), out.ctypes.data_as(C.POINTER(C.c_double))),
                  "sph_cfl")
        return {"dt": float(out[0]), "v_max": float(out[1]), "a_max": float(out[2])}

    def snapshot_save(self, path: str, step: int = 0):
        self._chk(self._L.sph_snapshot_save(self._h, os.fsencode(path), int(step)),
                  "sph_snapshot_save")

    def snapshot_load(self, path: str) -> int:
        st = C.c_int64()
        self._chk(self._L.sph_snapshot_load(self._h, os.fsencode(path), C.byref(st)),
                  "sph_snapshot_load")
        return int(st.value)

    def export_state(self, dev_ptr: int):
        """State (rx | ry | vx | vy, particle-id order) into device memory at dev_ptr."""
        self._chk(self._L.sph_export_state(self._h, C.c_void_p(dev_ptr)), "sph_export_state")

    def upload_device(self, n: int, rx_ptr: int, ry_ptr: int, vx_ptr: int, vy_ptr: int):
        self._chk(self._L.sph_upload_device(self._h, int(n), C.c_void_p(rx_ptr),
                                            C.c_void_p(ry_ptr), C.c_void_p(vx_ptr),
                                            C.c_void_p(vy_ptr)), "sph_upload_device")

    def device_view(self) -> DeviceView:
        v = DeviceView()
        self._chk(self._L.sph_get_device_view(self._h, C.byref(v)), "sph_get_device_view")
        return v

    # -- timing
    def set_timing(self, enable=True):
        self._chk(self._L.sph_set_timing(self._h, int(bool(enable))), "sph_set_timing")

    def stage_times(self):
        ms = np.zeros(len(STAGES), np.float64)
        cnt = np.zeros(len(STAGES), np.int64)
        self._chk(self._L.sph_stage_times(self._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                          cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                          len(STAGES)), "sph_stage_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(STAGES)}

    def synchronize(self):
        self._chk(self._L.sph_synchronize(self._h), "sph_synchronize")


class Batch(Context):
    """Many independent scenes in one context (sph_batch_*): one radix sort and
    one launch per stage serve every scene. Arrays passed to upload() and
    returned by download()/fields() hold the scenes concatenated in scene order
    (download returns the sorted order plus global ids, as Context does)."""

    def __init__(self, params_list, counts, device: int = 0):
        self._L = lib()
        self.params = params_list[0]
        self.scene_params = list(params_list)
        arr = (SphParams * len(params_list))()
        for k, p in enumerate(params_list):
            p.fill(arr[k])
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        h = C.c_void_p()
        rc = self._L.sph_batch_create(arr, len(params_list),
                                      cnt.ctypes.data_as(C.POINTER(C.c_int64)), int(device),
                                      C.byref(h))
        if rc != SPH_OK:
            raise SphError(f"sph_batch_create -> {rc}: {self._L.sph_last_error(None).decode()}")
        self._h = h
        self.n_max = int(cnt.sum())
        self.counts = cnt
        self.first = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)

    @property
    def n_scenes(self) -> int:
        return int(self._L.sph_batch_scenes(self._h))

    def scene_info(self, scene: int):
        f, n, g = C.c_int64(), C.c_int64(), GridInfo()
        self._chk(self._L.sph_batch_scene_info(self._h, int(scene), C.byref(f), C.byref(n),
                                               C.byref(g)), "sph_batch_scene_info")
        return int(f.value), int(n.value), g

    def set_tuning(self, cell_factor=1.0, tile=256, resort=RESORT_INCREMENTAL):
        t = BatchTuning(float(cell_factor), int(tile), int(resort))
        self._chk(self._L.sph_batch_set_tuning(self._h, C.byref(t)), "sph_batch_set_tuning")

    def tuning(self):
        t = BatchTuning()
        self._chk(self._L.sph_batch_get_tuning(self._h, C.byref(t)), "sph_batch_get_tuning")
        return t.as_tuple()

    @staticmethod
    def default_candidates():
        L = lib()
        m = L.sph_batch_default_candidates(None, 0)
        arr = (BatchTuning * m)()
        L.sph_batch_default_candidates(arr, m)
        return [arr[k].as_tuple() for k in range(m)]

    def autotune(self, candidates=None, trial_steps: int = 10):
        """Run the library's autotuner; returns (best tuning, {tuning: ms/step})."""
        cands = candidates if candidates is not None else self.default_candidates()
        arr = (BatchTuning * len(cands))()
        for k, (f, t, r) in enumerate(cands):
            arr[k] = BatchTuning(float(f), int(t), int(r))
        ms = np.zeros(len(cands), np.float64)
        best = C.c_int32(-1)
        self._chk(self._L.sph_batch_autotune(self._h, arr, len(cands), int(trial_steps),
                                             ms.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.byref(best)), "sph_batch_autotune")
        return tuple(cands[best.value]), {tuple(c): float(m) for c, m in zip(cands, ms)}

    def measure_scenes(self) -> np.ndarray:
        out = np.zeros((self.n_scenes, 5), np.float64)
        self._chk(self._L.sph_batch_measure(self._h, out.ctypes.data_as(C.POINTER(C.c_double))),
                  "sph_batch_measure")
        return out


def snapshot_info(path: str):
    """(Params, n, step) stored in a binary snapshot (sph_snapshot_info)."""
    L = lib()
    sp = SphParams()
    n, st = C.c_int64(), C.c_int64()
    rc = L.sph_snapshot_info(os.fsencode(path), C.byref(sp), C.byref(n), C.byref(st))
    if rc != SPH_OK:
        raise SphError(f"sph_snapshot_info -> {rc}: {L.sph_last_error(None).decode()}")
    p = Params(lo=(sp.lo[0], sp.lo[1]), hi=(sp.hi[0], sp.hi[1]), h=sp.h, cell_size=sp.cell_size,
               dt=sp.dt, mass=sp.mass, rho0=sp.rho0, gravity_y=sp.gravity_y, eos_k=sp.eos_k,
               mu=sp.mu, c0=sp.c0, alpha=sp.alpha, eps_visc=sp.eps_visc, gamma=sp.gamma,
               physics=sp.physics, boundary=sp.boundary, order=sp.order,
               ell_width=sp.ell_width)
    return p, int(n.value), int(st.value)
