"""Learned surrogate of the SPH acceleration (ctypes front end of sph_nn_* and
sph_surrogate_run in libsph_amd.so).

The network, its training kernels and the surrogate-in-the-loop runner are in
the HIP library; this module builds configurations, initial weights and the
training data pipeline (solver snapshots kept in device memory that the library
allocates and fills).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .capi import (BC_PERIODIC, SPH_OK, Batch, Context, DeviceBuffer, NnConfig, Params,
                   SphError, SurrogateStats, lib)
from . import scenes as _scenes

WIDTH = 128
NODE_F = 8
EDGE_F = 8


def param_layout(width=WIDTH, layers=3):
    """(name, shape) of every parameter tensor in the library's flat order."""
    W = width
    out = [("enc_w1", (W, NODE_F)), ("enc_b1", (W,)), ("enc_w2", (W, W)), ("enc_b2", (W,))]
    for l in range(layers):
        out += [(f"mp{l}_A", (W, W)), (f"mp{l}_B", (W, W)), (f"mp{l}_C", (W, EDGE_F)),
                (f"mp{l}_c", (W,)), (f"mp{l}_D", (W, W)), (f"mp{l}_d", (W,)),
                (f"mp{l}_E", (W, 2 * W)), (f"mp{l}_e", (W,)), (f"mp{l}_F", (W, W)),
                (f"mp{l}_f", (W,))]
    out += [("ro_A", (W, W)), ("ro_B", (W, W)), ("ro_C", (W, EDGE_F)), ("ro_c", (W,)),
            ("ro_w", (1, W)), ("ro_b", (1,)), ("ro_u", (EDGE_F,))]
    return out


def init_params(layers=3, seed=0, width=WIDTH) -> np.ndarray:
    """Deterministic initial weights: Glorot-uniform matrices from SplitMix64
    (tensor k seeded seed * 1000003 + k), halved for the residual branch F and
    the readout vector; zero biases."""
    out = []
    for k, (name, shape) in enumerate(param_layout(width, layers)):
        n = int(np.prod(shape))
        if len(shape) == 1:
            out.append(np.zeros(n))
            continue
        fan_out, fan_in = shape
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        if name.endswith("_F") or name == "ro_w":
            lim *= 0.5
        u = _scenes.uniform(seed * 1000003 + k, n)
        out.append((2.0 * u - 1.0) * lim)
    return np.concatenate(out).astype(np.float32)


def nn_config(params: Params, max_nodes: int, layers=3, train=True, v_scale=None,
              a_scale=None) -> NnConfig:
    """Network configuration for scenes with the physics of `params`: relative
    velocities scaled by 0.1 c0 and the readout by 0.1 c0^2 / h (the order of one
    neighbor's pressure acceleration) unless given."""
    c0 = params.c0 if params.c0 > 0 else 10.0
    cfg = NnConfig()
    cfg.width = WIDTH
    cfg.layers = layers
    cfg.train = int(bool(train))
    cfg.periodic = int(params.boundary == BC_PERIODIC)
    cfg.max_nodes = int(max_nodes)
    cfg.h = params.h
    cfg.rho0 = params.rho0
    cfg.c0 = c0
    cfg.v_scale = v_scale if v_scale else 0.1 * c0
    cfg.a_scale = a_scale if a_scale else 0.1 * c0 * c0 / params.h
    cfg.gravity_x = 0.0
    cfg.gravity_y = params.gravity_y
    cfg.len_x = params.hi[0] - params.lo[0]
    cfg.len_y = params.hi[1] - params.lo[1]
    return cfg


class Network:
    """One network on one GPU (owns an sph_nn handle)."""

    def __init__(self, params: Params, max_nodes: int, layers=3, train=True, device=0,
                 v_scale=None, a_scale=None, seed=0):
        self._L = lib()
        self.layers = layers
        self.config = nn_config(params, max_nodes, layers, train, v_scale, a_scale)
        h = C.c_void_p()
        rc = self._L.sph_nn_create(C.byref(self.config), int(device), C.byref(h))
        if rc != SPH_OK:
            raise SphError(f"sph_nn_create -> {rc}: {self._L.sph_nn_last_error(None).decode()}")
        self._h = h
        self.n_params = int(self._L.sph_nn_num_params(h))
        self.set_params(init_params(layers, seed))

    def _chk(self, rc, what):
        if rc != SPH_OK:
            raise SphError(f"{what} -> {rc}: {self._L.sph_nn_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.sph_nn_destroy(self._h)
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

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        self._chk(self._L.sph_nn_set_params(self._h, t.ctypes.data_as(C.c_void_p), t.size),
                  "sph_nn_set_params")

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._chk(self._L.sph_nn_get_params(self._h, out.ctypes.data_as(C.c_void_p), out.size),
                  "sph_nn_get_params")
        return out

    def grad(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._chk(self._L.sph_nn_get_grad(self._h, out.ctypes.data_as(C.c_void_p), out.size),
                  "sph_nn_get_grad")
        return out

    def reset_optimizer(self):
        self._chk(self._L.sph_nn_reset_optimizer(self._h), "sph_nn_reset_optimizer")

    def set_clip(self, max_norm: float):
        """Global gradient-norm limit of the Adam step (0 turns it off)."""
        self._chk(self._L.sph_nn_set_clip(self._h, float(max_norm)), "sph_nn_set_clip")

    def forward(self, ctx: Context):
        n = ctx.n
        ax = np.empty(n, np.float32)
        ay = np.empty(n, np.float32)
        self._chk(self._L.sph_nn_forward(self._h, ctx.handle, ax.ctypes.data_as(C.c_void_p),
                                         ay.ctypes.data_as(C.c_void_p)), "sph_nn_forward")
        return ax, ay

    def loss_grad(self, ctx: Context):
        out = np.zeros(3, np.float64)
        self._chk(self._L.sph_nn_loss_grad(self._h, ctx.handle,
                                           out.ctypes.data_as(C.POINTER(C.c_double))),
                  "sph_nn_loss_grad")
        return {"loss": float(out[0]), "rms_rel": float(out[1]), "rms_target": float(out[2])}

    def adam(self, ctx: Context, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(self._L.sph_nn_adam(self._h, ctx.handle, float(lr), float(beta1),
                                      float(beta2), float(eps)), "sph_nn_adam")

    def train_step(self, ctx: Context, lr=1e-3):
        out = np.zeros(3, np.float64)
        self._chk(self._L.sph_nn_train_step(self._h, ctx.handle, float(lr),
                                            out.ctypes.data_as(C.POINTER(C.c_double))),
                  "sph_nn_train_step")
        return {"loss": float(out[0]), "rms_rel": float(out[1]), "rms_target": float(out[2])}

    def save(self, path):
        np.savez(path, params=self.params(), layers=np.int64(self.layers))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            if int(z["layers"]) != self.layers:
                raise SphError("saved network has a different layer count")
            self.set_params(z["params"])


def surrogate_run(ctx: Context, net: Network, steps: int, check_every=10, threshold=0.02,
                  mode=1) -> dict:
    """sph_surrogate_run: network accelerations in the loop with a plain-SPH check
    every `check_every` steps; mode 1 rolls a failed window back to plain SPH,
    mode 0 replaces only the failing check step."""
    st = SurrogateStats()
    rc = lib().sph_surrogate_run(ctx.handle, net.handle, int(steps), int(check_every),
                                 float(threshold), int(mode), C.byref(st))
    if rc != SPH_OK:
        raise SphError(f"sph_surrogate_run -> {rc}: {lib().sph_last_error(ctx.handle).decode()}")
    return st.as_dict()


# ------------------------------------------------------------------ data
def training_scenes(n_scenes=64, n_particles=65536, seed=0x5EED0003, box=(4.0, 2.0)):
    """Dam breaks of n_particles each with the 16k dam-break physics (dx =
    0.25/128) in `box`. Scene s picks its block shape (nx x ny = n_particles, nx a
    power of two, block at most half the box wide and as tall as the box) from
    uniform(seed)[s], and has its own +-0.01 dx jitter seeded seed + 1 + s."""
    dx = 0.25 / 128
    u = _scenes.uniform(seed, n_scenes)
    shapes = []
    nx = 1
    while nx <= n_particles:
        ny = n_particles // nx
        if nx * ny == n_particles and nx * dx <= box[0] / 2 and ny * dx <= box[1] and ny >= 8:
            shapes.append((nx, ny))
        nx *= 2
    if not shapes:
        raise ValueError("no block shape fits the box")
    out = []
    for s in range(n_scenes):
        nx, ny = shapes[int(u[s] * len(shapes))]
        out.append(_scenes.dam_break(nx, ny, box=box, seed=seed + 1 + s, steps=0,
                                     name=f"train_{s}"))
    return out


class Dataset:
    """Solver snapshots in device memory: snapshot t of scene s holds rx, ry, vx,
    vy (particle-id order) of n particles at element offset ((t * S + s) * 4) * n."""

    def __init__(self, snapshots: int, scenes: int, n: int, device: int = 0):
        self.T, self.S, self.n = int(snapshots), int(scenes), int(n)
        self.buf = DeviceBuffer(self.T * self.S * 4 * self.n, device)

    def offset(self, t: int, s: int, field: int = 0) -> int:
        if not (0 <= t < self.T and 0 <= s < self.S and 0 <= field < 4):
            raise IndexError("snapshot index out of range")
        return ((t * self.S + s) * 4 + field) * self.n

    def ptr(self, t: int, s: int, field: int = 0) -> int:
        return self.buf.ptr(self.offset(t, s, field))

    def host(self, t: int, s: int):
        a = self.buf.to_host(self.offset(t, s), 4 * self.n).reshape(4, self.n)
        return dict(rx=a[0], ry=a[1], vx=a[2], vy=a[3])

    def close(self):
        self.buf.close()


def generate_dataset(scenes_list, steps=2000, every=10, device=0) -> Dataset:
    """Advance all scenes together (one batch context) and keep the state every
    `every` steps; snapshot 0 is the initial state."""
    n = scenes_list[0].n
    if any(sc.n != n for sc in scenes_list):
        raise ValueError("training scenes must have equal particle counts")
    S = len(scenes_list)
    counts = np.full(S, n, np.int64)
    snaps = steps // every + 1
    data = Dataset(snaps, S, n, device)
    # the batch exports [rx | ry | vx | vy] of all scenes; scene s of snapshot t
    # is then copied field by field into its slot
    stage = DeviceBuffer(4 * S * n, device)
    b = Batch([sc.params for sc in scenes_list], counts, device)
    L = lib()

    def keep(t):
        b.export_state(stage.ptr(0))
        for s in range(S):
            for f in range(4):
                rc = L.sph_device_copy(C.c_void_p(data.ptr(t, s, f)),
                                       C.c_void_p(stage.ptr(f * S * n + s * n)), n * 4, 2)
                if rc != SPH_OK:
                    raise SphError(f"sph_device_copy -> {rc}")

    try:
        b.upload(np.concatenate([sc.rx for sc in scenes_list]),
                 np.concatenate([sc.ry for sc in scenes_list]),
                 np.concatenate([sc.vx for sc in scenes_list]),
                 np.concatenate([sc.vy for sc in scenes_list]))
        keep(0)
        b.prepare()
        b.set_graph(True)
        for t in range(1, snaps):
            b.step(every)
            keep(t)
        b.synchronize()
    finally:
        b.close()
        stage.close()
    return data


class Trainer:
    """Training loop over a snapshot tensor: each iteration uploads one snapshot
    into a plain-SPH context, computes the plain forces (the target) and runs one
    forward + backward + Adam step of the network."""

    def __init__(self, net: Network, params: Params, data: Dataset, seed=1):
        self.net = net
        self.data = data
        self.n = data.n
        self.ctx = Context(params, self.n)
        self.rng = np.random.default_rng(seed)

    def load(self, t: int, s: int):
        d = self.data
        self.ctx.upload_device(self.n, d.ptr(t, s, 0), d.ptr(t, s, 1), d.ptr(t, s, 2),
                               d.ptr(t, s, 3))
        self.ctx.prepare()
        self.ctx.forces(1)  # plain SPH accelerations from the fixed-width list

    def step(self, t: int, s: int, lr=1e-3):
        self.load(t, s)
        return self.net.train_step(self.ctx, lr)

    def train(self, iters: int, lr=1e-3, snapshots=None, scenes=None, lr_final=None,
              log_every=0, log=None):
        T = self.data.T if snapshots is None else snapshots
        S = self.data.S if scenes is None else scenes
        hist = []
        for it in range(iters):
            t = int(self.rng.integers(0, T))
            s = int(self.rng.integers(0, S))
            lr_it = lr if lr_final is None else lr * (lr_final / lr) ** (it / max(iters - 1, 1))
            r = self.step(t, s, lr_it)
            hist.append(r["rms_rel"])
            if log and log_every and (it % log_every == 0 or it == iters - 1):
                log(f"iter {it}: loss {r['loss']:.4e} rms_rel {r['rms_rel']:.4f}")
        return hist

    def evaluate(self, pairs):
        """Mean RMS relative error of the network on snapshots (t, s), and of the
        zero-order hold that reuses the plain acceleration of snapshot t-1 (the
        previous plain-SPH check, `every` steps earlier) for the same particles."""
        e_nn, e_hold = [], []
        for t, s in pairs:
            self.load(t - 1, s)
            f = self.ctx.fields()
            pid = self.ctx.download()["pid"]
            prev = np.empty((self.n, 2), np.float64)
            prev[pid, 0] = f["ax"]
            prev[pid, 1] = f["ay"]
            self.load(t, s)
            f = self.ctx.fields()
            pid = self.ctx.download()["pid"]
            ax, ay = self.net.forward(self.ctx)
            tx = f["ax"].astype(np.float64)
            ty = f["ay"].astype(np.float64)
            t2 = float(np.sum(tx * tx + ty * ty))
            e_nn.append(math.sqrt(float(np.sum((ax - tx) ** 2 + (ay - ty) ** 2)) / t2))
            hx, hy = prev[pid, 0], prev[pid, 1]
            e_hold.append(math.sqrt(float(np.sum((hx - tx) ** 2 + (hy - ty) ** 2)) / t2))
        return float(np.mean(e_nn)), float(np.mean(e_hold))

    def close(self):
        self.ctx.close()


def train_network(n_scenes=8, n_particles=65536, steps=400, every=10, iters=1500, lr=2e-3,
                  lr_final=1e-4, layers=3, seed=0x5EED0003, held_out=1, log=None, clip=1.0):
    """Generate solver snapshots of `n_scenes` dam breaks, train a network on all
    but the last `held_out` scenes and evaluate it on those. `clip` limits the
    gradient's global norm in each Adam step (0: no limit). Returns (network,
    report); the caller closes the network."""
    import time

    t0 = time.perf_counter()
    scs = training_scenes(n_scenes=n_scenes, n_particles=n_particles, seed=seed)
    data = generate_dataset(scs, steps=steps, every=every)
    t_data = time.perf_counter() - t0
    net = Network(scs[0].params, n_particles, layers=layers, train=True)
    net.set_clip(clip)
    tr = Trainer(net, scs[0].params, data, seed=seed)
    train_scenes = n_scenes - held_out
    pairs = [(t, s) for s in range(train_scenes, n_scenes)
             for t in range(1, data.T, max(1, data.T // 12))]
    e_before, e_hold = tr.evaluate(pairs)
    t1 = time.perf_counter()
    hist = tr.train(iters, lr=lr, scenes=train_scenes, lr_final=lr_final,
                    log_every=max(1, iters // 10), log=log)
    t_train = time.perf_counter() - t1
    e_after, _ = tr.evaluate(pairs)
    report = {
        "scenes": n_scenes, "particles_per_scene": n_particles, "solver_steps": steps,
        "snapshot_every": every, "snapshots": data.T * data.S, "held_out_scenes": held_out,
        "iterations": iters, "batch_particles": n_particles, "lr": [lr, lr_final], "clip": clip,
        "data_s": round(t_data, 3), "train_s": round(t_train, 3),
        "ms_per_iteration": round(t_train * 1e3 / max(iters, 1), 3),
        "train_rms_rel_last50": float(np.mean(hist[-50:])) if hist else None,
        "held_out_rms_rel": {"trained": e_after, "untrained": e_before, "zero_order_hold": e_hold},
        "params": net.n_params,
    }
    tr.close()
    data.close()
    return net, report
