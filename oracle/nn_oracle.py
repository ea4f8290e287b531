"""CPU oracle of the surrogate message-passing network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network's forward pass and gradients. The reference has no learned component;
this module restates the network defined by the GPU library (the parameter
layout, features and layers below), in FP64 with analytic backpropagation.

Two modes:
  * emulate_bf16=False: plain FP64 (the mathematical network). Pinned in the
    tests by central finite differences and by torch.autograd in float64.
  * emulate_bf16=True: rounds every GEMM operand (weights and activations) and
    the edge-layer weight matrices to bfloat16 (round to nearest even) exactly
    where the GPU kernels do, accumulating in FP64; the GPU result then differs
    only by FP32 accumulation order.

Network (width W, L message-passing layers, 8 node and 8 edge features):
  h0   = W2 silu(W1 x + b1) + b2                                   (encoder)
  z_ij = A h_i + B h_j + C f_ij + c         over the list j of i   (edge MLP,
  S_i  = sum_j silu(z_ij)                                            layer 1)
  M_i  = D S_i + deg_i d                     (edge MLP layer 2, after the sum:
                                               linear layers commute with it)
  h'   = h + F silu(E [h, M] + e) + f                              (node MLP)
  a_i  = g + a_s sum_j (r_i - r_j)/h (w . silu(Ao h_i + Bo h_j + Co f_ij + co) + bo
                                     + u . f_ij)
The readout is an equivariant edge readout: invariant edge scalars times the
separation vector, so the predicted acceleration rotates with the particles;
u . f_ij is a linear path from the edge features (one of them has the form of
the SPH pressure term). Loss: sum |a - t|^2 / sum |t|^2 over the batch (the
squared RMS relative error).
"""
from __future__ import annotations

import numpy as np

NODE_F = 8
EDGE_F = 8
RHO_SCALE = 50.0


def param_layout(width=128, layers=3):
    """(name, shape) in the flat parameter order of the GPU library."""
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


def unflatten(theta, width=128, layers=3):
    p, o = {}, 0
    for name, shape in param_layout(width, layers):
        k = int(np.prod(shape))
        p[name] = np.asarray(theta[o:o + k], np.float64).reshape(shape)
        o += k
    assert o == len(theta)
    return p


def flatten(p, width=128, layers=3):
    return np.concatenate([np.asarray(p[n], np.float64).ravel()
                           for n, _ in param_layout(width, layers)])


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    a = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


MAX_NEIGHBORS = 32


class Graph:
    """Edges (receiver i, sender j) from a fixed-width list (first
    min(count, w, 32) slots of each particle) and the per-particle inputs."""

    def __init__(self, rx, ry, vx, vy, rho, pr, counts, slots, cfg):
        n, w = slots.shape
        deg = np.minimum(counts, min(w, MAX_NEIGHBORS))
        recv = np.repeat(np.arange(n), deg)
        send = np.concatenate([slots[i, :deg[i]] for i in range(n)]).astype(np.int64) \
            if n else np.zeros(0, np.int64)
        self.n = n
        self.recv, self.send = recv, send
        self.deg = deg.astype(np.float64)
        f64 = lambda a: np.asarray(a, np.float64)
        self.rx, self.ry, self.vx, self.vy = f64(rx), f64(ry), f64(vx), f64(vy)
        self.rho, self.pr = f64(rho), f64(pr)
        self.cfg = cfg
        dx = self.rx[recv] - self.rx[send]
        dy = self.ry[recv] - self.ry[send]
        if cfg.get("periodic"):
            Lx, Ly = cfg["len"]
            dx = np.where(dx > Lx / 2, dx - Lx, np.where(dx < -Lx / 2, dx + Lx, dx))
            dy = np.where(dy > Ly / 2, dy - Ly, np.where(dy < -Ly / 2, dy + Ly, dy))
        self.dx, self.dy = dx, dy
        self.x = node_features(self, cfg)
        self.f = edge_features(self, cfg)


def node_features(g, cfg):
    rs = RHO_SCALE
    ps = RHO_SCALE * cfg["rho0"] / (cfg["c0"] ** 2)
    x = np.zeros((g.n, NODE_F))
    x[:, 0] = rs * (g.rho / cfg["rho0"] - 1.0)
    x[:, 1] = ps * g.pr
    x[:, 2] = g.deg / 8.0
    x[:, 3] = 1.0
    return x


def edge_features(g, cfg):
    h, vs = cfg["h"], cfg["v_scale"]
    rs = RHO_SCALE
    ps = RHO_SCALE * cfg["rho0"] / (cfg["c0"] ** 2)
    i, j = g.recv, g.send
    r = np.sqrt(g.dx * g.dx + g.dy * g.dy)
    q = r / h
    w = (1.0 - q) ** 2
    dvx = g.vx[i] - g.vx[j]
    dvy = g.vy[i] - g.vy[j]
    f = np.zeros((i.size, EDGE_F))
    f[:, 0] = q
    f[:, 1] = w
    f[:, 2] = w / np.maximum(q, 0.05)
    f[:, 3] = (dvx * g.dx + dvy * g.dy) / (np.maximum(r, 0.05 * h) * vs)
    f[:, 4] = np.sqrt(dvx * dvx + dvy * dvy) / vs
    f[:, 5] = rs * (g.rho[j] / cfg["rho0"] - 1.0)
    f[:, 6] = ps * g.pr[j]
    f[:, 7] = f[:, 2] * (ps * g.pr[i] + ps * g.pr[j])
    return f


def _segsum(v, idx, n):
    out = np.zeros((n,) + v.shape[1:])
    np.add.at(out, idx, v)
    return out


class Net:
    def __init__(self, theta, width=128, layers=3, emulate_bf16=False):
        self.W, self.L = width, layers
        self.p = unflatten(theta, width, layers)
        self.emu = emulate_bf16
        self.r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))

    def G(self, X, Wt):
        """GEMM of the GPU kernels: round(X) @ round(W)^T."""
        return self.r(X) @ self.r(Wt).T

    def forward(self, g: Graph):
        p, r = self.p, self.r
        cfg = g.cfg
        c = {}
        c["pre1"] = self.G(g.x, p["enc_w1"]) + p["enc_b1"]
        c["a1"] = silu(c["pre1"])
        h = self.G(c["a1"], p["enc_w2"]) + p["enc_b2"]
        c["layers"] = []
        for l in range(self.L):
            q = f"mp{l}_"
            P = self.G(h, p[q + "A"])
            Q = self.G(h, p[q + "B"])
            z = P[g.recv] + Q[g.send] + g.f @ r(p[q + "C"]).T + p[q + "c"]
            S = _segsum(silu(z), g.recv, g.n)
            M = self.G(S, p[q + "D"]) + g.deg[:, None] * p[q + "d"]
            HM = np.concatenate([h, M], axis=1)
            preU = self.G(HM, p[q + "E"]) + p[q + "e"]
            U = silu(preU)
            c["layers"].append(dict(h=h, z=z, S=S, M=M, HM=HM, preU=preU, U=U))
            h = h + self.G(U, p[q + "F"]) + p[q + "f"]
        c["h"] = h
        Po = self.G(h, p["ro_A"])
        Qo = self.G(h, p["ro_B"])
        zo = Po[g.recv] + Qo[g.send] + g.f @ r(p["ro_C"]).T + p["ro_c"]
        so = silu(zo)
        phi = so @ r(p["ro_w"])[0] + p["ro_b"][0] + g.f @ p["ro_u"]
        s = cfg["a_scale"] / cfg["h"]
        ax = cfg.get("gx", 0.0) + s * _segsum(g.dx * phi, g.recv, g.n)
        ay = cfg["gy"] + s * _segsum(g.dy * phi, g.recv, g.n)
        c.update(zo=zo, so=so)
        return np.stack([ax, ay], axis=1), c

    def loss(self, g, target):
        a, c = self.forward(g)
        t = np.asarray(target, np.float64)
        e = a - t
        return float((e * e).sum() / (t * t).sum()), a, c

    def backward(self, g, da, c):
        """Gradient of sum(da * a) with respect to the flat parameters."""
        p, r, G = self.p, self.r, self.G
        cfg = g.cfg
        grads = {}
        s = cfg["a_scale"] / cfg["h"]
        # readout
        dphi = s * (da[g.recv, 0] * g.dx + da[g.recv, 1] * g.dy)
        grads["ro_w"] = (dphi[:, None] * c["so"]).sum(0)[None, :]
        grads["ro_b"] = np.array([dphi.sum()])
        grads["ro_u"] = dphi @ g.f
        dzo = dphi[:, None] * r(p["ro_w"])[0][None, :] * dsilu(c["zo"])
        dPo = _segsum(dzo, g.recv, g.n)
        dQo = _segsum(dzo, g.send, g.n)
        grads["ro_C"] = dzo.T @ g.f
        grads["ro_c"] = dzo.sum(0)
        h = c["h"]
        grads["ro_A"] = r(dPo).T @ r(h)
        grads["ro_B"] = r(dQo).T @ r(h)
        dh = r(dPo) @ r(p["ro_A"]) + r(dQo) @ r(p["ro_B"])
        for l in reversed(range(self.L)):
            q = f"mp{l}_"
            L = c["layers"][l]
            grads[q + "F"] = r(dh).T @ r(L["U"])
            grads[q + "f"] = dh.sum(0)
            dU = r(dh) @ r(p[q + "F"])
            dpreU = dU * dsilu(L["preU"])
            grads[q + "E"] = r(dpreU).T @ r(L["HM"])
            grads[q + "e"] = dpreU.sum(0)
            dHM = r(dpreU) @ r(p[q + "E"])
            dM = dHM[:, self.W:]
            grads[q + "D"] = r(dM).T @ r(L["S"])
            grads[q + "d"] = (g.deg[:, None] * dM).sum(0)
            dS = r(dM) @ r(p[q + "D"])
            dz = dS[g.recv] * dsilu(L["z"])
            dP = _segsum(dz, g.recv, g.n)
            dQ = _segsum(dz, g.send, g.n)
            grads[q + "C"] = dz.T @ g.f
            grads[q + "c"] = dz.sum(0)
            hin = L["h"]
            grads[q + "A"] = r(dP).T @ r(hin)
            grads[q + "B"] = r(dQ).T @ r(hin)
            dh = dh + dHM[:, :self.W] + r(dP) @ r(p[q + "A"]) + r(dQ) @ r(p[q + "B"])
        grads["enc_w2"] = r(dh).T @ r(c["a1"])
        grads["enc_b2"] = dh.sum(0)
        da1 = r(dh) @ r(p["enc_w2"])
        dpre1 = da1 * dsilu(c["pre1"])
        grads["enc_w1"] = r(dpre1).T @ r(g.x)
        grads["enc_b1"] = dpre1.sum(0)
        return flatten(grads, self.W, self.L)

    def loss_grad(self, g, target):
        L, a, c = self.loss(g, target)
        t = np.asarray(target, np.float64)
        da = 2.0 * (a - t) / (t * t).sum()
        return L, self.backward(g, da, c), a

