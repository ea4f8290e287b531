"""Pins of the surrogate network's CPU oracle: its analytic gradient against
central finite differences and against torch.autograd (float64), and its
bfloat16 rounding against torch's conversion."""
import numpy as np
import pytest
import torch

from oracle import nn_oracle as nno


def small_graph(n=48, seed=3, periodic=False, w=12):
    rng = np.random.default_rng(seed)
    h = 0.1
    rx = rng.random(n) * 0.35
    ry = rng.random(n) * 0.35
    vx = rng.normal(0, 0.5, n)
    vy = rng.normal(0, 0.5, n)
    rho = 1000.0 * (1 + rng.normal(0, 0.01, n))
    pr = rng.normal(0, 1e-3, n)
    counts = np.zeros(n, np.int32)
    slots = np.zeros((n, w), np.int32)
    L = 0.35
    for i in range(n):
        k = 0
        for j in range(n):
            if i == j:
                continue
            dx, dy = rx[i] - rx[j], ry[i] - ry[j]
            if periodic:
                dx -= L * np.round(dx / L)
                dy -= L * np.round(dy / L)
            if dx * dx + dy * dy <= h * h:
                if k < w:
                    slots[i, k] = j
                k += 1
        counts[i] = k
    cfg = dict(h=h, rho0=1000.0, c0=20.0, v_scale=2.0, a_scale=10.0, gy=-9.81,
               periodic=periodic, len=(L, L))
    g = nno.Graph(rx, ry, vx, vy, rho, pr, counts, slots, cfg)
    return g


def rand_theta(width, layers, seed=1, scale=0.6):
    rng = np.random.default_rng(seed)
    n = sum(int(np.prod(s)) for _, s in nno.param_layout(width, layers))
    return rng.normal(0, scale, n) / np.sqrt(width)


@pytest.mark.parametrize("periodic", [False, True])
def test_gradient_vs_finite_differences(periodic):
    W, L = 8, 2
    g = small_graph(periodic=periodic)
    assert g.recv.size > 50
    theta = rand_theta(W, L)
    target = np.random.default_rng(5).normal(0, 5, (g.n, 2))
    net = nno.Net(theta, W, L)
    loss, grad, _ = net.loss_grad(g, target)
    assert 1e-3 < loss < 1e3  # well-conditioned probe (no cancellation in the differences)
    rng = np.random.default_rng(7)
    idx = rng.choice(theta.size, 80, replace=False)
    # every parameter tensor gets at least one probe
    o = 0
    for _, shape in nno.param_layout(W, L):
        idx = np.append(idx, o)
        o += int(np.prod(shape))
    eps = 1e-5
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (nno.Net(tp, W, L).loss(g, target)[0] - nno.Net(tm, W, L).loss(g, target)[0]) / (2 * eps)
        assert abs(fd - grad[k]) <= 1e-6 * max(1e-2, abs(fd)), (k, fd, grad[k])


def torch_forward(theta, g, W, L):
    """The same network written with torch ops (float64); its gradient comes
    from autograd, independent of the oracle's hand-written backward."""
    T = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)
    p, o = {}, 0
    for name, shape in nno.param_layout(W, L):
        k = int(np.prod(shape))
        p[name] = theta[o:o + k].reshape(shape)
        o += k
    recv, send = torch.as_tensor(g.recv), torch.as_tensor(g.send)
    x, f, deg = T(g.x), T(g.f), T(g.deg)
    dx, dy = T(g.dx), T(g.dy)
    act = torch.nn.functional.silu
    h = act(x @ p["enc_w1"].T + p["enc_b1"]) @ p["enc_w2"].T + p["enc_b2"]
    for l in range(L):
        q = f"mp{l}_"
        z = (h @ p[q + "A"].T)[recv] + (h @ p[q + "B"].T)[send] + f @ p[q + "C"].T + p[q + "c"]
        S = torch.zeros(g.n, W, dtype=torch.float64).index_add(0, recv, act(z))
        M = S @ p[q + "D"].T + deg[:, None] * p[q + "d"]
        U = act(torch.cat([h, M], 1) @ p[q + "E"].T + p[q + "e"])
        h = h + U @ p[q + "F"].T + p[q + "f"]
    zo = (h @ p["ro_A"].T)[recv] + (h @ p["ro_B"].T)[send] + f @ p["ro_C"].T + p["ro_c"]
    phi = act(zo) @ p["ro_w"][0] + p["ro_b"][0] + f @ p["ro_u"]
    s = g.cfg["a_scale"] / g.cfg["h"]
    ax = s * torch.zeros(g.n, dtype=torch.float64).index_add(0, recv, dx * phi)
    ay = g.cfg["gy"] + s * torch.zeros(g.n, dtype=torch.float64).index_add(0, recv, dy * phi)
    return torch.stack([ax, ay], 1)


def test_gradient_vs_torch_autograd():
    W, L = 16, 3
    g = small_graph(n=60, seed=11)
    theta = rand_theta(W, L, seed=2)
    target = np.random.default_rng(9).normal(0, 5, (g.n, 2))
    net = nno.Net(theta, W, L)
    loss, grad, a = net.loss_grad(g, target)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    at = torch_forward(th, g, W, L)
    tt = torch.as_tensor(target)
    lt = ((at - tt) ** 2).sum() / (tt ** 2).sum()
    lt.backward()
    np.testing.assert_allclose(at.detach().numpy(), a, rtol=1e-12, atol=1e-12)
    assert abs(lt.item() - loss) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(grad, th.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -(1.0 + 2 ** -8)]]
                       ).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(nno.bf16(x), ref)
    assert nno.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even (down)
    assert nno.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6  # tie to even (up)


def test_readout_is_rotation_equivariant():
    """Rotating positions and velocities rotates the predicted acceleration
    minus gravity (invariant features times separation vectors)."""
    W, L = 8, 2
    g = small_graph(n=40, seed=4)
    theta = rand_theta(W, L, seed=8)
    a, _ = nno.Net(theta, W, L).forward(g)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    pos = np.stack([g.rx, g.ry], 1) @ R.T
    vel = np.stack([g.vx, g.vy], 1) @ R.T
    n, w = g.n, 12
    counts = np.bincount(g.recv, minlength=n)
    slots = np.zeros((n, w), np.int32)
    k = np.zeros(n, np.int64)
    for i, j in zip(g.recv, g.send):
        slots[i, k[i]] = j
        k[i] += 1
    g2 = nno.Graph(pos[:, 0], pos[:, 1], vel[:, 0], vel[:, 1], g.rho, g.pr, counts, slots, g.cfg)
    a2, _ = nno.Net(theta, W, L).forward(g2)
    grav = np.array([0.0, g.cfg["gy"]])
    np.testing.assert_allclose(a2 - grav, (a - grav) @ R.T, rtol=1e-9, atol=1e-9)
