"""Pins the CPU oracle of the learned seed (oracle/mlp_oracle.py) against
independent references: torch's tanh-approximate GELU, binary cross-entropy
with logits and autograd in float64, central finite differences, torch's
BF16 rounding, and closed forms of both kinds of pair features."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(M.NPARAM)
    for off, fo, fi in ((M.OFF_W1, M.NH, M.NIN), (M.OFF_W2, M.NH, M.NH), (M.OFF_W3, M.NH, M.NH),
                        (M.OFF_W4, M.NOUT, M.NH)):
        a = np.sqrt(6.0 / (fo + fi)) * scale
        w[off:off + fo * fi] = rng.uniform(-a, a, fo * fi)
    for off, k in ((M.OFF_B1, M.NH), (M.OFF_B2, M.NH), (M.OFF_B3, M.NH), (M.OFF_B4, M.NOUT)):
        w[off:off + k] = rng.normal(0, 0.1, k)
    return w


def _targets(rng, n):
    """(u, v, h) targets, about half of them hits."""
    return np.concatenate([rng.uniform(0, 1, (n, 2)), (rng.uniform(size=(n, 1)) < 0.5).astype(np.float64)], 1)


def _torch_loss(w, X, T, hit_weight=0.05):
    P = M.unpack(w)
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in P.items()}
    x = torch.tensor(X, dtype=torch.float64)
    g = lambda z: torch.nn.functional.gelu(z, approximate="tanh")  # noqa: E731
    h = g(x @ t["W1"].T + t["b1"])
    h = g(h @ t["W2"].T + t["b2"])
    h = g(h @ t["W3"].T + t["b3"])
    z = h @ t["W4"].T + t["b4"]
    T = torch.tensor(T)
    hit = T[:, 2]
    se = (((torch.sigmoid(z[:, :2]) - T[:, :2]) ** 2).sum(1) * hit).sum()
    bce = torch.nn.functional.binary_cross_entropy_with_logits(z[:, 2], hit, reduction="sum")
    loss = (se + hit_weight * bce) / X.shape[0]
    loss.backward()
    return loss.item(), M.pack({k: t[k].grad.numpy() for k in t})


def test_gelu_matches_torch_tanh_gelu():
    x = np.linspace(-8, 8, 2001)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(xt, approximate="tanh")
    y.sum().backward()
    np.testing.assert_allclose(M.gelu(x), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(M.gelu_grad(x), xt.grad.numpy(), rtol=1e-12, atol=1e-14)


def test_gelu_infer_follows_tanh_gelu():
    # the inference activation is a polynomial fit of the tanh GELU: within
    # its stated bound everywhere, the identity gelu(x) - gelu(-x) = x of any
    # x * sigmoid-like form, and exactly x / 0 beyond the clamp at |x| = 3.4
    x = np.linspace(-40, 40, 800001)
    dev = np.abs(M.gelu_infer(x) - M.gelu(x))
    assert dev.max() < 1.08e-3
    assert dev.max() > 1.0e-3  # a minimax fit: the bound is reached (equioscillation)
    inner = np.abs(x) < 3.4
    np.testing.assert_allclose(M.gelu_infer(x[inner]) - M.gelu_infer(-x[inner]), x[inner], atol=1e-14)
    big = x >= 3.4
    assert np.array_equal(M.gelu_infer(x[big]), x[big])
    assert np.all(M.gelu_infer(-x[big]) == 0)
    q = M.GELU_INFER_Q
    assert 3.4 * sum(a * 3.4 ** (2 * i) for i, a in enumerate(q)) == pytest.approx(0.5, abs=1e-12)
    assert M.gelu_infer(0.0) == 0.0


def test_gradient_matches_autograd():
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (37, M.NIN))
    T = _targets(rng, 37)
    w = _params(5)
    for lam in (0.05, 1.0):
        l0, g0 = M.loss_and_grad(w, X, T, hit_weight=lam)
        l1, g1 = _torch_loss(w, X, T, hit_weight=lam)
        assert abs(l0 - l1) < 1e-14
        np.testing.assert_allclose(g0, g1, rtol=1e-10, atol=1e-15)
    # misses carry no (u, v) error: their (u, v) targets do not matter
    T2 = T.copy()
    T2[T[:, 2] == 0, :2] = 0.123
    np.testing.assert_array_equal(M.loss_and_grad(w, X, T2)[1], M.loss_and_grad(w, X, T)[1])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (11, M.NIN))
    T = _targets(rng, 11)
    w = _params(6, scale=2.0)
    _, g = M.loss_and_grad(w, X, T)
    # a few entries of every tensor
    idx = []
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_W4), (M.OFF_W4, M.OFF_B4), (M.OFF_B4, M.NPARAM)):
        idx += list(rng.choice(np.arange(a, b), size=min(6, b - a), replace=False))
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss_and_grad(wp, X, T)[0] - M.loss_and_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_quantized_gradient_close_to_fp64():
    """BF16 rounding at the matrix-core inputs perturbs the gradient by a few
    parts in a thousand, not more."""
    rng = np.random.default_rng(8)
    X = rng.normal(0, 1, (200, M.NIN))
    T = _targets(rng, 200)
    w = _params(9)
    _, g = M.loss_and_grad(w, X, T)
    _, gq = M.loss_and_grad(w, X, T, quantize=True)
    assert np.linalg.norm(gq - g) / np.linalg.norm(g) < 3e-2


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 10, 5000), [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, 65504.0]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16(x), ref)
    assert M.bf16(1 + 2 ** -8) == 1.0 and M.bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6


def _flat_net(p, q, x0=0.5, y0=1.5, z=3.0, w=None):
    """A flat (p, q) net over [x0, x0+1] x [y0, y0+1] at height z, rows along x."""
    net = np.zeros((p + 1, q + 1, 4))
    for i in range(p + 1):
        for j in range(q + 1):
            net[i, j] = [x0 + i / p, y0 + j / q, z, 1.0]
    if w is not None:  # rational: scale the homogeneous point, same position
        net *= w[..., None]
    return net.reshape(-1, 4)


def test_features_closed_forms():
    # ray planes: for a ray along +z, n1 = (0, 1, 0), n2 = n1 x z = (1, 0, 0)
    n1, n2 = M.ray_planes(np.array([[0.0, 0.0, 2.0], [3.0, 0.0, 0.0]]))
    np.testing.assert_allclose(n1[0], [0, 1, 0])
    np.testing.assert_allclose(n2[0], [1, 0, 0])
    np.testing.assert_allclose(n1[1], [0, -1, 0])   # x-dominant: (d.y, -d.x, 0)/|.|
    np.testing.assert_allclose(n2[1], [0, 0, 1])    # (0, -1, 0) x (1, 0, 0)
    # flat unit net centred at (1, 2, 3), box half extent 0.5; a ray along +z
    # through (1, 2): the samples are the net's corners, edge middles and
    # centre at (n1, n2) = (y - 2, x - 1) / 0.5
    bounds = np.array([[0.5, 1.5, 3.0, 1.5, 2.5, 3.0]])
    o, d = np.array([[1.0, 2.0, -7.0]]), np.array([[0.0, 0.0, 2.0]])
    for p, q in ((2, 2), (3, 2), (3, 5), (4, 3)):
        f = M.features(o, d, bounds, [_flat_net(p, q)], [p], [q])[0]
        s = f[:18].reshape(3, 3, 2)
        for a, xa in enumerate((-1.0, 0.0, 1.0)):
            for b, yb in enumerate((-1.0, 0.0, 1.0)):
                np.testing.assert_allclose(s[a, b], [yb, xa], atol=1e-12)
        np.testing.assert_allclose(f[18:], [(p - 3.5) / 1.5, (q - 3.5) / 1.5])
    # weights change the homogeneous net, not the dehomogenised points
    w = np.random.default_rng(3).uniform(0.5, 2.0, (4, 3))
    np.testing.assert_allclose(M.features(o, d, bounds, [_flat_net(3, 2, w=w)], [3], [2]),
                               M.features(o, d, bounds, [_flat_net(3, 2)], [3], [2]), atol=1e-12)
    # moving the ray by (dx, dy) shifts every sample by -(dy, dx) / 0.5
    f0 = M.features(o, d, bounds, [_flat_net(3, 3)], [3], [3])[0]
    f1 = M.features(o + [0.1, -0.2, 0.0], d, bounds, [_flat_net(3, 3)], [3], [3])[0]
    np.testing.assert_allclose((f1 - f0)[:18].reshape(9, 2), np.tile([0.4, -0.2], (9, 1)), atol=1e-12)
    # odd degree: the middle sample averages the two central rows
    net = _flat_net(3, 2).reshape(4, 3, 4)
    net[1, :, 2] += 0.3   # lift row 1 along z: no effect on the projection along z
    net[2, :, 0] += 0.2   # move row 2 along x by 0.2: the middle row moves by 0.1
    f2 = M.features(o, d, bounds, [net.reshape(-1, 4)], [3], [2])[0]
    g0 = M.features(o, d, bounds, [_flat_net(3, 2)], [3], [2])[0]
    np.testing.assert_allclose((f2 - g0)[:18].reshape(3, 3, 2)[1, :, 1], 0.1 / 0.5, atol=1e-12)
    np.testing.assert_allclose((f2 - g0)[:18].reshape(3, 3, 2)[[0, 2]], 0.0, atol=1e-12)
    # similarity invariance: translate and scale the whole configuration
    rng = np.random.default_rng(2)
    o, d = rng.normal(size=(5, 3)), rng.normal(size=(5, 3))
    lo = rng.normal(size=(5, 3))
    b = np.concatenate([lo, lo + rng.uniform(0.1, 2, (5, 3))], axis=1)
    p = np.array([2, 3, 4, 5, 3])
    nets = [np.concatenate([rng.normal(size=((k + 1) * (k + 1), 3)), np.ones(((k + 1) * (k + 1), 1))], 1) for k in p]
    a = M.features(o, d, b, nets, p, p)
    s, t = 3.7, np.array([1.0, -2.0, 0.5])
    nets2 = [np.concatenate([n[:, :3] * s + t, n[:, 3:]], 1) for n in nets]
    a2 = M.features(o * s + t, d, b * s + np.tile(t, 2), nets2, p, p)
    np.testing.assert_allclose(a, a2, rtol=1e-12, atol=1e-12)


def test_forward_known_answers():
    w = np.zeros(M.NPARAM)
    w[M.OFF_B4:] = [0.0, np.log(3.0), -np.log(4.0)]   # sigmoid(0) = 0.5, sigmoid(log 3) = 0.75, 0.2
    y, _ = M.forward(w, np.ones((3, M.NIN)))
    np.testing.assert_allclose(y, [[0.5, 0.75, 0.2]] * 3)
    # the loss at those outputs: a hit at (0.5, 0.5) costs 0.0625 + lambda (-log 0.2), a miss lambda (-log 0.8)
    T = np.array([[0.5, 0.5, 1.0], [0.9, 0.1, 0.0], [0.5, 0.75, 1.0]])
    lam = 0.3
    want = (0.0625 + lam * -np.log(0.2) + lam * -np.log(0.8) + lam * -np.log(0.2)) / 3
    assert M.loss_and_grad(w, np.ones((3, M.NIN)), T, hit_weight=lam)[0] == pytest.approx(want, rel=1e-12)
    # one path: x0 -> unit 0 of each layer -> u logit, weight 1 everywhere
    w = np.zeros(M.NPARAM)
    w[M.OFF_W1 + 0] = 1.0
    w[M.OFF_W2 + 0] = 1.0
    w[M.OFF_W3 + 0] = 1.0
    w[M.OFF_W4 + 0] = 1.0
    x = np.zeros((1, M.NIN))
    x[0, 0] = 1.5
    y, _ = M.forward(w, x)
    assert y[0, 0] == pytest.approx(M.sigmoid(M.gelu(M.gelu(M.gelu(1.5)))))
    assert y[0, 1] == 0.5 and y[0, 2] == 0.5


def test_box_features_closed_forms():
    # a ray along +z through the centre (1, 2) of a flat unit net at z = 3:
    # it enters the (flat) box at its centre; corners at (+-1, +-1, 0) / 0.5 scale
    bounds = np.array([[0.5, 1.5, 3.0, 1.5, 2.5, 3.0]])
    o, d = np.array([[1.0, 2.0, -7.0]]), np.array([[0.0, 0.0, 2.0]])
    for p, q in ((2, 2), (3, 5), (5, 4)):
        f = M.box_features(o, d, bounds, [_flat_net(p, q)], [p], [q])[0]
        np.testing.assert_allclose(f[:3], 0.0, atol=1e-12)
        np.testing.assert_allclose(f[3:6], [0, 0, 1])
        np.testing.assert_allclose(f[6:18].reshape(4, 3), [[-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 1, 0]])
        np.testing.assert_allclose(f[18:], [(p - 3.5) / 1.5, (q - 3.5) / 1.5])
    # the origin moved along the ray (outside the box) leaves the features
    # unchanged; an origin inside the box is its own entry point
    rng = np.random.default_rng(7)
    lo = rng.normal(size=(6, 3))
    b = np.concatenate([lo, lo + rng.uniform(0.2, 2, (6, 3))], 1)
    c = 0.5 * (b[:, :3] + b[:, 3:])
    d = rng.normal(size=(6, 3))
    o = c - 10.0 * d  # far in front of each box, aimed at its centre
    p = np.array([2, 3, 3, 4, 5, 2])
    nets = [np.concatenate([rng.normal(size=((k + 1) * (k + 1), 3)), rng.uniform(0.5, 2, ((k + 1) * (k + 1), 1))], 1)
            for k in p]
    f0 = M.box_features(o, d, b, nets, p, p)
    f1 = M.box_features(o + 3.0 * d, d, b, nets, p, p)
    np.testing.assert_allclose(f0, f1, atol=1e-12)
    hs = 0.5 * (b[:, 3:] - b[:, :3]).max(1)
    entry = f0[:, :3] * hs[:, None] + c  # on the box surface: one coordinate at a face
    face = np.minimum(np.abs(entry - b[:, :3]), np.abs(entry - b[:, 3:])).min(1)
    np.testing.assert_allclose(face, 0.0, atol=1e-9)
    inside = M.box_features(c + 0.1 * d, d, b, nets, p, p)
    np.testing.assert_allclose(inside[:, :3], 0.1 * d / hs[:, None], atol=1e-12)
    # the corners are the dehomogenised net corners; weights scale nothing else
    k0 = nets[0].reshape(3, 3, 4)
    np.testing.assert_allclose(f0[0, 6:9], (k0[0, 0, :3] / k0[0, 0, 3] - c[0]) / hs[0], atol=1e-12)
    np.testing.assert_allclose(f0[0, 15:18], (k0[2, 2, :3] / k0[2, 2, 3] - c[0]) / hs[0], atol=1e-12)
    # similarity invariance (translate + scale), direction scale invariance
    s, t = 2.5, np.array([-1.0, 4.0, 0.5])
    nets2 = [np.concatenate([n[:, :3] * s + t * n[:, 3:], n[:, 3:]], 1) for n in nets]  # homogeneous: s x w + t w
    f2 = M.box_features(o * s + t, 3.0 * d, b * s + np.tile(t, 2), nets2, p, p)
    np.testing.assert_allclose(f0, f2, atol=1e-10)
    # a zero direction component: the slab test drops that axis
    f3 = M.box_features(np.array([[1.0, 2.0, 10.0]]), np.array([[0.0, 0.0, -1.0]]), bounds, [_flat_net(2, 2)], [2], [2])
    np.testing.assert_allclose(f3[0, :3], 0.0, atol=1e-12)
    # parallel to a slab it lies outside: an infinite entry counts as 0, and
    # the entry point is clamped to 4 half extents
    f4 = M.box_features(np.array([[-9.0, 2.0, 10.0]]), np.array([[0.0, 0.0, -1.0]]), bounds, [_flat_net(2, 2)], [2], [2])
    np.testing.assert_allclose(f4[0, :3], [-4.0, 0.0, 4.0], atol=1e-12)
