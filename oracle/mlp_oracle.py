"""CPU oracle of the learned Newton seed (test infrastructure only).

The seed network has no counterpart in the reference: it is this project's
design, and this module restates it in numpy FP64 so the GPU kernels can be
checked against an independent implementation:

  ray_planes(d)       the two unit ray-plane normals (the reference's
                      ray-plane construction, math_utils.h:33-54)
  features(...)       the 20 pair features of kind 0 (a 3 x 3 sample of the
                      control net projected on the ray planes, degrees)
  box_features(...)   the 20 pair features of kind 1 (ray entry point and
                      direction, four corners, degrees, in the patch box's frame)
  forward(...)        20 -> 64 -> 64 -> 64 -> 3, GELU (tanh form; with
                      infer=True its polynomial fit gelu_infer, the GPU's
                      inference activation), sigmoid out (u, v, hit probability)
  loss_and_grad(...)  mean over pairs of h |uv - target|^2 + lambda BCE(hit, h)
                      and its gradient

With quantize=True the functions round to BF16 (round to nearest even) at
exactly the points where the GPU kernels feed the matrix cores: features,
weights, hidden activations, and the backward-pass error signals of hidden
layers 1..3 that enter products; the arithmetic between those points is
FP64.  With quantize=False everything is FP64.

Pinning (tests/test_mlp_oracle.py): the FP64 gradient is checked against
torch.autograd in float64 (torch's own binary cross-entropy with logits) and
against central finite differences; GELU against torch's tanh-approximate
GELU; gelu_infer against the tanh GELU (deviation bound, exact tails beyond
the clamp, odd/even structure); both feature kinds against closed forms (axis-aligned rays over flat
nets, odd/even middle indices, entry points, similarity invariance).
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 20, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NOUT * NH
NPARAM = OFF_B4 + NOUT

GK0 = np.sqrt(2.0 / np.pi)
GK1 = 0.044715


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    rnd = ((b >> 16) & 1) + 0x7FFF
    r = ((b + rnd) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def unpack(w):
    w = np.asarray(w, np.float64)
    assert w.size == NPARAM
    return dict(W1=w[OFF_W1:OFF_B1].reshape(NH, NIN), b1=w[OFF_B1:OFF_W2],
                W2=w[OFF_W2:OFF_B2].reshape(NH, NH), b2=w[OFF_B2:OFF_W3],
                W3=w[OFF_W3:OFF_B3].reshape(NH, NH), b3=w[OFF_B3:OFF_W4],
                W4=w[OFF_W4:OFF_B4].reshape(NOUT, NH), b4=w[OFF_B4:NPARAM])


def pack(P):
    return np.concatenate([P["W1"].ravel(), P["b1"], P["W2"].ravel(), P["b2"], P["W3"].ravel(), P["b3"],
                           P["W4"].ravel(), P["b4"]])


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GK0 * (x + GK1 * x ** 3)))


# inference activation: x * clip(0.5 + x Q(x^2), 0, 1), Q of degree 4 fitted
# (minimax) to gelu(x)/x - 0.5 on [0, 3.4] with 3.4 Q(3.4^2) = 0.5
GELU_INFER_Q = (3.940125455699649e-01, -5.970189424328037e-02, 6.703262281516497e-03, -4.152066351606772e-04,
                1.0574128852131707e-05)


def gelu_infer(x):
    x = np.asarray(x, np.float64)
    x2 = x * x
    t = np.zeros_like(x)
    for a in GELU_INFER_Q[::-1]:
        t = t * x2 + a
    return x * np.clip(0.5 + x * t, 0.0, 1.0)


def gelu_grad(x):
    t = np.tanh(GK0 * (x + GK1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GK0 * (1.0 + 3.0 * GK1 * x * x)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def ray_planes(d):
    """Unit normals n1, n2 ((n,3) each) of the two planes through a ray with
    direction d (n,3): n1 from the two largest-magnitude components' axis
    pair, n2 = n1 x d/|d| (math_utils.h:33-54)."""
    d = np.asarray(d, np.float64)
    ax, ay, az = np.abs(d).T
    xd = (ax > ay) & (ax > az)
    b0 = np.where(xd, d[:, 0], d[:, 1])
    b1 = np.where(xd, d[:, 1], d[:, 2])
    den = 1.0 / np.sqrt(b0 * b0 + b1 * b1)
    z = np.zeros_like(den)
    n1 = np.where(xd[:, None], np.stack([d[:, 1] * den, -d[:, 0] * den, z], 1),
                  np.stack([z, d[:, 2] * den, -d[:, 1] * den], 1))
    n2 = np.cross(n1, d / np.linalg.norm(d, axis=1, keepdims=True))
    return n1, n2


def features(o, d, bounds, nets, p, q):
    """o, d: (n,3) ray origin/direction; bounds: (n,6) min xyz, max xyz of the
    patch; nets: n homogeneous control nets, each ((p+1)(q+1), 4) row-major
    (rows = u index); p, q: (n,).  For sample rows {0, p/2, p} x columns
    {0, q/2, q} (an odd degree's middle = its two central indices), the mean
    of the dehomogenised points at those indices, x, gives
    (n1.(x - o), n2.(x - o)) / hs, hs = largest half extent of the box;
    then the degrees (p - 3.5)/1.5, (q - 3.5)/1.5."""
    o, d, bounds = (np.asarray(a, np.float64) for a in (o, d, bounds))
    p, q = np.asarray(p), np.asarray(q)
    n1, n2 = ray_planes(d)
    hs = 0.5 * np.maximum((bounds[:, 3:] - bounds[:, :3]).max(axis=1), 2e-20)
    out = np.zeros((len(o), NIN))
    for i in range(len(o)):
        pi, qi = int(p[i]), int(q[i])
        net = np.asarray(nets[i], np.float64).reshape(pi + 1, qi + 1, 4)
        x = net[..., :3] / net[..., 3:]
        rows = ((0, 0), (pi // 2, (pi + 1) // 2), (pi, pi))
        cols = ((0, 0), (qi // 2, (qi + 1) // 2), (qi, qi))
        for a, (r0, r1) in enumerate(rows):
            for b, (c0, c1) in enumerate(cols):
                m = 0.25 * (x[r0, c0] + x[r0, c1] + x[r1, c0] + x[r1, c1]) - o[i]
                out[i, 2 * (3 * a + b)] = m @ n1[i] / hs[i]
                out[i, 2 * (3 * a + b) + 1] = m @ n2[i] / hs[i]
        out[i, 18] = (pi - 3.5) / 1.5
        out[i, 19] = (qi - 3.5) / 1.5
    return out


def box_features(o, d, bounds, nets, p, q):
    """Kind 1: o, d, bounds, nets, p, q as features().  With c the box centre
    and hs its largest half extent, the ray enters the box at
    t_e = max(0, max_k min((lo_k - o_k)/d_k, (hi_k - o_k)/d_k)) (NaN terms of a
    zero direction component dropped; an infinite t_e counts as 0); then
    (o + t_e d - c)/hs clamped to [-4, 4], d/|d|, (corner - c)/hs for the
    dehomogenised corners P00, Pp0, P0q, Ppq, and the degrees (p - 3.5)/1.5,
    (q - 3.5)/1.5."""
    o, d, bounds = (np.asarray(a, np.float64) for a in (o, d, bounds))
    p, q = np.asarray(p), np.asarray(q)
    lo, hi = bounds[:, :3], bounds[:, 3:]
    with np.errstate(divide="ignore", invalid="ignore"):
        a, b = (lo - o) / d, (hi - o) / d
    near = np.fmin(a, b)
    te = np.fmax(np.fmax(np.fmax(near[:, 0], near[:, 1]), near[:, 2]), 0.0)
    te = np.where(np.isfinite(te), te, 0.0)
    c = 0.5 * (lo + hi)
    hs = np.maximum(0.5 * (hi - lo).max(axis=1), 2e-20)
    out = np.zeros((len(o), NIN))
    out[:, 0:3] = np.clip((o + te[:, None] * d - c) / hs[:, None], -4.0, 4.0)
    out[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    for i in range(len(o)):
        pi, qi = int(p[i]), int(q[i])
        net = np.asarray(nets[i], np.float64).reshape(pi + 1, qi + 1, 4)
        x = net[..., :3] / net[..., 3:]
        for k, (r, cc) in enumerate(((0, 0), (pi, 0), (0, qi), (pi, qi))):
            out[i, 6 + 3 * k:9 + 3 * k] = (x[r, cc] - c[i]) / hs[i]
        out[i, 18] = (pi - 3.5) / 1.5
        out[i, 19] = (qi - 3.5) / 1.5
    return out


def forward(w, X, quantize=False, infer=False):
    """X: (n, 20) -> (y (n,3): u, v, hit probability; cache).  infer=True
    uses the inference activation gelu_infer (the GPU's inference passes),
    otherwise the tanh GELU of training."""
    P = unpack(w)
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    act = gelu_infer if infer else gelu
    x = q(X)
    W = {k: q(P[k]) for k in ("W1", "W2", "W3", "W4")}
    z1 = x @ W["W1"].T + P["b1"]
    h1 = q(act(z1))
    z2 = h1 @ W["W2"].T + P["b2"]
    h2 = q(act(z2))
    z3 = h2 @ W["W3"].T + P["b3"]
    h3 = q(act(z3))
    z4 = h3 @ W["W4"].T + P["b4"]
    y = sigmoid(z4)
    return y, dict(x=x, W=W, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, y=y)


def loss_and_grad(w, X, T, quantize=False, hit_weight=0.05):
    """T: (n, 3) targets (u, v, h), h = 1 for a hit.  Mean over pairs of
    h |y_uv - (u,v)|^2 + hit_weight BCE(y_h, h) and its gradient (flat, NPARAM)."""
    n = X.shape[0]
    y, c = forward(w, X, quantize)
    q = bf16 if quantize else (lambda a: a)
    T = np.asarray(T, np.float64)
    h = (T[:, 2] > 0.5).astype(np.float64)
    e = (y[:, :2] - T[:, :2]) * h[:, None]
    yh = y[:, 2]
    bce = -(h * np.log(yh) + (1.0 - h) * np.log(1.0 - yh))
    loss = ((e * e).sum() + hit_weight * bce.sum()) / n
    dz4 = np.concatenate([2.0 * e * y[:, :2] * (1.0 - y[:, :2]), (hit_weight * (yh - h))[:, None]], 1) / n
    g = {}
    g["W4"] = dz4.T @ c["h3"]
    g["b4"] = dz4.sum(0)
    dz3 = (dz4 @ c["W"]["W4"]) * gelu_grad(c["z3"])
    g["b3"] = dz3.sum(0)
    g["W3"] = q(dz3).T @ c["h2"]
    dz2 = (q(dz3) @ c["W"]["W3"]) * gelu_grad(c["z2"])
    g["b2"] = dz2.sum(0)
    g["W2"] = q(dz2).T @ c["h1"]
    dz1 = (q(dz2) @ c["W"]["W2"]) * gelu_grad(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = q(dz1).T @ c["x"]
    return loss, pack(g)
