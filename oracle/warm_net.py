"""CPU oracle of the learned WLS warm start (lapstack_amd/csrc/warm.hip).

Test infrastructure: only tests/ call it, as the checker of the device's
forward and backward passes; nothing in the product imports it.

The network has no counterpart in the reference (whose PCG starts from the FGS
sweeps, crates/lapstack-core/src/depth.rs:1019-1034); this file states it in
numpy f64: zero-padded 3x3 convolutions, ReLU, 2x2 average pooling over the
cells inside the finer grid, nearest upsampling, channel concatenation, the
loss of the normalised prediction (mean squared error, or the error's energy
norm under the WLS operator of depth.rs:653-672/994), and the backward pass by
the chain rule.  With ``emulate=True`` it rounds to bf16 exactly where the device stores
bf16 (weights, the packed input, every activation and every activation
gradient), so the two differ only by the order of f32 accumulation.

Pinned by: central finite differences of the f64 loss, both losses, and the
energy norm against e^T A e with A assembled as a scipy sparse matrix
(tests/test_oracle_warm.py),
the convolution against a direct 7-loop sum, and pooling/upsampling as
adjoints (<P x, y> = <x, P* y>).
"""
from __future__ import annotations

import numpy as np

SPECS = [(3, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
         (64, 64, True), (64, 64, True), (128, 64, True), (96, 32, True), (48, 16, True), (16, 1, False)]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2", "d1", "d0", "out"]


def nparams() -> int:
    return sum(ci * co * 9 + co for ci, co, _ in SPECS)


def unflatten(p):
    """flat params -> [(W (co, ci, 3, 3), b (co,))] in network order."""
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for ci, co, _ in SPECS:
        W = p[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = p[o:o + co]
        o += co
        out.append((W, b))
    assert o == p.size
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init(seed: int = 0, scale: float = 1.0):
    """He-normal weights, small random biases (a test initialisation)."""
    rng = np.random.default_rng(seed)
    layers = []
    for ci, co, _ in SPECS:
        W = rng.standard_normal((co, ci, 3, 3)) * np.sqrt(2.0 / (9 * ci)) * scale
        b = rng.standard_normal(co) * 0.05
        layers.append((W, b))
    return flatten(layers)


def bf16(x):
    """Round to bf16 (to nearest, ties to even), via f32, returned as f64."""
    x32 = np.ascontiguousarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv(x, W, b=None):
    """x (B, H, W, ci), W (co, ci, 3, 3): zero-padded 3x3 correlation."""
    B, H, Wd, ci = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    cols = np.stack([xp[:, ky:ky + H, kx:kx + Wd] for ky in range(3) for kx in range(3)], axis=3)
    Wr = W.transpose(2, 3, 1, 0).reshape(9, ci, -1)  # [tap][ci][co]
    out = np.tensordot(cols, Wr, axes=([3, 4], [0, 1]))
    if b is not None:
        out = out + b
    return out


def conv_grads(x, W, dz):
    """dW, db of conv(x, W) + b for the upstream gradient dz; dx separately (conv_dx)."""
    B, H, Wd, ci = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros_like(W)
    for ky in range(3):
        for kx in range(3):
            dW[:, :, ky, kx] = np.tensordot(dz, xp[:, ky:ky + H, kx:kx + Wd], axes=([0, 1, 2], [0, 1, 2]))
    return dW, dz.sum(axis=(0, 1, 2))


def conv_dx(dz, W):
    """The input gradient: the correlation of dz with the flipped, transposed kernel."""
    return conv(dz, W[:, :, ::-1, ::-1].transpose(1, 0, 2, 3))


def pool(x):
    B, H, Wd, C = x.shape
    h, w = (H + 1) // 2, (Wd + 1) // 2
    s = np.zeros((B, h, w, C))
    cnt = np.zeros((h, w))
    for dy in range(2):
        for dx in range(2):
            part = x[:, dy::2, dx::2]
            s[:, :part.shape[1], :part.shape[2]] += part
            cnt[:part.shape[1], :part.shape[2]] += 1
    return s / cnt[None, :, :, None]


def pool_bwd(dp, shape):
    """The adjoint of pool onto the finer grid of ``shape`` (B, H, W, C)."""
    B, H, Wd, C = shape
    h, w = dp.shape[1], dp.shape[2]
    cnt = np.zeros((h, w))
    for dy in range(2):
        for dx in range(2):
            cnt[:len(range(dy, H, 2)), :len(range(dx, Wd, 2))] += 1
    g = (dp / cnt[None, :, :, None])
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)[:, :H, :Wd]


def up(x, shape):
    H, Wd = shape
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :H, :Wd]


def up_bwd(du, shape):
    """The adjoint of up onto the coarser grid of ``shape`` (h, w)."""
    h, w = shape
    B, H, Wd, C = du.shape
    s = np.zeros((B, h, w, C))
    for dy in range(2):
        for dx in range(2):
            part = du[:, dy::2, dx::2]
            s[:, :part.shape[1], :part.shape[2]] += part
    return s


def _inputs(d, conf, g, dscale, emulate):
    d, conf, g = (np.asarray(a, np.float32) for a in (d, conf, g))
    if d.ndim == 2:
        d, conf, g = d[None], conf[None], g[None]
    dn = (d * np.float32(1.0 / dscale)).astype(np.float32)
    x = np.stack([dn, conf, g], axis=-1).astype(np.float64)
    return (bf16(x) if emulate else x), dn.astype(np.float64)


def forward(params, d, conf, g, dscale, emulate=False):
    """The normalised prediction d / dscale + out (B, H, W) and the activations."""
    R = bf16 if emulate else (lambda a: a)
    L = [(R(W) if emulate else W, b) for W, b in unflatten(params)]
    relu = lambda a: np.maximum(a, 0.0)
    X0, dn = _inputs(d, conf, g, dscale, emulate)
    c = {"X0": X0}
    c["A0a"] = R(relu(conv(X0, *L[0])))
    c["S0"] = R(relu(conv(c["A0a"], *L[1])))
    c["P0"] = R(pool(c["S0"]))
    c["A1a"] = R(relu(conv(c["P0"], *L[2])))
    c["S1"] = R(relu(conv(c["A1a"], *L[3])))
    c["P1"] = R(pool(c["S1"]))
    c["A2a"] = R(relu(conv(c["P1"], *L[4])))
    c["S2"] = R(relu(conv(c["A2a"], *L[5])))
    c["P2"] = R(pool(c["S2"]))
    c["A3a"] = R(relu(conv(c["P2"], *L[6])))
    c["B3"] = R(relu(conv(c["A3a"], *L[7])))
    c["C2"] = np.concatenate([up(c["B3"], c["S2"].shape[1:3]), c["S2"]], axis=-1)
    c["D2"] = R(relu(conv(c["C2"], *L[8])))
    c["C1"] = np.concatenate([up(c["D2"], c["S1"].shape[1:3]), c["S1"]], axis=-1)
    c["D1"] = R(relu(conv(c["C1"], *L[9])))
    c["C0"] = np.concatenate([up(c["D1"], c["S0"].shape[1:3]), c["S0"]], axis=-1)
    c["D0"] = R(relu(conv(c["C0"], *L[10])))
    out = conv(c["D0"], *L[11])[..., 0]
    if emulate:
        out = out.astype(np.float32).astype(np.float64)
    c["L"] = L
    c["dn"] = dn
    return dn + out, c


def energy(e, conf, g, lam, sigma):
    """The energy norm per cell of the error planes e (B, H, W) under each
    image's WLS operator (data weights max(conf, 0) + 1e-4, edges
    lam * exp(-|dg| / sigma)), and its gradient 2 (A e) / N."""
    e = np.asarray(e, np.float64)
    conf = np.asarray(conf, np.float64).reshape(e.shape)
    g = np.asarray(g, np.float64).reshape(e.shape)
    w = np.maximum(conf, 0.0) + 1e-4
    ax = lam * np.exp(-np.abs(g[:, :, 1:] - g[:, :, :-1]) / sigma)
    ay = lam * np.exp(-np.abs(g[:, 1:] - g[:, :-1]) / sigma)
    dx, dy = e[:, :, :-1] - e[:, :, 1:], e[:, :-1] - e[:, 1:]
    N = e.size
    loss = (np.sum(w * e * e) + np.sum(ax * dx * dx) + np.sum(ay * dy * dy)) / N
    Ae = w * e
    Ae[:, :, :-1] += ax * dx
    Ae[:, :, 1:] -= ax * dx
    Ae[:, :-1] += ay * dy
    Ae[:, 1:] -= ay * dy
    return loss, 2.0 * Ae / N


def loss_and_grads(params, d, conf, g, target, dscale, emulate=False, loss="mse", lam=0.0, sigma=1.0):
    """The loss of the normalised prediction against target / dscale (mean
    squared error, or the energy norm under the WLS operator) and its gradient
    with respect to the flat parameters."""
    R = bf16 if emulate else (lambda a: a)
    pred, c = forward(params, d, conf, g, dscale, emulate)
    L = c["L"]
    t = np.asarray(target, np.float32)
    if t.ndim == 2:
        t = t[None]
    tn = (t * np.float32(1.0 / dscale)).astype(np.float64)
    if emulate:
        pred = pred.astype(np.float32).astype(np.float64)
    e = pred - tn
    N = e.size
    if emulate:
        e = e.astype(np.float32).astype(np.float64)
    if loss == "energy":
        cf = np.asarray(conf, np.float32)
        gg = np.asarray(g, np.float32)
        loss_v, de = energy(e, cf.reshape(e.shape), gg.reshape(e.shape), lam, sigma)
    else:
        loss_v, de = float(np.mean(e * e)), (2.0 / N) * e
    grads = [None] * len(L)
    pos = lambda a: (a > 0).astype(np.float64)

    def layer(i, x, dz):
        grads[i] = conv_grads(x, L[i][0], dz)

    dOUT = R(de)[..., None]
    layer(11, c["D0"], dOUT)
    dD0 = R(conv_dx(dOUT, L[11][0]) * pos(c["D0"]))
    layer(10, c["C0"], dD0)
    dC0 = R(conv_dx(dD0, L[10][0]) * pos(c["C0"]))
    dD1 = R(up_bwd(dC0[..., :32], c["D1"].shape[1:3]))
    layer(9, c["C1"], dD1)
    dC1 = R(conv_dx(dD1, L[9][0]) * pos(c["C1"]))
    dD2 = R(up_bwd(dC1[..., :64], c["D2"].shape[1:3]))
    layer(8, c["C2"], dD2)
    dC2 = R(conv_dx(dD2, L[8][0]) * pos(c["C2"]))
    dB3 = R(up_bwd(dC2[..., :64], c["B3"].shape[1:3]))
    layer(7, c["A3a"], dB3)
    dA3a = R(conv_dx(dB3, L[7][0]) * pos(c["A3a"]))
    layer(6, c["P2"], dA3a)
    dP2 = R(conv_dx(dA3a, L[6][0]))
    dS2 = R(dC2[..., 64:] + pool_bwd(dP2, c["S2"].shape) * pos(c["S2"]))
    layer(5, c["A2a"], dS2)
    dA2a = R(conv_dx(dS2, L[5][0]) * pos(c["A2a"]))
    layer(4, c["P1"], dA2a)
    dP1 = R(conv_dx(dA2a, L[4][0]))
    dS1 = R(dC1[..., 64:] + pool_bwd(dP1, c["S1"].shape) * pos(c["S1"]))
    layer(3, c["A1a"], dS1)
    dA1a = R(conv_dx(dS1, L[3][0]) * pos(c["A1a"]))
    layer(2, c["P0"], dA1a)
    dP0 = R(conv_dx(dA1a, L[2][0]))
    dS0 = R(dC0[..., 32:] + pool_bwd(dP0, c["S0"].shape) * pos(c["S0"]))
    layer(1, c["A0a"], dS0)
    dA0a = R(conv_dx(dS0, L[1][0]) * pos(c["A0a"]))
    layer(0, c["X0"], dA0a)
    return float(loss_v), flatten(grads)
