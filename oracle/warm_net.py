"""CPU oracle of the learned WLS warm start (lapstack_amd/csrc/warm.hip).

Test infrastructure: only tests/ call it, as the checker of the device's
forward and backward passes; nothing in the product imports it.

The network has no counterpart in the reference (whose PCG starts from the FGS
sweeps, crates/lapstack-core/src/depth.rs:1019-1034); this file states it in
numpy f64.  Its input: the WLS system's planes averaged over 8 x 8 blocks of
the grid (partial blocks at the right and bottom edges average the cells that
exist).  With dm the block's mean depth weighted by the data weight
max(conf, 0) + 1e-4 (depth.rs:994), depth enters as differences from dm in
units of 1/32: the plain mean depth, and per push-pull level (4 x 4, 16 x 16
and 64 x 64 blocks of blocks) the weighted mean resampled bilinearly to the
block grid; then the mean data weight and its log, the mean edge conductances
exp(-|dg| / sigma_c) (depth.rs:653-672; 0 past the last column or row), and
the first two levels' log mean weight, resampled.
The network: zero-padded 3x3 convolutions, ReLU, 2x2 average pooling over the
cells inside the finer grid, nearest upsampling, channel concatenation; its
prediction of the solution's block means is the weighted mean plus 1/32 of
the output, and the start is that prediction resampled bilinearly to the WLS
grid.  The loss: per system the mean squared error against the target's block
means, relative to the mean square of the target's own deviation from the
weighted mean (at least 1e-4, in units of 1/32), averaged over the systems;
the backward pass by the chain rule.  With ``emulate=True`` it rounds to bf16
exactly where the device stores bf16 (weights, the packed input, every
activation and every activation gradient), so the two differ only by the
order of f32 accumulation and the f32 arithmetic of the input planes.

Pinned by (tests/test_oracle_warm.py): central finite differences of the f64
loss, the convolution against a direct 7-loop sum, pooling/upsampling as
adjoints (<P x, y> = <x, P* y>), the block planes against a loop over the
cells of each block, and the bilinear resampling against
torch.nn.functional.interpolate (bilinear, align_corners=False).
"""
from __future__ import annotations

import numpy as np

SPECS = [(10, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
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


WB = 8                # block edge, in WLS cells
KAPPA = 1.0 / 32.0    # the output's scale, in depth / dscale
PP = (4, 16, 64)      # push-pull levels, in blocks


def block_means(a, k):
    """Means of a (B, H, W) over k x k blocks; partial blocks at the right and
    bottom edges average the cells that exist."""
    a = np.asarray(a, np.float64)
    B, H, W = a.shape
    h, w = -(-H // k), -(-W // k)
    s = np.zeros((B, h, w))
    cnt = np.zeros((h, w))
    for dy in range(k):
        for dx in range(k):
            part = a[:, dy::k, dx::k]
            s[:, :part.shape[1], :part.shape[2]] += part
            cnt[:part.shape[1], :part.shape[2]] += 1
    return s / cnt


def resample(p, oh, ow):
    """Bilinear resampling of p (B, ih, iw) onto an (oh, ow) grid over the same
    area: source position (o + 1/2) in/out - 1/2, clamped at 0, the
    right/lower neighbour clamped to the last cell."""
    p = np.asarray(p, np.float64)
    _, ih, iw = p.shape
    sy = np.maximum((np.arange(oh) + 0.5) * (ih / oh) - 0.5, 0.0)
    sx = np.maximum((np.arange(ow) + 0.5) * (iw / ow) - 0.5, 0.0)
    y0, x0 = sy.astype(np.int64), sx.astype(np.int64)
    y1, x1 = np.minimum(y0 + 1, ih - 1), np.minimum(x0 + 1, iw - 1)
    ly, lx = (sy - y0)[None, :, None], (sx - x0)[None, None, :]
    top = (1 - lx) * p[:, y0][:, :, x0] + lx * p[:, y0][:, :, x1]
    bot = (1 - lx) * p[:, y1][:, :, x0] + lx * p[:, y1][:, :, x1]
    return (1 - ly) * top + ly * bot


def _planes(*a):
    a = [np.asarray(x, np.float32) for x in a]
    return [x[None] if x.ndim == 2 else x for x in a]


def features(d, conf, g, dscale, sigma=0.04, t=None):
    """The input planes X (B, ch, cw, 10), the weighted mean dm (B, ch, cw), and
    with a target t its block means' deviation tc = mean(t) / dscale - dm."""
    d, conf, g = (x.astype(np.float64) for x in _planes(d, conf, g))
    wd = np.maximum(conf, 0.0) + 1e-4
    pw = block_means(wd, WB)
    dm = block_means(wd * d, WB) / pw / dscale
    ax = np.zeros_like(g)
    ax[:, :, :-1] = np.exp(-np.abs(g[:, :, :-1] - g[:, :, 1:]) / sigma)
    ay = np.zeros_like(g)
    ay[:, :-1] = np.exp(-np.abs(g[:, :-1] - g[:, 1:]) / sigma)
    ch, cw = pw.shape[1:]
    feats = [(block_means(d, WB) / dscale - dm) / KAPPA]
    lpk = []
    for k in PP:
        pk = block_means(pw, k)
        mk = block_means(pw * dm, k) / pk
        feats.append((resample(mk, ch, cw) - dm) / KAPPA)
        lpk.append(resample(np.log10(pk), ch, cw) / 4 + 1)
    feats += [pw, np.log10(pw) / 4 + 1, block_means(ax, WB), block_means(ay, WB)] + lpk[:2]
    tc = None
    if t is not None:
        (t,) = _planes(t)
        tc = block_means(t, WB) / dscale - dm
    return np.stack(feats, axis=-1), dm, tc


def forward(params, d, conf, g, dscale, sigma=0.04, emulate=False):
    """The predicted block means dm + out / 32 (B, ch, cw), in depth / dscale,
    and the activations."""
    R = bf16 if emulate else (lambda a: a)
    L = [(R(W) if emulate else W, b) for W, b in unflatten(params)]
    relu = lambda a: np.maximum(a, 0.0)
    X, dm, _ = features(d, conf, g, dscale, sigma)
    X0 = bf16(X) if emulate else X
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
    c["dm"] = dm
    c["out"] = out
    return dm + KAPPA * out, c


def predict(params, d, conf, g, dscale, sigma=0.04, emulate=False):
    """The start on the WLS grid (B, H, W): dscale times the predicted block
    means resampled bilinearly."""
    (d,) = _planes(d)
    pred, _ = forward(params, d, conf, g, dscale, sigma, emulate)
    return dscale * resample(pred, d.shape[1], d.shape[2])


def loss_and_grads(params, d, conf, g, target, dscale, sigma=0.04, emulate=False):
    """The loss (per system the mean squared error of out against tc * 32,
    divided by max(mean((tc * 32)^2), 1e-4), averaged over the systems) and its
    gradient with respect to the flat parameters."""
    R = bf16 if emulate else (lambda a: a)
    _, c = forward(params, d, conf, g, dscale, sigma, emulate)
    L = c["L"]
    _, _, tc = features(d, conf, g, dscale, sigma, t=target)
    that = tc / KAPPA
    e = c["out"] - that
    if emulate:
        e = e.astype(np.float32).astype(np.float64)
    B, ch, cw = e.shape
    wimg = 1.0 / np.maximum(np.mean(that * that, axis=(1, 2)), 1e-4)
    loss_v = float(np.mean(wimg * np.mean(e * e, axis=(1, 2))))
    de = 2.0 * wimg[:, None, None] * e / (B * ch * cw)
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
