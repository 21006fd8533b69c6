"""Procedural fixtures restated from the reference's own unit tests (they hold no
data files): xorshift noise (pyramid.rs:235-250), the LCG (depth.rs:1206-1209),
checkerboards (fuse.rs:547-558, depth.rs:1343-1354), repeated 5-tap blurs
(fuse.rs:561-569, depth.rs:1356-1364); and the WLS system (W + lambda L) as a
scipy sparse matrix in f64, the textbook form of depth.rs:653-672/786-810."""
import numpy as np
import scipy.sparse


def noise_img(w, h, seed):
    s = seed & 0xFFFFFFFF
    out = np.zeros((3, h, w), np.float32)
    flat = out.reshape(3, -1)
    for c in range(3):
        for i in range(w * h):
            s ^= (s << 13) & 0xFFFFFFFF
            s ^= s >> 17
            s ^= (s << 5) & 0xFFFFFFFF
            flat[c, i] = np.float32(s % 10000) / np.float32(10000.0)
    return out


def lcg_seq(seed, n):
    s = seed
    out = np.zeros(n, np.float32)
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = np.float32((s >> 33) % 10000) / np.float32(10000.0)
    return out


def checker(w, h, cell, factors=(1.0, 0.9, 0.5)):
    y, x = np.mgrid[0:h, 0:w]
    v = np.where(((x // cell) + (y // cell)) % 2 == 0, np.float32(0.2), np.float32(0.8)).astype(np.float32)
    return np.stack([(v * np.float32(f)).astype(np.float32) for f in factors])


def blur(img3, passes):
    from oracle import ref
    k = np.array([1 / 16, 4 / 16, 6 / 16, 4 / 16, 1 / 16], np.float32)
    o = img3.copy()
    for _ in range(passes):
        o = np.stack([ref.window_sum(o[c], k) for c in range(3)])
    return o


def rms(a, b):
    d = (a.astype(np.float32) - b.astype(np.float32)).astype(np.float64)
    return float(np.sqrt((d * d).mean()))


def wls_matrix(guide, conf, lam, sigma):
    h, w = guide.shape
    g = guide.astype(np.float32)
    inv = np.float32(-1.0) / np.float32(max(sigma, 1e-6))
    ex = np.exp(np.abs(g[:, :-1] - g[:, 1:]) * inv).astype(np.float64) * lam
    ey = np.exp(np.abs(g[:-1] - g[1:]) * inv).astype(np.float64) * lam
    n = w * h
    idx = np.arange(n).reshape(h, w)
    rows, cols, vals = [], [], []
    diag = np.maximum(conf.astype(np.float64), 0) + 1e-4
    diag = diag.ravel().copy()
    for (a, b, wgt) in [(idx[:, :-1], idx[:, 1:], ex), (idx[:-1], idx[1:], ey)]:
        a, b, wgt = a.ravel(), b.ravel(), wgt.ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [-wgt, -wgt]
        np.add.at(diag, a, wgt)
        np.add.at(diag, b, wgt)
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    A = scipy.sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A
