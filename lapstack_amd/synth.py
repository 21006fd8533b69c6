"""Seeded synthetic focus brackets (the workloads the tests and the benchmark run).

A stack is made from a scene and a depth map:

* scene: 1/f band-limited noise (white Gaussian noise on the power-of-two
  canvas that holds w x h, filtered by 1/|f| for 1/min(w,h) <= |f| <= 0.25
  cycles/px, cropped to w x h, normalised to max |v| = 1, times 0.3, plus
  0.5), twelve random step edges (half-planes, +-0.1), RGB gains
  (1.0, 0.9, 0.8) and a +-0.02 smooth colour field per channel, clipped to [0, 1];
* depth: a tilted plane plus six Gaussians (centres uniform, sigma 5-25 % of
  the short side, amplitudes +-[0.2, 0.6]), rescaled to [0, 1];
* frame i of n focuses at f_i = i/(n-1): every pixel is blurred by a Gaussian
  of sigma = 6 px * |depth - f_i|.  The per-pixel blur is drawn from a stack of
  the scene blurred at sigma = 0, 0.5, ..., 6.5 px (by its spectrum: the scene
  reflect-padded by 3 sigma_max, then edge-extended to power-of-two sizes),
  linearly between the two nearest sigmas;
* jitter (optional): frame i is resampled (bicubic) through the similarity
  A_i(p) = c + s_i R(theta_i) (p - c) + t_i about the centre c, with
  s_i = 1 + 0.002 i, t_i uniform in +-3 px per axis and theta_i uniform in
  +-0.05 deg; frame 0 is left at the identity.  A_i is exactly the Sim that
  registers frame i onto frame 0 (xoff = t_x/w, yoff = t_y/h, scale s_i,
  rot theta_i);
* noise N(0, 0.5*256/65535) is added and the frame quantised to u16 as
  round(clip(v) * 65535), interleaved RGB (h, w, 3).

Seeds: scene s0, depth s0+1, noise s0+100+i, jitter s0+1000+i.  The generator
runs in torch on whatever device it is given; frames come back as device
tensors of torch.uint16.  Every FFT is of power-of-two sizes, so stacks of many
different frame sizes share a few transform plans (on the GPU a plan of an
unusual size is compiled at run time, seconds each).
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

SIGMA_STEP = 0.5
SIGMA_MAX = 6.5
NOISE_SIGMA = 0.5 * 256.0 / 65535.0


def _gen(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def _pow2(n):
    return 1 << max(0, (int(n) - 1).bit_length())


def _one_over_f(h, w, seed, device):
    g = _gen(seed, device)
    ch, cw = _pow2(h), _pow2(w)
    white = torch.randn(ch, cw, generator=g, device=device)
    spec = torch.fft.rfft2(white)
    fy = torch.fft.fftfreq(ch, device=device)[:, None]
    fx = torch.fft.rfftfreq(cw, device=device)[None, :]
    f = torch.sqrt(fy * fy + fx * fx)
    fmin = 1.0 / min(w, h)
    filt = torch.where((f >= fmin) & (f <= 0.25), 1.0 / f.clamp_min(1e-12), torch.zeros_like(f))
    v = torch.fft.irfft2(spec * filt, s=(ch, cw))[:h, :w]
    return v / v.abs().max().clamp_min(1e-12)


def scene(h, w, seed, device="cpu"):
    """(3, h, w) float32 in [0, 1]."""
    base = _one_over_f(h, w, seed, device) * 0.3 + 0.5
    g = _gen(seed + 7, device)
    yy = torch.arange(h, device=device, dtype=torch.float32)[:, None]
    xx = torch.arange(w, device=device, dtype=torch.float32)[None, :]
    for _ in range(12):
        ang = float(torch.rand((), generator=g, device=device)) * 2 * math.pi
        px = float(torch.rand((), generator=g, device=device)) * w
        py = float(torch.rand((), generator=g, device=device)) * h
        step = 0.1 if float(torch.rand((), generator=g, device=device)) < 0.5 else -0.1
        side = (xx - px) * math.cos(ang) + (yy - py) * math.sin(ang) > 0
        base = base + step * side.float()
    out = []
    for c, gain in enumerate((1.0, 0.9, 0.8)):
        col = _one_over_f(max(h // 8, 8), max(w // 8, 8), seed + 11 + c, device)[None, None]
        col = F.interpolate(col, size=(h, w), mode="bilinear", align_corners=False)[0, 0]
        out.append(base * gain + 0.02 * col)
    return torch.stack(out).clamp(0.0, 1.0).float()


def depth_map(h, w, seed, device="cpu"):
    """(h, w) float32 in [0, 1]."""
    g = _gen(seed, device)
    r = lambda: float(torch.rand((), generator=g, device=device))
    yy = torch.arange(h, device=device, dtype=torch.float32)[:, None]
    xx = torch.arange(w, device=device, dtype=torch.float32)[None, :]
    d = (r() - 0.5) * xx / w + (r() - 0.5) * yy / h
    short = min(w, h)
    for _ in range(6):
        cx, cy = r() * w, r() * h
        s = (0.05 + 0.20 * r()) * short
        amp = (0.2 + 0.4 * r()) * (1.0 if r() < 0.5 else -1.0)
        d = d + amp * torch.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    d = d - d.min()
    return (d / d.max().clamp_min(1e-12)).float()


class _Blurrer:
    """Gaussian blurs of one image by its spectrum: reflect-padded by 3 sigma_max,
    then edge-extended on the bottom and right to power-of-two sizes (the
    circular wrap of a blur of at most 3 sigma_max reaches only the padding)."""

    def __init__(self, img):
        self.pad = int(math.ceil(3 * SIGMA_MAX)) + 1
        c, h, w = img.shape
        self.h, self.w = h, w
        mode = "reflect" if self.pad < min(h, w) else "replicate"
        x = F.pad(img[None], (self.pad,) * 4, mode=mode)
        ph, pw = _pow2(x.shape[2]), _pow2(x.shape[3])
        x = F.pad(x, (0, pw - x.shape[3], 0, ph - x.shape[2]), mode="replicate")[0]
        self.ph, self.pw = ph, pw
        self.spec = torch.fft.rfft2(x)
        fy = torch.fft.fftfreq(self.ph, device=img.device)[:, None]
        fx = torch.fft.rfftfreq(self.pw, device=img.device)[None, :]
        self.f2 = (fy * fy + fx * fx).float()

    def blur(self, sigma):
        tf = torch.exp(-2.0 * math.pi * math.pi * sigma * sigma * self.f2)
        x = torch.fft.irfft2(self.spec * tf, s=(self.ph, self.pw))
        p = self.pad
        return x[:, p:p + self.h, p:p + self.w].contiguous()


class Bracket:
    """A seeded focus bracket of n frames of w x h; frames are made on demand."""

    def __init__(self, w, h, n, seed=1, jitter=False, device="cpu"):
        self.w, self.h, self.n, self.seed, self.jitter, self.device = w, h, n, seed, jitter, device
        sc = scene(h, w, seed, device)
        self.depth = depth_map(h, w, seed + 1, device)
        self.levels = [sc]
        bl = _Blurrer(sc)
        s = SIGMA_STEP
        while s <= SIGMA_MAX + 1e-9:
            self.levels.append(bl.blur(s))
            s += SIGMA_STEP
        del bl, sc
        self.sims = [self.true_sim(i) for i in range(n)]

    def true_sim(self, i):
        """The Sim (xoff, yoff, scale, rot, aspect, shear, px, py) registering frame i onto frame 0."""
        if not self.jitter or i == 0:
            return (0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0)
        g = _gen(self.seed + 1000 + i, "cpu")
        tx = (float(torch.rand((), generator=g)) * 2 - 1) * 3.0
        ty = (float(torch.rand((), generator=g)) * 2 - 1) * 3.0
        th = math.radians((float(torch.rand((), generator=g)) * 2 - 1) * 0.05)
        return (tx / self.w, ty / self.h, 1.0 + 0.002 * i, th, 1.0, 0.0, 0.0, 0.0)

    def frame(self, i):
        """Frame i: (h, w, 3) torch.uint16 on the generator's device."""
        fi = i / (self.n - 1) if self.n > 1 else 0.0
        sigma = (6.0 * (self.depth - fi).abs()).clamp(0, SIGMA_MAX - 1e-4)
        pos = sigma / SIGMA_STEP
        lo = pos.floor().long()
        t = (pos - lo.float())[None]
        img = torch.zeros(3, self.h, self.w, device=self.device)
        for k in range(len(self.levels) - 1):
            m = (lo == k)[None]
            if not bool(m.any()):
                continue
            a, b = self.levels[k].float(), self.levels[k + 1].float()
            img = torch.where(m, a * (1 - t) + b * t, img)
        if self.jitter and i > 0:
            xoff, yoff, s, th = self.sims[i][:4]
            cx, cy = self.w / 2.0, self.h / 2.0
            yy, xx = torch.meshgrid(torch.arange(self.h, device=self.device, dtype=torch.float32),
                                    torch.arange(self.w, device=self.device, dtype=torch.float32), indexing="ij")
            c, sn = math.cos(th), math.sin(th)
            qx = cx + s * (c * (xx - cx) - sn * (yy - cy)) + xoff * self.w
            qy = cy + s * (sn * (xx - cx) + c * (yy - cy)) + yoff * self.h
            grid = torch.stack([2 * qx / (self.w - 1) - 1, 2 * qy / (self.h - 1) - 1], -1)[None]
            del xx, yy, qx, qy
            img = F.grid_sample(img[None], grid, mode="bicubic", padding_mode="border", align_corners=True)[0]
            del grid
        g = _gen(self.seed + 100 + i, self.device)
        img = img + NOISE_SIGMA * torch.randn(img.shape, generator=g, device=self.device)
        q = torch.round(img.clamp(0.0, 1.0) * 65535.0).to(torch.int32)
        return q.permute(1, 2, 0).contiguous().to(torch.uint16)
