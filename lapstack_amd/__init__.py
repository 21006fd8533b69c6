"""lapstack_amd -- the MI355X (gfx950) backend of lapstack-core's stacking engine.

The host side mirrors the reference engine's backend seam
(crates/lapstack-core/src/stack.rs:122-230, align.rs:27-49) over the C ABI of
lapstack_amd/liblsh.so (declared in include/lsh.h):

* ``available()``          -- the backend probe (gpu.rs:307-310).
* ``Fuser``                -- fuse.rs:349-521 Fuser / gpu.rs GpuFuser: new, levels,
                              push, push_frames (the device handoff), shares, finish.
* ``HipFrames``            -- gpu.rs:636-900 GpuFrames: process (register, warp,
                              brightness, focus slice), ingest (registered stacks).
* ``depth_from_slices``    -- depth.rs:1119-1200.
* ``depth_from_source``    -- the same with its slice iterator and log callback.
* ``depth_grid``           -- depth.rs:1125-1159: the WLS system's inputs.
* ``WarmNet``              -- the learned warm start of the WLS solve (no
                              reference counterpart), trained on the device.
* ``wls_solve``            -- depth.rs:1081 wls_solve, in three precisions
                              (reference FP32, FP64-refined FP32, FP64).
* ``fuse_and_depth``       -- stack.rs:740-853: the per-frame loop over a stack.
* ``stack_stream``         -- the same loop with no host round trip per frame,
                              the per-frame step replayed as a hipGraph.
* ``batch_fuse``           -- many uneven stacks through the fuse path in one
                              persistent launch, tile shape autotuned.

Buffers are numpy arrays (host memory) or torch tensors on the HIP device
(device memory, passed by pointer, no copy).  There is no CPU fallback: every
call runs on the device, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import LSH_DEVICE, LSH_HOST, HipError

__all__ = ["available", "Backend", "FuseParams", "DepthParams", "AlignParams", "Fuser", "HipFrames", "DepthMap",
           "depth_from_slices", "depth_from_source", "depth_grid", "stack_stream", "batch_fuse", "WavParams", "WeightedAverage",
           "weighted_average", "slab_ranges", "fuse_slabs",
           "common_area", "crop", "resize", "restretch", "to_u16", "gauss_levels", "wls_solve", "wls_solve_noalloc", "image_info", "load_rgb", "save_rgb", "save_rgb_u16", "save_gray16", "save_gray", "stack_stream_files", "WarmNet", "WLS_MODES", "fuse_and_depth", "Interp", "TopRule", "HipError", "SIM_ID"]

SIM_ID = (0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0)


class Backend:
    """align.rs:29-36 Backend, with the third engine this package provides."""

    CPU = "cpu"
    CUDA = "cuda"
    WGPU = "wgpu"
    HIP = "hip"


class Interp:
    """align.rs:128-143, numbered as Interp::id (the order of Interp::ALL)."""

    NEAREST, BILINEAR, BICUBIC, SPLINE4X4, SPLINE6X6, LANCZOS3 = range(6)
    NAMES = {"nearest": 0, "nn": 0, "bilinear": 1, "linear": 1, "bicubic": 2, "cubic": 2, "spline4x4": 3,
             "spline16": 3, "spline6x6": 4, "spline36": 4, "lanczos3": 5, "lanczos": 5}

    @classmethod
    def parse(cls, s) -> int:
        if isinstance(s, int):
            return s
        v = cls.NAMES.get(str(s).strip().lower())
        if v is None:
            raise ValueError(f"unknown interpolation kernel {s!r}")
        return v


class TopRule:
    """fuse.rs:48-67."""

    DEV_ENTROPY, DEVIATION, AVERAGE = 0, 1, 2
    NAMES = {"de": 0, "dev-entropy": 0, "dev": 1, "deviation": 1, "avg": 2, "average": 2}

    @classmethod
    def parse(cls, s) -> int:
        return s if isinstance(s, int) else cls.NAMES[s]


@dataclasses.dataclass
class FuseParams:
    """fuse.rs:69-107 with its defaults."""

    levels: Optional[int] = None
    energy_radius: int = 1
    top_rule: object = TopRule.DEV_ENTROPY
    top_radius: int = 2
    entropy_bins: int = 256
    use_chroma: bool = False
    depth_level: int = 2
    halo: float = 0.0

    def _c(self):
        return _lib.FuseParamsC(0 if self.levels is None else int(self.levels), self.energy_radius,
                                TopRule.parse(self.top_rule), self.top_radius, self.entropy_bins,
                                int(self.use_chroma), self.depth_level, float(self.halo))


@dataclasses.dataclass
class DepthParams:
    """depth.rs:68-109 with its defaults; ``tol`` is the PCG stop (1e-5 in depth.rs:1054)."""

    scale: int = 1
    focus: str = "rdf"
    r_in: int = 1
    r_out: int = 3
    sml_step: int = 1
    agg_radius: int = 3
    agg_eps: float = 1e-4
    lam: float = 3.0
    sigma_c: float = 0.04
    cg_iters: int = 200
    median: bool = True
    upsample: str = "guided"
    up_radius: int = 2
    up_eps: float = 1e-2
    gate: float = 1.0
    robust: float = 1.0
    tol: float = 1e-5
    precision: str = "reference"  # the WLS mode: "reference" (FP32 MG-PCG) or "refined" (inside FP64 refinement)
    # a WarmNet: the first WLS solve starts from its prediction, the reweighted one from the first's result
    warm: Optional["WarmNet"] = None

    def _c(self):
        if self.precision not in ("reference", "refined"):
            raise ValueError(f"depth: unknown WLS precision {self.precision!r}")
        return _lib.DepthParamsC(self.scale, 0 if self.focus == "rdf" else 1, self.r_in, self.r_out, self.sml_step,
                                 self.agg_radius, self.agg_eps, self.lam, self.sigma_c, self.cg_iters,
                                 int(self.median), 1 if self.upsample == "guided" else 0, self.up_radius,
                                 self.up_eps, self.gate, self.robust, self.tol, WLS_MODES[self.precision],
                                 self.warm._h if self.warm is not None else None)

    def grid(self, w, h):
        k = 1 << self.scale
        return -(-w // k), -(-h // k)


@dataclasses.dataclass
class AlignParams:
    """align.rs:51-77 (similarity model by default: shift, scale, rotation free)."""

    shift: bool = True
    scale: bool = True
    rotation: bool = True
    model: str = "similarity"
    coarsen: int = 0
    interp: object = "spline4x4"

    def free(self):
        """align.rs:115-119 free_mask."""
        affine = self.model != "similarity"
        proj = self.model == "projective"
        return [self.shift, self.shift, self.scale, self.rotation, affine, affine, proj, proj]


def available() -> Optional[str]:
    """None when a gfx950 device and the library are usable, else the reason."""
    try:
        lib = _lib.load()
    except OSError as e:
        return str(e)
    return _lib.take_err(lib.lsh_available())


def _buf(a, dtype):
    """(pointer, lsh_mem, keepalive) of a numpy array or a device tensor."""
    if isinstance(a, np.ndarray):
        a = np.ascontiguousarray(a, dtype=dtype)
        return a.ctypes.data_as(C.c_void_p), LSH_HOST, a
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(a, torch.Tensor):
        if not a.is_cuda:
            return _buf(a.numpy(), dtype)
        tdt = {np.float32: torch.float32, np.float64: torch.float64, np.uint16: torch.uint16}[dtype]
        if a.dtype != tdt:
            raise TypeError(f"device tensor must be {tdt}, got {a.dtype}")
        if not a.is_contiguous():
            raise ValueError("device tensor must be contiguous")
        return C.c_void_p(a.data_ptr()), LSH_DEVICE, a
    raise TypeError(f"unsupported buffer {type(a)}")


class HipFrames:
    """gpu.rs:636-900 GpuFrames: one frame at a time through the device."""

    def __init__(self, w: int, h: int, interp="spline4x4"):
        self._lib = _lib.load()
        self.w, self.h = w, h
        self.interp = Interp.parse(interp)
        h_ = C.c_void_p()
        _lib.check(self._lib.lsh_frames_new(w, h, self.interp, C.byref(h_)))
        self._h = h_

    def _slice_buf(self, measure, slice_out):
        if measure is None:
            return None, LSH_HOST, None
        if slice_out is None:
            dw, dh = measure.grid(self.w, self.h)
            slice_out = np.zeros((dh, dw), np.float32)
        p, mem, keep = _buf(slice_out, np.float32)
        return p, mem, (keep, slice_out)

    def process(self, rgb, guess=SIM_ID, free=(1, 1, 1, 1, 0, 0, 0, 0), coarsen=0, known=None, brightness=True,
                measure: Optional[DepthParams] = None, slice_out=None):
        """Register (against the previous frame), warp, equalize and measure one frame.

        Returns (sim, gains, slice) -- slice is the output buffer (or None)."""
        rp, rmem, rkeep = _buf(rgb, np.uint16)
        g = (C.c_double * 8)(*guess)
        fr = (C.c_uint8 * 8)(*[int(bool(v)) for v in free])
        ks = kg = None
        if known is not None:
            ks = (C.c_double * 8)(*known[0])
            kg = (C.c_float * 3)(*known[1])
        sp, smem, skeep = self._slice_buf(measure, slice_out)
        mp = C.byref(measure._c()) if measure is not None else None
        sim = (C.c_double * 8)()
        gains = (C.c_float * 3)()
        _lib.check(self._lib.lsh_frames_process(self._h, rp, rmem, g, fr, coarsen, ks, kg, int(brightness), mp, sim,
                                                gains, sp, smem))
        del rkeep
        return tuple(sim), tuple(gains), (skeep[1] if skeep else None)

    def ingest(self, rgb, brightness=True, measure: Optional[DepthParams] = None, slice_out=None):
        """stack.rs:361-419 for registered stacks: decode-equivalent planes, gains to frame 0's means."""
        rp, rmem, rkeep = _buf(rgb, np.uint16)
        sp, smem, skeep = self._slice_buf(measure, slice_out)
        mp = C.byref(measure._c()) if measure is not None else None
        gains = (C.c_float * 3)()
        _lib.check(self._lib.lsh_frames_ingest(self._h, rp, rmem, int(brightness), mp, gains, sp, smem))
        del rkeep
        return tuple(gains), (skeep[1] if skeep else None)

    @property
    def last_evals(self) -> int:
        return int(self._lib.lsh_frames_last_evals(self._h))

    def search_levels(self):
        """The last search's cost evaluations per registration level (level 0 = full resolution)."""
        ev = (C.c_uint64 * 16)()
        _lib.check(self._lib.lsh_frames_search_stats(self._h, ev, None, 16))
        return [int(v) for v in ev]

    def trace_costs(self, cap: int):
        """Record the costs of the following searches (up to cap per search; 0: off)."""
        _lib.check(self._lib.lsh_frames_search_trace(self._h, int(cap)))

    def search_costs(self) -> np.ndarray:
        """The costs the last search counted, in the reference's order (tracing on)."""
        n = C.c_uint32()
        _lib.check(self._lib.lsh_frames_search_costs(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        _lib.check(self._lib.lsh_frames_search_costs(self._h, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return out

    def search_blocks(self) -> np.ndarray:
        """Per block that took part in the last search: its digest of every sum it folded."""
        n = C.c_uint32()
        _lib.check(self._lib.lsh_frames_search_blocks(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint64)
        _lib.check(self._lib.lsh_frames_search_blocks(self._h, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return out

    def download(self) -> np.ndarray:
        out = np.zeros((3, self.h, self.w), np.float32)
        _lib.check(self._lib.lsh_frames_download(self._h, out.ctypes.data_as(C.c_void_p), LSH_HOST))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lsh_frames_free(self._h)
            self._h = None

    __del__ = close


class Fuser:
    """fuse.rs:349-521 Fuser on the device (gpu.rs GpuFuser's contract)."""

    def __init__(self, w: int, h: int, params: Optional[FuseParams] = None, **kw):
        self._lib = _lib.load()
        self.params = params or FuseParams(**kw)
        self.w, self.h = w, h
        h_ = C.c_void_p()
        _lib.check(self._lib.lsh_fuser_new(w, h, C.byref(self.params._c()), C.byref(h_)))
        self._h = h_
        lv, dl, dw, dh = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _lib.check(self._lib.lsh_fuser_levels(self._h, C.byref(lv), C.byref(dl), C.byref(dw), C.byref(dh)))
        self.levels, self.depth_level, self.dw, self.dh = lv.value, dl.value, dw.value, dh.value
        self.count = 0

    def push(self, img3):
        """Fold one aligned frame (3 planes in [0, 1])."""
        p, mem, keep = _buf(img3, np.float32)
        shape = tuple(img3.shape)
        if shape != (3, self.h, self.w):
            raise ValueError(f"frame size mismatch: {shape} vs (3, {self.h}, {self.w})")
        _lib.check(self._lib.lsh_fuser_push(self._h, p, mem))
        self.count += 1
        del keep

    def push_frames(self, frames: HipFrames):
        """gpu.rs:451 push_device: the frame the frames handle holds, without a copy."""
        _lib.check(self._lib.lsh_fuser_push_frames(self._h, frames._h))
        self.count += 1

    def shares(self) -> np.ndarray:
        out = np.zeros(max(self.count, 1), np.float32)
        n = C.c_uint32()
        _lib.check(self._lib.lsh_fuser_shares(self._h, out.ctypes.data_as(C.c_void_p), len(out), C.byref(n)))
        return out[: n.value]

    def finish(self, img_out=None, winner_out=None, want_level=False):
        """(image (3,h,w), winner map (h,w) f32[, winner at depth_level u16])."""
        img_out = np.zeros((3, self.h, self.w), np.float32) if img_out is None else img_out
        winner_out = np.zeros((self.h, self.w), np.float32) if winner_out is None else winner_out
        ip, imem, ik = _buf(img_out, np.float32)
        wp, wmem, wk = _buf(winner_out, np.float32)
        wl = np.zeros((self.dh, self.dw), np.uint16) if want_level else None
        _lib.check(self._lib.lsh_fuser_finish(self._h, ip, imem, wp, wmem,
                                              wl.ctypes.data_as(C.c_void_p) if want_level else None))
        del ik, wk
        return (img_out, winner_out, wl) if want_level else (img_out, winner_out)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lsh_fuser_free(self._h)
            self._h = None

    __del__ = close


@dataclasses.dataclass
class DepthMap:
    """depth.rs:140-156, plus the WLS report."""

    depth: object
    conf: object
    w: int
    h: int
    dw: int
    dh: int
    floor: object
    info: dict


def depth_from_slices(slices, n: int, fused, params: Optional[DepthParams] = None, depth_out=None, conf_out=None,
                      floor_out=None) -> DepthMap:
    """depth.rs:1119-1200 on the device.  ``slices``: (n, dh, dw) contiguous, host or device."""
    lib = _lib.load()
    p = params or DepthParams()
    if n <= 0:
        raise HipError("depth: no slices")
    _, h, w = tuple(fused.shape)
    dw, dh = p.grid(w, h)
    if tuple(slices.shape) != (n, dh, dw):
        raise HipError(f"depth: slices have shape {tuple(slices.shape)}, the working grid is {n}x{dh}x{dw}")
    sp, smem, sk = _buf(slices, np.float32)
    fp, fmem, fk = _buf(fused, np.float32)
    if depth_out is None:
        depth_out = np.zeros((h, w), np.float32)
        conf_out = np.zeros((h, w), np.float32)
        floor_out = np.zeros((dh, dw), np.float32)
    dp, dmem, dk = _buf(depth_out, np.float32)
    cp, cmem, ck = _buf(conf_out, np.float32)
    flp = _buf(floor_out, np.float32)[0] if floor_out is not None else None
    info = _lib.DepthInfoC()
    _lib.check(lib.lsh_depth_from_slices(sp, smem, n, fp, fmem, w, h, C.byref(p._c()), dp, cp, dmem, flp,
                                         C.byref(info)))
    del sk, fk, dk, ck
    d = {k: getattr(info, k) for k, _ in info._fields_}
    return DepthMap(depth_out, conf_out, w, h, dw, dh, floor_out, d)


def depth_from_source(slices, n: int, fused, params: Optional[DepthParams] = None, log=None, depth_out=None,
                      conf_out=None, floor_out=None) -> DepthMap:
    """depth.rs:1119-1200 depth_from_slices with its slice iterator and log
    (include/lsh.h lsh_depth_from_source).  ``slices``: any iterable of the n
    focus slices in frame order, each dw*dh float32 cells (numpy, host or
    device tensor); an item that is an Exception is the iterator's Err and
    fails the pass with its message, as does an exception the iterator raises.
    ``log``: a callable taking each progress line, or None.  A device slice is
    read before the next one is asked for, so one buffer may serve them all."""
    lib = _lib.load()
    p = params or DepthParams()
    _, h, w = tuple(fused.shape)
    dw, dh = p.grid(w, h)
    it = iter(slices)
    held = {}
    failed = []

    def next_slice(_user, m, slice_p, cells_p, mem_p):
        try:
            try:
                s = next(it)
            except StopIteration:
                slice_p[0] = None
                return None
            if isinstance(s, BaseException):
                raise s
            ptr, mem, keep = _buf(s, np.float32)
            if mem == LSH_DEVICE:
                import torch

                torch.cuda.current_stream().synchronize()  # the slice's producer has finished
            held["slice"] = keep
            slice_p[0] = ptr.value if isinstance(ptr, C.c_void_p) else ptr
            cells_p[0] = int(np.prod(keep.shape)) if hasattr(keep, "shape") else 0
            mem_p[0] = mem
            return None
        except BaseException as e:  # noqa: BLE001 -- handed to the library as the source's error
            failed.append(e)
            held["msg"] = C.create_string_buffer(str(e).encode() or type(e).__name__.encode())
            return C.addressof(held["msg"])

    logged = []

    def log_line(_user, line):
        try:
            log(line.decode())
        except BaseException as e:  # noqa: BLE001 -- raised once the pass has returned
            logged.append(e)

    nfn = _lib.SLICE_FN(next_slice)
    lfn = _lib.LOG_FN(log_line) if log is not None else _lib.LOG_FN()
    fp, fmem, fk = _buf(fused, np.float32)
    if depth_out is None:
        depth_out = np.zeros((h, w), np.float32)
        conf_out = np.zeros((h, w), np.float32)
        floor_out = np.zeros((dh, dw), np.float32)
    dp, dmem, dk = _buf(depth_out, np.float32)
    cp, cmem, ck = _buf(conf_out, np.float32)
    flp = _buf(floor_out, np.float32)[0] if floor_out is not None else None
    info = _lib.DepthInfoC()
    err = lib.lsh_depth_from_source(nfn, None, n, fp, fmem, w, h, C.byref(p._c()), dp, cp, dmem, flp, C.byref(info),
                                    lfn, None)
    held.clear()
    if err and failed:
        raise HipError(_lib.take_err(err)) from failed[0]
    _lib.check(err)
    if logged:
        raise logged[0]
    del fk, dk, ck
    d = {k: getattr(info, k) for k, _ in info._fields_}
    return DepthMap(depth_out, conf_out, w, h, dw, dh, floor_out, d)


@dataclasses.dataclass
class WavParams:
    """wav.rs:60-80 WavParams (defaults wav.rs:82-86)."""

    power: float = 2.0
    smooth: int = 3
    gate: float = 0.5
    edge: bool = True

    def _c(self):
        return _lib.WavParamsC(float(self.power), int(self.smooth), float(self.gate), int(bool(self.edge)))


class WeightedAverage:
    """wav.rs:97-161 weighted_average as a device accumulator (include/lsh.h
    lsh_wav_*): made with the depth pass's floor and the fused image (the guide),
    fed every aligned frame, finished into the average."""

    def __init__(self, w: int, h: int, floor, fused=None, depth: Optional[DepthParams] = None,
                 params: Optional[WavParams] = None):
        self._lib = _lib.load()
        self.w, self.h = w, h
        self.depth = depth or DepthParams()
        self.params = params or WavParams()
        dw, dh = self.depth.grid(w, h)
        fshape = tuple(floor.shape) if hasattr(floor, "shape") else tuple(np.shape(floor))
        if fshape not in ((dh, dw), (dh * dw,)):
            raise HipError(f"weighted average: the noise floor has shape {fshape}, the grid is {dh}x{dw}")
        fp, fmem, fk = _buf(floor, np.float32)
        gp, gmem, gk = _buf(fused, np.float32) if fused is not None else (None, LSH_HOST, None)
        h_ = C.c_void_p()
        _lib.check(self._lib.lsh_wav_new(w, h, C.byref(self.depth._c()), C.byref(self.params._c()), fp, fmem, gp,
                                         gmem, C.byref(h_)))
        del fk, gk
        self._h = h_
        self.count = 0

    def push(self, img3, slice_=None):
        """One aligned frame ((3, h, w) planes) and, optionally, its focus slice as the depth pass took it."""
        ip, imem, ik = _buf(img3, np.float32)
        sp, smem, sk = _buf(slice_, np.float32) if slice_ is not None else (None, LSH_HOST, None)
        _lib.check(self._lib.lsh_wav_push(self._h, ip, imem, sp, smem))
        del ik, sk
        self.count += 1

    def push_frames(self, frames: HipFrames):
        """The frame a frames handle holds (after process/ingest), without a copy."""
        _lib.check(self._lib.lsh_wav_push_frames(self._h, frames._h))
        self.count += 1

    def finish(self, out=None):
        if out is None:
            out = np.zeros((3, self.h, self.w), np.float32)
        op, omem, ok = _buf(out, np.float32)
        _lib.check(self._lib.lsh_wav_finish(self._h, op, omem))
        del ok
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lsh_wav_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def weighted_average(frames, depth: DepthParams, params: WavParams, floor, guide=None, slices=None, out=None):
    """wav.rs:97 weighted_average(src, dp, wp, floor, guide): the aligned
    frames ((3, h, w) each, host or device) averaged by their contrast above
    the depth pass's noise floor; ``slices`` (the depth pass's focus slices,
    optional) spare measuring them again."""
    frames = list(frames)
    if not frames:
        raise HipError("weighted average: no frames")
    _, h, w = tuple(frames[0].shape)
    acc = WeightedAverage(w, h, floor, guide, depth, params)
    try:
        for i, f in enumerate(frames):
            acc.push(f, slices[i] if slices is not None else None)
        return acc.finish(out)
    finally:
        acc.close()


def slab_ranges(count: int, size: int, overlap: int):
    """stack.rs:76-92 slab_ranges: the (lo, hi) frame ranges of the slabs."""
    lib = _lib.load()
    ns = lib.lsh_slab_ranges(count, size, overlap, None, 0)
    buf = (C.c_uint32 * max(2 * ns, 2))()
    lib.lsh_slab_ranges(count, size, overlap, buf, ns)
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(ns)]


def fuse_slabs(frames, size: int, overlap: int, fuse: Optional[FuseParams] = None, sims=None, gains=None,
               interp="spline4x4", brightness: bool = True):
    """stack.rs:873-893 fuse_slabs: every slab of slab_ranges fused on its own
    over the same frames, returned as [(lo, hi, image, winner)].  Each frame is
    taken once -- registered with its Sim and gains from a run over the whole
    stack when ``sims``/``gains`` are given (a re-pass, no search), as decoded
    otherwise -- and folded into every slab that holds it."""
    frames = list(frames)
    n = len(frames)
    ranges = slab_ranges(n, size, overlap)
    if not ranges:
        return []
    fuse = fuse or FuseParams()
    h, w, _ = tuple(frames[0].shape)
    fr = HipFrames(w, h, interp)
    fusers = [Fuser(w, h, fuse) for _ in ranges]
    try:
        for i, f in enumerate(frames):
            if sims is not None:
                fr.process(f, SIM_ID, (0,) * 8, 0, (sims[i], gains[i]), brightness, None)
            else:
                fr.ingest(f, brightness, None)
            for (lo, hi), fu in zip(ranges, fusers):
                if lo <= i <= hi:
                    fu.push_frames(fr)
        return [(lo, hi) + tuple(fu.finish()) for (lo, hi), fu in zip(ranges, fusers)]
    finally:
        for fu in fusers:
            fu.close()
        fr.close()


def gauss_levels(y):
    """align.rs:741-754 gauss_levels on the device: the registration pyramid's
    levels above the plane y ((h, w) float32, host), as a list of numpy arrays."""
    lib = _lib.load()
    y = np.ascontiguousarray(y, np.float32)
    h, w = y.shape
    n = C.c_uint32()
    _lib.check(lib.lsh_gauss_levels(y.ctypes.data_as(C.c_void_p), LSH_HOST, w, h, None, LSH_HOST, C.byref(n)))
    dims, cw, ch = [], w, h
    for _ in range(n.value):
        cw, ch = (cw + 1) // 2, (ch + 1) // 2
        dims.append((ch, cw))
    packed = np.zeros(max(1, sum(a * b for a, b in dims)), np.float32)
    if dims:
        _lib.check(lib.lsh_gauss_levels(y.ctypes.data_as(C.c_void_p), LSH_HOST, w, h,
                                        packed.ctypes.data_as(C.c_void_p), LSH_HOST, C.byref(n)))
    out, off = [], 0
    for a, b in dims:
        out.append(packed[off:off + a * b].reshape(a, b))
        off += a * b
    return out


def common_area(sims, w: int, h: int, interp="spline4x4"):
    """align.rs:344-403 common_area: (x, y, w, h) of the largest rectangle every frame covers."""
    lib = _lib.load()
    s = np.ascontiguousarray(np.asarray(sims, np.float64).reshape(-1, 8))
    rect = (C.c_uint32 * 4)()
    _lib.check(lib.lsh_common_area(s.ctypes.data_as(C.c_void_p) if len(s) else None, len(s), w, h,
                                   Interp.parse(interp), rect))
    return tuple(int(v) for v in rect)


def _planes_of(img):
    shape = tuple(img.shape)
    return (1,) + shape if len(shape) == 2 else shape


def _like(img, shape):
    import torch

    if isinstance(img, torch.Tensor) and img.is_cuda:
        return torch.empty(shape, dtype=torch.float32, device=img.device)
    return np.zeros(shape, np.float32)


def crop(img, rect, out=None):
    """Img3::crop: the rect (x, y, w, h) of an image ((c, h, w) or (h, w)), host or device."""
    c, h, w = _planes_of(img)
    rx, ry, rw, rh = (int(v) for v in rect)
    shape = (c, rh, rw) if len(tuple(img.shape)) == 3 else (rh, rw)
    if out is None:
        out = _like(img, shape)
    sp, smem, sk = _buf(img, np.float32)
    dp, dmem, dk = _buf(out, np.float32)
    r4 = (C.c_uint32 * 4)(rx, ry, rw, rh)
    _lib.check(_lib.load().lsh_crop(sp, smem, c, w, h, r4, dp, dmem))
    del sk, dk
    return out


def resize(img, w: int, h: int, interp="spline4x4", out=None):
    """prep.rs:46-71 resize / resize_plane on the device: an image ((c, ph, pw) or (ph, pw)) resampled to w x h."""
    c, ph, pw = _planes_of(img)
    shape = (c, h, w) if len(tuple(img.shape)) == 3 else (h, w)
    if out is None:
        out = _like(img, shape)
    sp, smem, sk = _buf(img, np.float32)
    dp, dmem, dk = _buf(out, np.float32)
    _lib.check(_lib.load().lsh_resize(sp, smem, c, pw, ph, w, h, Interp.parse(interp), dp, dmem))
    del sk, dk
    return out


def restretch(image, rect, w: int, h: int, interp="spline4x4"):
    """stack.rs:951-974 restretched, for one image cropped to rect: back to the frame's w x h (each axis on its
    own), or unchanged when the crop kept the whole frame."""
    if rect is None or (int(rect[2]), int(rect[3])) == (w, h):
        return image
    return resize(image, w, h, interp)


def to_u16(img3, out=None):
    """io.rs:283-288: (3, h, w) planes to interleaved RGB u16 (h, w, 3)."""
    _, h, w = tuple(img3.shape)
    if out is None:
        out = np.zeros((h, w, 3), np.uint16)
    sp, smem, sk = _buf(img3, np.float32)
    dp, dmem, dk = _buf(out, np.uint16)
    _lib.check(_lib.load().lsh_to_u16(sp, smem, w, h, dp, dmem))
    del sk, dk
    return out


def depth_grid(slices, n: int, fused, params: Optional[DepthParams] = None, out_device: bool = False):
    """depth.rs:1125-1159 on the device: the depth pass up to the WLS solve.
    Returns (guide, depth, conf, info), each plane (dh, dw) -- the WLS system's
    inputs -- as numpy arrays, or as device tensors with ``out_device``."""
    lib = _lib.load()
    p = params or DepthParams()
    if n <= 0:
        raise HipError("depth: no slices")
    _, h, w = tuple(fused.shape)
    dw, dh = p.grid(w, h)
    if tuple(slices.shape) != (n, dh, dw):
        raise HipError(f"depth: slices have shape {tuple(slices.shape)}, the working grid is {n}x{dh}x{dw}")
    sp, smem, sk = _buf(slices, np.float32)
    fp, fmem, fk = _buf(fused, np.float32)
    if out_device:
        import torch

        outs = [torch.zeros((dh, dw), dtype=torch.float32, device="cuda") for _ in range(3)]
    else:
        outs = [np.zeros((dh, dw), np.float32) for _ in range(3)]
    ptrs = [_buf(o, np.float32) for o in outs]
    info = _lib.DepthInfoC()
    _lib.check(lib.lsh_depth_grid(sp, smem, n, fp, fmem, w, h, C.byref(p._c()), ptrs[0][0], ptrs[1][0], ptrs[2][0],
                                  ptrs[0][1], C.byref(info)))
    del sk, fk
    return outs[0], outs[1], outs[2], {k: getattr(info, k) for k, _ in info._fields_}


class WarmNet:
    """The learned warm start of the WLS solve (include/lsh.h lsh_warm_*): a
    bf16 convolutional encoder-decoder on the grid of 8 x 8 blocks that
    predicts the solution's block means from the system's planes (peak depth,
    confidence, guide) averaged over the blocks; the start is their bilinear
    prolongation.  Run and trained on the device.  Planes are (h, w) or (nimg,
    h, w), numpy or device tensors."""

    def __init__(self, seed: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.lsh_warm_new(seed, C.byref(h)))
        self._h = h
        n = C.c_uint32()
        _lib.check(self._lib.lsh_warm_nparams(h, C.byref(n)))
        self.nparams = n.value

    def get(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        _lib.check(self._lib.lsh_warm_get(self._h, p.ctypes.data_as(C.c_void_p), self.nparams))
        return p

    def set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.nparams:
            raise HipError(f"warm: {p.size} parameters, the network has {self.nparams}")
        _lib.check(self._lib.lsh_warm_set(self._h, p.ctypes.data_as(C.c_void_p), self.nparams))

    def set_sigma(self, sigma_c: float):
        """The guide scale of the WLS systems served (edge conductances
        exp(-|dg| / sigma_c) are among the inputs); 0.04 by default."""
        _lib.check(self._lib.lsh_warm_set_sigma(self._h, float(sigma_c)))

    def save(self, path: str):
        self.get().tofile(path)

    def load(self, path: str):
        self.set(np.fromfile(path, np.float32))

    @staticmethod
    def _shape(a):
        s = tuple(a.shape)
        return (1,) + s if len(s) == 2 else s

    def _in(self, *planes):
        shape = self._shape(planes[0])
        bufs = [_buf(a, np.float32) for a in planes]
        if any(self._shape(a) != shape for a in planes) or len({b[1] for b in bufs}) != 1:
            raise HipError("warm: planes must share one shape and one memory")
        return shape, bufs

    def predict(self, d, conf, guide, dscale: float, out=None):
        (nimg, h, w), bufs = self._in(d, conf, guide)
        if out is None:
            out = np.zeros(tuple(d.shape), np.float32)
        op, omem, ok = _buf(out, np.float32)
        _lib.check(self._lib.lsh_warm_predict(self._h, bufs[0][0], bufs[1][0], bufs[2][0], bufs[0][1], nimg, w, h,
                                              float(dscale), op, omem))
        del bufs, ok
        return out

    def grads(self, d, conf, guide, target, dscale: float):
        (nimg, h, w), bufs = self._in(d, conf, guide, target)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        _lib.check(self._lib.lsh_warm_grads(self._h, *[b[0] for b in bufs], bufs[0][1], nimg, w, h, float(dscale),
                                            C.byref(loss), g.ctypes.data_as(C.c_void_p), self.nparams))
        return loss.value, g

    def train_step(self, d, conf, guide, target, dscale: float, lr: float = 1e-3) -> float:
        (nimg, h, w), bufs = self._in(d, conf, guide, target)
        loss = C.c_double()
        _lib.check(self._lib.lsh_warm_train_step(self._h, *[b[0] for b in bufs], bufs[0][1], nimg, w, h,
                                                 float(dscale), float(lr), C.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lsh_warm_free(self._h)
            self._h = None

    __del__ = close


WLS_MODES = {"reference": 0, "refined": 1, "fp64": 2}


def wls_solve(guide, conf, d, lam: float, sigma_c: float, mode: str = "reference", tol: float = 1e-5,
              max_iters: int = 200, warm=None, want_err: bool = False, u_out=None, u64_out=None):
    """depth.rs:1081 wls_solve on the device: solve (W + lam L) u = W d with W =
    max(conf, 0) + 1e-4 and L the guide's edge-weighted Laplacian.  ``mode``:
    "reference" (depth.rs:1006-1077, FP32), "refined" (FP32 MG-PCG inside FP64
    iterative refinement) or "fp64" (the FP64 oracle).  Planes are (h, w),
    numpy or device tensors (all in one memory).  Returns (u, report, u64);
    u64 is the FP64 iterate ("refined" and "fp64", else None)."""
    lib = _lib.load()
    if mode not in WLS_MODES:
        raise ValueError(f"unknown WLS mode {mode!r}")
    h, w = tuple(guide.shape)
    for name, a in (("conf", conf), ("d", d), ("warm", warm)):
        if a is not None and tuple(a.shape) != (h, w):
            raise HipError(f"wls: {name} has shape {tuple(a.shape)}, the grid is {h}x{w}")
    gp, gmem, gk = _buf(guide, np.float32)
    cp, cmem, ck = _buf(conf, np.float32)
    dp, dmem, dk = _buf(d, np.float32)
    wp, wk = None, None
    if warm is not None:
        wp, wmem, wk = _buf(warm, np.float32)
        if wmem != gmem:
            raise HipError("wls: inputs must share one memory")
    if not (gmem == cmem == dmem):
        raise HipError("wls: inputs must share one memory")
    if u_out is None:
        u_out = np.zeros((h, w), np.float32)
        if mode != "reference" and u64_out is None:
            u64_out = np.zeros((h, w), np.float64)
    up, umem, uk = _buf(u_out, np.float32)
    u64p = None
    if u64_out is not None:
        u64p, u64mem, u64k = _buf(u64_out, np.float64)
        if u64mem != umem:
            raise HipError("wls: outputs must share one memory")
    rep = _lib.WlsReportC()
    _lib.check(lib.lsh_wls_solve(gp, cp, dp, wp, gmem, w, h, lam, sigma_c, WLS_MODES[mode], tol, max_iters,
                                 int(want_err), up, u64p, umem, C.byref(rep)))
    del gk, ck, dk, wk, uk
    return u_out, {k: getattr(rep, k) for k, _ in rep._fields_}, u64_out


def wls_solve_noalloc(guide, conf, d, lam: float, sigma_c: float, tol: float = 1e-5, max_iters: int = 200, warm=None,
                      u_out=None):
    """wls_solve in the "reference" mode with no device allocation
    (include/lsh.h lsh_wls_solve_noalloc): every buffer of the solver in the
    library's per-GPU scratch array (grids of up to about 230 000 cells).
    Planes are (h, w) float32 device tensors; returns (u, report)."""
    import torch

    lib = _lib.load()
    h, w = tuple(guide.shape)
    planes = [("guide", guide), ("conf", conf), ("d", d)] + ([("warm", warm)] if warm is not None else [])
    for name, a in planes:
        if not (isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float32 and a.is_contiguous()):
            raise HipError(f"wls no-alloc: {name} must be a contiguous float32 device tensor")
        if tuple(a.shape) != (h, w):
            raise HipError(f"wls no-alloc: {name} has shape {tuple(a.shape)}, the grid is {h}x{w}")
    if u_out is None:
        u_out = torch.empty((h, w), dtype=torch.float32, device=guide.device)
    rep = _lib.WlsReportC()
    _lib.check(lib.lsh_wls_solve_noalloc(C.c_void_p(guide.data_ptr()), C.c_void_p(conf.data_ptr()),
                                         C.c_void_p(d.data_ptr()),
                                         C.c_void_p(warm.data_ptr()) if warm is not None else None, w, h, lam,
                                         sigma_c, tol, max_iters, C.c_void_p(u_out.data_ptr()), C.byref(rep)))
    return u_out, {k: getattr(rep, k) for k, _ in rep._fields_}


def stack_stream(frames: Sequence, fuse: Optional[FuseParams] = None, align: Optional[AlignParams] = None,
                 depth: Optional[DepthParams] = None, brightness: bool = True, graph: bool = True, slices_out=None,
                 img_out=None, winner_out=None):
    """stack.rs:740-826 over a whole stack with no host round trip per frame
    (include/lsh.h lsh_stack_stream): frame 0 through the ordinary path, every
    later frame through the device-only step, captured once as a hipGraph and
    replayed per frame (``graph``), then the collapse.  ``frames``: (h, w, 3)
    uint16 frames, all numpy or all device tensors.  Returns a dict with image,
    winner, sims, gains, evals and slices (when ``depth`` is given)."""
    lib = _lib.load()
    fuse = fuse or FuseParams()
    align = align or AlignParams()
    n = len(frames)
    if n == 0:
        raise HipError("no input images")
    h, w, _ = tuple(frames[0].shape)
    fr = HipFrames(w, h, align.interp)
    fu = Fuser(w, h, fuse)
    bufs = [_buf(f, np.uint16) for f in frames]
    if len({b[1] for b in bufs}) != 1:
        raise HipError("stream: frames must share one memory")
    ptrs = (C.c_void_p * n)(*[b[0].value if isinstance(b[0], C.c_void_p) else b[0] for b in bufs])
    free = (C.c_uint8 * 8)(*align.free())
    dp = depth._c() if depth is not None else None
    sp, smem, sk = None, LSH_HOST, None
    if depth is not None:
        dw, dh = depth.grid(w, h)
        if slices_out is None:
            slices_out = np.zeros((n, dh, dw), np.float32)
        sp, smem, sk = _buf(slices_out, np.float32)
    sims = np.zeros((n, 8), np.float64)
    gains = np.zeros((n, 3), np.float32)
    evals = np.zeros(n, np.uint64)
    _lib.check(lib.lsh_stack_stream(fr._h, fu._h, ptrs, n, bufs[0][1], free, align.coarsen, int(brightness),
                                    C.byref(dp) if dp is not None else None, sp, smem, int(graph),
                                    sims.ctypes.data_as(C.c_void_p), gains.ctypes.data_as(C.c_void_p),
                                    evals.ctypes.data_as(C.c_void_p)))
    image, winner = fu.finish(img_out, winner_out)
    fr.close()
    fu.close()
    del bufs, sk
    return dict(image=image, winner=winner, sims=sims, gains=gains, evals=evals, slices=slices_out)


def image_info(path: str):
    """(w, h, bits) of a PNG or TIFF file."""
    lib = _lib.load()
    w, h, b = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _lib.check(lib.lsh_image_info(path.encode(), C.byref(w), C.byref(h), C.byref(b)))
    return w.value, h.value, b.value


def load_rgb(path: str, out=None):
    """io.rs:62-97 for PNG and TIFF: the frame as interleaved RGB u16 (h, w, 3)
    (8-bit samples v as v*257) and the file's bits per sample.  ``out``: a
    numpy uint16 array of that shape (or a pinned host tensor's numpy view)."""
    lib = _lib.load()
    w, h, b = image_info(path)
    if out is None:
        out = np.empty((h, w, 3), np.uint16)
    if out.dtype != np.uint16 or out.size < w * h * 3 or not out.flags["C_CONTIGUOUS"]:
        raise HipError(f"load_rgb: the output must be contiguous uint16 with room for {h}x{w}x3")
    ww, hh, bb = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _lib.check(lib.lsh_image_load(path.encode(), out.ctypes.data_as(C.c_void_p), out.size, C.byref(ww), C.byref(hh),
                                  C.byref(bb)))
    return out.reshape(-1)[:w * h * 3].reshape(h, w, 3), bb.value


def save_rgb(img3, path: str, bits: int = 16):
    """io.rs:273-315 save_rgb: (3, h, w) planes, host or device, as a 16- or 8-bit TIFF or PNG."""
    lib = _lib.load()
    _, h, w = tuple(img3.shape)
    p, mem, keep = _buf(img3, np.float32)
    _lib.check(lib.lsh_image_save(path.encode(), p, mem, w, h, bits))
    del keep


def save_rgb_u16(rgb, path: str):
    """Interleaved RGB u16 (h, w, 3), host or device, as a 16-bit TIFF or PNG."""
    lib = _lib.load()
    h, w, _ = tuple(rgb.shape)
    p, mem, keep = _buf(rgb, np.uint16)
    _lib.check(lib.lsh_image_save_u16(path.encode(), p, mem, w, h))
    del keep


def save_gray16(plane, path: str, scale: float):
    """io.rs:325-335: a plane in [0, scale] as a 16-bit gray PNG with the fixed mapping scale -> 65535."""
    lib = _lib.load()
    h, w = tuple(plane.shape)
    p, mem, keep = _buf(plane, np.float32)
    _lib.check(lib.lsh_save_gray16(path.encode(), p, mem, w, h, scale))
    del keep


def save_gray(plane, path: str):
    """io.rs:337-346: a plane min-max normalised to an 8-bit gray PNG."""
    lib = _lib.load()
    h, w = tuple(plane.shape)
    p, mem, keep = _buf(plane, np.float32)
    _lib.check(lib.lsh_save_gray(path.encode(), p, mem, w, h))
    del keep


def stack_stream_files(paths: Sequence[str], fuse: Optional[FuseParams] = None, align: Optional[AlignParams] = None,
                       depth: Optional[DepthParams] = None, brightness: bool = True, graph: bool = True,
                       threads: int = 8, slices_out=None, img_out=None, winner_out=None):
    """stack_stream over image files: decoded by ``threads`` host threads ahead
    of their use into pinned buffers and uploaded beside the device's work
    (include/lsh.h lsh_stack_stream_files).  Returns what stack_stream returns."""
    lib = _lib.load()
    fuse = fuse or FuseParams()
    align = align or AlignParams()
    n = len(paths)
    if n == 0:
        raise HipError("no input images")
    w, h, _ = image_info(paths[0])
    fr = HipFrames(w, h, align.interp)
    fu = Fuser(w, h, fuse)
    cpaths = (C.c_char_p * n)(*[p.encode() for p in paths])
    free = (C.c_uint8 * 8)(*align.free())
    dp = depth._c() if depth is not None else None
    sp, smem, sk = None, LSH_HOST, None
    if depth is not None:
        dw, dh = depth.grid(w, h)
        if slices_out is None:
            slices_out = np.zeros((n, dh, dw), np.float32)
        sp, smem, sk = _buf(slices_out, np.float32)
    sims = np.zeros((n, 8), np.float64)
    gains = np.zeros((n, 3), np.float32)
    evals = np.zeros(n, np.uint64)
    try:
        _lib.check(lib.lsh_stack_stream_files(fr._h, fu._h, cpaths, n, threads, free, align.coarsen, int(brightness),
                                              C.byref(dp) if dp is not None else None, sp, smem, int(graph),
                                              sims.ctypes.data_as(C.c_void_p), gains.ctypes.data_as(C.c_void_p),
                                              evals.ctypes.data_as(C.c_void_p)))
        image, winner = fu.finish(img_out, winner_out)
    finally:
        fr.close()
        fu.close()
    del sk
    return dict(image=image, winner=winner, sims=sims, gains=gains, evals=evals, slices=slices_out)


def batch_fuse(stacks: Sequence[Sequence], levels: Optional[Sequence[int]] = None, fuse: Optional[FuseParams] = None,
               tune: int = 1, tile_h: int = 0, blocks_per_cu: int = 0, frames_per_round: int = 0):
    """Many stacks of uneven sizes and frame counts through the fuse path in
    one persistent launch (include/lsh.h lsh_batch_fuse).  ``stacks``: lists
    of (h, w, 3) uint16 device tensors; ``levels``: pyramid levels per stack
    (None: ``fuse.levels`` for all).  ``tune``: 0 the configuration given by
    ``tile_h``, ``blocks_per_cu`` and ``frames_per_round`` (0: defaults), 1 the
    autotuner (cached per shape class), 2 the autotuner again.  Returns
    ([(image, winner)] per stack as numpy arrays, report dict)."""
    import torch

    lib = _lib.load()
    fuse = fuse or FuseParams()
    k = len(stacks)
    if k == 0:
        raise HipError("batch: no stacks")
    fusers, descs, keep = [], (_lib.BatchStackC * k)(), []
    try:
        for i, frames in enumerate(stacks):
            if len(frames) == 0:
                raise HipError("batch: a stack without frames")
            h, w, _ = tuple(frames[0].shape)
            for f in frames:
                if not (isinstance(f, torch.Tensor) and f.is_cuda and f.dtype == torch.uint16 and f.is_contiguous()):
                    raise HipError("batch: frames must be contiguous uint16 device tensors")
                if tuple(f.shape) != (h, w, 3):
                    raise HipError("batch: a stack's frames must share one size")
            fp = dataclasses.replace(fuse, levels=levels[i]) if levels is not None else fuse
            fusers.append(Fuser(w, h, fp))
            ptrs = (C.c_void_p * len(frames))(*[f.data_ptr() for f in frames])
            keep.append((frames, ptrs))
            descs[i].w, descs[i].h, descs[i].n = w, h, len(frames)
            descs[i].frames = C.cast(ptrs, C.c_void_p)
        handles = (C.c_void_p * k)(*[fu._h.value for fu in fusers])
        opt = _lib.BatchOptionsC(int(tune), int(tile_h), int(blocks_per_cu), int(frames_per_round))
        rep = _lib.BatchReportC()
        _lib.check(lib.lsh_batch_fuse(handles, descs, k, C.byref(opt), C.byref(rep)))
        out = [fu.finish() for fu in fusers]
    finally:
        for fu in fusers:
            fu.close()
        del keep
    r = {name: getattr(rep, name) for name, _ in rep._fields_ if name not in ("tried_cfg", "tried_us_per_mpf")}
    r["tried"] = [{"tile_h": c & 0xFF, "blocks_per_cu": (c >> 8) & 0xFF, "frames_per_round": c >> 16,
                   "us_per_mpix_frame": round(float(u), 4)}
                  for c, u in zip(list(rep.tried_cfg)[:rep.ntried], list(rep.tried_us_per_mpf)[:rep.ntried])]
    return out, r


def fuse_and_depth(frames: Sequence, fuse: Optional[FuseParams] = None, align: Optional[AlignParams] = None,
                   depth: Optional[DepthParams] = None, brightness: bool = True, log=None):
    """stack.rs:740-853 on the device: the per-frame loop (register against the
    previous aligned frame, warp, equalize, fold, measure), the collapse, then
    the depth pass.  ``frames``: a sequence of interleaved RGB u16 frames
    (h, w, 3), numpy or device tensors.  Returns a dict like stack.rs Output."""
    fuse = fuse or FuseParams()
    n = len(frames)
    if n == 0:
        raise HipError("no input images")
    h, w, _ = tuple(frames[0].shape)
    fr = HipFrames(w, h, align.interp if align else "spline4x4")
    fu = Fuser(w, h, fuse)
    sims, gains, slices = [], [], None
    if depth is not None:
        dw, dh = depth.grid(w, h)
        slices = np.zeros((n, dh, dw), np.float32)
    aligned = align is not None and n > 1
    last = SIM_ID
    for i, f in enumerate(frames):
        so = slices[i] if slices is not None else None
        if aligned:
            sim, g, _ = fr.process(f, last, align.free(), align.coarsen, None, brightness, depth, so)
            last = sim
        else:
            g, _ = fr.ingest(f, brightness, depth, so)
            sim = SIM_ID
        sims.append(sim)
        gains.append(g)
        fu.push_frames(fr)
        if log:
            log(f"  frame {i + 1:>3}/{n} folded")
    shares = fu.shares()
    image, winner = fu.finish()
    out = dict(image=image, winner=winner, sims=np.array(sims), gains=np.array(gains), shares=shares,
               levels=fu.levels, slices=slices)
    if depth is not None:
        dm = depth_from_slices(slices, n, image, depth)
        out.update(depth=dm.depth, conf=dm.conf, floor=dm.floor, info=dm.info)
    fr.close()
    fu.close()
    return out
