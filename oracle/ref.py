"""ctypes binding of the CPU oracle (oracle/liblsh_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP backend is compared against and
the CPU baseline bench.py times beside it.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product package
(lapstack_amd) never does.

Arrays are numpy, row-major; an RGB image is a (3, h, w) float32 array of
planes, the layout of the reference's ``Img3`` (crates/lapstack-core/src/
pyramid.rs:18-24).  A Sim is 8 float64: xoff, yoff, scale, rot, aspect, shear,
px, py (align.rs:199-253).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liblsh_oracle.so")

SIM_ID = np.array([0, 0, 1, 0, 1, 0, 0, 0], dtype=np.float64)

INTERP = {"nearest": 0, "bilinear": 1, "bicubic": 2, "spline4x4": 3, "spline6x6": 4, "lanczos3": 5}
TOP_RULE = {"dev-entropy": 0, "deviation": 1, "average": 2}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def _load():
    if not os.path.exists(_SO):
        build()
    return C.CDLL(_SO)


_L = _load()
_p = C.c_void_p
_sz = C.c_size_t


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _sig(name, res, *args):
    fn = getattr(_L, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


class FuseParams(C.Structure):
    """fuse.rs:69-107 FuseParams (levels <= 0: auto_levels(w, h, 32))."""

    _fields_ = [("levels", C.c_int32), ("energy_radius", C.c_uint32), ("top_rule", C.c_uint32),
                ("top_radius", C.c_uint32), ("entropy_bins", C.c_uint32), ("use_chroma", C.c_uint32),
                ("depth_level", C.c_uint32), ("halo", C.c_float)]

    @classmethod
    def make(cls, levels=None, energy_radius=1, top_rule="dev-entropy", top_radius=2, entropy_bins=256,
             use_chroma=False, depth_level=2, halo=0.0):
        return cls(0 if levels is None else int(levels), energy_radius, TOP_RULE.get(top_rule, top_rule), top_radius,
                   entropy_bins, int(use_chroma), depth_level, halo)


class DepthParams(C.Structure):
    """depth.rs:68-109 DepthParams plus the PCG tolerance (depth.rs:1054 fixes 1e-5)."""

    _fields_ = [("scale", C.c_uint32), ("focus_kind", C.c_uint32), ("r_in", C.c_uint32), ("r_out", C.c_uint32),
                ("sml_step", C.c_uint32), ("agg_radius", C.c_uint32), ("agg_eps", C.c_float),
                ("lambda_", C.c_float), ("sigma_c", C.c_float), ("cg_iters", C.c_uint32), ("median", C.c_uint32),
                ("upsample_kind", C.c_uint32), ("up_radius", C.c_uint32), ("up_eps", C.c_float),
                ("gate", C.c_float), ("robust", C.c_float), ("tol", C.c_float)]

    @classmethod
    def make(cls, scale=1, focus="rdf", r_in=1, r_out=3, sml_step=1, agg_radius=3, agg_eps=1e-4, lam=3.0,
             sigma_c=0.04, cg_iters=200, median=True, upsample="guided", up_radius=2, up_eps=1e-2, gate=1.0,
             robust=1.0, tol=1e-5):
        return cls(scale, 0 if focus == "rdf" else 1, r_in, r_out, sml_step, agg_radius, agg_eps, lam, sigma_c,
                   cg_iters, int(median), 1 if upsample == "guided" else 0, up_radius, up_eps, gate, robust, tol)


class StackParams(C.Structure):
    _fields_ = [("align", C.c_uint32), ("free8", C.c_uint8 * 8), ("coarsen", C.c_uint32), ("interp", C.c_uint32),
                ("brightness", C.c_uint32), ("depth", C.c_uint32)]

    @classmethod
    def make(cls, align=False, free=(1, 1, 1, 1, 0, 0, 0, 0), coarsen=0, interp="spline4x4", brightness=True,
             depth=True):
        return cls(int(align), (C.c_uint8 * 8)(*free), coarsen, INTERP.get(interp, interp), int(brightness),
                   int(depth))


class DepthOut(C.Structure):
    _fields_ = [("dw", C.c_uint32), ("dh", C.c_uint32), ("iters1", C.c_uint32), ("iters2", C.c_uint32),
                ("rel1", C.c_float), ("rel2", C.c_float), ("p90", C.c_float), ("floor", C.c_float),
                ("agg_s", C.c_double), ("peaks_s", C.c_double), ("wls_s", C.c_double), ("up_s", C.c_double)]


class StackTimes(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("register_s", "warp_s", "fold_s", "slice_s", "finish_s", "depth_s",
                                          "total_s")]


COST_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)

_sig("ora_reflect", _sz, C.c_long, _sz)
_sig("ora_reduce", None, _p, _sz, _sz, _p)
_sig("ora_expand", None, _p, _sz, _sz, _sz, _sz, _p)
_sig("ora_auto_levels", _sz, _sz, _sz, _sz)
_sig("ora_pyr_size", _sz, _sz, _sz, _sz)
_sig("ora_pyr_build", None, _p, _sz, _sz, _sz, _p)
_sig("ora_pyr_collapse", None, _p, _sz, _sz, _sz, _p)
_sig("ora_binomial", None, _sz, _p)
_sig("ora_window_sum", None, _p, _sz, _sz, _p, _sz, _p)
_sig("ora_region_energy", None, _p, _sz, _sz, _sz, C.c_int, _p)
_sig("ora_deviation_entropy", None, _p, _sz, _sz, _sz, _sz, _p, _p)
_sig("ora_fuse_residuals", None, _p, _sz, _sz, _sz, C.POINTER(FuseParams), _p)
_sig("ora_winner_shares", None, _p, _p, _sz, _sz, _p)
_sig("ora_halo_weight", C.c_float, C.c_float, C.c_float)
_sig("ora_fuser_new", _p, _sz, _sz, C.POINTER(FuseParams))
_sig("ora_fuser_levels", _sz, _p)
_sig("ora_fuser_level_dims", None, _p, _sz, C.POINTER(_sz), C.POINTER(_sz))
_sig("ora_fuser_push", None, _p, _p)
_sig("ora_fuser_shares", _sz, _p, _p)
_sig("ora_fuser_finish", None, _p, _p, _p, _p)
_sig("ora_fuser_free", None, _p)
_sig("ora_sim_matrix", None, _p, _sz, _sz, _p)
_sig("ora_inverse", None, _p, _p)
_sig("ora_map", None, _p, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("ora_luma", None, _p, _sz, _p)
_sig("ora_warp_plane", None, _p, _sz, _sz, _p, _sz, _sz, C.c_uint32, _p, _p)
_sig("ora_warp_cost", C.c_double, _p, _sz, _p, _sz, _sz, _p)
_sig("ora_reduce_burt", None, _p, _sz, _sz, _p)
_sig("ora_gauss_count", _sz, _sz, _sz, _p)
_sig("ora_gauss_levels", None, _p, _sz, _sz, _p)
_sig("ora_level_steps", None, _p, _sz, _sz, _sz, _sz, _p, _p)
_sig("ora_nelder_mead", None, COST_FN, _p, _sz, _p, _p, _p, _p, _p, _p, C.POINTER(C.c_uint32))
_sig("ora_multiscale_align", None, _p, _p, _sz, _sz, _p, _p, _sz, _p, C.POINTER(C.c_uint32))
_sig("ora_multiscale_align_trace", None, _p, _p, _sz, _sz, _p, _p, _sz, _p, C.POINTER(C.c_uint32), _p, _p, _sz,
     C.POINTER(C.c_size_t))
_sig("ora_sincos", None, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("ora_huber_weights", None, _p, _p, _p, _sz, C.c_float, _p)
_sig("ora_depth_grid", C.c_int, _p, _sz, _p, _sz, _sz, C.POINTER(DepthParams), _p, _p, _p, C.POINTER(C.c_float))
_sig("ora_common_area", None, _p, _sz, _sz, _sz, C.c_uint32, _p)
_sig("ora_bright_grid", None, _p, _sz, _sz, _p)
_sig("ora_bright_gains", None, _p, _p, _sz, _sz, _p, _p)
_sig("ora_bright_ratio", None, _p, _p, _sz, _p)
_sig("ora_bright_means", None, _p, _sz, _sz, _p)
_sig("ora_bright_gains_to", None, _p, _p, _sz, _sz, _p)
_sig("ora_bright_apply", None, _p, _sz, _p)
_sig("ora_block_mean", None, _p, _sz, _sz, _sz, _p)
_sig("ora_upsample_bilinear", None, _p, _sz, _sz, _sz, _sz, _sz, _p)
_sig("ora_box_mean", None, _p, _sz, _sz, _sz, _p)
_sig("ora_guided_filter", None, _p, _p, _sz, _sz, _sz, C.c_float, _p)
_sig("ora_guided_coeffs", None, _p, _p, _sz, _sz, _sz, C.c_float, _p, _p)
_sig("ora_rdf_taps", _sz, _sz, _sz, _p, _p, _p)
_sig("ora_focus_measure", None, _p, _sz, _sz, C.POINTER(DepthParams), _p)
_sig("ora_focus_slice", None, _p, _sz, _sz, C.POINTER(DepthParams), _p)
_sig("ora_peaks_new", _p, _sz)
_sig("ora_peaks_push", None, _p, _p)
_sig("ora_peaks_finish", None, _p, C.c_float, _p, _p, _p, C.POINTER(C.c_float))
_sig("ora_peaks_free", None, _p)
_sig("ora_normalize_conf", C.c_float, _p, _sz)
_sig("ora_median3", None, _p, _sz, _sz, _p)
_sig("ora_solve_rows", None, _p, _p, _p, _p, _sz, _sz, C.c_float)
_sig("ora_solve_cols", None, _p, _p, _p, _p, _sz, _sz, C.c_float)
_sig("ora_wls_new", _p, _p, _sz, _sz, C.c_float, C.c_float)
_sig("ora_wls_levels", _sz, _p)
_sig("ora_wls_set_weights", None, _p, _p)
_sig("ora_wls_solve", C.c_uint32, _p, _p, C.c_uint32, C.c_float, _p, C.POINTER(C.c_float))
_sig("ora_wls_vcycle", None, _p, _p, _p)
_sig("ora_wls_apply", None, _p, _p, _p)
_sig("ora_wls_free", None, _p)
_sig("ora_wav_weight", C.c_float, C.c_float, C.c_float, C.c_float, C.c_float)
_sig("ora_weighted_average", None, _p, _sz, _sz, _sz, C.POINTER(DepthParams), C.c_float, _sz, C.c_float, C.c_int, _p, _p,
     _p)
_sig("ora_slab_ranges", _sz, _sz, _sz, _sz, _p, _sz)
_sig("ora_depth_from_slices", C.c_int, _p, _sz, _p, _sz, _sz, C.POINTER(DepthParams), _p, _p, _p,
     C.POINTER(DepthOut))
_sig("ora_stack_run", C.c_int, _p, _sz, _sz, _sz, C.POINTER(FuseParams), C.POINTER(StackParams),
     C.POINTER(DepthParams), _p, _p, _p, _p, _p, _p, _p, C.POINTER(DepthOut), C.POINTER(StackTimes), _p)
_sig("ora_u16_to_planes", None, _p, _sz, _p)
_sig("ora_planes_to_u16", None, _p, _sz, _p)


def half(n):
    return (n + 1) // 2


# ---------------------------------------------------------------- pyramid.rs
def reflect(i, n):
    return _L.ora_reflect(i, n)


def reduce(src):
    src = _f32(src)
    h, w = src.shape
    out = np.zeros((half(h), half(w)), np.float32)
    _L.ora_reduce(_ptr(src), w, h, _ptr(out))
    return out


def expand(coarse, ow, oh):
    coarse = _f32(coarse)
    ch, cw = coarse.shape
    out = np.zeros((oh, ow), np.float32)
    _L.ora_expand(_ptr(coarse), cw, ch, ow, oh, _ptr(out))
    return out


def auto_levels(w, h, min_top=32):
    return _L.ora_auto_levels(w, h, min_top)


def level_dims(w, h, levels):
    dims = []
    for _ in range(levels + 1):
        dims.append((w, h))
        w, h = half(w), half(h)
    return dims


def pyr_build(img3, levels):
    img3 = _f32(img3)
    _, h, w = img3.shape
    packed = np.zeros(_L.ora_pyr_size(w, h, levels), np.float32)
    _L.ora_pyr_build(_ptr(img3), w, h, levels, _ptr(packed))
    out, off = [], 0
    for (lw, lh) in level_dims(w, h, levels):
        out.append(packed[off:off + 3 * lw * lh].reshape(3, lh, lw))
        off += 3 * lw * lh
    return out


def pyr_collapse(pyr):
    w, h = pyr[0].shape[2], pyr[0].shape[1]
    packed = np.concatenate([_f32(p).ravel() for p in pyr])
    out = np.zeros((3, h, w), np.float32)
    _L.ora_pyr_collapse(_ptr(packed), w, h, len(pyr) - 1, _ptr(out))
    return out


# ------------------------------------------------------------------- fuse.rs
def binomial(r):
    out = np.zeros(2 * r + 1, np.float32)
    _L.ora_binomial(r, _ptr(out))
    return out


def window_sum(src, wt):
    src, wt = _f32(src), _f32(wt)
    h, w = src.shape
    out = np.zeros_like(src)
    _L.ora_window_sum(_ptr(src), w, h, _ptr(wt), len(wt), _ptr(out))
    return out


def region_energy(l3, radius=1, use_chroma=False):
    l3 = _f32(l3)
    _, h, w = l3.shape
    out = np.zeros((h, w), np.float32)
    _L.ora_region_energy(_ptr(l3), w, h, radius, int(use_chroma), _ptr(out))
    return out


def deviation_entropy(y, radius=2, bins=256):
    y = _f32(y)
    h, w = y.shape
    dev, ent = np.zeros_like(y), np.zeros_like(y)
    _L.ora_deviation_entropy(_ptr(y), w, h, radius, bins, _ptr(dev), _ptr(ent))
    return dev, ent


def fuse_residuals(tops, params=None):
    tops = _f32(np.stack(tops))
    n, _, h, w = tops.shape
    p = params or FuseParams.make()
    out = np.zeros((3, h, w), np.float32)
    _L.ora_fuse_residuals(_ptr(tops), n, w, h, C.byref(p), _ptr(out))
    return out


def winner_shares(win, best, frames):
    win = np.ascontiguousarray(win, np.uint16)
    best = _f32(best)
    out = np.zeros(frames, np.float32)
    _L.ora_winner_shares(_ptr(win), _ptr(best), win.size, frames, _ptr(out))
    return out


def halo_weight(re, p):
    return _L.ora_halo_weight(re, p)


class Fuser:
    """fuse.rs:349-521 Fuser, restated."""

    def __init__(self, w, h, params=None, **kw):
        self.params = params or FuseParams.make(**kw)
        self.w, self.h = w, h
        self._f = _L.ora_fuser_new(w, h, C.byref(self.params))
        self.levels = _L.ora_fuser_levels(self._f)
        self.count = 0

    def level_dims(self, l):
        a, b = _sz(), _sz()
        _L.ora_fuser_level_dims(self._f, l, C.byref(a), C.byref(b))
        return a.value, b.value

    @property
    def depth_level(self):
        return min(self.params.depth_level, self.levels - 1)

    def push(self, img3):
        img3 = _f32(img3)
        assert img3.shape == (3, self.h, self.w), "frame size mismatch"
        _L.ora_fuser_push(self._f, _ptr(img3))
        self.count += 1

    def shares(self):
        out = np.zeros(max(self.count, 1), np.float32)
        n = _L.ora_fuser_shares(self._f, _ptr(out))
        return out[:n]

    def finish(self):
        img = np.zeros((3, self.h, self.w), np.float32)
        win = np.zeros((self.h, self.w), np.float32)
        dw, dh = self.level_dims(self.depth_level)
        wl = np.zeros((dh, dw), np.uint16)
        _L.ora_fuser_finish(self._f, _ptr(img), _ptr(win), _ptr(wl))
        return img, win, wl

    def __del__(self):
        if getattr(self, "_f", None):
            _L.ora_fuser_free(self._f)
            self._f = None


# ------------------------------------------------------------------ align.rs
def sim_vec(xoff=0.0, yoff=0.0, scale=1.0, rot=0.0, aspect=1.0, shear=0.0, px=0.0, py=0.0):
    return np.array([xoff, yoff, scale, rot, aspect, shear, px, py], np.float64)


def sim_matrix(sim, w, h):
    sim = _f64(sim)
    m = np.zeros(9, np.float64)
    _L.ora_sim_matrix(_ptr(sim), w, h, _ptr(m))
    return m.reshape(3, 3)


def inverse(m):
    m = _f64(m).ravel()
    out = np.zeros(9, np.float64)
    _L.ora_inverse(_ptr(m), _ptr(out))
    return out.reshape(3, 3)


def map_point(inv, x, y):
    inv = _f64(inv).ravel()
    sx, sy = C.c_double(), C.c_double()
    _L.ora_map(_ptr(inv), x, y, C.byref(sx), C.byref(sy))
    return sx.value, sy.value


def luma(img3):
    img3 = _f32(img3)
    _, h, w = img3.shape
    out = np.zeros((h, w), np.float32)
    _L.ora_luma(_ptr(img3), w * h, _ptr(out))
    return out


def warp_plane(src, sim, interp="spline4x4", ow=None, oh=None):
    src, sim = _f32(src), _f64(sim)
    h, w = src.shape
    ow, oh = ow or w, oh or h
    out = np.zeros((oh, ow), np.float32)
    valid = np.zeros((oh, ow), np.uint8)
    _L.ora_warp_plane(_ptr(src), w, h, _ptr(sim), ow, oh, INTERP.get(interp, interp), _ptr(out), _ptr(valid))
    return out, valid


def resize_plane(p, w, h, interp="spline4x4"):
    """prep.rs:63-71 resize_plane: the plane warped onto w x h with the Sim that
    scales it over the output about its center, pixel centers onto pixel centers."""
    p = _f32(p)
    ph, pw = p.shape
    if (pw, ph) == (w, h):
        return p.copy()
    sw, sh = float(pw), float(ph)
    kx, ky = w / sw, h / sh
    sim = [(kx - 1.0) * (sw + 1.0) / (2.0 * sw), (ky - 1.0) * (sh + 1.0) / (2.0 * sh), kx, 0.0, ky / kx, 0.0, 0.0, 0.0]
    return warp_plane(p, sim, interp, w, h)[0]


def warp_cost(rf, tgt, sim):
    rf, tgt, sim = _f32(rf), _f32(tgt), _f64(sim)
    th, tw = tgt.shape
    return _L.ora_warp_cost(_ptr(rf), rf.shape[1], _ptr(tgt), tw, th, _ptr(sim))


def reduce_burt(src):
    src = _f32(src)
    h, w = src.shape
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), np.float32)
    _L.ora_reduce_burt(_ptr(src), w, h, _ptr(out))
    return out


def gauss_levels(y):
    y = _f32(y)
    h, w = y.shape
    n = _L.ora_gauss_count(w, h, None)
    dims = np.zeros(2 * max(n, 1), np.uintp)
    _L.ora_gauss_count(w, h, _ptr(dims))
    tot = sum(int(dims[2 * i]) * int(dims[2 * i + 1]) for i in range(n))
    packed = np.zeros(max(tot, 1), np.float32)
    _L.ora_gauss_levels(_ptr(y), w, h, _ptr(packed))
    out, off = [], 0
    for i in range(n):
        lw, lh = int(dims[2 * i]), int(dims[2 * i + 1])
        out.append(packed[off:off + lw * lh].reshape(lh, lw))
        off += lw * lh
    return out


def level_steps(free_idx, lvl, w, h):
    fi = np.ascontiguousarray(free_idx, np.uintp)
    step, tol = np.zeros(len(fi)), np.zeros(len(fi))
    _L.ora_level_steps(_ptr(fi), len(fi), lvl, w, h, _ptr(step), _ptr(tol))
    return step, tol


def nelder_mead(f, x0, lo, hi, step, tol):
    n = len(x0)

    def cb(x, _ctx):
        return float(f(np.ctypeslib.as_array(x, shape=(n,)).copy()))

    fn = COST_FN(cb)
    out = np.zeros(n)
    ev = C.c_uint32()
    _L.ora_nelder_mead(fn, None, n, _ptr(_f64(x0)), _ptr(_f64(lo)), _ptr(_f64(hi)), _ptr(_f64(step)),
                       _ptr(_f64(tol)), _ptr(out), C.byref(ev))
    return out, ev.value


def multiscale_align(rf, tg, init=None, free=(1, 1, 1, 1, 0, 0, 0, 0), coarsen=0):
    rf, tg = _f32(rf), _f32(tg)
    h, w = rf.shape
    init = _f64(SIM_ID if init is None else init)
    fr = np.ascontiguousarray(free, np.uint8)
    out = np.zeros(8)
    ev = C.c_uint32()
    _L.ora_multiscale_align(_ptr(rf), _ptr(tg), w, h, _ptr(init), _ptr(fr), coarsen, _ptr(out), C.byref(ev))
    return out, ev.value


MAXL = 16


def multiscale_align_trace(rf, tg, init=None, free=(1, 1, 1, 1, 0, 0, 0, 0), coarsen=0, cap=1 << 16):
    """multiscale_align with its evaluations per level (level 0 = full resolution)
    and every cost in the order nelder_mead evaluates them."""
    rf, tg = _f32(rf), _f32(tg)
    h, w = rf.shape
    init = _f64(SIM_ID if init is None else init)
    fr = np.ascontiguousarray(free, np.uint8)
    out = np.zeros(8)
    ev = C.c_uint32()
    lev = np.zeros(MAXL, np.uint32)
    costs = np.zeros(cap)
    n = C.c_size_t()
    _L.ora_multiscale_align_trace(_ptr(rf), _ptr(tg), w, h, _ptr(init), _ptr(fr), coarsen, _ptr(out), C.byref(ev),
                                  _ptr(lev), _ptr(costs), cap, C.byref(n))
    return out, ev.value, lev, costs[:min(n.value, cap)].copy()


def sincos(x):
    """The oracle's sin and cos of a rotation (correctly rounded for |x| <= pi/4)."""
    s, c = C.c_double(), C.c_double()
    _L.ora_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def common_area(sims, w, h, interp="spline4x4"):
    s = _f64(np.array(sims, np.float64).reshape(-1, 8)) if len(sims) else np.zeros((0, 8))
    rect = np.zeros(4, np.uintp)
    _L.ora_common_area(_ptr(s) if len(sims) else None, len(sims), w, h, INTERP.get(interp, interp), _ptr(rect))
    return tuple(int(v) for v in rect)


# ------------------------------------------------------------- brightness.rs
def bright_gains(reference3, frame3, sim):
    r, f, sim = _f32(reference3), _f32(frame3), _f64(sim)
    _, h, w = r.shape
    gw, gh = -(-w // 4), -(-h // 4)
    grid = np.zeros(3 * gw * gh, np.float32)
    _L.ora_bright_grid(_ptr(r), w, h, _ptr(grid))
    g = np.zeros(3, np.float32)
    _L.ora_bright_gains(_ptr(grid), _ptr(f), w, h, _ptr(sim), _ptr(g))
    return g


def bright_ratio(ref, frm, n):
    g = np.zeros(3, np.float32)
    _L.ora_bright_ratio(_ptr(_f64(ref)), _ptr(_f64(frm)), n, _ptr(g))
    return g


def bright_means(img3):
    img3 = _f32(img3)
    m = np.zeros(3)
    _L.ora_bright_means(_ptr(img3), img3.shape[2], img3.shape[1], _ptr(m))
    return m


def bright_gains_to(ref_means, img3):
    img3 = _f32(img3)
    g = np.zeros(3, np.float32)
    _L.ora_bright_gains_to(_ptr(_f64(ref_means)), _ptr(img3), img3.shape[2], img3.shape[1], _ptr(g))
    return g


def bright_apply(img3, g):
    img3 = _f32(img3).copy()
    _L.ora_bright_apply(_ptr(img3), img3.shape[1] * img3.shape[2], _ptr(_f32(g)))
    return img3


# ------------------------------------------------------------------ depth.rs
def block_mean(src, k):
    src = _f32(src)
    h, w = src.shape
    out = np.zeros((-(-h // k), -(-w // k)), np.float32)
    _L.ora_block_mean(_ptr(src), w, h, k, _ptr(out))
    return out


def upsample_bilinear(g, w, h, k):
    g = _f32(g)
    dh, dw = g.shape
    out = np.zeros((h, w), np.float32)
    _L.ora_upsample_bilinear(_ptr(g), dw, dh, w, h, k, _ptr(out))
    return out


def huber_weights(conf, d, u, tau):
    """depth.rs:1168: conf * min(tau / |d - u|, 1)."""
    conf, d, u = _f32(conf), _f32(d), _f32(u)
    out = np.zeros_like(conf)
    _L.ora_huber_weights(_ptr(conf), _ptr(d), _ptr(u), conf.size, tau, _ptr(out))
    return out


def depth_grid(slices, fused3, params=None, **kw):
    """depth.rs:1125-1159: (guide, peak depth after the median, p90-normalised conf, p90) on the working grid."""
    slices = _f32(np.stack(slices))
    fused3 = _f32(fused3)
    n, dh, dw = slices.shape
    _, h, w = fused3.shape
    p = params or DepthParams.make(**kw)
    g, d, c = (np.zeros((dh, dw), np.float32) for _ in range(3))
    p90 = C.c_float()
    if _L.ora_depth_grid(_ptr(slices), n, _ptr(fused3), w, h, C.byref(p), _ptr(g), _ptr(d), _ptr(c),
                         C.byref(p90)) != 0:
        raise ValueError("depth: no slices")
    return g, d, c, p90.value


def box_mean(src, r):
    src = _f32(src)
    h, w = src.shape
    out = np.zeros_like(src)
    _L.ora_box_mean(_ptr(src), w, h, r, _ptr(out))
    return out


def guided_filter(guide, p, r, eps):
    guide, p = _f32(guide), _f32(p)
    h, w = p.shape
    out = np.zeros_like(p)
    _L.ora_guided_filter(_ptr(guide), _ptr(p), w, h, r, eps, _ptr(out))
    return out


def guided_coeffs(guide, p, r, eps):
    guide, p = _f32(guide), _f32(p)
    h, w = p.shape
    a, b = np.zeros_like(p), np.zeros_like(p)
    _L.ora_guided_coeffs(_ptr(guide), _ptr(p), w, h, r, eps, _ptr(a), _ptr(b))
    return a, b


def rdf_taps(r_in=1, r_out=3):
    n = _L.ora_rdf_taps(r_in, r_out, None, None, None)
    dy, dx, wt = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
    _L.ora_rdf_taps(r_in, r_out, _ptr(dy), _ptr(dx), _ptr(wt))
    return list(zip(dy.tolist(), dx.tolist(), wt.tolist()))


def focus_measure(y, params=None, **kw):
    y = _f32(y)
    h, w = y.shape
    p = params or DepthParams.make(**kw)
    out = np.zeros_like(y)
    _L.ora_focus_measure(_ptr(y), w, h, C.byref(p), _ptr(out))
    return out


def focus_slice(img3, params=None, **kw):
    img3 = _f32(img3)
    _, h, w = img3.shape
    p = params or DepthParams.make(**kw)
    k = 1 << p.scale
    out = np.zeros((-(-h // k), -(-w // k)), np.float32)
    _L.ora_focus_slice(_ptr(img3), w, h, C.byref(p), _ptr(out))
    return out


class PeakTracker:
    """depth.rs:533-611 PeakTracker, restated."""

    def __init__(self, m):
        self.m = m
        self._t = _L.ora_peaks_new(m)

    def push(self, c):
        c = _f32(c).ravel()
        assert c.size == self.m
        _L.ora_peaks_push(self._t, _ptr(c))

    def finish(self, gate):
        d, c, mins = (np.zeros(self.m, np.float32) for _ in range(3))
        fl = C.c_float()
        _L.ora_peaks_finish(self._t, gate, _ptr(d), _ptr(c), _ptr(mins), C.byref(fl))
        return d, c, mins, fl.value

    def __del__(self):
        if getattr(self, "_t", None):
            _L.ora_peaks_free(self._t)
            self._t = None


def normalize_conf(conf):
    conf = _f32(conf).copy()
    p90 = _L.ora_normalize_conf(_ptr(conf), conf.size)
    return conf, p90


def median3(src):
    src = _f32(src)
    h, w = src.shape
    out = np.zeros_like(src)
    _L.ora_median3(_ptr(src), w, h, _ptr(out))
    return out


def solve_rows(f, wd, a, lam):
    f, wd, a = _f32(f), _f32(wd), _f32(a)
    h, w = f.shape
    u = np.zeros_like(f)
    _L.ora_solve_rows(_ptr(u), _ptr(f), _ptr(wd), _ptr(a), w, h, lam)
    return u


def solve_cols(f, wd, a, lam):
    f, wd, a = _f32(f), _f32(wd), _f32(a)
    h, w = f.shape
    u = np.zeros_like(f)
    _L.ora_solve_cols(_ptr(u), _ptr(f), _ptr(wd), _ptr(a), w, h, lam)
    return u


class WlsSolver:
    """depth.rs:958-1078 WlsSolver, restated (tolerance a parameter)."""

    def __init__(self, guide, lam, sigma_c):
        guide = _f32(guide)
        self.h, self.w = guide.shape
        self._s = _L.ora_wls_new(_ptr(guide), self.w, self.h, lam, sigma_c)
        self.levels = _L.ora_wls_levels(self._s)

    def set_weights(self, conf):
        _L.ora_wls_set_weights(self._s, _ptr(_f32(conf)))

    def solve(self, d, max_iters=200, tol=1e-5, u=None):
        d = _f32(d)
        u = np.zeros_like(d) if u is None else _f32(u).copy()
        rel = C.c_float()
        it = _L.ora_wls_solve(self._s, _ptr(d), max_iters, tol, _ptr(u), C.byref(rel))
        return u, rel.value, it

    def vcycle(self, r):
        r = _f32(r)
        z = np.zeros_like(r)
        _L.ora_wls_vcycle(self._s, _ptr(r), _ptr(z))
        return z

    def apply(self, x):
        x = _f32(x)
        y = np.zeros_like(x)
        _L.ora_wls_apply(self._s, _ptr(x), _ptr(y))
        return y

    def __del__(self):
        if getattr(self, "_s", None):
            _L.ora_wls_free(self._s)
            self._s = None


def wls_solve(d, conf, guide, lam, sigma_c, max_iters, tol=1e-5):
    s = WlsSolver(guide, lam, sigma_c)
    s.set_weights(conf)
    u, rel, it = s.solve(d, max_iters, tol)
    return u, rel, it


def depth_from_slices(slices, fused3, params=None, **kw):
    slices = _f32(np.stack(slices))
    fused3 = _f32(fused3)
    n = slices.shape[0]
    _, h, w = fused3.shape
    p = params or DepthParams.make(**kw)
    depth, conf = np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)
    floor = np.zeros(slices.shape[1:], np.float32)
    info = DepthOut()
    rc = _L.ora_depth_from_slices(_ptr(slices), n, _ptr(fused3), w, h, C.byref(p), _ptr(depth), _ptr(conf),
                                  _ptr(floor), C.byref(info))
    if rc != 0:
        raise ValueError("depth: no slices")
    return depth, conf, floor, info


# -------------------------------------------------------------- wav.rs, slabs
def wav_weight(c, f, gate, power):
    """wav.rs:88-91 weight(), one cell."""
    return float(_L.ora_wav_weight(float(c), float(f), float(gate), float(power)))


def weighted_average(frames, floor, guide=None, power=2.0, smooth=3, gate=0.5, edge=True, params=None, **kw):
    """wav.rs:97-161 weighted_average over aligned frames ((3, h, w) each), the
    depth pass's grid floor, and the fused image as guide (or None)."""
    frames = _f32(np.stack(frames))
    n, _, h, w = frames.shape
    p = params or DepthParams.make(**kw)
    k = 1 << p.scale
    floor = _f32(floor)
    assert floor.size == (-(-h // k)) * (-(-w // k)), "floor: not the working grid"
    g = _f32(guide) if guide is not None else None
    out = np.zeros((3, h, w), np.float32)
    _L.ora_weighted_average(_ptr(frames), n, w, h, C.byref(p), float(power), int(smooth), float(gate), int(bool(edge)),
                            _ptr(floor), _ptr(g) if g is not None else None, _ptr(out))
    return out


def slab_ranges(count, size, overlap):
    """stack.rs:76-92: the (lo, hi) frame ranges of the slabs."""
    cap = max(1, count)
    out = np.zeros(2 * cap, np.uintp)
    ns = _L.ora_slab_ranges(count, size, overlap, out.ctypes.data_as(C.c_void_p), cap)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(min(ns, cap))]


# -------------------------------------------------------------- io / stack
def u16_to_planes(rgb):
    rgb = np.ascontiguousarray(rgb, np.uint16)
    h, w, _ = rgb.shape
    out = np.zeros((3, h, w), np.float32)
    _L.ora_u16_to_planes(_ptr(rgb), w * h, _ptr(out))
    return out


def planes_to_u16(img3):
    img3 = _f32(img3)
    _, h, w = img3.shape
    out = np.zeros((h, w, 3), np.uint16)
    _L.ora_planes_to_u16(_ptr(img3), w * h, _ptr(out))
    return out


def stack_run(frames, fuse=None, stack=None, depth=None, want_slices=False):
    """The reference CPU path over interleaved RGB u16 frames (n, h, w, 3)."""
    frames = np.ascontiguousarray(frames, np.uint16)
    n, h, w, _ = frames.shape
    fp = fuse or FuseParams.make()
    sp = stack or StackParams.make()
    dp = depth or DepthParams.make()
    img = np.zeros((3, h, w), np.float32)
    win = np.zeros((h, w), np.float32)
    sims = np.zeros((n, 8))
    gains = np.zeros((n, 3), np.float32)
    dmap = np.zeros((h, w), np.float32) if sp.depth else None
    conf = np.zeros((h, w), np.float32) if sp.depth else None
    k = 1 << dp.scale
    slices = np.zeros((n, -(-h // k), -(-w // k)), np.float32) if (want_slices and sp.depth) else None
    info, times = DepthOut(), StackTimes()
    evals = np.zeros((n, 1 + MAXL), np.uint32)
    rc = _L.ora_stack_run(_ptr(frames), n, w, h, C.byref(fp), C.byref(sp), C.byref(dp), _ptr(img), _ptr(win),
                          _ptr(sims), _ptr(gains), _ptr(dmap), _ptr(conf), _ptr(slices), C.byref(info),
                          C.byref(times), _ptr(evals))
    if rc != 0:
        raise ValueError("stack run failed")
    return dict(image=img, winner=win, sims=sims, gains=gains, depth=dmap, conf=conf, slices=slices, info=info,
                times=times, evals=evals[:, 0].astype(np.int64), level_evals=evals[:, 1:].astype(np.int64))
