"""Python mirror of deedsBCV's hot-path functions over the deeds_hip C ABI.

Same names and argument meaning as the reference functions (volumes are numpy
arrays of shape (o, n, m) in C order = the reference's voxel order; the
reference's image_m/n/o globals are taken from the array shapes).  Each call
runs on the GPU through libdeeds_hip.so; there is no CPU path.

    descriptor            MINDSSCbox.h:127
    dataCostCL            dataCostD.h:108
    primsGraph            primsMST.h:52
    regularisationCL      regularisation.h:115
    upsampleDeformationsCL transformations.h:245
    consistentMappingCL   transformations.h:189
    jacobian              transformations.h:84
    warpAffine            dataCostD.h:267
    warpImageCL           dataCostD.h:213
    warpAffineS           dataCostD.h:241
    register              deedsBCV0.cpp:55 (the full symmetric registration)
    dice                  diceMulti.cpp:17 (per-label Dice overlap)
    linear_register       linearBCV.cpp:144 (affine / rigid pre-alignment)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DeedsError, Pair, Timing, make_schedule  # noqa: F401

IDENTITY_X = np.eye(4, dtype=np.float32).T.ravel().copy()  # column-wise, deedsBCV0.cpp:161-172


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _dims(vol):
    if vol.ndim != 3:
        raise ValueError("volumes must be 3-D arrays of shape (o, n, m)")
    o, n, m = vol.shape
    return m, n, o


def descriptor(im, qs):
    """MIND-SSC codes (uint64, 12 x 5-bit thermometer fields) of image `im`."""
    im = _f32(im)
    m, n, o = _dims(im)
    out = np.empty(im.shape, np.uint64)
    _lib.call("deeds_descriptor", _p(out), _p(im), m, n, o, int(qs))
    return out


def dataCostCL(fixed_mind, moving_mind, step, hw, quant, alpha=1.6, randnum=1):
    """Cost volume (sz1, (2hw+1)^3) float32, node-major."""
    F = np.ascontiguousarray(fixed_mind, np.uint64)
    M = np.ascontiguousarray(moving_mind, np.uint64)
    if F.shape != M.shape:
        raise ValueError("fixed and moving descriptors must have the same shape")
    m, n, o = _dims(F)
    sz1 = (m // step) * (n // step) * (o // step)
    cost = np.empty((sz1, (2 * hw + 1) ** 3), np.float32)
    _lib.call("deeds_datacost", _p(F), _p(M), _p(cost), m, n, o, int(step), int(hw), float(quant), float(alpha),
              int(randnum))
    return cost


def primsGraph(im, step):
    """(ordered, parents, edgemst) of the MST over the control grid of `im`."""
    im = _f32(im)
    m, n, o = _dims(im)
    sz1 = (m // step) * (n // step) * (o // step)
    ordered = np.empty(sz1, np.int32)
    parents = np.empty(sz1, np.int32)
    edgemst = np.empty(sz1, np.float32)
    _lib.call("deeds_prims_graph", _p(im), _p(ordered), _p(parents), _p(edgemst), int(step), m, n, o)
    return ordered, parents, edgemst


def regularisationCL(cost, u0, v0, w0, hw, step, quant, ordered, parents, edgemst, image_shape):
    """Tree BP; returns (u1, v1, w1) on the control grid.  image_shape = (o, n, m)."""
    cost = _f32(cost).copy()
    u0, v0, w0 = _f32(u0), _f32(v0), _f32(w0)
    o, n, m = image_shape
    out = [np.empty(u0.shape, np.float32) for _ in range(3)]
    _lib.call("deeds_regularisation", _p(cost), _p(u0), _p(v0), _p(w0), *[_p(x) for x in out], int(hw), int(step),
              float(quant), _p(np.ascontiguousarray(ordered, np.int32)),
              _p(np.ascontiguousarray(parents, np.int32)), _p(_f32(edgemst)), m, n, o)
    return tuple(out)


def upsampleDeformationsCL(u0, v0, w0, shape):
    """Trilinear resampling of a field (o2, n2, m2) to `shape` = (o, n, m)."""
    u0, v0, w0 = _f32(u0), _f32(v0), _f32(w0)
    m2, n2, o2 = _dims(u0)
    o, n, m = shape
    out = [np.empty(shape, np.float32) for _ in range(3)]
    _lib.call("deeds_upsample_deformations", *[_p(x) for x in out], _p(u0), _p(v0), _p(w0), m, n, o, m2, n2, o2)
    return tuple(out)


def consistentMappingCL(u, v, w, u2, v2, w2, factor):
    """Symmetric inverse-consistent fields; returns new copies (u, v, w, u2, v2, w2)."""
    a = [_f32(x).copy() for x in (u, v, w, u2, v2, w2)]
    m, n, o = _dims(a[0])
    _lib.call("deeds_consistent_mapping", *[_p(x) for x in a], m, n, o, int(factor))
    return tuple(a)


def jacobian(u, v, w, factor):
    """(std of det J, fraction of det J < 0) of x + u/factor (diagnostic)."""
    u, v, w = _f32(u), _f32(v), _f32(w)
    m, n, o = _dims(u)
    s = np.zeros(1, np.float32)
    neg = np.zeros(1, np.float32)
    _lib.call("deeds_jacobian", _p(u), _p(v), _p(w), m, n, o, int(factor), _p(s), _p(neg))
    return float(s[0]), float(neg[0])


def warpAffine(inp, im1b, X, u, v, w):
    """Trilinear affine+field warp; returns (warped, (ssd_before, ssd_after))."""
    inp = _f32(inp)
    m, n, o = _dims(inp)
    out = np.empty(inp.shape, np.float32)
    ssd = np.zeros(2, np.float32)
    im1b = None if im1b is None else _f32(im1b)
    _lib.call("deeds_warp_affine", _p(out), _p(inp), _p(im1b), _p(_f32(X if X is not None else IDENTITY_X)),
              _p(_f32(u)), _p(_f32(v)), _p(_f32(w)), m, n, o, _p(ssd) if im1b is not None else None)
    return out, (float(ssd[0]), float(ssd[1]))


def warpImageCL(im1, im1b, u, v, w):
    im1 = _f32(im1)
    m, n, o = _dims(im1)
    out = np.empty(im1.shape, np.float32)
    ssd = np.zeros(2, np.float32)
    im1b = None if im1b is None else _f32(im1b)
    _lib.call("deeds_warp_image", _p(out), _p(im1), _p(im1b), _p(_f32(u)), _p(_f32(v)), _p(_f32(w)), m, n, o,
              _p(ssd) if im1b is not None else None)
    return out, (float(ssd[0]), float(ssd[1]))


def warpAffineS(seg, X, u, v, w):
    seg = np.ascontiguousarray(seg, np.int16)
    m, n, o = _dims(seg)
    out = np.empty(seg.shape, np.int16)
    _lib.call("deeds_warp_affine_s", _p(out), _p(seg), _p(_f32(X if X is not None else IDENTITY_X)), _p(_f32(u)),
              _p(_f32(v)), _p(_f32(w)), m, n, o)
    return out


def register(fixed, moving, grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4), quant=(5, 4, 3, 2, 1), alpha=1.6,
             X=None, want_warped=True):
    """Full symmetric registration (HU volumes).  Returns (flow, warped, timing dict).

    flow is the _displacements.dat payload [u; v; w] over the final control grid.
    """
    fixed, moving = _f32(fixed), _f32(moving)
    if fixed.shape != moving.shape:
        raise ValueError("Inconsistent image sizes (must have same dimensions)")
    m, n, o = _dims(fixed)
    g = grid[-1]
    sz1 = (m // g) * (n // g) * (o // g)
    flow = np.empty(3 * sz1, np.float32)
    warped = np.empty(fixed.shape, np.float32) if want_warped else None
    sched = make_schedule(grid, radius, quant, alpha)
    t = Timing()
    _lib.call("deeds_register", _p(fixed), _p(moving), _p(_f32(X) if X is not None else None), m, n, o,
              ctypes.byref(sched), _p(flow), _p(warped), ctypes.byref(t))
    return flow, warped, t.as_dict()


def register_batch(pairs, grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4), quant=(5, 4, 3, 2, 1), alpha=1.6,
                   concurrency=16, want_warped=True):
    """Multi-atlas batch: `pairs` is a list of (fixed, moving) or (fixed, moving, X) volumes,
    each with its own shape, registered level-synchronously in groups of `concurrency`
    pairs.  Returns a list of (flow, warped, timing); every result equals register() of
    that pair alone."""
    keep, arr = [], (Pair * max(1, len(pairs)))()
    outs = []
    for i, p in enumerate(pairs):
        fixed, moving = _f32(p[0]), _f32(p[1])
        if fixed.shape != moving.shape:
            raise ValueError("Inconsistent image sizes (must have same dimensions)")
        X = _f32(p[2]) if len(p) > 2 and p[2] is not None else None
        m, n, o = _dims(fixed)
        g = grid[-1]
        flow = np.empty(3 * (m // g) * (n // g) * (o // g), np.float32)
        warped = np.empty(fixed.shape, np.float32) if want_warped else None
        keep += [fixed, moving, X]
        arr[i] = Pair(fixed.ctypes.data, moving.ctypes.data, X.ctypes.data if X is not None else None, m, n, o,
                      flow.ctypes.data, warped.ctypes.data if warped is not None else None)
        outs.append((flow, warped))
    sched = make_schedule(grid, radius, quant, alpha)
    times = (Timing * max(1, len(pairs)))()
    _lib.call("deeds_register_batch", arr, len(pairs), ctypes.byref(sched), int(concurrency), times)
    return [(f, w, times[i].as_dict()) for i, (f, w) in enumerate(outs)]


def register_batch_device(pairs, device=0, grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4), quant=(5, 4, 3, 2, 1),
                          alpha=1.6, concurrency=16):
    """Batch on device data: `pairs` is a list of dicts with torch CUDA tensors
    fixed, moving (o, n, m), flow (3*sz1) and optional warped, X (numpy, host)."""
    arr = (Pair * max(1, len(pairs)))()
    keep = []
    for i, p in enumerate(pairs):
        o, n, m = p["fixed"].shape
        X = _f32(p["X"]) if p.get("X") is not None else None
        keep.append(X)
        w = p.get("warped")
        arr[i] = Pair(p["fixed"].data_ptr(), p["moving"].data_ptr(), X.ctypes.data if X is not None else None,
                      m, n, o, p["flow"].data_ptr(), w.data_ptr() if w is not None else None)
    sched = make_schedule(grid, radius, quant, alpha)
    times = (Timing * max(1, len(pairs)))()
    _lib.call("deeds_register_batch_device", int(device), arr, len(pairs), ctypes.byref(sched), int(concurrency),
              times)
    return [times[i].as_dict() for i in range(len(pairs))]


def linear_register(fixed, moving, grid=(7, 6, 5, 4), radius=(5, 4, 3, 2), quant=(4, 3, 2, 1), rigid=False):
    """linearBCV (linearBCV.cpp:144-263): the 16-float affine (or rigid) matrix, column-wise,
    aligning moving to fixed (HU volumes), and the point pairs fitted per level."""
    fixed, moving = _f32(fixed), _f32(moving)
    if fixed.shape != moving.shape:
        raise ValueError("Inconsistent image sizes (must have same dimensions)")
    m, n, o = _dims(fixed)
    sched = make_schedule(grid, radius, quant, 1.0)
    X = np.zeros(16, np.float32)
    used = np.zeros(len(grid), np.int32)
    _lib.call("deeds_linear_register", _p(fixed), _p(moving), m, n, o, ctypes.byref(sched), int(bool(rigid)), _p(X),
              _p(used))
    return X, used


def dice(seg1, seg2):
    """diceMulti (diceMulti.cpp:17-79): (labels, dice) for every label of seg1 except its
    smallest value, in increasing order; dice[i] = 2|A & B| / (|A| + |B|) in float."""
    a = np.ascontiguousarray(seg1, np.int16)
    b = np.ascontiguousarray(seg2, np.int16)
    if a.shape != b.shape:
        raise ValueError("Inconsistent image sizes (must have same dimensions)")
    count = ctypes.c_int()
    _lib.call("deeds_dice", _p(a), _p(b), a.size, 0, None, None, ctypes.byref(count))
    labels = np.empty(count.value, np.int16)
    values = np.empty(count.value, np.float32)
    _lib.call("deeds_dice", _p(a), _p(b), a.size, count.value, _p(labels), _p(values), ctypes.byref(count))
    return labels, values


def apply_seg(seg, flow, grid_step, X=None):
    """applyBCV: warp a label volume with a _displacements.dat field (nearest neighbour)."""
    seg = np.ascontiguousarray(seg, np.int16)
    m, n, o = _dims(seg)
    out = np.empty(seg.shape, np.int16)
    _lib.call("deeds_apply_seg", _p(out), _p(seg), _p(_f32(X) if X is not None else None), _p(_f32(flow)),
              int(grid_step), m, n, o)
    return out


def apply_float(img, flow, grid_step, X=None):
    """applyBCVfloat: warp an image with a _displacements.dat field (trilinear)."""
    img = _f32(img)
    m, n, o = _dims(img)
    out = np.empty(img.shape, np.float32)
    _lib.call("deeds_apply_float", _p(out), _p(img), _p(_f32(X) if X is not None else None), _p(_f32(flow)),
              int(grid_step), m, n, o)
    return out


class Context:
    """A deeds_hip context for device-resident work (torch CUDA tensors in, out)."""

    def __init__(self, device=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.deeds_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise DeedsError("deeds_ctx_create", rc, lib.deeds_last_error().decode())
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _lib.load().deeds_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        _lib.ctx_call(self.h, "deeds_ctx_set_stream", ctypes.c_void_p(stream_handle))

    def synchronize(self):
        _lib.ctx_call(self.h, "deeds_ctx_synchronize")

    def prims_graph(self, im, step):
        """primsGraph (primsMST.h:52) of a float32 CUDA tensor image (o, n, m): host arrays
        (ordered, parents, edgemst), the same tree as primsGraph() of the host copy.  The
        caller makes sure the image is ready (e.g. by synchronizing torch's stream)."""
        o, n, m = im.shape
        sz1 = (m // step) * (n // step) * (o // step)
        ordered = np.empty(sz1, np.int32)
        parents = np.empty(sz1, np.int32)
        edgemst = np.empty(sz1, np.float32)
        _lib.ctx_call(self.h, "deeds_ctx_prims_graph", ctypes.c_void_p(im.data_ptr()), _p(ordered), _p(parents),
                      _p(edgemst), int(step), m, n, o)
        return ordered, parents, edgemst

    def register(self, fixed, moving, flow, warped=None, grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4),
                 quant=(5, 4, 3, 2, 1), alpha=1.6, X=None):
        """fixed/moving/flow/warped: torch.float32 CUDA tensors (fixed, moving shaped (o, n, m))."""
        o, n, m = fixed.shape
        sched = make_schedule(grid, radius, quant, alpha)
        t = Timing()
        Xa = _f32(X) if X is not None else None
        _lib.ctx_call(self.h, "deeds_ctx_register", ctypes.c_void_p(fixed.data_ptr()),
                      ctypes.c_void_p(moving.data_ptr()), _p(Xa), m, n, o, ctypes.byref(sched),
                      ctypes.c_void_p(flow.data_ptr()),
                      ctypes.c_void_p(warped.data_ptr()) if warped is not None else None, ctypes.byref(t))
        return t.as_dict()
