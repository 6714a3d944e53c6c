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


def _dev(t, name, dtype, numel, device):
    """Device pointer of a CUDA tensor after checking what the C ABI assumes of it:
    dtype, contiguity, device and at least `numel` elements.  A mismatch raises
    ValueError before anything is launched."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"{name} must be on cuda:{device}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < numel:
        raise ValueError(f"{name} holds {t.numel()} elements, {numel} needed")
    return ctypes.c_void_p(t.data_ptr())


def _torch_stream(device):
    import torch
    return torch.cuda.current_stream(device)


def tuning_report():
    """The autotuner's choices so far in this process (dict)."""
    import json
    lib = _lib.load()
    n = lib.deeds_tuning_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.deeds_tuning_report(buf, n + 1)
    return json.loads(buf.value.decode())


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
    import torch
    arr = (Pair * max(1, len(pairs)))()
    keep = []
    g = grid[-1]
    for i, p in enumerate(pairs):
        if p["fixed"].dim() != 3 or p["moving"].shape != p["fixed"].shape:
            raise ValueError(f"pair {i}: fixed and moving must be 3-D tensors of the same shape")
        o, n, m = p["fixed"].shape
        sz, sz1 = m * n * o, (m // g) * (n // g) * (o // g)
        X = _f32(p["X"]) if p.get("X") is not None else None
        if X is not None and X.size != 16:
            raise ValueError(f"pair {i}: X must hold 16 floats")
        keep.append(X)
        w = p.get("warped")
        arr[i] = Pair(_dev(p["fixed"], "fixed", torch.float32, sz, device).value,
                      _dev(p["moving"], "moving", torch.float32, sz, device).value,
                      X.ctypes.data if X is not None else None, m, n, o,
                      _dev(p["flow"], "flow", torch.float32, 3 * sz1, device).value,
                      _dev(w, "warped", torch.float32, sz, device).value if w is not None else None)
    # the batch runs on the library's own streams: the inputs must be complete
    _torch_stream(device).synchronize()
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
    """A deeds_hip context for device-resident work (torch CUDA tensors in, out).

    The context runs its work on torch's current stream of `device` at the
    time of each call (so it is ordered after the torch work that produced
    its inputs, and torch work queued afterwards sees its outputs).  The
    asynchronous methods return before the GPU has finished; the
    registration, the tree and the BP return when their results are complete.
    """

    def __init__(self, device=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.deeds_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise DeedsError("deeds_ctx_create", rc, lib.deeds_last_error().decode())
        self.h = h
        self.device = device
        self._bound = None

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
        """Run the context's work on an existing HIP stream; None or 0 restores its own stream."""
        _lib.ctx_call(self.h, "deeds_ctx_set_stream", ctypes.c_void_p(stream_handle or None))
        self._bound = stream_handle or None

    def _call(self, name, *args):
        """Call deeds_ctx_<...> `name` ordered with torch's current stream.  On a torch
        stream the context simply runs on it.  torch's default stream (handle 0) is the
        legacy null stream, which the library's own non-blocking stream does not order
        against, so there the call waits for torch's pending work first and for its own
        work afterwards."""
        stream = _torch_stream(self.device)
        handle = stream.cuda_stream
        if handle:
            if handle != self._bound:
                self.set_stream(handle)
            _lib.ctx_call(self.h, name, *args)
            return
        if self._bound is not None:
            self.set_stream(None)
        stream.synchronize()
        _lib.ctx_call(self.h, name, *args)
        self.synchronize()

    def synchronize(self):
        _lib.ctx_call(self.h, "deeds_ctx_synchronize")

    # ---- hot-path functions on device tensors --------------------------------
    def descriptor(self, im, qs, out=None):
        """MINDSSCbox.h:127 descriptor() of a float32 image (o, n, m): int64 tensor of 60-bit codes."""
        import torch
        o, n, m = im.shape
        out = torch.empty(im.shape, dtype=torch.int64, device=im.device) if out is None else out
        self._call("deeds_ctx_descriptor", _dev(out, "out", torch.int64, m * n * o, self.device),
                      _dev(im, "im", torch.float32, m * n * o, self.device), m, n, o, int(qs))
        return out

    def datacost(self, F, M, step, hw, quant, alpha=1.6, randnum=1, out=None):
        """dataCostD.h:108 dataCostCL() of two int64 code volumes (o, n, m): (sz1, len^3) float32."""
        import torch
        if F.shape != M.shape or F.dim() != 3:
            raise ValueError("fixed and moving descriptors must be 3-D and of the same shape")
        o, n, m = F.shape
        sz1 = (m // step) * (n // step) * (o // step)
        L3 = (2 * hw + 1) ** 3
        out = torch.empty((sz1, L3), dtype=torch.float32, device=F.device) if out is None else out
        self._call("deeds_ctx_datacost", _dev(F, "F", torch.int64, m * n * o, self.device),
                      _dev(M, "M", torch.int64, m * n * o, self.device),
                      _dev(out, "cost", torch.float32, sz1 * L3, self.device), m, n, o, int(step), int(hw),
                      float(quant), float(alpha), int(randnum))
        return out

    def regularisation(self, cost, u0, hw, step, quant, tree, image_shape, allinds=None):
        """regularisation.h:115 regularisationCL(): `cost` (sz1, len^3) float32 is consumed;
        u0 (3, sz1) float32; tree = (ordered, parents, edgemst) host arrays; image_shape =
        (o, n, m).  Returns u1 (3, sz1); `allinds` (sz1, len^3) int16, if given, receives the
        argmin tables.  Returns when complete."""
        import torch
        o, n, m = image_shape
        sz1 = (m // step) * (n // step) * (o // step)
        L3 = (2 * hw + 1) ** 3
        ordered, parents, edgemst = (np.ascontiguousarray(tree[0], np.int32), np.ascontiguousarray(tree[1], np.int32),
                                     _f32(tree[2]))
        if ordered.size != sz1 or parents.size != sz1 or edgemst.size != sz1:
            raise ValueError(f"tree arrays must hold {sz1} entries")
        u0 = u0.reshape(3, sz1)
        u1 = torch.empty((3, sz1), dtype=torch.float32, device=cost.device)
        self._call("deeds_ctx_regularisation", _dev(cost, "cost", torch.float32, sz1 * L3, self.device),
                      _dev(u0[0], "u0", torch.float32, sz1, self.device),
                      _dev(u0[1], "v0", torch.float32, sz1, self.device),
                      _dev(u0[2], "w0", torch.float32, sz1, self.device), ctypes.c_void_p(u1[0].data_ptr()),
                      ctypes.c_void_p(u1[1].data_ptr()), ctypes.c_void_p(u1[2].data_ptr()), int(hw), int(step),
                      float(quant), _p(ordered), _p(parents), _p(edgemst), m, n, o,
                      _dev(allinds, "allinds", torch.int16, sz1 * L3, self.device) if allinds is not None else None)
        return u1

    def upsample_deformations(self, f0, shape):
        """transformations.h:245 of a (3, o2, n2, m2) float32 field to shape = (o, n, m): (3, o, n, m)."""
        import torch
        _, o2, n2, m2 = f0.shape
        o, n, m = shape
        out = torch.empty((3, o, n, m), dtype=torch.float32, device=f0.device)
        s2 = m2 * n2 * o2
        self._call("deeds_ctx_upsample_deformations", *[ctypes.c_void_p(out[i].data_ptr()) for i in range(3)],
                      *[_dev(f0[i], "field", torch.float32, s2, self.device) for i in range(3)], m, n, o, m2, n2, o2)
        return out

    def consistent_mapping(self, f, factor):
        """transformations.h:189 in place on a (6, o, n, m) float32 tensor (u, v, w, u2, v2, w2)."""
        import torch
        _, o, n, m = f.shape
        self._call("deeds_ctx_consistent_mapping",
                      *[_dev(f[i], "field", torch.float32, m * n * o, self.device) for i in range(6)], m, n, o,
                      int(factor))
        return f

    def warp_affine(self, inp, X, f, im1b=None):
        """dataCostD.h:267 warpAffine(): (warped, (ssd_before, ssd_after)) of a float32 image by
        X (16 floats, column-wise, or None) and the full-resolution field f (3, o, n, m)."""
        import torch
        o, n, m = inp.shape
        sz = m * n * o
        out = torch.empty_like(inp)
        ssd = np.zeros(2, np.float32)
        Xa = _f32(X) if X is not None else None
        self._call("deeds_ctx_warp_affine", _dev(out, "out", torch.float32, sz, self.device),
                      _dev(inp, "in", torch.float32, sz, self.device),
                      _dev(im1b, "im1b", torch.float32, sz, self.device) if im1b is not None else None, _p(Xa),
                      *[_dev(f[i], "field", torch.float32, sz, self.device) for i in range(3)], m, n, o,
                      _p(ssd) if im1b is not None else None)
        return out, (float(ssd[0]), float(ssd[1]))

    def warp_image(self, im1, f, im1b=None):
        """dataCostD.h:213 warpImageCL(): trilinear warp of im1 by x + f (3, o, n, m)."""
        import torch
        o, n, m = im1.shape
        sz = m * n * o
        out = torch.empty_like(im1)
        ssd = np.zeros(2, np.float32)
        self._call("deeds_ctx_warp_image", _dev(out, "out", torch.float32, sz, self.device),
                      _dev(im1, "im1", torch.float32, sz, self.device),
                      _dev(im1b, "im1b", torch.float32, sz, self.device) if im1b is not None else None,
                      *[_dev(f[i], "field", torch.float32, sz, self.device) for i in range(3)], m, n, o,
                      _p(ssd) if im1b is not None else None)
        return out, (float(ssd[0]), float(ssd[1]))

    def warp_affine_s(self, seg, X, f):
        """dataCostD.h:241 warpAffineS(): nearest-neighbour warp of an int16 label volume."""
        import torch
        o, n, m = seg.shape
        sz = m * n * o
        out = torch.empty_like(seg)
        Xa = _f32(X) if X is not None else None
        self._call("deeds_ctx_warp_affine_s", _dev(out, "out", torch.int16, sz, self.device),
                      _dev(seg, "seg", torch.int16, sz, self.device), _p(Xa),
                      *[_dev(f[i], "field", torch.float32, sz, self.device) for i in range(3)], m, n, o)
        return out

    def apply_float(self, img, flow, grid_step, X=None):
        """applyBCVfloat: trilinear warp of a float32 image (o, n, m) by a _displacements.dat
        field (3*sz1 float32 tensor on grid `grid_step`), the field upsampled in the same pass."""
        import torch
        o, n, m = img.shape
        sz1 = (m // grid_step) * (n // grid_step) * (o // grid_step)
        out = torch.empty_like(img)
        Xa = _f32(X) if X is not None else None
        self._call("deeds_ctx_apply_float", _dev(out, "out", torch.float32, m * n * o, self.device),
                      _dev(img, "img", torch.float32, m * n * o, self.device), _p(Xa),
                      _dev(flow, "flow", torch.float32, 3 * sz1, self.device), int(grid_step), m, n, o)
        return out

    def apply_seg(self, seg, flow, grid_step, X=None):
        """applyBCV: nearest-neighbour warp of an int16 label volume by a _displacements.dat field."""
        import torch
        o, n, m = seg.shape
        sz1 = (m // grid_step) * (n // grid_step) * (o // grid_step)
        out = torch.empty_like(seg)
        Xa = _f32(X) if X is not None else None
        self._call("deeds_ctx_apply_seg", _dev(out, "out", torch.int16, m * n * o, self.device),
                      _dev(seg, "seg", torch.int16, m * n * o, self.device), _p(Xa),
                      _dev(flow, "flow", torch.float32, 3 * sz1, self.device), int(grid_step), m, n, o)
        return out

    def jacobian(self, f, factor):
        """transformations.h:84 jacobian() of a (3, o, n, m) float32 field: (std, J<0 fraction)."""
        import torch
        _, o, n, m = f.shape
        s = np.zeros(1, np.float32)
        neg = np.zeros(1, np.float32)
        self._call("deeds_ctx_jacobian",
                      *[_dev(f[i], "field", torch.float32, m * n * o, self.device) for i in range(3)], m, n, o,
                      int(factor), _p(s), _p(neg))
        return float(s[0]), float(neg[0])

    def prims_graph(self, im, step):
        """primsGraph (primsMST.h:52) of a float32 CUDA tensor image (o, n, m): host arrays
        (ordered, parents, edgemst), the same tree as primsGraph() of the host copy."""
        import torch
        o, n, m = im.shape
        sz1 = (m // step) * (n // step) * (o // step)
        ordered = np.empty(sz1, np.int32)
        parents = np.empty(sz1, np.int32)
        edgemst = np.empty(sz1, np.float32)
        self._call("deeds_ctx_prims_graph", _dev(im, "im", torch.float32, m * n * o, self.device),
                      _p(ordered), _p(parents), _p(edgemst), int(step), m, n, o)
        return ordered, parents, edgemst

    def register(self, fixed, moving, flow, warped=None, grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4),
                 quant=(5, 4, 3, 2, 1), alpha=1.6, X=None):
        """fixed/moving/flow/warped: torch.float32 CUDA tensors (fixed, moving shaped (o, n, m)).
        Returns the timing dict when the registration is complete."""
        import torch
        if fixed.dim() != 3 or moving.shape != fixed.shape:
            raise ValueError("fixed and moving must be 3-D tensors of the same shape")
        o, n, m = fixed.shape
        sz, g = m * n * o, grid[-1]
        sz1 = (m // g) * (n // g) * (o // g)
        sched = make_schedule(grid, radius, quant, alpha)
        t = Timing()
        Xa = _f32(X) if X is not None else None
        if Xa is not None and Xa.size != 16:
            raise ValueError("X must hold 16 floats")
        args = (_dev(fixed, "fixed", torch.float32, sz, self.device),
                _dev(moving, "moving", torch.float32, sz, self.device), _p(Xa), m, n, o, ctypes.byref(sched),
                _dev(flow, "flow", torch.float32, 3 * sz1, self.device),
                _dev(warped, "warped", torch.float32, sz, self.device) if warped is not None else None,
                ctypes.byref(t))
        self._call("deeds_ctx_register", *args)
        return t.as_dict()
