"""Differentiable soft variant of one deeds registration level.

torch.autograd.Function wrappers over the deeds_hip C ABI (the
``deeds_ctx_soft_*`` entry points); forward and backward both run the
library's HIP kernels, float32 or float64 as the inputs are.  There is no
counterpart in deedsBCV; the model relaxes the plain level step
(MINDSSCbox.h:127 descriptor -> dataCostD.h:108 dataCostCL ->
regularisation.h:115 regularisationCL) into smooth functions:

    SoftFeatures  image (o, n, m) -> phi (o, n, m, 12): the twelve SSC values
                  that descriptor() thresholds, before thresholding
                  (MINDSSCbox.h:184-201)
    SoftCost      phi_F, phi_M -> C (N, L^3) = a1 * sum over the block samples
                  of ||phi_F(s) - phi_M(clamp(s + d_l))||^2 (samples, clamping
                  and a1 of dataCostD.h:141-196)
    SoftBP        C, u0 -> u1 (3, N): sum-product BP at temperature T on the
                  plain path's MST for E = sum_n C_n + sum_c
                  |l_p - l_c + (u0_p - u0_c)/quant|^2 / edgemst[c]; u1 = u0 +
                  the expected label displacement of every node's marginal

Gradients are exact vector-Jacobian products computed by hand-written adjoint
kernels; the BP keeps every node's upward belief, upward message and full
belief and recomputes the distance transforms in its backward sweeps.  The
tree and u0 are constants (u0 must not require grad).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

REAL_F32, REAL_F64 = 0, 1
_contexts = {}


def _torch():
    import torch
    return torch


def _real(t):
    torch = _torch()
    if t.dtype == torch.float32:
        return REAL_F32
    if t.dtype == torch.float64:
        return REAL_F64
    raise TypeError(f"soft deeds supports float32 and float64 tensors, got {t.dtype}")


def _context(device):
    """The deeds_hip context of a CUDA device (created on first use, kept for the process)."""
    idx = device.index if device.index is not None else _torch().cuda.current_device()
    h = _contexts.get(idx)
    if h is None:
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.deeds_ctx_create(int(idx), ctypes.byref(h))
        if rc != 0:
            raise _lib.DeedsError("deeds_ctx_create", rc, lib.deeds_last_error().decode())
        _contexts[idx] = h
    return h


def _check(*tensors):
    dev = tensors[0].device
    dt = tensors[0].dtype
    for t in tensors:
        if not t.is_cuda:
            raise ValueError("soft deeds works on CUDA tensors")
        if t.device != dev or t.dtype != dt:
            raise ValueError("all tensors of a soft deeds call must share device and dtype")
    return dev


def _run(device, name, *args):
    """Call a deeds_ctx_soft_* entry point on the library's stream, ordered after
    torch's pending work on the current stream and completed before returning."""
    torch = _torch()
    h = _context(device)
    torch.cuda.current_stream(device).synchronize()
    conv = [ctypes.c_void_p(a.data_ptr()) if isinstance(a, torch.Tensor) else a for a in args]
    _lib.ctx_call(h, name, *conv)
    _lib.ctx_call(h, "deeds_ctx_synchronize")


def _vol_dims(t, ndim, what):
    if t.dim() != ndim:
        raise ValueError(f"{what} must have {ndim} dimensions")
    o, n, m = t.shape[:3]
    return m, n, o


def mind_step(quant):
    """The self-similarity spacing qs of a level with quantisation quant (deedsBCV0.cpp:185)."""
    return int(math.floor(0.5 * float(quant) + 1.0))


def _make_functions():
    torch = _torch()

    class _SoftFeatures(torch.autograd.Function):
        @staticmethod
        def forward(ctx, im, qs):
            im = im.contiguous()
            dev = _check(im)
            m, n, o = _vol_dims(im, 3, "image")
            phi = torch.empty((o, n, m, 12), dtype=im.dtype, device=dev)
            _run(dev, "deeds_ctx_soft_features", _real(im), phi, im, m, n, o, int(qs))
            ctx.save_for_backward(im)
            ctx.qs = int(qs)
            return phi

        @staticmethod
        def backward(ctx, gphi):
            (im,) = ctx.saved_tensors
            gphi = gphi.contiguous()
            m, n, o = _vol_dims(im, 3, "image")
            gim = torch.empty_like(im)
            _run(im.device, "deeds_ctx_soft_features_backward", _real(im), gim, gphi, im, m, n, o, ctx.qs)
            return gim, None

    class _SoftCost(torch.autograd.Function):
        @staticmethod
        def forward(ctx, F, M, step, hw, quant, alpha):
            F, M = F.contiguous(), M.contiguous()
            dev = _check(F, M)
            if F.shape != M.shape or F.dim() != 4 or F.shape[3] != 12:
                raise ValueError("features must both be shaped (o, n, m, 12)")
            m, n, o = _vol_dims(F, 4, "features")
            N = (m // step) * (n // step) * (o // step)
            L3 = (2 * hw + 1) ** 3
            cost = torch.empty((N, L3), dtype=F.dtype, device=dev)
            _run(dev, "deeds_ctx_soft_cost", _real(F), cost, F, M, m, n, o, int(step), int(hw), float(quant),
                 float(alpha))
            ctx.save_for_backward(F, M)
            ctx.args = (m, n, o, int(step), int(hw), float(quant), float(alpha))
            return cost

        @staticmethod
        def backward(ctx, gcost):
            F, M = ctx.saved_tensors
            gF, gM = torch.empty_like(F), torch.empty_like(M)
            _run(F.device, "deeds_ctx_soft_cost_backward", _real(F), gF, gM, gcost.contiguous(), F, M, *ctx.args)
            return gF, gM, None, None, None, None

    class _SoftBP(torch.autograd.Function):
        @staticmethod
        def forward(ctx, cost, u0, tree, dims, step, hw, quant, T):
            if u0.requires_grad:
                raise ValueError("SoftBP treats u0 as a constant; detach it first")
            cost, u0 = cost.contiguous(), u0.contiguous()
            dev = _check(cost, u0)
            m, n, o = dims
            N = (m // step) * (n // step) * (o // step)
            L3 = (2 * hw + 1) ** 3
            if cost.shape != (N, L3) or u0.shape != (3, N):
                raise ValueError(f"cost must be ({N}, {L3}) and u0 (3, {N})")
            ordered, parents, edgemst = (np.ascontiguousarray(tree[0], np.int32),
                                         np.ascontiguousarray(tree[1], np.int32),
                                         np.ascontiguousarray(tree[2], np.float32))
            if not (ordered.shape == parents.shape == edgemst.shape == (N,)):
                raise ValueError(f"tree arrays must have {N} entries")
            lib = _lib.load()
            saved = torch.empty(int(lib.deeds_soft_saved_count(N, hw)), dtype=cost.dtype, device=dev)
            u1 = torch.empty_like(u0)
            targs = (ordered.ctypes.data_as(ctypes.c_void_p), parents.ctypes.data_as(ctypes.c_void_p),
                     edgemst.ctypes.data_as(ctypes.c_void_p))
            _run(dev, "deeds_ctx_soft_bp", _real(cost), u1, saved, cost, u0, *targs, m, n, o, int(step), int(hw),
                 float(quant), float(T))
            ctx.save_for_backward(cost, u0, saved)
            ctx.tree = (ordered, parents, edgemst)
            ctx.args = (m, n, o, int(step), int(hw), float(quant), float(T))
            return u1

        @staticmethod
        def backward(ctx, gu1):
            cost, u0, saved = ctx.saved_tensors
            gcost = torch.empty_like(cost)
            ordered, parents, edgemst = ctx.tree
            targs = (ordered.ctypes.data_as(ctypes.c_void_p), parents.ctypes.data_as(ctypes.c_void_p),
                     edgemst.ctypes.data_as(ctypes.c_void_p))
            _run(cost.device, "deeds_ctx_soft_bp_backward", _real(cost), gcost, gu1.contiguous(), saved, cost, u0,
                 *targs, *ctx.args)
            return gcost, None, None, None, None, None, None, None

    return _SoftFeatures, _SoftCost, _SoftBP


SoftFeatures, SoftCost, SoftBP = _make_functions()


def soft_features(im, qs):
    """phi (o, n, m, 12) of an image (o, n, m)."""
    return SoftFeatures.apply(im, qs)


def soft_cost(phi_fixed, phi_moving, step, hw, quant, alpha=1.6):
    """Cost volume (N, (2hw+1)^3), node-major."""
    return SoftCost.apply(phi_fixed, phi_moving, step, hw, quant, alpha)


def soft_bp(cost, u0, tree, image_shape, step, hw, quant, T=0.5):
    """Expected displacements (3, N) (u, v, w) of sum-product BP on `tree`
    = (ordered, parents, edgemst).  image_shape = (o, n, m)."""
    o, n, m = image_shape
    return SoftBP.apply(cost, u0, tree, (m, n, o), step, hw, quant, T)


_mst_contexts = {}


def _mst_context(device):
    """A second context per device for tree builds that overlap the soft kernels."""
    from . import api
    idx = device.index if device.index is not None else _torch().cuda.current_device()
    c = _mst_contexts.get(idx)
    if c is None:
        c = _mst_contexts[idx] = api.Context(idx)
    return c


def _tree_job(fixed, step):
    """A callable building the plain path's MST (primsMST.h:52) of `fixed` as host
    arrays: from the device image when it is a CUDA tensor (edge costs and image
    statistics on the GPU), else from the host array."""
    torch = _torch()
    if isinstance(fixed, torch.Tensor) and fixed.is_cuda:
        im = fixed.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream(im.device).synchronize()  # the library runs on its own streams
        ctx = _mst_context(im.device)
        return lambda: ctx.prims_graph(im, step)
    from . import api
    img = fixed.detach().to("cpu", torch.float32).numpy() if isinstance(fixed, torch.Tensor) else fixed
    return lambda: api.primsGraph(img, step)


def tree_of(fixed, step):
    """The plain path's MST of the fixed image (primsMST.h:52) as host arrays."""
    return _tree_job(fixed, step)()


_tree_pool = None


def soft_register(fixed, moving, grid=6, radius=6, quant=2, alpha=1.6, T=0.5, tree=None, u0=None, tree_image=None):
    """One soft registration level: expected displacements (3, o1, n1, m1) of
    the control grid, differentiable with respect to both images.  Without
    `tree`, the MST of the fixed image (or of the host array `tree_image`) is
    built on a host thread while the features and the cost run on the GPU."""
    global _tree_pool
    torch = _torch()
    o, n, m = fixed.shape
    qs = mind_step(quant)
    pending = None
    if tree is None:
        import concurrent.futures as cf
        if _tree_pool is None:
            _tree_pool = cf.ThreadPoolExecutor(max_workers=1)
        pending = _tree_pool.submit(_tree_job(tree_image if tree_image is not None else fixed, grid))
    F = soft_features(fixed, qs)
    M = soft_features(moving, qs)
    C = soft_cost(F, M, grid, radius, quant, alpha)
    N = C.shape[0]
    if pending is not None:
        tree = pending.result()
    if u0 is None:
        u0 = torch.zeros((3, N), dtype=fixed.dtype, device=fixed.device)
    u1 = soft_bp(C, u0.reshape(3, N), tree, (o, n, m), grid, radius, quant, T)
    return u1.reshape(3, o // grid, n // grid, m // grid)
