"""ctypes binding of libdeeds_hip.so (the C ABI declared in include/deeds_hip.h).

The library is built in-tree by ``python -m deedsbcv_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libdeeds_hip.so"

MAX_LEVELS = 10
c_int, c_float, c_void_p, c_char_p = ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p


class Schedule(ctypes.Structure):
    _fields_ = [("levels", c_int), ("grid", c_int * MAX_LEVELS), ("radius", c_int * MAX_LEVELS),
                ("quant", c_int * MAX_LEVELS), ("alpha", c_float)]


class Timing(ctypes.Structure):
    _fields_ = [("levels", c_int), ("mind", ctypes.c_double * MAX_LEVELS), ("data", ctypes.c_double * MAX_LEVELS),
                ("mst_reg", ctypes.c_double * MAX_LEVELS), ("transf", ctypes.c_double * MAX_LEVELS),
                ("total", ctypes.c_double), ("hamming_evals", ctypes.c_double), ("ssd_before", c_float),
                ("ssd_after", c_float), ("jac_std", c_float * MAX_LEVELS), ("jac_neg", c_float * MAX_LEVELS),
                ("datacost_kernel_s", ctypes.c_double), ("datacost_launches", c_int),
                ("datacost_level_s", ctypes.c_double * MAX_LEVELS),
                ("datacost_level_evals", ctypes.c_double * MAX_LEVELS),
                ("tree_build_s", ctypes.c_double), ("tree_wait_s", ctypes.c_double),
                ("tree_build_level_s", ctypes.c_double * MAX_LEVELS),
                ("tree_wait_level_s", ctypes.c_double * MAX_LEVELS), ("stats_s", ctypes.c_double),
                ("bp_idle_level_s", ctypes.c_double * MAX_LEVELS), ("bp_level_s", ctypes.c_double * MAX_LEVELS),
                ("tree_ready_level_s", ctypes.c_double * MAX_LEVELS),
                ("tree_shape_level_s", ctypes.c_double * MAX_LEVELS)]

    def as_dict(self):
        L = self.levels
        return {"levels": L, "mind": list(self.mind[:L]), "data": list(self.data[:L]),
                "mst_reg": list(self.mst_reg[:L]), "transf": list(self.transf[:L]), "total": self.total,
                "hamming_evals": self.hamming_evals, "ssd_before": self.ssd_before, "ssd_after": self.ssd_after,
                "jac_std": list(self.jac_std[:L]), "jac_neg": list(self.jac_neg[:L]),
                "datacost_kernel_s": self.datacost_kernel_s, "datacost_launches": self.datacost_launches,
                "datacost_level_s": list(self.datacost_level_s[:L]),
                "datacost_level_evals": list(self.datacost_level_evals[:L]),
                "tree_build_s": self.tree_build_s, "tree_wait_s": self.tree_wait_s,
                "tree_build_level_s": list(self.tree_build_level_s[:L]),
                "tree_wait_level_s": list(self.tree_wait_level_s[:L]), "stats_s": self.stats_s,
                "bp_idle_level_s": list(self.bp_idle_level_s[:L]), "bp_level_s": list(self.bp_level_s[:L]),
                "tree_ready_level_s": list(self.tree_ready_level_s[:L]),
                "tree_shape_level_s": list(self.tree_shape_level_s[:L])}


class Pair(ctypes.Structure):
    _fields_ = [("fixed", c_void_p), ("moving", c_void_p), ("X", c_void_p), ("m", c_int), ("n", c_int),
                ("o", c_int), ("flow", c_void_p), ("warped", c_void_p)]


# name -> argtypes (all return int status unless listed in _RESTYPES)
I, F, V = c_int, c_float, c_void_p
SIGNATURES = {
    "deeds_version": [],
    "deeds_last_error": [],
    "deeds_set_default_device": [I],
    "deeds_ctx_create": [I, ctypes.POINTER(c_void_p)],
    "deeds_ctx_destroy": [V],
    "deeds_ctx_last_error": [V],
    "deeds_ctx_set_stream": [V, V],
    "deeds_ctx_synchronize": [V],
    "deeds_ctx_trim": [V],
    "deeds_inject_fault": [I, ctypes.c_longlong],
    "deeds_tuning_report": [V, ctypes.c_size_t],
    "deeds_set_datacost_tiles": [I, I],
    "deeds_set_batch_graphs": [I],
    "deeds_descriptor": [V, V, I, I, I, I],
    "deeds_datacost": [V, V, V, I, I, I, I, I, F, F, I],
    "deeds_prims_graph": [V, V, V, V, I, I, I, I],
    "deeds_ctx_prims_graph": [V, V, V, V, V, I, I, I, I],
    "deeds_image_stats": [V, ctypes.c_longlong, V, V],
    "deeds_dice": [V, V, ctypes.c_longlong, I, V, V, V],
    "deeds_linear_register": [V, V, I, I, I, V, I, V, V],
    "deeds_ctx_linear_register": [V, V, V, I, I, I, V, I, V, V],
    "deeds_ctx_dice": [V, V, V, ctypes.c_longlong, I, V, V, V],
    "deeds_ctx_dice_noalloc": [V, V, V, ctypes.c_longlong, I, V, V, V],
    "deeds_regularisation": [V, V, V, V, V, V, V, I, I, F, V, V, V, I, I, I],
    "deeds_upsample_deformations": [V, V, V, V, V, V, I, I, I, I, I, I],
    "deeds_consistent_mapping": [V, V, V, V, V, V, I, I, I, I],
    "deeds_jacobian": [V, V, V, I, I, I, I, V, V],
    "deeds_warp_affine": [V, V, V, V, V, V, V, I, I, I, V],
    "deeds_warp_image": [V, V, V, V, V, V, I, I, I, V],
    "deeds_warp_affine_s": [V, V, V, V, V, V, I, I, I],
    "deeds_register": [V, V, V, I, I, I, ctypes.POINTER(Schedule), V, V, ctypes.POINTER(Timing)],
    "deeds_register_batch": [ctypes.POINTER(Pair), I, ctypes.POINTER(Schedule), I, ctypes.POINTER(Timing)],
    "deeds_register_batch_device": [I, ctypes.POINTER(Pair), I, ctypes.POINTER(Schedule), I,
                                    ctypes.POINTER(Timing)],
    "deeds_apply_seg": [V, V, V, V, I, I, I, I],
    "deeds_apply_float": [V, V, V, V, I, I, I, I],
    "deeds_ctx_descriptor": [V, V, V, I, I, I, I],
    "deeds_ctx_image_stats": [V, V, ctypes.c_longlong, V, V],
    "deeds_ctx_datacost": [V, V, V, V, I, I, I, I, I, F, F, I],
    "deeds_ctx_regularisation": [V, V, V, V, V, V, V, V, I, I, F, V, V, V, I, I, I, V],
    "deeds_ctx_register": [V, V, V, V, I, I, I, ctypes.POINTER(Schedule), V, V, ctypes.POINTER(Timing)],
    "deeds_ctx_upsample_deformations": [V, V, V, V, V, V, V, I, I, I, I, I, I],
    "deeds_ctx_consistent_mapping": [V, V, V, V, V, V, V, I, I, I, I],
    "deeds_ctx_jacobian": [V, V, V, V, I, I, I, I, V, V],
    "deeds_ctx_warp_affine": [V, V, V, V, V, V, V, V, I, I, I, V],
    "deeds_ctx_warp_image": [V, V, V, V, V, V, V, I, I, I, V],
    "deeds_ctx_warp_affine_s": [V, V, V, V, V, V, V, I, I, I],
    "deeds_ctx_apply_seg": [V, V, V, V, V, I, I, I, I],
    "deeds_ctx_apply_float": [V, V, V, V, V, I, I, I, I],
    "deeds_soft_saved_count": [I, I],
    "deeds_ctx_soft_features": [V, I, V, V, I, I, I, I],
    "deeds_ctx_soft_features_backward": [V, I, V, V, V, I, I, I, I],
    "deeds_ctx_soft_cost": [V, I, V, V, V, I, I, I, I, I, F, F],
    "deeds_ctx_soft_cost_backward": [V, I, V, V, V, V, V, I, I, I, I, I, F, F],
    "deeds_ctx_soft_bp": [V, I, V, V, V, V, V, V, V, I, I, I, I, I, F, F],
    "deeds_ctx_soft_bp_backward": [V, I, V, V, V, V, V, V, V, V, I, I, I, I, I, F, F],
}
_RESTYPES = {"deeds_version": c_char_p, "deeds_last_error": c_char_p, "deeds_ctx_last_error": c_char_p,
             "deeds_soft_saved_count": ctypes.c_size_t, "deeds_tuning_report": ctypes.c_size_t}

_lib = None


class DeedsError(RuntimeError):
    """A deeds_hip call returned a nonzero status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def _init_torch_first():
    """PyTorch-ROCm ships its own HIP runtime next to the system one this
    library links; both sit on the same HSA runtime (so device pointers are
    shared), but torch's fails to find the GPU if the system runtime was
    initialised first.  When torch is importable, bring its runtime up first."""
    try:
        import torch
    except Exception:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def load(path: Path | str | None = None):
    """Load (once) and return the ctypes library with all signatures set."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"deeds_hip library not found at {p}; build it with `python -m deedsbcv_amd.build`")
    _init_torch_first()
    lib = ctypes.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    if path is None:
        _lib = lib
    return lib


def call(name, *args):
    """Call `name`, raising DeedsError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.deeds_last_error()
        raise DeedsError(name, rc, msg.decode() if msg else "")
    return rc


def ctx_call(ctx, name, *args):
    lib = load()
    rc = getattr(lib, name)(ctx, *args)
    if rc != 0:
        msg = lib.deeds_ctx_last_error(ctx)
        raise DeedsError(name, rc, msg.decode() if msg else "")
    return rc


def make_schedule(grid, radius, quant, alpha=1.6):
    if not (len(grid) == len(radius) == len(quant)) or not 1 <= len(grid) <= MAX_LEVELS:
        raise ValueError("grid, radius and quant must have the same length (1..10)")
    s = Schedule()
    s.levels = len(grid)
    for i, (g, r, q) in enumerate(zip(grid, radius, quant)):
        s.grid[i], s.radius[i], s.quant[i] = int(g), int(r), int(q)
    s.alpha = float(alpha)
    return s
