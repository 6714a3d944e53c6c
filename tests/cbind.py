"""ctypes front-ends with one numpy API over the three C libraries the tests use.

* ``Oracle``  -- oracle/liboracle.so, the C restatement (checker).
* ``Ref``     -- oracle/_ref/libdeedsref.so, the unmodified reference headers
                 compiled where they lie (present only where it was built).
Both expose the same methods, so a test can run one case through either.
Volumes are numpy arrays of shape (o, n, m), C order (reference voxel order).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libdeedsref.so"

c_int, c_float, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_void_p


def P(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def dims_of(vol):
    o, n, m = vol.shape
    return m, n, o


class _Lib:
    prefix = ""

    def __init__(self, path):
        self.path = Path(path)
        self.lib = ctypes.CDLL(str(self.path))
        self._setup()

    def f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _setup(self):
        I, F, V = c_int, c_float, c_void_p
        sig = {
            "boxfilter": None,
            "descriptor": [V, V, I, I, I, I],
            "datacost": [V, V, V, I, I, I, I, I, F, F, I],
            "prims": [V, V, V, V, I, I, I, I],
            "message_dt": None,
            "regularisation": None,
            "interp3": [V, V, V, V, V, I, I, I, I, I, I, I],
            "upsample": [V, V, V, V, V, V, I, I, I, I, I, I],
            "consistent": [V, V, V, V, V, V, I, I, I, I],
            "warp_affine": [V, V, V, V, V, V, V, I, I, I, V],
            "warp_image": [V, V, V, V, V, V, I, I, I, V],
            "warp_affine_s": [V, V, V, V, V, V, I, I, I],
        }
        for k, a in sig.items():
            if a is not None:
                fn = self.f(k)
                fn.argtypes = a
                fn.restype = None

    # ---- MIND -----------------------------------------------------------
    def descriptor(self, im, qs):
        im = _f32(im)
        out = np.zeros(im.shape, np.uint64)
        m, n, o = dims_of(im)
        self.f("descriptor")(P(out), P(im), m, n, o, qs)
        return out

    # ---- data cost ------------------------------------------------------
    def datacost(self, F, M, step, hw, quant, alpha=1.6, randnum=1):
        F = np.ascontiguousarray(F, np.uint64)
        M = np.ascontiguousarray(M, np.uint64)
        m, n, o = dims_of(F)
        len3 = (2 * hw + 1) ** 3
        sz1 = (m // step) * (n // step) * (o // step)
        cost = np.zeros((sz1, len3), np.float32)
        self.f("datacost")(P(F), P(M), P(cost), m, n, o, step, hw, float(quant), float(alpha), randnum)
        return cost

    # ---- MST ------------------------------------------------------------
    def prims(self, im, step):
        im = _f32(im)
        m, n, o = dims_of(im)
        sz1 = (m // step) * (n // step) * (o // step)
        ordered = np.zeros(sz1, np.int32)
        parents = np.zeros(sz1, np.int32)
        edgemst = np.zeros(sz1, np.float32)
        self.f("prims")(P(im), P(ordered), P(parents), P(edgemst), step, m, n, o)
        return ordered, parents, edgemst

    # ---- transforms -----------------------------------------------------
    def interp3(self, vol, x1, y1, z1, add_identity):
        vol = _f32(vol)
        x1, y1, z1 = _f32(x1), _f32(y1), _f32(z1)
        m2, n2, o2 = dims_of(vol)
        m, n, o = dims_of(x1)
        out = np.zeros(x1.shape, np.float32)
        self.f("interp3")(P(out), P(vol), P(x1), P(y1), P(z1), m, n, o, m2, n2, o2, int(add_identity))
        return out

    def upsample(self, u0, v0, w0, shape_mno):
        u0, v0, w0 = _f32(u0), _f32(v0), _f32(w0)
        m2, n2, o2 = dims_of(u0)
        m, n, o = shape_mno
        out = [np.zeros((o, n, m), np.float32) for _ in range(3)]
        self.f("upsample")(P(out[0]), P(out[1]), P(out[2]), P(u0), P(v0), P(w0), m, n, o, m2, n2, o2)
        return out

    def consistent(self, u, v, w, u2, v2, w2, factor):
        a = [_f32(x).copy() for x in (u, v, w, u2, v2, w2)]
        m, n, o = dims_of(a[0])
        self.f("consistent")(*[P(x) for x in a], m, n, o, factor)
        return a

    def warp_affine(self, inp, im1b, X, u, v, w):
        inp, im1b, X = _f32(inp), _f32(im1b), _f32(X)
        u, v, w = _f32(u), _f32(v), _f32(w)
        m, n, o = dims_of(inp)
        out = np.zeros(inp.shape, np.float32)
        ssd = np.zeros(2, np.float32)
        self.f("warp_affine")(P(out), P(inp), P(im1b), P(X), P(u), P(v), P(w), m, n, o, P(ssd))
        return out, ssd

    def warp_image(self, im1, im1b, u, v, w):
        im1, im1b = _f32(im1), _f32(im1b)
        u, v, w = _f32(u), _f32(v), _f32(w)
        m, n, o = dims_of(im1)
        out = np.zeros(im1.shape, np.float32)
        ssd = np.zeros(2, np.float32)
        self.f("warp_image")(P(out), P(im1), P(im1b), P(u), P(v), P(w), m, n, o, P(ssd))
        return out, ssd

    def warp_affine_s(self, seg, X, u, v, w):
        seg = np.ascontiguousarray(seg, np.int16)
        X = _f32(X)
        u, v, w = _f32(u), _f32(v), _f32(w)
        m, n, o = dims_of(seg)
        out = np.zeros(seg.shape, np.int16)
        self.f("warp_affine_s")(P(out), P(seg), P(X), P(u), P(v), P(w), m, n, o)
        return out


class _LinearFns:
    """lin_warp / qrsolve / affine_robust / estimate_affine shared by the oracle and
    the reference harness (same signatures up to the prefix)."""

    def _lin_setup(self):
        I, F, V = c_int, c_float, c_void_p
        p = self.prefix
        for name, args in (("lin_warp", [V, V, V, I, I, I]), ("qrsolve", [V, V, V, I, I]),
                           ("affine_robust", [V, V, V, I, I])):
            fn = getattr(self.lib, p + name)
            fn.argtypes = args
            fn.restype = None
        fn = getattr(self.lib, p + "estimate_affine")
        fn.argtypes = [V, V, V, V, V, V, I, I, I, I, F, I, I] + ([V] if p == "orc_" else [])
        fn.restype = None

    def lin_warp(self, im, X):
        im = _f32(im)
        m, n, o = dims_of(im)
        out = np.zeros(im.shape, np.float32)
        getattr(self.lib, self.prefix + "lin_warp")(P(out), P(im), P(_f32(X)), m, n, o)
        return out

    def qrsolve(self, A, b, length, len2):
        X = np.zeros(4 * len2, np.float32)
        getattr(self.lib, self.prefix + "qrsolve")(P(X), P(_f32(A)), P(_f32(b)), length, len2)
        return X

    def affine_robust(self, pts1, pts2, length, rigid):
        RT = np.zeros(16, np.float32)
        getattr(self.lib, self.prefix + "affine_robust")(P(RT), P(_f32(pts1)), P(_f32(pts2)), length, int(rigid))
        return RT

    def estimate_affine(self, Xprev, im1, im2, cost, cost2, step, quant, hw, rigid):
        im1, im2 = _f32(im1), _f32(im2)
        m, n, o = dims_of(im1)
        c1, c2 = _f32(cost).copy(), _f32(cost2).copy()
        X = np.zeros(16, np.float32)
        extra = [None] if self.prefix == "orc_" else []
        getattr(self.lib, self.prefix + "estimate_affine")(P(X), P(_f32(Xprev)), P(im1), P(im2), P(c1), P(c2), m, n,
                                                           o, step, float(quant), hw, int(rigid), *extra)
        return X, c1, c2


class RefLinear(_LinearFns):
    """oracle/_ref/libdeedsreflin.so: the unmodified linearBCV.cpp functions."""
    prefix = "ref_"

    def __init__(self, path):
        self.lib = ctypes.CDLL(str(path))
        self._lin_setup()


class Oracle(_Lib, _LinearFns):
    prefix = "orc_"

    def _setup(self):
        super()._setup()
        I, F, V = c_int, c_float, c_void_p
        L = self.lib
        L.orc_boxfilter.argtypes = [V, I, I, I, I]
        L.orc_message_dt.argtypes = [V, V, I, F, F, F]
        L.orc_regularisation.argtypes = [V, V, V, V, V, V, V, I, I, F, V, V, V, I, I, I, V, V]
        L.orc_mind_thresholds.argtypes = [V]
        L.orc_datacost_scale.argtypes = [I, F, F, I, V]
        L.orc_datacost_scale.restype = F
        L.orc_edge_sad.argtypes = [V, V, V, I, I, I, I]
        L.orc_image_stats.argtypes = [V, ctypes.c_longlong, V, V]
        L.orc_jacobian.argtypes = [V, V, V, I, I, I, I, V]
        L.orc_jacobian.restype = F
        L.orc_register.argtypes = [V, V, V, I, I, I, I, V, V, V, F, V, V, I]
        L.orc_register.restype = I
        L.orc_ssd_channel.argtypes = [V, V, I, I, I, I, I]
        L.orc_dice.argtypes = [V, V, ctypes.c_longlong, I, V, V]
        L.orc_dice.restype = I
        L.orc_linear_register.argtypes = [V, V, I, I, I, I, V, V, V, I, V]
        L.orc_linear_register.restype = None
        L.orc_estimate_affine.argtypes = [V, V, V, V, V, V, I, I, I, I, F, I, I, V]
        L.orc_estimate_affine.restype = None
        for k in ("boxfilter", "message_dt", "regularisation", "mind_thresholds", "edge_sad",
                  "image_stats", "ssd_channel"):
            getattr(L, "orc_" + k).restype = None
        self._lin_setup()

    def boxfilter(self, vol, hw):
        v = _f32(vol).copy()
        m, n, o = dims_of(v)
        self.lib.orc_boxfilter(P(v), hw, m, n, o)
        return v

    def mind_thresholds(self):
        t = np.zeros(5, np.float32)
        self.lib.orc_mind_thresholds(P(t))
        return t

    def datacost_scale(self, step, quant, alpha, randnum=1):
        sk = np.zeros(3, np.int32)
        s = self.lib.orc_datacost_scale(step, float(quant), float(alpha), randnum, P(sk))
        return np.float32(s), tuple(int(x) for x in sk)

    def edge_sad(self, im, step):
        im = _f32(im)
        m, n, o = dims_of(im)
        sz1 = (m // step) * (n // step) * (o // step)
        ec = np.zeros(6 * sz1, np.float32)
        nb = np.zeros(6 * sz1, np.int32)
        self.lib.orc_edge_sad(P(im), P(ec), P(nb), step, m, n, o)
        return ec, nb

    def image_stats(self, im):
        im = _f32(im)
        a = np.zeros(1, np.float32)
        b = np.zeros(1, np.float32)
        self.lib.orc_image_stats(P(im), im.size, P(a), P(b))
        return a[0], b[0]

    def message_dt(self, cube, offx, offy, offz):
        c = _f32(cube).copy()
        ln = round(c.size ** (1.0 / 3.0))
        inds = np.zeros(c.size, np.int16)
        self.lib.orc_message_dt(P(c), P(inds), ln, float(offx), float(offy), float(offz))
        return c, inds

    def regularisation(self, cost, u0, v0, w0, hw, step, quant, ordered, parents, edgemst, dims_mno,
                       want_inds=False):
        cost = _f32(cost).copy()
        u0, v0, w0 = _f32(u0), _f32(v0), _f32(w0)
        sz1 = u0.size
        u1, v1, w1 = (np.zeros(u0.shape, np.float32) for _ in range(3))
        ordered = np.ascontiguousarray(ordered, np.int32)
        parents = np.ascontiguousarray(parents, np.int32)
        edgemst = _f32(edgemst)
        inds = np.zeros(cost.shape, np.int16) if want_inds else None
        sel = np.zeros(sz1, np.int32) if want_inds else None
        m, n, o = dims_mno
        self.lib.orc_regularisation(P(cost), P(u0), P(v0), P(w0), P(u1), P(v1), P(w1), hw, step, float(quant),
                                    P(ordered), P(parents), P(edgemst), m, n, o, P(inds), P(sel))
        if want_inds:
            return u1, v1, w1, inds, sel
        return u1, v1, w1

    def jacobian(self, u, v, w, factor):
        u, v, w = _f32(u), _f32(v), _f32(w)
        m, n, o = dims_of(u)
        neg = np.zeros(1, np.float32)
        s = self.lib.orc_jacobian(P(u), P(v), P(w), m, n, o, factor, P(neg))
        return np.float32(s), neg[0]

    def linear_register(self, fixed, moving, grid, radius, quant, rigid=False):
        fixed, moving = _f32(fixed), _f32(moving)
        m, n, o = dims_of(fixed)
        X = np.zeros(16, np.float32)
        self.lib.orc_linear_register(P(fixed), P(moving), m, n, o, len(grid), P(np.asarray(grid, np.int32)),
                                     P(np.asarray(radius, np.int32)), P(np.asarray(quant, np.int32)), int(rigid),
                                     P(X))
        return X

    def dice(self, seg1, seg2):
        a = np.ascontiguousarray(seg1, np.int16)
        b = np.ascontiguousarray(seg2, np.int16)
        k = self.lib.orc_dice(P(a), P(b), a.size, 0, None, None)
        labels = np.zeros(k, np.int16)
        dice = np.zeros(k, np.float32)
        self.lib.orc_dice(P(a), P(b), a.size, k, P(labels), P(dice))
        return labels, dice

    def register(self, fixed, moving, grid, radius, quant, alpha=1.6, X=None, want_warped=True):
        fixed, moving = _f32(fixed), _f32(moving)
        m, n, o = dims_of(fixed)
        if X is None:
            X = np.eye(4, dtype=np.float32).T.ravel()  # column-wise identity
        X = _f32(X)
        L = len(grid)
        g = np.asarray(grid, np.int32)
        r = np.asarray(radius, np.int32)
        q = np.asarray(quant, np.int32)
        step = grid[-1]
        sz1 = (m // step) * (n // step) * (o // step)
        flow = np.zeros(3 * sz1, np.float32)
        warped = np.zeros(fixed.shape, np.float32) if want_warped else None
        rc = self.lib.orc_register(P(fixed), P(moving), P(X), m, n, o, L, P(g), P(r), P(q), float(alpha),
                                   P(flow), P(warped), 0)
        assert rc == sz1
        return flow, warped


class Ref(_Lib):
    prefix = "ref_"

    def _setup(self):
        super()._setup()
        I, F, V = c_int, c_float, c_void_p
        L = self.lib
        L.ref_boxfilter.argtypes = [V, I, I, I, I]
        L.ref_message_dt.argtypes = [I, V, V, I, F, F, F]
        L.ref_regularisation.argtypes = [V, V, V, V, V, V, V, I, I, F, V, V, V, I, I, I]
        L.ref_jacobian.argtypes = [V, V, V, I, I, I, I]
        L.ref_jacobian.restype = F
        L.ref_distances.argtypes = [V, V, I, I, I, I, I]
        for k in ("boxfilter", "message_dt", "regularisation", "distances"):
            getattr(L, "ref_" + k).restype = None

    def boxfilter(self, vol, hw):
        v = _f32(vol).copy()
        m, n, o = dims_of(v)
        self.lib.ref_boxfilter(P(v), hw, m, n, o)
        return v

    def message_dt(self, cube, offx, offy, offz):
        c = _f32(cube).copy()
        ln = round(c.size ** (1.0 / 3.0))
        inds = np.zeros(c.size, np.int16)
        self.lib.ref_message_dt(0, P(c), P(inds), ln, float(offx), float(offy), float(offz))
        return c, inds

    def regularisation(self, cost, u0, v0, w0, hw, step, quant, ordered, parents, edgemst, dims_mno):
        cost = _f32(cost).copy()
        u0, v0, w0 = _f32(u0), _f32(v0), _f32(w0)
        u1, v1, w1 = (np.zeros(u0.shape, np.float32) for _ in range(3))
        ordered = np.ascontiguousarray(ordered, np.int32).copy()
        parents = np.ascontiguousarray(parents, np.int32).copy()
        edgemst = _f32(edgemst).copy()
        m, n, o = dims_mno
        self.lib.ref_regularisation(P(cost), P(u0), P(v0), P(w0), P(u1), P(v1), P(w1), hw, step, float(quant),
                                    P(ordered), P(parents), P(edgemst), m, n, o)
        return u1, v1, w1

    def jacobian(self, u, v, w, factor):
        u, v, w = _f32(u), _f32(v), _f32(w)
        m, n, o = dims_of(u)
        return np.float32(self.lib.ref_jacobian(P(u), P(v), P(w), m, n, o, factor))


def load_oracle():
    return Oracle(ORACLE_SO)


def load_ref():
    return Ref(REF_SO) if REF_SO.exists() else None


REF_LIN_SO = ROOT / "oracle" / "_ref" / "libdeedsreflin.so"


def load_ref_linear():
    return RefLinear(REF_LIN_SO) if REF_LIN_SO.exists() else None
