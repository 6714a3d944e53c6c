"""Dense float64 restatement of the soft deeds variant -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's CPU baseline use this module, as
the checker; the product path (deedsbcv_amd) never imports it.

The soft variant has no counterpart in deedsBCV, so this restates its
definition (the docstring of deedsbcv_amd/soft.py) directly with torch ops in
float64 on the CPU, written differently from the HIP kernels wherever a shared
mistake could hide: the box filter is a sum of zero-padded shifted slices, the
cost a full gather, and each BP message a log-sum-exp over the dense
L^3 x L^3 pairwise matrix (no separable passes, no minimum shifts).  Autograd
through it gives gradients independent of the hand-written adjoints.

How it is pinned (tests/test_soft_oracle.py):
  features  thresholding phi with the thresholds of MINDSSCbox.h:176-180
            reproduces the oracle's (reference-pinned) descriptor codes except
            at threshold boundaries; phi sums to 12 per voxel
  cost      an explicit per-sample loop on a tiny case
  BP        exact marginals by enumerating every labeling of tiny trees; the
            T -> 0 limit with unit edge weights equals the plain min-sum
            labels of the oracle's regularisation (reference-pinned)
  gradients central finite differences of this module in float64
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import torch

# (di, dj, dk) of the six SSD shifts in units of qs (MINDSSCbox.h:110-112: dy, dx, dz)
SSD_SHIFT = [(+1, +1, 0), (-1, +1, 0), (0, -1, +1), (-1, 0, +1), (0, +1, +1), (+1, 0, +1)]
# (di, dj, dk, channel) of the twelve SSC values (MINDSSCbox.h:136-140: sy, sx, sz, index)
SSC = [(0, -1, 0, 0), (-1, 0, 0, 0), (0, -1, 0, 1), (+1, 0, 0, 1), (0, 0, -1, 2), (0, +1, 0, 2),
       (0, 0, -1, 3), (+1, 0, 0, 3), (0, 0, -1, 4), (0, -1, 0, 4), (0, 0, -1, 5), (-1, 0, 0, 5)]
MIN_EDGE_WEIGHT = np.float32(1e-6)


def _shift_or_self(x, di, dj, dk):
    """y[k, j, i] = x[k+dk, j+dj, i+di] where that voxel exists, else x[k, j, i]."""
    o, n, m = x.shape
    k = torch.arange(o).view(o, 1, 1)
    j = torch.arange(n).view(1, n, 1)
    i = torch.arange(m).view(1, 1, m)
    kk, jj, ii = k + dk, j + dj, i + di
    ok = (kk >= 0) & (kk < o) & (jj >= 0) & (jj < n) & (ii >= 0) & (ii < m)
    kk = torch.where(ok, kk, k).expand(o, n, m)
    jj = torch.where(ok, jj, j).expand(o, n, m)
    ii = torch.where(ok, ii, i).expand(o, n, m)
    return x[kk, jj, ii]


def _box1(x, r, dim):
    """Window sum of radius r along dim, the window clipped to the volume."""
    L = x.shape[dim]
    zs = list(x.shape)
    zs[dim] = r
    z = torch.zeros(zs, dtype=x.dtype)
    xp = torch.cat([z, x, z], dim=dim)
    out = xp.narrow(dim, 0, L)
    for t in range(1, 2 * r + 1):
        out = out + xp.narrow(dim, t, L)
    return out


def features(im, qs):
    """phi (o, n, m, 12) of a float64 image (o, n, m)."""
    D = []
    for di, dj, dk in SSD_SHIFT:
        S = (_shift_or_self(im, di * qs, dj * qs, dk * qs) - im) ** 2
        D.append(_box1(_box1(_box1(S, qs, 2), qs, 1), qs, 0))
    d = torch.stack([_shift_or_self(D[c], di * qs, dj * qs, dk * qs) for di, dj, dk, c in SSC], dim=-1)
    mn = d.min(dim=-1, keepdim=True).values
    noise = ((d - mn).sum(dim=-1, keepdim=True) / 12.0).clamp_min(1e-6)
    return (d - mn) / noise


def cost_plan(step, quant, alpha, randnum=1):
    """(skip_i, skip_j, skip_k, scale) of dataCostD.h:141-163 with the float
    rounding of the scale a1 (float32 arithmetic, as the C code)."""
    sz = sx = sy = 1
    if step > 4:
        if randnum > 0:
            sz = sx = 2
        if randnum > 1:
            sy = 2
    if randnum > 1 and step > 7:
        sz = sx = sy = 3
    if step == 4 and randnum > 1:
        sz = 2
    f = np.float32
    maxsamp = f(math.ceil(step / sx) * math.ceil(step / sz) * math.ceil(step / sy))
    alphai = f(step) / (f(alpha) * f(quant))
    scale = f(0.5 * float(alphai) / float(maxsamp))
    return sy, sx, sz, float(scale)


def cost(F, M, step, hw, quant, alpha=1.6):
    """C (N, L^3) from features (o, n, m, 12) by a full gather."""
    o, n, m, _ = F.shape
    si, sj, sk, scale = cost_plan(step, quant, alpha)
    m1, n1, o1 = m // step, n // step, o // step
    L = 2 * hw + 1
    q = int(quant)
    y, x, z = np.meshgrid(np.arange(m1), np.arange(n1), np.arange(o1), indexing="ij")
    y, x, z = (a.transpose(2, 1, 0).ravel() for a in (y, x, z))  # node = y + x*m1 + z*m1*n1
    ii, jj, kk = np.meshgrid(np.arange(0, step, si), np.arange(0, step, sj), np.arange(0, step, sk), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    ly, lx, lz = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    ly, lx, lz = (a.transpose(2, 1, 0).ravel() for a in (ly, lx, lz))  # l = ly + lx*L + lz*L^2
    fi = (y * step)[:, None] + ii[None, :]
    fj = (x * step)[:, None] + jj[None, :]
    fk = (z * step)[:, None] + kk[None, :]
    fidx = fi + fj * m + fk * m * n  # (N, S)
    mi = np.clip(fi[:, None, :] + ((ly - hw) * q)[None, :, None], 0, m - 1)
    mj = np.clip(fj[:, None, :] + ((lx - hw) * q)[None, :, None], 0, n - 1)
    mk = np.clip(fk[:, None, :] + ((lz - hw) * q)[None, :, None], 0, o - 1)
    midx = mi + mj * m + mk * m * n  # (N, L3, S)
    Ff = F.reshape(-1, 12)
    Mf = M.reshape(-1, 12)
    diff = Ff[torch.from_numpy(fidx)][:, None, :, :] - Mf[torch.from_numpy(midx)]
    return scale * (diff ** 2).sum(dim=(2, 3))


def label_disp(hw):
    """(L^3, 3) label displacement in label units, columns (u, v, w) = (xs, ys, zs) - hw."""
    L = 2 * hw + 1
    l = np.arange(L ** 3)
    return np.stack([(l // L) % L - hw, l % L - hw, l // (L * L) - hw], axis=1).astype(np.float64)


def pairwise(u0, c, p, w, hw, quant):
    """R[a, b] = |disp_a - disp_b + (u0[:, p] - u0[:, c]) / quant|^2 / w for
    parent label a and child label b (regularisation.h:208-210 offsets)."""
    disp = torch.from_numpy(label_disp(hw))
    off = (u0[:, p] - u0[:, c]) / quant
    diff = disp[:, None, :] - disp[None, :, :] + off[None, None, :]
    return (diff ** 2).sum(-1) / w


def bp(C, u0, ordered, parents, edgemst, hw, quant, T):
    """u1 (3, N) of sum-product BP on the tree, dense messages."""
    N, L3 = C.shape
    ordered = [int(v) for v in ordered]
    parents = [int(v) for v in parents]
    children = {v: [] for v in range(N)}
    for c in ordered[1:]:
        children[parents[c]].append(c)
    R = {}
    for c in ordered[1:]:
        w = float(max(np.float32(edgemst[c]), MIN_EDGE_WEIGHT))
        R[c] = pairwise(u0, c, parents[c], w, hw, quant)
    h, msg = {}, {}
    for c in reversed(ordered):
        hc = C[c]
        for k in children[c]:
            hc = hc + msg[k]
        h[c] = hc
        if parents[c] >= 0:
            msg[c] = -T * torch.logsumexp(-(hc[None, :] + R[c]) / T, dim=1)
    g = {ordered[0]: h[ordered[0]]}
    for c in ordered[1:]:
        y = g[parents[c]] - msg[c]
        g[c] = h[c] - T * torch.logsumexp(-(y[:, None] + R[c]) / T, dim=0)
    disp = torch.from_numpy(label_disp(hw)) * quant
    E = torch.stack([torch.softmax(-g[c] / T, dim=0) @ disp for c in range(N)], dim=1)  # (3, N)
    return u0 + E


def brute_force_bp(C, u0, parents, edgemst, hw, quant, T):
    """u1 (3, N) by enumerating every labeling (tiny N and hw only), numpy float64."""
    C = np.asarray(C, np.float64)
    u0 = np.asarray(u0, np.float64)
    N, L3 = C.shape
    disp = label_disp(hw)
    logp = np.zeros((L3,) * N)
    for labels in itertools.product(range(L3), repeat=N):
        e = sum(C[v, labels[v]] for v in range(N))
        for c in range(N):
            p = int(parents[c])
            if p < 0:
                continue
            w = float(max(np.float32(edgemst[c]), MIN_EDGE_WEIGHT))
            d = disp[labels[p]] - disp[labels[c]] + (u0[:, p] - u0[:, c]) / quant
            e += float((d ** 2).sum()) / w
        logp[labels] = -e / T
    P = np.exp(logp - logp.max())
    P /= P.sum()
    u1 = u0.copy()
    for v in range(N):
        axes = tuple(a for a in range(N) if a != v)
        marg = P.sum(axis=axes)
        u1[:, v] += marg @ disp * quant
    return u1
