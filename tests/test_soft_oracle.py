"""Pins of the dense float64 soft-variant checker (oracle/soft_dense.py), on the CPU.

The soft variant has no reference counterpart; these tests tie its checker
to what does pin it: the oracle's reference-pinned descriptor codes and
plain regularisation, explicit loops, exhaustive enumeration of tiny trees,
and central finite differences.
"""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import soft_dense as sd  # noqa: E402

from deedsbcv_amd import phantom  # noqa: E402

THRESH = np.array([np.float32(-math.log((c + 1.5) / 6.0)) for c in range(5)], np.float32)


def quantize(phi):
    """MINDSSCbox.h:202-217 applied to float features (o, n, m, 12)."""
    v = (THRESH[None, None, None, None, :] > phi[..., None]).sum(-1).astype(np.uint64)
    fields = (np.uint64(1) << v) - np.uint64(1)
    shifts = (np.uint64(5) * np.arange(12, dtype=np.uint64))
    return np.bitwise_or.reduce(fields << shifts, axis=-1)


@pytest.mark.parametrize("qs", [1, 2])
def test_features_threshold_to_oracle_codes(oracle, qs):
    f, _ = phantom.make_pair((30, 26, 22), 11)
    phi = sd.features(torch.from_numpy(f.astype(np.float64)), qs).numpy()
    codes = oracle.descriptor(f, qs)
    # the oracle sums in float with sequential prefix sums; the checker in float64,
    # so only values within rounding of a threshold may quantise differently
    agree = (quantize(phi.astype(np.float32)) == codes).mean()
    assert agree > 0.999, agree
    assert np.allclose(phi.sum(-1), 12.0, atol=1e-9)
    assert phi.min() >= 0


def test_cost_matches_explicit_loop():
    rng = np.random.default_rng(3)
    m, n, o, step, hw, q, alpha = 11, 10, 10, 5, 1, 2, 1.6
    F = rng.random((o, n, m, 12))
    M = rng.random((o, n, m, 12))
    C = sd.cost(torch.from_numpy(F), torch.from_numpy(M), step, hw, q, alpha).numpy()
    si, sj, sk, scale = sd.cost_plan(step, q, alpha)
    assert (si, sj, sk) == (1, 2, 2)  # dataCostD.h:141-147 with one random sample level
    L = 2 * hw + 1
    m1, n1, o1 = m // step, n // step, o // step
    want = np.zeros_like(C)
    for z in range(o1):
        for x in range(n1):
            for y in range(m1):
                node = y + x * m1 + z * m1 * n1
                for lz in range(L):
                    for lx in range(L):
                        for ly in range(L):
                            acc = 0.0
                            for k in range(0, step, sk):
                                for j in range(0, step, sj):
                                    for i in range(0, step, si):
                                        fi, fj, fk = y * step + i, x * step + j, z * step + k
                                        mi = min(max(fi + (ly - hw) * q, 0), m - 1)
                                        mj = min(max(fj + (lx - hw) * q, 0), n - 1)
                                        mk = min(max(fk + (lz - hw) * q, 0), o - 1)
                                        acc += ((F[fk, fj, fi] - M[mk, mj, mi]) ** 2).sum()
                            want[node, ly + lx * L + lz * L * L] = scale * acc
    assert np.allclose(C, want, rtol=1e-13, atol=0)


def _tiny_trees():
    # (parents, ordered) of a chain and a star, root first
    return [([-1, 0, 1], [0, 1, 2]), ([1, -1, 1, 1], [1, 0, 2, 3])]


@pytest.mark.parametrize("tree", _tiny_trees())
@pytest.mark.parametrize("T", [0.5, 2.0])
def test_bp_matches_enumeration(tree, T):
    parents, ordered = tree
    N, hw, quant = len(parents), 1, 2
    rng = np.random.default_rng(N + int(T * 10))
    C = rng.random((N, 27)) * 4
    u0 = rng.normal(size=(3, N)) * 1.5
    w = rng.uniform(0.3, 1.0, N).astype(np.float32)
    got = sd.bp(torch.from_numpy(C), torch.from_numpy(u0), ordered, parents, w, hw, quant, T).numpy()
    want = sd.brute_force_bp(C, u0, parents, w, hw, quant, T)
    assert np.allclose(got, want, rtol=0, atol=1e-12)


def test_bp_zero_temperature_is_plain_minsum(oracle):
    """With unit edge weights the plain energy (regularisation.h:197-232) and
    the soft energy coincide, so as T -> 0 the expected displacement is the
    plain path's argmin displacement."""
    f, _ = phantom.make_pair((20, 20, 20), 5)
    step, hw, quant = 4, 2, 2
    ordered, parents, _ = oracle.prims(f, step)
    N = ordered.size
    ones = np.ones(N, np.float32)
    rng = np.random.default_rng(9)
    C = (rng.random((N, 125)) * 20).astype(np.float32)
    u0 = (rng.normal(size=(3, N)) * 2).astype(np.float32)
    u1 = oracle.regularisation(C, u0[0], u0[1], u0[2], hw, step, quant, ordered, parents, ones, (20, 20, 20))
    got = sd.bp(torch.from_numpy(C.astype(np.float64)), torch.from_numpy(u0.astype(np.float64)), ordered, parents,
                ones, hw, quant, 1e-5).numpy()
    want = np.stack(u1).astype(np.float64)
    assert np.abs(got - want).max() < 1e-4


def test_dense_gradients_match_finite_differences():
    """Autograd through the checker agrees with central differences (h = 1e-4)."""
    f, mv = phantom.make_pair((9, 10, 11), 21)
    fx = torch.from_numpy(f.astype(np.float64)).requires_grad_(True)
    mo = torch.from_numpy(mv.astype(np.float64)).requires_grad_(True)
    step, hw, quant, T = 4, 1, 1, 0.5
    ordered, parents, edgemst = (np.array([0, 1, 2, 3, 4, 5, 6, 7]),
                                 np.array([-1, 0, 0, 1, 2, 3, 4, 5]),
                                 np.linspace(0.4, 1.0, 8).astype(np.float32))
    rng = np.random.default_rng(1)
    wgt = torch.from_numpy(rng.normal(size=(3, 8)))

    def loss(a, b):
        C = sd.cost(sd.features(a, 1), sd.features(b, 1), step, hw, quant)
        return (sd.bp(C, torch.zeros(3, 8, dtype=torch.float64), ordered, parents, edgemst, hw, quant, T) * wgt).sum()

    L = loss(fx, mo)
    gf, gm = torch.autograd.grad(L, (fx, mo))
    h = 1e-4
    for trial in range(2):
        df = torch.from_numpy(rng.normal(size=f.shape))
        dm = torch.from_numpy(rng.normal(size=f.shape))
        with torch.no_grad():
            fd = (loss(fx + h * df, mo + h * dm) - loss(fx - h * df, mo - h * dm)) / (2 * h)
        an = (gf * df).sum() + (gm * dm).sum()
        assert abs(float(fd - an)) <= 1e-6 * max(1.0, abs(float(an))), (float(fd), float(an))
