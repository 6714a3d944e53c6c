"""GPU parity of the per-label Dice (diceMulti.cpp:17-79): the library against
the oracle bit for bit, including negative labels, a label set too large for
the LDS counters, one label only and disjoint volumes; the diceMulti command
line against the unmodified reference binary's output."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deedsbcv_amd import api, nifti
from test_oracle import _dice_text, _label_pair

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _eq(got, want):
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


@pytest.mark.parametrize("case", [(1, 6, 0), (2, 13, -40), (3, 2, 1000)])
def test_dice_vs_oracle(oracle, case):
    seed, k, offset = case
    print(f"case: dice seed={seed} labels={k} offset={offset}", flush=True)
    a, b = _label_pair((61, 47, 39), seed, k, offset)
    _eq(api.dice(a, b), oracle.dice(a, b))


def test_dice_edge_cases(oracle):
    print("case: dice edge cases", flush=True)
    rng = np.random.default_rng(9)
    # more labels than the LDS counters hold (global-atomic path), full int16 range
    a = rng.integers(-32768, 32767, (40, 40, 40), endpoint=True).astype(np.int16)
    b = a.copy()
    b[rng.random(a.shape) < 0.3] = 7
    _eq(api.dice(a, b), oracle.dice(a, b))
    # a single label: nothing but the background
    z = np.zeros((9, 8, 7), np.int16)
    labels, d = api.dice(z, z)
    assert labels.size == 0 and d.size == 0
    # disjoint
    c, _ = _label_pair((30, 20, 10), 4, 5, 0)
    _eq(api.dice(c, (c + 50).astype(np.int16)), oracle.dice(c, (c + 50).astype(np.int16)))


def test_dice_cli_matches_reference_binary(oracle, tmp_path):
    print("case: diceMulti cli", flush=True)
    exe = ROOT / "deedsbcv_amd" / "bin" / "diceMulti"
    if not exe.exists():
        pytest.skip("tools not built")
    a, b = _label_pair((45, 37, 29), 11, 9, 0)
    nifti.write_nifti(tmp_path / "a.nii.gz", a)
    nifti.write_nifti(tmp_path / "b.nii.gz", b)
    r = subprocess.run([str(exe), str(tmp_path / "a.nii.gz"), str(tmp_path / "b.nii.gz")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout == _dice_text(oracle.dice(a, b))
    ref = ROOT / "oracle" / "_ref" / "diceMulti"
    if ref.exists():
        rr = subprocess.run([str(ref), str(tmp_path / "a.nii.gz"), str(tmp_path / "b.nii.gz")], capture_output=True,
                            text=True, timeout=120)
        assert r.stdout == rr.stdout
