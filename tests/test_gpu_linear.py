"""GPU parity of the affine pre-alignment (linearBCV.cpp:144-263): the whole
multi-level estimation through the C ABI against the oracle bit for bit
(affine and rigid, a MIND radius change between levels, non-cubic volumes),
and the linearBCV command line against the unmodified reference binary's
_matrix.txt and _deformed_seg.nii.gz."""
import gzip
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deedsbcv_amd import api, nifti, phantom

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
CASES = [((64, 56, 48), 17, (6, 5, 4), (3, 2, 2), (2, 2, 1)),
         ((57, 61, 45), 23, (7, 5), (2, 2), (3, 1))]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("rigid", [0, 1])
def test_linear_vs_oracle(oracle, case, rigid):
    shape, seed, g, r, q = CASES[case]
    print(f"case: linear {shape} seed={seed} rigid={rigid}", flush=True)
    f, mv = phantom.make_pair(shape, seed)
    X, used = api.linear_register(f, mv, g, r, q, rigid)
    want = oracle.linear_register(f.astype(np.float32), mv.astype(np.float32), g, r, q, rigid)
    assert np.array_equal(X.view(np.uint32), want.view(np.uint32)), (X, want)
    assert np.all(used > 0)


def test_linear_errors():
    print("case: linear errors", flush=True)
    f, mv = phantom.make_pair((40, 40, 40), 1)
    with pytest.raises(ValueError):
        api.linear_register(f, mv[:, :, :30])
    with pytest.raises(api.DeedsError):
        api.linear_register(f[:3, :3, :3], mv[:3, :3, :3], (2,), (1,), (4,))


def test_linearBCV_cli_matches_reference_binary(tmp_path):
    exe = ROOT / "deedsbcv_amd" / "bin" / "linearBCV"
    ref = ROOT / "oracle" / "_ref" / "linearBCV"
    if not exe.exists():
        pytest.skip("tools not built")
    if not ref.exists():
        pytest.skip("reference linearBCV not built")
    print("case: linearBCV cli", flush=True)
    f, mv = phantom.make_pair((64, 56, 48), 29)
    seg = (mv > 0).astype(np.int16) + (mv > 150).astype(np.int16)
    nifti.write_nifti(tmp_path / "f.nii.gz", f)
    nifti.write_nifti(tmp_path / "m.nii.gz", mv)
    nifti.write_nifti(tmp_path / "s.nii.gz", seg)
    for rigid in ("0", "1"):
        outs = []
        for tag, prog in (("ours", exe), ("ref", ref)):
            stem = tmp_path / f"{tag}{rigid}"
            r = subprocess.run([str(prog), "-F", str(tmp_path / "f.nii.gz"), "-M", str(tmp_path / "m.nii.gz"), "-O",
                                str(stem), "-G", "6x5x4", "-L", "3x2x2", "-Q", "2x2x1", "-R", rigid, "-S",
                                str(tmp_path / "s.nii.gz")], capture_output=True, text=True, timeout=300)
            assert r.returncode == 0, r.stdout + r.stderr
            with gzip.open(f"{stem}_deformed_seg.nii.gz", "rb") as fh:
                outs.append(((tmp_path / f"{tag}{rigid}_matrix.txt").read_text(), fh.read()))
        assert outs[0][0] == outs[1][0]
        assert outs[0][1] == outs[1][1]
