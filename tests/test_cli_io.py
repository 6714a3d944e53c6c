"""CPU test of the command lines' NIfTI writer (deedsbcv_amd/csrc/tool_common.h):
the output is several gzip members compressed on host threads, and zlib's
gzread (the reference's reader, imageIOgzType.h:7-105) and Python's gzip both
read back exactly the header and voxel bytes written."""
import gzip
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

DRIVER = r"""
#include "tool_common.h"
int main(int argc, char** argv) {
    int m = atoi(argv[2]), n = atoi(argv[3]), o = atoi(argv[4]);
    std::vector<char> h(352, 0);
    h[0] = 92; h[1] = 1;  /* sizeof_hdr 348 */
    std::vector<float> d((size_t)m * n * o);
    for (size_t i = 0; i < d.size(); i++) d[i] = (float)((i * 2654435761u) % 1000) / 7.0f;
    if (!tool::write_nifti(argv[1], h, d.data(), m, n, o, 16, 32, 4)) return 1;
    tool::Volume v;                     /* read back through zlib's gzread */
    if (!tool::read_nifti(argv[1], v, false)) return 2;
    if (v.m != m || v.n != n || v.o != o) return 3;
    for (size_t i = 0; i < d.size(); i++) if (v.f[i] != d[i]) return 4;
    return 0;
}
"""


@pytest.mark.parametrize("dims", [(256, 256, 128), (7, 5, 3)])
def test_parallel_gzip_members_round_trip(tmp_path, dims):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    (tmp_path / "t.cpp").write_text(DRIVER)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "deedsbcv_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(tmp_path / "t.cpp"), "-o", str(exe), "-lz", "-lpthread"], check=True)
    out = tmp_path / "o.nii.gz"
    r = subprocess.run([str(exe), str(out), *map(str, dims)])
    assert r.returncode == 0, f"gzread round trip failed ({r.returncode})"
    raw = out.read_bytes()
    m, n, o = dims
    members = raw.count(b"\x1f\x8b\x08")  # gzip member headers (a lower bound check only)
    if m * n * o * 4 > 4 << 20:
        assert members >= 2
    data = gzip.decompress(raw)
    assert struct.unpack_from("<5h", data, 40) == (3, m, n, o, 1)
    assert struct.unpack_from("<hh", data, 70) == (16, 32)
    i = np.arange(m * n * o, dtype=np.uint64)
    want = ((i * np.uint64(2654435761)) % np.uint64(1000)).astype(np.float32) / np.float32(7.0)
    assert np.array_equal(np.frombuffer(data[352:], np.float32), want)
