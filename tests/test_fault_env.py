"""CPU check of the environment form of the fault hook: DEEDS_FAULT_INJECT is
parsed when the library loads, and deeds_inject_fault rejects unknown kinds.
No GPU call is made."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_fault_hook_api_without_gpu():
    from deedsbcv_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    code = ("import ctypes,sys; sys.path.insert(0, %r); from deedsbcv_amd import _lib; lib=_lib.load(); "
            "assert lib.deeds_inject_fault(1, 5) == 0; assert lib.deeds_inject_fault(2, -1) == 0; "
            "assert lib.deeds_inject_fault(9, 0) == 1; print('ok')" % str(ROOT))
    env = {"DEEDS_FAULT_INJECT": "alloc:2", "PATH": "/usr/bin:/bin", "HIP_VISIBLE_DEVICES": ""}
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
