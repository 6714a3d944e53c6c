"""CPU checks of the C ABI: the library loads and exports every function the
header declares (no GPU call is made)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "deeds_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(deeds_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for fn in ("deeds_descriptor", "deeds_datacost", "deeds_prims_graph", "deeds_regularisation",
               "deeds_upsample_deformations", "deeds_consistent_mapping", "deeds_warp_affine", "deeds_warp_image",
               "deeds_warp_affine_s", "deeds_jacobian", "deeds_register"):
        assert fn in names


def test_library_exports_every_declared_symbol():
    from deedsbcv_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a Python signature
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_version_string_without_gpu():
    from deedsbcv_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    assert b"gfx950" in _lib.load().deeds_version()


def test_no_oracle_in_product_sources():
    """The product must never route through the checker."""
    for p in list((ROOT / "deedsbcv_amd").rglob("*.py")) + list((ROOT / "deedsbcv_amd" / "csrc").glob("*")):
        if p.is_file():
            t = p.read_text(errors="ignore")
            assert "liboracle" not in t and "deeds_oracle" not in t and "orc_" not in t, p
