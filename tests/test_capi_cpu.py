"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public header declares, and validates arguments without touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pysift_amd.h")).read()
    return sorted(set(re.findall(r"^int\s+(msift_\w+)\s*\(", src, re.M)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for n in ["msift_create", "msift_detect_describe", "msift_match_u8", "msift_ransac_homography",
              "msift_warp_feather"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from pysift_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in _declared():
        assert hasattr(L, n), n
    assert set(_declared()) == set(_lib.EXPORTS)


def test_create_validates_parameters():
    from pysift_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    good = _lib.MsiftParams(5, 3, 1.6, 0.04, 10.0, 1, 1, 16_000_000, 25000, 8000, 5, 100, 1)
    assert L.msift_create(ctypes.byref(good), ctypes.byref(h)) == 0
    n = ctypes.c_size_t()
    assert L.msift_workspace_bytes(h, 480, 640, ctypes.byref(n)) == 0
    # the doubled octave 0 alone holds S+3 = 6 float levels (24 B per pixel)
    assert n.value > 480 * 640 * 4 * 24
    assert L.msift_destroy(h) == 0
    bad = _lib.MsiftParams(0, 3, 1.6, 0.04, 10.0, 1, 1, 16_000_000, 25000, 8000, 5, 100, 1)
    assert L.msift_create(ctypes.byref(bad), ctypes.byref(h)) == _lib.MSIFT_EINVAL
    assert "n_octaves" in _lib.last_error()
    huge = _lib.MsiftParams(5, 3, 9.0, 0.04, 10.0, 1, 1, 16_000_000, 25000, 8000, 5, 100, 1)
    assert L.msift_create(ctypes.byref(huge), ctypes.byref(h)) == _lib.MSIFT_EINVAL


def test_pysift_constructor_errors():
    from pysift_amd import PySIFT
    with pytest.raises(ValueError):
        PySIFT(dsp_n_scales=7)


def test_ransac_rejects_lo_with_fewer_than_four_inliers():
    # the LO refit solves a DLT on the inliers; min_inliers < 4 would hand it
    # 0-3 points (the reference's _compute_homography asserts >= 4)
    from pysift_amd import _lib
    L = _lib.lib()
    rc = L.msift_ransac_homography(None, None, 10, 42, 16, 3.0, 3.0, 0, 1, 2, None, 0, None, None, None, None, None)
    assert rc == _lib.MSIFT_EINVAL
    assert "min_inliers" in _lib.last_error()
    rc = L.msift_ransac_homography(None, None, 3, 42, 16, 3.0, 3.0, 0, 1, 10, None, 0, None, None, None, None, None)
    assert rc == _lib.MSIFT_EINVAL and "4 correspondences" in _lib.last_error()


def test_stitch_int8_matching_mapping():
    # int8_matching selects the exact INT8 matchers; the approximate float-path
    # INT8 similarity is refused rather than silently ignored
    from pysift_amd.stitch import GPUPyStitch
    with pytest.raises(NotImplementedError):
        GPUPyStitch(int8_matching=True, match_method="float", verbose=False, save_output=False)


def test_match_tile_setter_and_method_validation():
    # the two-pass tile setter accepts shapes 0-4 and 1-64 chunks; the matcher
    # rejects unknown methods before touching the device
    from pysift_amd import _lib
    L = _lib.lib()
    for shape, cpw in [(5, 4), (-1, 4), (0, 0), (0, 65)]:
        assert L.msift_match_set_tile(shape, cpw) == _lib.MSIFT_EINVAL
        assert "chunks per workgroup" in _lib.last_error()
    assert L.msift_match_set_tile(2, 8) == 0
    assert L.msift_match_set_tile(0, 4) == 0   # back to the default
    rc = L.msift_match_u8(None, 10, None, 10, 0.75, 3, None, 0, None, None, None, None, None)
    assert rc == _lib.MSIFT_EINVAL
