"""CLAHE (core.py:2562-2565, the cv2.createCLAHE(...).apply that runs ahead of
detection in _extract_features): the HIP kernels against the CPU oracle,
bit for bit, over tile grids, clip limits, ragged and tiny images, row
pitches, in-place use, a natural image, and the stitcher's use of it."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "imc")


def _apply(img, clip, tiles):
    import torch
    from pysift_amd.clahe import createCLAHE
    out = createCLAHE(clip, tiles).apply(torch.from_numpy(np.array(img)).cuda())
    return out.cpu().numpy()


@pytest.mark.parametrize("H,W,clip,tiles", [
    (768, 1024, 2.0, (8, 8)),      # the stitch views' size, the stitcher's settings
    (97, 131, 2.0, (8, 8)),        # ragged: padded both ways
    (61, 60, 0.0, (4, 4)),         # ragged in H only (W still padded), no clipping
    (5, 7, 2.0, (8, 8)),           # smaller than the tile grid
    (1, 1, 2.0, (8, 8)),
    (300, 200, 40.0, (3, 5)),      # OpenCV's default clip, non-square grid
    (640, 480, 1.0, (16, 16)),
    (1000, 1500, 3.0, (1, 1)),     # one tile
    (257, 4099, 2.0, (128, 2)),    # widest grid, narrow tiles
    (1000, 4099, 2.0, (128, 128)), # > 64 LUTs per apply block: read from global memory
])
def test_clahe_bitexact(H, W, clip, tiles):
    img = syn.synth_gray(H, W, H + 7 * W) if min(H, W) >= 8 else np.random.default_rng(H * W).integers(
        0, 256, (H, W)).astype(np.uint8)
    print(f"[case] clahe {H}x{W} clip {clip} tiles {tiles}", flush=True)
    np.testing.assert_array_equal(_apply(img, clip, tiles), O.clahe(img, clip, tiles))


def test_clahe_low_contrast_and_constant():
    rng = np.random.default_rng(4)
    dim = (rng.normal(60, 4, (480, 640)).clip(0, 255)).astype(np.uint8)   # heavy clipping
    np.testing.assert_array_equal(_apply(dim, 2.0, (8, 8)), O.clahe(dim, 2.0, (8, 8)))
    const = np.full((64, 64), 100, np.uint8)
    out = _apply(const, 2.0, (8, 8))
    assert (out == 108).all()


def test_clahe_natural_image():
    from PIL import Image
    im = Image.open(os.path.join(GOLDEN, "imc_british_museum.jpg"))
    g = np.asarray(im.convert("L"))
    print(f"[case] clahe natural image {g.shape}", flush=True)
    np.testing.assert_array_equal(_apply(g, 2.0, (8, 8)), O.clahe(g, 2.0, (8, 8)))


def test_clahe_pitch_and_in_place():
    import torch
    from pysift_amd.clahe import CLAHE
    img = syn.synth_gray(203, 131, 77)
    exp = O.clahe(img, 2.0, (8, 8))
    c = CLAHE(2.0, (8, 8))
    for pad in (1, 3, 61):   # odd pitches take the byte path, 4-aligned ones the word path
        big = torch.zeros((203, 131 + pad), dtype=torch.uint8, device="cuda")
        big[:, :131] = torch.from_numpy(img).cuda()
        view = big[:, :131]
        assert view.stride(0) == 131 + pad
        print(f"[case] clahe pitch {131 + pad}", flush=True)
        np.testing.assert_array_equal(c.apply(view).cpu().numpy(), exp)
    t = torch.from_numpy(img).cuda()
    r = c.apply(t, dst=t)
    assert r.data_ptr() == t.data_ptr()
    np.testing.assert_array_equal(t.cpu().numpy(), exp)
    # a second image through the same object (the workspace is reused)
    img2 = syn.synth_gray(480, 640, 78)
    np.testing.assert_array_equal(c.apply(torch.from_numpy(img2).cuda()).cpu().numpy(), O.clahe(img2, 2.0, (8, 8)))


def test_clahe_c_abi_errors():
    import torch
    from pysift_amd import _lib
    L = _lib.lib()
    n = ctypes.c_size_t()
    assert L.msift_clahe_workspace_bytes(8, 8, ctypes.byref(n)) == 0 and n.value >= 64 * 256 * 5
    assert L.msift_clahe_workspace_bytes(0, 8, ctypes.byref(n)) == _lib.MSIFT_EINVAL
    src = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    s = _lib.stream_ptr()
    call = lambda H, W, pitch, clip, tx, ty, wsb: L.msift_clahe(_lib.ptr(src), H, W, pitch, clip, tx, ty, _lib.ptr(ws),
                                                               wsb, _lib.ptr(dst), s)
    assert call(16, 16, 15, 2.0, 8, 8, n.value) == _lib.MSIFT_EINVAL            # pitch < W
    assert call(16, 16, 16, float("nan"), 8, 8, n.value) == _lib.MSIFT_EINVAL   # clip NaN
    assert call(16, 16, 16, 2.0, 129, 8, n.value) == _lib.MSIFT_EINVAL          # grid too wide
    assert call(16, 16, 16, 2.0, 8, 8, 100) == _lib.MSIFT_EWORKSPACE
    assert L.msift_clahe(_lib.ptr(src), 16, 16, 16, 2.0, 8, 8, ctypes.c_void_p(ws.data_ptr() + 1), n.value - 1,
                         _lib.ptr(dst), s) == _lib.MSIFT_EINVAL                  # misaligned workspace
    assert call(16, 16, 16, 2.0, 8, 8, n.value) == 0
    torch.cuda.synchronize()
    from pysift_amd.clahe import CLAHE
    with pytest.raises(ValueError):
        CLAHE(2.0, (0, 8))
    with pytest.raises(ValueError):
        CLAHE(2.0, (8, 8)).apply(torch.zeros((4, 4, 3), dtype=torch.uint8, device="cuda"))


def test_stitch_features_use_clahe():
    # _extract_features: gray -> CLAHE(2.0, 8 x 8) -> detect, as the oracle
    import torch
    from pysift_amd import GPUPyStitch
    va, _, _ = syn.synth_pair(seed=1001, view_hw=(240, 320), canvas_hw=(480, 640))
    st = GPUPyStitch(verbose=False, save_output=False, n_octaves=4)
    kp, du, hw = st._extract_features(torch.from_numpy(va).cuda())
    rk, _, ru = O.detect_describe(O.clahe(O.bgr_to_gray(va), 2.0, (8, 8)), O.params(n_octaves=4))
    assert hw == (240, 320)
    np.testing.assert_array_equal(kp.cpu().numpy(), rk)
    np.testing.assert_array_equal(du.cpu().numpy(), ru)
