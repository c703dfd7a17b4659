"""The reference's composition (gain + DP seam + multiband, core.py:3686-3714,
3783-3802) on the GPU against the CPU oracle, bit for bit: panorama, valid
mask, softened seam mask, gains, seam orientation and seam RMSE."""
import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _gpu(va, vb, H, cw, ch, ox, oy, levels=6, color_match="gain"):
    import torch
    from pysift_amd import GPUPyStitch
    st = GPUPyStitch(verbose=False, save_output=False, blend="multiband", n_blend_levels=levels,
                     color_match=color_match)
    b1, b2 = torch.from_numpy(va).cuda(), torch.from_numpy(vb).cuda()
    pano, valid, sm, stats = st.composite_multiband(b1, b2, H, cw, ch, ox, oy, seam_mask=True)
    return pano.cpu().numpy(), valid.cpu().numpy().astype(bool), sm.cpu().numpy(), stats


def _check(va, vb, H, levels=6, tag="", color_match="gain"):
    cw, ch, ox, oy = O.canvas(va.shape[0], va.shape[1], vb.shape[0], vb.shape[1], H)
    print(f"[case] composite {tag} canvas {cw}x{ch} levels {levels} {color_match}", flush=True)
    p, v, sm, stats = _gpu(va, vb, H, cw, ch, ox, oy, levels, color_match)
    rp, rv, rsm, rg, rr, rvert = O.composite(va, vb, H, cw, ch, ox, oy, levels, color_match)
    np.testing.assert_array_equal(v, rv)
    np.testing.assert_array_equal(stats["gain"], rg)
    assert stats["vertical"] == rvert
    np.testing.assert_array_equal(sm, rsm)
    np.testing.assert_array_equal(p, rp)
    if np.isnan(rr):
        assert np.isnan(stats["rmse"])
    else:
        assert stats["rmse"] == rr
    return stats


@pytest.mark.parametrize("color_match", ["gain", "hist"])
def test_composite_stitch_pair(color_match):
    va, vb, Hgt = syn.synth_pair(seed=1000, view_hw=(384, 512), canvas_hw=(768, 1024))
    st = _check(va, vb, Hgt, 6, "side by side", color_match)
    assert st["vertical"] == 1 and st["n_overlap"] > 10000


def test_composite_stacked_views_horizontal_seam():
    rng = np.random.default_rng(3)
    canvas = syn.synth_gray(500, 300, 11, as_float=True)
    tint = rng.uniform(0.8, 1.2, 3)
    img = np.clip(canvas[..., None] * tint, 0, 255).astype(np.uint8)
    va, vb = img[:300].copy(), img[180:].copy()
    H = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, -180.0], [0.0, 0.0, 1.0]])   # view a's pixels in view b's frame
    st = _check(va, vb, H, 5, "stacked")
    assert st["vertical"] == 0
    _check(va, vb, H, 5, "stacked", "hist")


@pytest.mark.parametrize("levels", [1, 3, 8])
def test_composite_levels(levels):
    va, vb, Hgt = syn.synth_pair(seed=1001, view_hw=(200, 260), canvas_hw=(400, 520))
    _check(va, vb, Hgt, levels, "levels")


def test_composite_without_overlap():
    rng = np.random.default_rng(4)
    va = rng.integers(0, 256, (60, 80, 3)).astype(np.uint8)
    vb = rng.integers(0, 256, (60, 80, 3)).astype(np.uint8)
    H = np.array([[1.0, 0.0, 200.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])   # far apart: no overlap
    st = _check(va, vb, H, 4, "disjoint")
    assert st["vertical"] == -1 and np.isnan(st["rmse"])


def test_stitch_multiband_matches_oracle():
    from pysift_amd import GPUPyStitch
    va, vb, Hgt = syn.synth_pair(seed=1000, view_hw=(384, 512), canvas_hw=(768, 1024))
    st = GPUPyStitch(inlier_thresh=3.0, ransac_iters=2000, verbose=False, save_output=False, blend="multiband")
    pano = st.stitch(va, vb)
    H = st.last_homography
    cw, ch, ox, oy = O.canvas(384, 512, 384, 512, H)
    rp, rv, _, _, rr, _ = O.composite(va, vb, H, cw, ch, ox, oy, 6)
    np.testing.assert_array_equal(pano, O.autocrop(rp, rv))
    assert st.last_metrics["seam_rmse"] == rr
