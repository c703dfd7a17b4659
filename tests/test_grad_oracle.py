"""Pins the FP64 PyTorch restatement of the differentiable descriptor (CPU).

The restatement (oracle/desc_torch64.py) is the reference gradient for the
device backward, so it is checked here against the C oracle's float32
pyramid and descriptor (same algorithm, other arithmetic) and against float64
central finite differences of its own loss.
"""
import math

import numpy as np
import pytest
import torch

from oracle import desc_torch64 as T
from oracle import oracle as O
from pysift_amd import synthetic as syn


def test_pyramid_matches_c_oracle_single_size_base():
    img = syn.synth_gray(72, 88, seed=11)
    p = O.params(n_octaves=3, double_image=False)
    octs = T.pyramid(torch.from_numpy(img.astype(np.float64) / 255.0), 3)
    for o in range(3):
        g, _ = O.pyramid_octave(img, p, o)
        ref = octs[o].numpy()
        assert ref.shape == g[:4].shape
        assert np.abs(g[:4] - ref).max() < 3e-6, o


def test_pyramid_matches_c_oracle_doubled_base():
    img = syn.synth_gray(40, 52, seed=12)
    p = O.params(n_octaves=3, double_image=True)
    octs = T.pyramid(torch.from_numpy(img.astype(np.float64) / 255.0), 3, doubled=True)
    for o in range(3):
        g, _ = O.pyramid_octave(img, p, o)
        ref = octs[o].numpy()
        assert ref.shape == g[:4].shape
        assert np.abs(g[:4] - ref).max() < 3e-6, o


def test_reflect_blur_is_scipy_reflect():
    import scipy.ndimage as nd
    x = np.random.default_rng(2).random((9, 30))
    for sigma in (0.8, 2.4, 3.3):
        ref = nd.gaussian_filter1d(nd.gaussian_filter1d(x, sigma, axis=1, mode="reflect", truncate=4.0),
                                   sigma, axis=0, mode="reflect", truncate=4.0)
        got = T.blur(torch.from_numpy(x), sigma).numpy()
        assert np.abs(got - ref).max() < 1e-12


@pytest.mark.parametrize("si,ang", [(1, 0.0), (2, 37.5), (3, 301.0)])
def test_descriptor_matches_c_oracle(si, ang):
    img = syn.synth_gray(96, 96, seed=12)
    p = O.params(n_octaves=1, double_image=False)
    g, _ = O.pyramid_octave(img, p, 0)
    ph = T.patch_half_sizes()[si]
    ri, ci = np.float32(47.6), np.float32(44.2)
    ang_rad = float(np.float32(np.float32(ang) * np.float32(0.0174532924)))
    ref, _ = O.descriptor(g[si], ri, ci, ang_rad, ph)
    got = T.describe(torch.from_numpy(g[si].astype(np.float64)), ph, ri, ci, np.float32(ang)).numpy()
    assert np.abs(got - ref).max() < 2e-5
    assert abs(np.linalg.norm(got) - 1.0) < 1e-7   # RootSIFT: 1 - 1e-7 / (|v|_1 + 1e-7)


def test_patch_half_sizes_match_reference_formula():
    # max(4, round(3 sigma sqrt2 2.5)) with sigma = 1.6 k^si: 17, 21, 27, 34
    assert T.patch_half_sizes() == [17, 21, 27, 34]


def test_autograd_matches_float64_central_differences():
    # the descriptor is piecewise smooth (clip at 0.2, orientation-bin edges):
    # at eps 1e-5 the central difference stays inside one piece and pins the
    # autograd gradient tightly; at eps 1e-3 a perturbation can straddle a kink
    # and the difference quotient itself moves by up to a few percent there
    img = syn.synth_gray(64, 64, seed=13).astype(np.float64) / 255.0
    kps = syn.random_keypoints(1, 10, 64, 64, n_oct=3, seed=14, margin=3)[0]
    Wt = np.random.default_rng(15).standard_normal((10, 128))
    _, g = T.loss_and_grad(img, kps, Wt, n_oct=3)
    rng = np.random.default_rng(16)
    pix = [(int(y), int(x)) for y, x in zip(rng.integers(0, 64, 24), rng.integers(0, 64, 24))]
    an = np.array([float(g[y, x]) for y, x in pix])
    fd5 = T.central_difference(img, kps, Wt, pix, eps=1e-5, n_oct=3)
    assert np.linalg.norm(an - fd5) <= 1e-5 * np.linalg.norm(fd5)
    fd3 = T.central_difference(img, kps, Wt, pix, eps=1e-3, n_oct=3)
    assert np.linalg.norm(an - fd3) <= 0.1 * np.linalg.norm(fd3)


def test_central_difference_locality_radius_is_exact():
    # keypoints far from the pixel do not change: the restricted sum equals the full one
    img = syn.synth_gray(64, 64, seed=17).astype(np.float64) / 255.0
    kps = syn.random_keypoints(1, 8, 64, 64, n_oct=2, seed=18, margin=3)[0]
    Wt = np.random.default_rng(19).standard_normal((8, 128))
    pix = [(10, 12), (40, 50)]
    full = T.central_difference(img, kps, Wt, pix, n_oct=2)
    local = T.central_difference(img, kps, Wt, pix, n_oct=2, radius=200)
    assert np.array_equal(full, local)


@pytest.mark.parametrize("shape", [(3, 3), (2, 5), (17, 29), (64, 80), (101, 67)])
def test_fp64_spline_upsample_matches_scipy_zoom(shape):
    # the doubled base of the gradient path: scipy.ndimage.zoom(x, 2, order=3)
    # (mirror-boundary cubic B-spline, grid_mode=False) in float64
    import scipy.ndimage as nd
    x = np.random.default_rng(shape[0] * 7 + shape[1]).random(shape)
    up = T.upsample2x(torch.from_numpy(x)).numpy()
    np.testing.assert_allclose(up, nd.zoom(x, 2.0, order=3), rtol=0, atol=1e-13)


def test_doubled_pyramid_levels_and_fd_checked_pixels():
    # doubled base = blur(zoom(x), sqrt(1.6^2 - 1)); the autograd gradient of a
    # doubled-base descriptor loss agrees with central differences on the
    # checked pixels
    img = syn.synth_float_batch(1, 40, 48, seed=91)[0].astype(np.float64)
    octs = T.pyramid(torch.from_numpy(img), 3, doubled=True)
    assert octs[0].shape == (4, 80, 96) and octs[1].shape == (4, 40, 48)
    ref0 = T.blur(T.upsample2x(torch.from_numpy(img)), math.sqrt(1.6 ** 2 - 1.0))
    np.testing.assert_allclose(octs[0][0].numpy(), ref0.numpy(), rtol=0, atol=1e-15)
    kps = syn.random_keypoints(1, 12, 80, 96, n_oct=3, seed=92, margin=2)[0]
    Wt = np.random.default_rng(93).standard_normal((12, 128))
    _, g = T.loss_and_grad(img, kps, Wt, n_oct=3, rootsift=False, doubled=True)
    pix = [(5, 7), (20, 30), (33, 11), (12, 40), (25, 25), (38, 45)]
    fd, ck = T.central_difference(img, kps, Wt, pix, eps=1e-4, n_oct=3, rootsift=False, return_checked=True,
                                  doubled=True)
    an = np.array([g[y, x].item() for y, x in pix])
    assert ck.sum() >= 2
    np.testing.assert_allclose(an[ck], fd[ck], rtol=1e-5, atol=1e-9)
