"""Pins the CPU oracle to independent references (CPU only, no GPU).

Each check is chosen so that a plausible mistake in the oracle (wrong sign,
index, boundary, constant, dropped term) fails it: SciPy for the blur and the
spline zoom, closed forms for the entropy threshold, quadratic refinement,
ramp orientations and homographies, brute force NumPy restatements for the
extrema, descriptor histogram and matcher, the Random123 known-answer vectors
for Philox, and the reference's own RTX 3050 keypoint fixtures.
"""
import math
import os

import numpy as np
import pytest
import scipy.ndimage as nd

from oracle import oracle as O
from pysift_amd import synthetic as syn

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------- elementary --
def test_exp_log2_atan2_sincos_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-5, 5, 2000), [0.0, 1e-300, -1e-10]])
    for x in xs:
        ref = math.exp(x)
        assert abs(O.exp(x) - ref) <= 4e-16 * ref
    assert O.exp(-800.0) == 0.0
    for x in np.concatenate([rng.uniform(1e-300, 1.0, 500), rng.uniform(1, 1e6, 500)]):
        assert abs(O.log2(x) - math.log2(x)) <= 1e-15 * max(1.0, abs(math.log2(x)))
    for y, x in rng.normal(size=(3000, 2)) * rng.choice([1e-3, 1.0, 1e3], (3000, 1)):
        assert abs(O.atan2(y, x) - math.atan2(y, x)) <= 5e-16 * max(1.0, abs(math.atan2(y, x)))
    for y, x in [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (2.0, 2.0)]:
        assert O.atan2(y, x) == math.atan2(y, x)
    for x in np.concatenate([rng.uniform(0, 2 * math.pi, 2000), rng.uniform(-50, 50, 500)]):
        s, c = O.sincos(x)
        assert abs(s - math.sin(x)) <= 4e-16 and abs(c - math.cos(x)) <= 4e-16


def _ulp_err(a, ref):
    ref32 = np.abs(ref).astype(np.float32)
    return np.abs(a.astype(np.float64) - ref) / np.spacing(ref32).astype(np.float64)


def test_float_exp_within_one_ulp():
    rng = np.random.default_rng(5)
    x = np.concatenate([np.linspace(-87, 88, 1_000_001, dtype=np.float32),
                        rng.uniform(-10, 0, 500_000).astype(np.float32),
                        np.float32([0.0, -0.0, 1e-30, -1e-30, 88.0, -87.0])])
    y = O.exp_f(x)
    assert _ulp_err(y, np.exp(x.astype(np.float64))).max() <= 1.0
    assert O.exp_f(np.float32([0.0]))[0] == 1.0
    assert O.exp_f(np.float32([-87.5, -1e4]))[0] == 0.0
    assert np.isinf(O.exp_f(np.float32([89.0]))[0])
    # e^(a+b) = e^a e^b up to rounding: catches a wrong ln2 split or 2^n scale
    a = rng.uniform(-20, 20, 10000).astype(np.float32)
    assert np.allclose(O.exp_f(a) * O.exp_f(-a), 1.0, rtol=4e-7)


def test_float_atan2_within_two_ulp_and_octants():
    rng = np.random.default_rng(6)
    n = 1_000_000
    y = (rng.standard_normal(n) * rng.choice([1e-4, 1.0, 300.0], n)).astype(np.float32)
    x = (rng.standard_normal(n) * rng.choice([1e-4, 1.0, 300.0], n)).astype(np.float32)
    a = O.atan2_f(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert _ulp_err(a, ref).max() <= 2.0
    # axes, signed zeros, the |x| == |y| diagonals and infinities
    sy = np.float32([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 2.0, -2.0, np.inf, -np.inf, 1.0, np.inf])
    sx = np.float32([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, 2.0, 2.0, -2.0, -2.0, 1.0, 1.0, -np.inf, -np.inf])
    got = O.atan2_f(sy, sx)
    want = np.arctan2(sy.astype(np.float64), sx.astype(np.float64))
    assert np.all(np.signbit(got) == np.signbit(want))
    assert _ulp_err(got, want).max() <= 2.0


# ---------------------------------------------------------------- pyramid --
def test_taps_match_reference_formula():
    # core.py:561-567: radius int(4 sigma + 0.5); float32 exp taps normalised
    for s, r in [(1.2490, 5), (1.2263, 5), (1.5450, 6), (1.9466, 8), (2.4525, 10), (3.0900, 12), (1.5199, 6)]:
        assert O.radius(s) == r
        t = O.taps(s)
        x = np.arange(-r, r + 1, dtype=np.float32)
        k = np.exp(-0.5 * (x / np.float32(s)) ** 2).astype(np.float32)
        k /= k.sum()
        # the reference evaluates exp in float32 (a few ulp at the tails); the
        # oracle rounds the double-precision taps once
        np.testing.assert_allclose(t, k, rtol=2e-6, atol=0)
        x64 = np.arange(-r, r + 1) / s
        k64 = np.exp(-0.5 * x64 * x64); k64 /= k64.sum()
        np.testing.assert_array_equal(t, k64.astype(np.float32))
        assert abs(float(t.astype(np.float64).sum()) - 1.0) < 1e-6


@pytest.mark.parametrize("shape", [(37, 53), (9, 11), (3, 40), (1, 7), (64, 64)])
def test_blur_matches_scipy_reflect(shape):
    x = np.random.default_rng(1).random(shape).astype(np.float32)
    for s in [0.3, 1.2490, 1.5450, 3.0900]:
        ref = nd.gaussian_filter1d(nd.gaussian_filter1d(x.astype(np.float64), s, axis=1, mode="reflect", truncate=4.0),
                                   s, axis=0, mode="reflect", truncate=4.0)
        np.testing.assert_allclose(O.blur(x, s), ref, rtol=0, atol=4e-7)
    # sigma below 0.01 returns the input unchanged (core.py:559-560)
    np.testing.assert_array_equal(O.blur(x, 0.005), x)


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (3, 7), (5, 4), (31, 47), (48, 64)])
def test_upsample_matches_scipy_zoom(shape):
    x = np.random.default_rng(2).random(shape).astype(np.float32)
    ref = nd.zoom(x.astype(np.float64), 2.0, order=3)
    out = O.upsample2x(x)
    assert out.shape == ref.shape
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-7)


def test_pyramid_structure():
    img = syn.synth_gray(90, 120, 3)
    p = O.params(n_octaves=4)
    prev_g = None
    for o in range(4):
        g, d = O.pyramid_octave(img, p, o)
        assert g.shape == (6,) + ((180 + (1 << o) - 1) // (1 << o), (240 + (1 << o) - 1) // (1 << o))
        np.testing.assert_array_equal(d, g[1:] - g[:-1])       # DoG = G[s+1] - G[s]
        if prev_g is not None:
            np.testing.assert_array_equal(g[0], prev_g[3][::2, ::2])   # next base = G[S][::2, ::2]
        prev_g = g
    # level 0 of octave 0 = blur(zoom(u8/255), sqrt(1.6^2 - 1))
    g0, _ = O.pyramid_octave(img, p, 0)
    unit = img.astype(np.float32) / np.float32(255.0)
    np.testing.assert_array_equal(g0[0], O.blur(O.upsample2x(unit), math.sqrt(1.6 ** 2 - 1.0)))
    # incremental sigma of level 1: sigma0 sqrt(k^2 - 1)
    k = 2 ** (1 / 3)
    np.testing.assert_array_equal(g0[1], O.blur(g0[0], math.sqrt((1.6 * k) ** 2 - 1.6 ** 2)))


def test_adaptive_threshold_closed_forms():
    flat = np.full((16, 16), 9, np.uint8)                        # entropy 0 -> 0.8
    assert abs(O.adaptive_thresh(flat) - 0.04 * 0.8) < 1e-15
    two = np.zeros((16, 16), np.uint8); two[:8] = 200             # entropy 1 bit -> 0.825
    assert abs(O.adaptive_thresh(two) - 0.04 * 0.825) < 1e-15
    uni = np.arange(256, dtype=np.uint8).reshape(16, 16)           # entropy 8 -> min(1, 1.0)
    assert abs(O.adaptive_thresh(uni) - 0.04) < 1e-15
    q = np.repeat(np.arange(4, dtype=np.uint8), 64).reshape(16, 16)  # 4 equal bins -> 2 bits
    assert abs(O.adaptive_thresh(q) - 0.04 * 0.85) < 1e-15


# ---------------------------------------------------------------- extrema --
def test_extrema_match_brute_force():
    rng = np.random.default_rng(4)
    D = nd.gaussian_filter(rng.normal(size=(5, 60, 70)), (0, 1.2, 1.2)).astype(np.float32)
    thr = np.float32(0.02)
    got = O.extrema_octave(D, thr, border=5)
    mx = nd.maximum_filter(D, size=3, mode="nearest")
    mn = nd.minimum_filter(D, size=3, mode="nearest")
    ok = ((D >= mx) | (D <= mn)) & (np.abs(D) > thr)
    ok[[0, 4]] = False
    ok[:, :5] = False; ok[:, -5:] = False; ok[:, :, :5] = False; ok[:, :, -5:] = False
    ref = np.argwhere(ok).astype(np.int32)        # raster order (s, r, c)
    assert len(ref) > 20
    np.testing.assert_array_equal(got, ref)
    # the per-octave cap keeps the first candidates in raster order
    np.testing.assert_array_equal(O.extrema_octave(D, thr, cap=7), ref[:7])


# ----------------------------------------------------------------- refine --
def _bump(s0, r0, c0, a=0.2, b=(0.01, 0.004, 0.006), shape=(5, 40, 40)):
    s, r, c = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    return (a - b[0] * (s - s0) ** 2 - b[1] * (r - r0) ** 2 - b[2] * (c - c0) ** 2).astype(np.float32)


def test_refine_quadratic_bump_exact_offset():
    # for a quadratic D the Newton step lands on the analytic maximum at once
    for s0, r0, c0 in [(2.2, 20.3, 17.8), (1.9, 19.6, 22.4), (2.0, 20.0, 20.0)]:
        D = _bump(s0, r0, c0)
        cand = np.array([[int(round(s0)), int(round(r0)), int(round(c0))]])
        valid, out = O.refine_octave(D, cand, thr=0.01)
        assert valid[0]
        assert abs(out[0, 0] - r0) < 2e-5 and abs(out[0, 1] - c0) < 2e-5
        assert int(out[0, 2]) == int(round(s0))
        assert abs(out[0, 3] - 0.2) < 1e-5                   # response = peak value
    # far start: moves by whole pixels until converged
    D = _bump(2.0, 20.3, 17.8)
    valid, out = O.refine_octave(D, np.array([[2, 21, 19]]), thr=0.01)
    assert valid[0] and abs(out[0, 1] - 17.8) < 2e-5


def test_refine_rejects_ridge_and_low_contrast():
    # edge-like ridge: curvature ratio far above (r+1)^2/r = 12.1
    D = _bump(2.0, 20.0, 20.0, b=(0.01, 0.0002, 0.02))
    valid, _ = O.refine_octave(D, np.array([[2, 20, 20]]), thr=0.01)
    assert not valid[0]
    D = _bump(2.0, 20.0, 20.0, a=0.005)
    valid, _ = O.refine_octave(D, np.array([[2, 20, 20]]), thr=0.01)
    assert not valid[0]


# ------------------------------------------------------------ orientation --
@pytest.mark.parametrize("theta", [0.0, 12.5, 37.0, 95.0, 181.0, 270.0, 333.3])
def test_orientation_of_linear_ramp(theta):
    # G = c cos t + r sin t has gradient angle atan2(dy, dx) = t everywhere
    t = math.radians(theta)
    r, c = np.mgrid[0:64, 0:64].astype(np.float64)
    G = (0.01 * (c * math.cos(t) + r * math.sin(t))).astype(np.float32)
    ang = O.orientation(G, 32.0, 32.0, 2.0)
    assert len(ang) == 1
    d = (ang[0] - theta + 180) % 360 - 180
    assert abs(d) < 0.6, (ang, theta)


def test_orientation_two_directions_give_two_peaks():
    # G = c + |r - 32|: gradient direction -45 deg above the kink, +45 below,
    # with equal magnitude -> two peaks of equal height
    r, c = np.mgrid[0:64, 0:64].astype(np.float64)
    G = (0.01 * (c + np.abs(r - 32))).astype(np.float32)
    ang = O.orientation(G, 32.0, 32.0, 2.0)
    # 45 deg sits on the boundary of bins 4 and 5: both bins are equal, both
    # pass the non-strict local-maximum test (core.py:315-318) and each
    # parabola refines to 4.5 bins -> every direction is reported twice
    np.testing.assert_allclose(np.sort(ang), [45.0, 45.0, 315.0, 315.0], atol=1e-3)


# ------------------------------------------------------------- descriptor --
def _desc_hist64(G, ri, ci, ang_rad, ph):
    """Float64 NumPy restatement of one patch histogram of sift_descriptor_coop
    (core.py:693-792; one scale of sift_descriptor_dsp_coop, core.py:883-950)."""
    H, W = G.shape
    ro, co = np.mgrid[-ph:ph + 1, -ph:ph + 1]
    rr, cc = int(ri) + ro, int(ci) + co
    inb = (rr >= 1) & (rr < H - 1) & (cc >= 1) & (cc < W - 1)
    ca, sa = math.cos(ang_rad), math.sin(ang_rad)
    xr = ca * co - sa * ro
    yr = sa * co + ca * ro
    xb = xr * (4.0 / (2 * ph)) + 1.5
    yb = yr * (4.0 / (2 * ph)) + 1.5
    ok = inb & (xb >= -0.5) & (xb < 3.5) & (yb >= -0.5) & (yb < 3.5)
    rr, cc, xr, yr, xb, yb = rr[ok], cc[ok], xr[ok], yr[ok], xb[ok], yb[ok]
    Gd = G.astype(np.float64)
    dx = Gd[rr, cc + 1] - Gd[rr, cc - 1]
    dy = Gd[rr + 1, cc] - Gd[rr - 1, cc]
    mag = np.hypot(dx, dy)
    ori = np.fmod((np.arctan2(dy, dx) - ang_rad) * 57.29577951 + 360.0, 360.0)
    sd = ph * 0.5657
    wmag = mag * np.exp(-(xr * xr + yr * yr) / (2 * sd * sd + 1e-12))
    ob = ori / 45.0
    hist = np.zeros(128)
    oi0 = np.floor(ob).astype(int); df = ob - oi0
    xi0 = np.floor(xb).astype(int); yi0 = np.floor(yb).astype(int)
    dxb = xb - xi0; dyb = yb - yi0
    for yo, wy in ((0, 1 - dyb), (1, dyb)):
        for xo, wx in ((0, 1 - dxb), (1, dxb)):
            for oo, wo in ((0, 1 - df), (1, df)):
                idx = np.clip(yi0 + yo, 0, 3) * 32 + np.clip(xi0 + xo, 0, 3) * 8 + (oi0 + oo) % 8
                np.add.at(hist, idx, wy * wx * wo * wmag)
    return hist


DSP_PRESETS = {3: [1 / math.sqrt(2), 1.0, math.sqrt(2)], 4: [0.5, 1 / math.sqrt(2), 1.0, math.sqrt(2)],
               5: [0.5, 1 / math.sqrt(2), 1.0, math.sqrt(2), 2.0]}


def _desc_reference(G, ri, ci, ang_rad, ph, dsp_n_scales=0):
    """Float64 restatement of sift_descriptor_coop, or of
    sift_descriptor_dsp_coop (the patch half-sizes max(4, (int)(ph m + 0.5)) in
    float32, one histogram divided by the scale count; core.py:1039-1040,
    884-885, 953-957), + RootSIFT."""
    if dsp_n_scales:
        phs = [max(4, int(np.float32(ph) * np.float32(m) + np.float32(0.5))) for m in DSP_PRESETS[dsp_n_scales]]
        hist = sum(_desc_hist64(G, ri, ci, ang_rad, q) for q in phs) / len(phs)
    else:
        hist = _desc_hist64(G, ri, ci, ang_rad, ph)
    v = hist / math.sqrt((hist ** 2).sum() + 1e-14)
    v = np.minimum(v, 0.2)
    v = v / math.sqrt((v ** 2).sum() + 1e-14)
    v = v / (np.abs(v).sum() + 1e-7)
    return np.sqrt(np.maximum(v, 0))


@pytest.mark.parametrize("ph,ang", [(4, 0.3), (7, 2.1), (21, 5.0), (34, 1.0)])
def test_descriptor_matches_float64_restatement(ph, ang):
    G = nd.gaussian_filter(np.random.default_rng(ph).random((120, 130)), 1.5).astype(np.float32)
    d, u = O.descriptor(G, 60.4, 61.7, np.float32(ang), ph)
    ref = _desc_reference(G, 60.4, 61.7, float(np.float32(ang)), ph)
    np.testing.assert_allclose(d, ref, atol=2e-4)
    assert abs(float((d.astype(np.float64) ** 2).sum()) - 1.0) < 1e-4       # RootSIFT is unit L2
    np.testing.assert_array_equal(u, np.minimum(255, (d * 512 + 0.5).astype(np.int64)))


def test_descriptor_of_horizontal_ramp_uses_orientation_bin_zero():
    r, c = np.mgrid[0:80, 0:80].astype(np.float32)
    G = (0.01 * c).astype(np.float32)
    d, _ = O.descriptor(G, 40.0, 40.0, np.float32(0.0), 8)
    h = d.reshape(4, 4, 8)
    assert h[:, :, 1:].max() == 0.0 and h[:, :, 0].min() > 0.0
    z, _ = O.descriptor(np.zeros((80, 80), np.float32), 40.0, 40.0, np.float32(0.0), 8)
    assert not z.any()


@pytest.mark.parametrize("ph,ang,n", [(21, 0.7, 5), (9, 2.5, 3), (27, 4.0, 4), (4, 1.0, 5)])
def test_dsp_descriptor_matches_float64_restatement(ph, ang, n):
    G = nd.gaussian_filter(np.random.default_rng(ph + n).random((300, 320)), 2.0).astype(np.float32)
    d, u = O.descriptor(G, 150.3, 160.8, np.float32(ang), ph, dsp_n_scales=n)
    ref = _desc_reference(G, 150.3, 160.8, float(np.float32(ang)), ph, dsp_n_scales=n)
    np.testing.assert_allclose(d, ref, atol=2e-4)
    assert abs(float((d.astype(np.float64) ** 2).sum()) - 1.0) < 1e-4
    plain, _ = O.descriptor(G, 150.3, 160.8, np.float32(ang), ph)
    assert np.abs(plain - d).max() > 1e-3          # the pooling changes the descriptor


def test_f16_round_matches_numpy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32) * np.float32(s)
                        for s in (1e-8, 1e-6, 1e-4, 1e-2, 1, 1e2, 3e4)])
    bits = np.arange(0, 0x7bff, 7, dtype=np.uint16)
    h = bits.view(np.float16).astype(np.float32)
    h1 = (bits + 1).view(np.float16).astype(np.float32)
    ties = ((h.astype(np.float64) + h1) / 2).astype(np.float32)        # exact midpoints: ties to even
    x = np.concatenate([x, h, -h, np.nextafter(h, np.float32(np.inf)), ties, -ties,
                        np.float32([0.0, -0.0, 65504, 65519.996, 65520, 1e9, -1e9, 2.9e-8, 3e-8, 6e-8])])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(O.f16_round(x), ref)


def test_fp16_pyramid_levels_and_base():
    img = syn.synth_gray(120, 160, 5)
    p = O.params(n_octaves=4, fp16_pyramid=True)
    # octave 0 base: the binary16 u8/255 array, zoomed, cast to binary16 again,
    # then blurred in float32 (core.py:1256, 1274, 1286-1288)
    g0, _ = O.pyramid_octave(img, p, 0)
    up = O.f16_round(O.upsample2x(O.f16_round(img.astype(np.float32) / np.float32(255.0))))
    np.testing.assert_array_equal(g0[0], O.blur(up, math.sqrt(1.6 ** 2 - 1.0)))
    for o in range(4):
        g, d = O.pyramid_octave(img, p, o)
        np.testing.assert_array_equal(d, g[1:] - g[:-1])
        # octaves 0-1 stay float32, octaves >= 2 hold binary16 values
        assert np.array_equal(O.f16_round(g), g) == (o >= 2)
    # the chain is float32: octave 2 of the plain pyramid with the same
    # binary16 base, rounded, gives the fp16 mode's stored levels
    g2, _ = O.pyramid_octave(img, p, 2)
    g1, _ = O.pyramid_octave(img, p, 1)
    np.testing.assert_array_equal(g2[0], O.f16_round(g1[3][::2, ::2]))


@pytest.mark.parametrize("name", ["imc_british_museum", "imc_florence_cathedral_side", "imc_st_pauls_cathedral"])
def test_reference_fixture_dsp5_fp16_all_octaves(name):
    """The reference's RTX 3050 output in the mode it was recorded in (DSP-5
    descriptors, binary16 pyramid, 4 octaves; manifest.json): all octaves'
    keypoints within 0.5 px, and the descriptors of keypoints matched by
    position and angle (a location with several orientations holds several
    keypoints) at cosine >= 0.9999 (the reference's EQUIVALENT bar is 95 % and
    0.99, tests/cross_determinism_test.py:112-113).  Tolerance of the port:
    >= 99.8 % of the reference's keypoints within 1e-3 px and >= 99.9 % of the
    uint8 descriptor values within 1 LSB.  The rest differ through CUDA's
    atan2f/expf and cupyx zoom bits, which cannot be reproduced here."""
    from PIL import Image
    from scipy.spatial import cKDTree
    z = np.load(os.path.join(GOLDEN, "imc", name + ".npz"))
    im = Image.open(os.path.join(GOLDEN, "imc", name + ".jpg"))
    im.draft("L", im.size)
    g = np.asarray(im.convert("L"))
    kp, d, u = O.detect_describe(g, O.params_low_vram(dsp_n_scales=5))
    dist, _ = cKDTree(kp[:, :2]).query(z["points"])
    assert (dist <= 0.5).mean() > 0.995
    # the same keypoint within 1e-3 px
    print(f"[fixture] {name}: {100 * (dist <= 1e-3).mean():.3f}% of {len(dist)} within 1e-3 px", flush=True)
    assert (dist <= 1e-3).mean() >= 0.998
    assert abs(len(kp) - len(z["points"])) <= 0.002 * len(z["points"])
    a = np.radians(kp[:, 3]); ar = np.radians(z["angles"])
    X = np.c_[kp[:, :2], 50 * np.cos(a), 50 * np.sin(a)]
    Y = np.c_[z["points"], 50 * np.cos(ar), 50 * np.sin(ar)]
    dist2, idx = cKDTree(X).query(Y)
    m = dist2 < 0.5
    assert m.mean() > 0.995
    rd = z["descriptors"][m].astype(np.float64)
    nd = d[idx[m]].astype(np.float64)
    cos = (rd * nd).sum(1) / (np.linalg.norm(rd, axis=1) * np.linalg.norm(nd, axis=1) + 1e-12)
    assert cos.mean() > 0.9999 and cos.min() > 0.999
    # uint8 descriptors (512 scale) within 1 LSB of the reference's values
    ref_u8 = np.minimum(255, (z["descriptors"][m].astype(np.float32) * 512 + 0.5).astype(np.int64))
    assert np.mean(np.abs(ref_u8 - u[idx[m]].astype(np.int64)) <= 1) > 0.999


# --------------------------------------------------------- full detection --
def test_detect_invariants_config0():
    img = syn.synth_gray(480, 640, 0)
    kp, d, u = O.detect_describe(img)
    assert 500 < len(kp) < 4 * 8000
    order = np.lexsort((kp[:, 1], kp[:, 0]))
    np.testing.assert_array_equal(order, np.arange(len(kp)))            # sorted by (x, y)
    assert (kp[:, 3] >= 0).all() and (kp[:, 3] < 360).all()
    k = 2 ** (1 / 3)
    sizes = {round(1.6 * k ** s * 2 ** o * 0.5 * 2, 4) for o in range(5) for s in (1, 2, 3)}
    assert {round(float(x), 4) for x in kp[:, 2]} <= sizes
    np.testing.assert_allclose((d.astype(np.float64) ** 2).sum(1), 1.0, atol=1e-4)


@pytest.mark.parametrize("name", ["imc_british_museum", "imc_florence_cathedral_side", "imc_st_pauls_cathedral"])
def test_reference_fixture_keypoints(name):
    """The reference's RTX 3050 keypoints (determinism_reference, fp32 octaves
    0-1) are reproduced within 0.5 px (the EQUIVALENT bar of the reference's
    cross_determinism_test is 95 %); orientations agree to hundredths of a
    degree."""
    from PIL import Image
    from scipy.spatial import cKDTree
    z = np.load(os.path.join(GOLDEN, "imc", name + ".npz"))
    im = Image.open(os.path.join(GOLDEN, "imc", name + ".jpg"))
    im.draft("L", im.size)
    g = np.asarray(im.convert("L"))
    assert g.shape == tuple(z["image_shape"])
    kp, _, _ = O.detect_describe(g, O.params(n_octaves=4))
    sel = z["sizes"] < 7.0                      # octaves 0-1 (the reference stores octaves >= 2 in fp16)
    mine = kp[:, 2] < 7.0
    dist, idx = cKDTree(kp[mine, :2]).query(z["points"][sel])
    assert (dist < 0.5).mean() > 0.99
    ang = np.abs(((z["angles"][sel] - kp[mine][idx, 3]) + 180) % 360 - 180)
    assert np.median(ang[dist < 0.5]) < 0.05
    assert abs(int(mine.sum()) - int(sel.sum())) <= 0.01 * sel.sum()


# ----------------------------------------------------------------- matcher --
def _brute_match(a, b, ratio):
    d = ((a[:, None, :].astype(np.int64) - b[None, :, :].astype(np.int64)) ** 2).sum(-1)
    def top2(m):
        j = np.argmin(m, 1)   # first index on ties
        d1 = m[np.arange(len(m)), j]
        m2 = m.copy(); m2[np.arange(len(m)), j] = np.iinfo(np.int64).max
        return j, d1, m2.min(1)
    fj, f1, f2 = top2(d); bi, b1, b2 = top2(d.T)
    r2 = ratio * ratio
    pairs = [(i, fj[i]) for i in range(len(a))
             if f1[i] < r2 * f2[i] and bi[fj[i]] == i and b1[fj[i]] < r2 * b2[fj[i]]]
    return np.c_[fj, f1, f2], np.c_[bi, b1, b2], np.array(pairs, np.int64).reshape(-1, 2)


@pytest.mark.parametrize("N,M,ratio", [(50, 70, 0.75), (120, 90, 0.8), (30, 30, 0.6)])
def test_matcher_matches_brute_force(N, M, ratio):
    rng = np.random.default_rng(N + M)
    a = rng.integers(0, 256, (N, 128)).astype(np.uint8)
    b = rng.integers(0, 256, (M, 128)).astype(np.uint8)
    b[: N // 2] = np.clip(a[: N // 2].astype(int) + rng.integers(-5, 6, (N // 2, 128)), 0, 255)
    b[-1] = b[-2]                           # a duplicate row: ties
    f, bw, p = O.match_u8(a, b, ratio)
    rf, rb, rp = _brute_match(a, b, ratio)
    np.testing.assert_array_equal(f, rf)
    np.testing.assert_array_equal(bw, rb)
    np.testing.assert_array_equal(p, rp)
    assert len(p) > 0


# ----------------------------------------------------------------- Philox --
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10 (also reproduced by rocRAND's
    # philox4x32_10_engine::ten_rounds)
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        np.testing.assert_array_equal(O.philox(ctr, key), np.array(out, np.uint32))
    idx = O.ransac_sample(1000, 37, 42)
    assert idx.min() >= 0 and idx.max() < 37
    assert abs(np.bincount(idx.ravel(), minlength=37).std() / (4000 / 37)) < 0.2


# ------------------------------------------------------------- homography --
def _H():
    return np.array([[0.95, 0.08, 120.0], [-0.05, 1.02, -30.0], [1e-4, -5e-5, 1.0]])


def _proj(H, p):
    q = np.c_[p, np.ones(len(p))] @ H.T
    return q[:, :2] / q[:, 2:]


def test_dlt_recovers_exact_homography():
    rng = np.random.default_rng(7)
    for n in (4, 10, 300):
        p1 = rng.uniform(0, 1000, (n, 2))
        H = O.dlt_ls(p1, _proj(_H(), p1))
        np.testing.assert_allclose(H, _H(), rtol=1e-8, atol=1e-9)


def test_dlt_agrees_with_numpy_svd_on_noisy_points():
    rng = np.random.default_rng(8)
    p1 = rng.uniform(0, 1000, (200, 2))
    p2 = _proj(_H(), p1) + rng.normal(0, 0.7, (200, 2))
    def norm(p):
        c = p.mean(0); s = math.sqrt(2) / (np.sqrt(((p - c) ** 2).sum(1)).mean() + 1e-10)
        return np.array([[s, 0, -s * c[0]], [0, s, -s * c[1]], [0, 0, 1]])
    T1, T2 = norm(p1), norm(p2)
    a = _proj(T1, p1); b = _proj(T2, p2)
    A = []
    for (x1, y1), (x2, y2) in zip(a, b):
        A.append([-x1, -y1, -1, 0, 0, 0, x2 * x1, x2 * y1, x2])
        A.append([0, 0, 0, -x1, -y1, -1, y2 * x1, y2 * y1, y2])
    Hn = np.linalg.svd(np.array(A))[2][-1].reshape(3, 3)
    ref = np.linalg.inv(T2) @ Hn @ T1
    ref /= ref[2, 2]
    np.testing.assert_allclose(O.dlt_ls(p1, p2), ref, rtol=1e-7, atol=1e-8)


def test_ransac_recovers_inliers_and_scores():
    rng = np.random.default_rng(9)
    N = 400
    p1 = rng.uniform(0, 1000, (N, 2))
    p2 = _proj(_H(), p1) + rng.normal(0, 0.3, (N, 2))
    out = rng.random(N) < 0.35
    p2[out] = rng.uniform(0, 1000, (out.sum(), 2))
    H, mask, best, sc, Hs = O.ransac(p1, p2, seed=42, n_hyp=500, thr=3.0, return_all=True)
    true_in = ~out & (np.linalg.norm(_proj(_H(), p1) - p2, axis=1) < 2)
    assert (mask == (~out)).mean() > 0.99 and mask[true_in].all()
    assert np.abs(_proj(H, p1[~out]) - _proj(_H(), p1[~out])).max() < 1.0
    # the MAGSAC score of a few hypotheses, restated in NumPy
    for b in [best, 0, 17]:
        Hb = Hs[b]
        if sc[b] < 0:
            continue
        Hi = np.linalg.inv(Hb)

        def projc(H, p):   # z clamped at 1e-10 (core.py:2295, 2309)
            q = np.c_[p, np.ones(len(p))] @ H.T
            return q[:, :2] / np.maximum(q[:, 2:], 1e-10)
        ef = ((projc(Hb, p1) - p2) ** 2).sum(1)
        eb = ((projc(Hi, p2) - p1) ** 2).sum(1)
        ref = np.exp(-(ef + eb) / (2 * 9.0)).sum()
        assert abs(sc[b] - ref) < 1e-9 * max(1.0, ref)
    assert best == int(np.argmax(sc))


def test_ransac_needs_four_points():
    with pytest.raises(ValueError):
        O.ransac(np.zeros((3, 2)), np.zeros((3, 2)))


# ---------------------------------------------------------- warp / feather --
def test_warp_identity_and_translation():
    rng = np.random.default_rng(10)
    img = rng.integers(0, 256, (40, 50, 3)).astype(np.uint8)
    I = np.eye(3)
    cw, ch, ox, oy = O.canvas(40, 50, 40, 50, I)
    assert (cw, ch, ox, oy) == (51, 41, 0.0, 0.0)
    pano, valid = O.warp_feather(img, img, I, cw, ch, ox, oy)
    np.testing.assert_array_equal(pano[:40, :50], img)
    assert valid[:40, :50].all() and not valid[40:].any() and not valid[:, 50:].any()
    # integer translation: img1 lands shifted by (+7, +3), exact copy where only img1 is valid
    T = np.array([[1.0, 0, 7.0], [0, 1.0, 3.0], [0, 0, 1.0]])
    cw, ch, ox, oy = O.canvas(40, 50, 40, 50, T)
    pano, valid = O.warp_feather(img, np.zeros_like(img), T, cw, ch, ox, oy)
    # where both are valid the feather blend mixes in img2 (zeros); pixels only img1 covers
    np.testing.assert_array_equal(pano[40:43, 7:57], img[37:40, 0:50])
    np.testing.assert_array_equal(pano[3:43, 50:57], img[0:40, 43:50])


# ---- the reference's float matcher (core.py:2574-2743, 2840-2940) ----

def _gamma_desc(n, seed):
    rng = np.random.default_rng(seed)
    return rng.gamma(0.6, 1.0, (n, 128)).astype(np.float32)


def test_l2n_rows_against_float64():
    x = _gamma_desc(300, 1) * 7.0
    x[:20] *= 1e-7 / 7.0   # norms near 1e-6, where the +1e-7 outside the root shows
    y = O.l2n_rows(x)
    X = x.astype(np.float64)
    ref = X / (np.linalg.norm(X, axis=1, keepdims=True) + 1e-7)
    assert np.abs(y - ref).max() < 2e-7
    assert np.abs(np.linalg.norm(y[20:].astype(np.float64), axis=1) - 1.0).max() < 1e-6
    n = np.linalg.norm(X[:20], axis=1)
    assert np.abs(np.linalg.norm(y[:20].astype(np.float64), axis=1) - n / (n + 1e-7)).max() < 1e-6


@pytest.mark.parametrize("N,M,dims", [(700, 900, 64), (150, 20, 32), (1, 300, 100), (400, 0, 128)])
def test_pca_fit_against_numpy_eigh(N, M, dims):
    xa, xb = O.l2n_rows(_gamma_desc(N, N + 3)), O.l2n_rows(_gamma_desc(M, M + 5))
    mean, proj, ev, sweeps = O.pca_fit(xa, xb, dims)
    X = np.vstack([xa, xb]).astype(np.float64)
    mu = X.mean(0)
    C = (X - mu).T @ (X - mu) / len(X)
    w, V = np.linalg.eigh(C)
    assert 3 <= sweeps < 40
    np.testing.assert_allclose(mean, mu.astype(np.float32), rtol=0, atol=1e-8)
    scale = w.max()
    assert np.abs(ev - w[::-1]).max() < 1e-12 * scale
    P = proj.astype(np.float64)
    # orthonormal columns, in ascending eigenvalue order, each an eigenvector
    assert np.abs(P.T @ P - np.eye(dims)).max() < 1e-6
    lam = ev[:dims][::-1]
    assert np.abs(C @ P - P * lam[None, :]).max() < 1e-6 * scale + 1e-9
    # the same subspace as numpy's top-dims eigenvectors (where the gap allows)
    if dims < 128 and w[-dims] - w[-dims - 1] > 1e-3 * scale:
        Pn = V[:, -dims:]
        assert np.abs(P @ P.T - Pn @ Pn.T).max() < 1e-5
    # sign rule: the largest-magnitude entry of every column is positive
    am = np.abs(proj).argmax(0)
    assert (proj[am, np.arange(dims)] > 0).all()


def test_pca_fit_known_spectrum():
    # rows = 0.25 + 2.5 e1 u + e2 v with e1 = (+ + - -)..., e2 = (+ - + -)...:
    # zero-mean, exactly uncorrelated, every entry exact in float32, so the
    # covariance is 6.25 u u^T + v v^T: eigenvalues 6.25, 1, then 0
    u = np.zeros(128); u[3] = 0.6; u[70] = 0.8
    v = np.zeros(128); v[10] = 1.0
    e1 = np.tile([1.0, 1.0, -1.0, -1.0], 128); e2 = np.tile([1.0, -1.0, 1.0, -1.0], 128)
    X = (0.25 + 2.5 * e1[:, None] * u[None] + e2[:, None] * v[None]).astype(np.float32)
    mean, proj, ev, _ = O.pca_fit(X[:300], X[300:], 2)
    assert abs(ev[0] - 6.25) < 1e-12 and abs(ev[1] - 1.0) < 1e-12 and np.abs(ev[2:]).max() < 1e-12
    assert (mean == np.float32(0.25)).all()
    # ascending order: column 0 is v (smaller), column 1 is u, both sign-normalised
    assert np.abs(proj[:, 1] - u).max() < 1e-6 and np.abs(proj[:, 0] - v).max() < 1e-6


def test_pca_project_against_float64():
    xa, xb = O.l2n_rows(_gamma_desc(200, 1)), O.l2n_rows(_gamma_desc(250, 2))
    mean, proj, _, _ = O.pca_fit(xa, xb, 64)
    y = O.pca_project(xa, mean, proj)
    Y = (xa.astype(np.float64) - mean) @ proj.astype(np.float64)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True) + 1e-7
    assert np.abs(y - Y).max() < 2e-6


def test_sim16_against_numpy_binary16():
    # products of binary16 values are exact in float64 and partial sums of
    # unit rows stay below 2, so a float64 matmul is exact; then one rounding
    # to float32 and one to binary16
    a = O.l2n_rows(_gamma_desc(60, 3))[:, :64].copy()
    b = O.l2n_rows(_gamma_desc(70, 4))[:, :64].copy()
    a /= np.linalg.norm(a, axis=1, keepdims=True); b /= np.linalg.norm(b, axis=1, keepdims=True)
    A = a.astype(np.float16).astype(np.float64); B = b.astype(np.float16).astype(np.float64)
    ref = (A @ B.T).astype(np.float32).astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(O.sim16(a, b), ref)


def _unit2(deg):
    t = np.radians(deg)
    return np.array([np.cos(t), np.sin(t)], np.float32)


def test_match_sim16_ratio_closed_forms():
    # query e0 against directions at theta: d = 2 - 2 cos(theta) (binary16 cos)
    a = np.array([[1.0, 0.0]], np.float32)
    # best 30 deg (d 0.268), second 60 deg (d 1): 0.268 < 0.5625 -> passes
    fwd, bwd, pairs = O.match_sim16(a, np.stack([_unit2(60), _unit2(30)]), 0.75)
    assert fwd[0, 0] == 1 and len(pairs) == 0   # b has one query only: the backward test needs 2 rows
    d1 = fwd[0, 1:].view(np.float32)
    c30 = np.float32(np.float16(np.cos(np.radians(30))))
    assert d1[0] == np.float32(2.0) - np.float32(2.0) * c30
    # mutual with a second query far away: the pair passes both ways
    a2 = np.stack([a[0], _unit2(200)])
    fwd, bwd, pairs = O.match_sim16(a2, np.stack([_unit2(60), _unit2(30)]), 0.75)
    assert pairs.tolist() == [[0, 1]]
    # best 50 deg (d 0.714), second 60 deg (d 1): 0.714 > 0.5625 -> rejected
    fwd, bwd, pairs = O.match_sim16(a2, np.stack([_unit2(60), _unit2(50)]), 0.75)
    assert fwd[0, 0] == 1 and len(pairs) == 0
    # a row above unit length: its binary16 self-similarity 1.00098^2 rounds
    # to 1.00195, so 2 - 2 s < 0 and d clamps to 0
    big = np.array([[1.001, 0.0], [0.0, 1.0]], np.float32)
    fwd, _, _ = O.match_sim16(big, big, 0.75)
    assert fwd[0, 1].view(np.float32) == 0.0 and fwd[0, 0] == 0
    # the same query twice in b: best == second -> rejected
    fwd, bwd, pairs = O.match_sim16(a2, np.stack([_unit2(30), _unit2(30), _unit2(100)]), 0.75)
    assert fwd[0, 0] == 0 and len(pairs) == 0


def test_match_sim16_permutation_and_symmetry():
    a = O.l2n_rows(_gamma_desc(300, 6))
    perm = np.random.default_rng(1).permutation(300)
    b = a[perm]
    fwd, bwd, pairs = O.match_sim16(a, b, 0.75)
    inv = np.argsort(perm)
    assert len(pairs) == 300 and (pairs[:, 1] == inv[pairs[:, 0]]).all()
    # swapping the sets swaps the pairs
    c = O.l2n_rows(_gamma_desc(200, 7))
    c[:100] = a[:100] + 0.02 * O.l2n_rows(_gamma_desc(100, 8))
    c = O.l2n_rows(c)
    _, _, p1 = O.match_sim16(a, c, 0.75)
    _, _, p2 = O.match_sim16(c, a, 0.75)
    assert sorted(map(tuple, p1.tolist())) == sorted((j, i) for i, j in p2.tolist())
    assert len(p1) > 80


# ---- compositing: gain + DP seam + multiband (core.py:3211-3239, 3287-3553, 3686-3714) ----

def test_grid_warp_against_torch_grid_sample():
    # the reference's _warp_image computed with torch on the CPU
    import torch
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (120, 160, 3)).astype(np.uint8)
    H = np.array([[0.9, 0.1, 60.3], [-0.07, 1.05, -20.6], [2e-4, -1e-4, 1.0]])
    cw, ch, ox, oy = O.canvas(120, 160, 100, 150, H)
    u = torch.arange(cw, dtype=torch.float32); v = torch.arange(ch, dtype=torch.float32)
    uu, vv = torch.meshgrid(u, v, indexing="xy")
    flat = torch.stack([uu - ox, vv - oy, torch.ones_like(uu)], 0).reshape(3, -1)
    src = torch.tensor(np.linalg.inv(H), dtype=torch.float32) @ flat
    sx = src[0] / src[2].clamp(min=1e-10); sy = src[1] / src[2].clamp(min=1e-10)
    gx = (2.0 * sx / 159) - 1.0; gy = (2.0 * sy / 119) - 1.0
    grid = torch.stack([gx, gy], -1).reshape(1, ch, cw, 2)
    it = torch.tensor(img.astype(np.float32)).permute(2, 0, 1).unsqueeze(0)
    out = torch.nn.functional.grid_sample(it, grid, mode="bilinear", padding_mode="zeros", align_corners=True)
    ref = out.squeeze(0).permute(1, 2, 0).numpy().clip(0, 255).astype(np.uint8)
    mref = ((gx >= -1) & (gx <= 1) & (gy >= -1) & (gy <= 1)).reshape(ch, cw).numpy()
    w, m = O.grid_warp(img, H, cw, ch, ox, oy)
    assert (m == mref).all()
    assert np.mean(w == ref) > 0.9995 and np.abs(w.astype(int) - ref).max() <= 1


def test_pyramid_filters_against_scipy():
    rng = np.random.default_rng(1)
    x = (rng.random((123, 217)) * 255).astype(np.float32)
    # sigma-1 blur then [::2, ::2] (float32 accumulation vs SciPy's double)
    ref = nd.gaussian_filter(x, 1.0, mode="reflect", truncate=4.0)[::2, ::2]
    assert np.abs(O.gauss_down(x) - ref).max() < 1e-4
    # bilinear zoom with align-corners coordinates: exactly SciPy's order-1 zoom
    for (h, w, th, tw) in [(62, 109, 123, 217), (31, 55, 62, 109), (1, 3, 2, 5)]:
        y = (rng.random((h, w)) * 255).astype(np.float32)
        np.testing.assert_array_equal(O.zoom_up(y, th, tw), nd.zoom(y, (th / h, tw / w), order=1))
    # Sobel: exactly SciPy's
    for ax in (0, 1):
        np.testing.assert_array_equal(O.sobel(x, ax), nd.sobel(x, axis=ax, mode="reflect"))
    # cv2.GaussianBlur(sigma 3): 25 taps, BORDER_REFLECT_101 (SciPy's "mirror")
    m = (rng.random((90, 140)) > 0.5).astype(np.float32)
    assert np.abs(O.cv_blur25(m) - nd.gaussian_filter(m, 3.0, mode="mirror", truncate=4.0)).max() < 1e-6
    w = O.gauss_weights_cv()
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6 and w[12] == w.max()


def _brute_seam(cost, ov):
    # all one-column-per-row paths over the rows that hold overlap
    import itertools
    rows = [r for r in range(cost.shape[0]) if ov[r].any()]
    W = cost.shape[1]
    best = None
    for start in range(W):
        for steps in itertools.product((-1, 0, 1), repeat=len(rows) - 1):
            cols = [start]
            for st in steps:
                cols.append(cols[-1] + st)
            if min(cols) < 0 or max(cols) >= W:
                continue
            if not all(ov[r, c] for r, c in zip(rows, cols)):
                continue
            t = sum(float(cost[r, c]) for r, c in zip(rows, cols))
            if best is None or t < best:
                best = t
    return rows, best


@pytest.mark.parametrize("seed", range(6))
def test_dp_seam_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    H, W = 5, 5
    cost = rng.integers(0, 9, (H, W)).astype(np.float32)
    ov = np.ones((H, W), bool)
    if seed % 2:
        ov[2] = False          # a row without overlap is skipped
        ov[:, 0] = seed == 3   # a column outside the overlap
        cost[~ov] = 1e9
    rows, best = _brute_seam(cost, ov)
    seam = O.dp_seam(cost, ov, True)
    assert sum(float(cost[r, seam[r]]) for r in rows) == best
    # the horizontal seam is the vertical one of the transpose
    np.testing.assert_array_equal(O.dp_seam(np.ascontiguousarray(cost.T), np.ascontiguousarray(ov.T), False), seam)


def test_hist_lut_equals_numpy_interp():
    # _hist_match's LUT: np.interp of the source CDF into the reference CDF,
    # including empty bins (plateaus of the CDF, where np.interp's search rule
    # decides), against NumPy itself
    rng = np.random.default_rng(12)
    for trial in range(60):
        hs = rng.integers(0, 40, (3, 256)).astype(np.uint64)
        hr = rng.integers(0, 40, (3, 256)).astype(np.uint64)
        if trial % 2:
            hs[:, rng.integers(0, 256, 120)] = 0
            hr[:, rng.integers(0, 256, 150)] = 0
        lut = O.hist_lut(hs, hr)
        for c in range(3):
            s = np.cumsum(hs[c].astype(np.int64)).astype(np.float64)
            r = np.cumsum(hr[c].astype(np.int64)).astype(np.float64)
            s /= s[-1] + 1e-8
            r /= r[-1] + 1e-8
            np.testing.assert_array_equal(lut[c], np.interp(s, r, np.arange(256)).astype(np.float32))


def test_composite_hist_matching_uses_the_lut():
    # two flat views shifted by 40 px.  The reference CDF of a channel is a
    # single step at the second colour, so NumPy's rule (the last index of a
    # plateau at or below x) sends the top source value to 255 and a value
    # below it into the step; the composite applies exactly that LUT
    a = np.empty((60, 80, 3), np.uint8); a[:] = (100, 120, 60)
    b = np.empty((60, 80, 3), np.uint8); b[:] = (110, 126, 66)
    Ht = np.array([[1.0, 0, 40.0], [0, 1.0, 0], [0, 0, 1.0]])
    cw, ch, ox, oy = O.canvas(60, 80, 60, 80, Ht)
    pano, valid, seam, gain, rmse, vert = O.composite(a, b, Ht, cw, ch, ox, oy, 6, "hist")
    w, m1 = O.grid_warp(a, Ht, cw, ch, ox, oy)
    m2 = np.zeros((ch, cw), bool); m2[:60, :80] = True
    ov = m1 & m2
    hs = np.stack([np.bincount(w[ov][:, c], minlength=256) for c in range(3)]).astype(np.uint64)
    hr = np.stack([np.bincount(np.broadcast_to(b[0, 0, c], ov.sum()), minlength=256) for c in range(3)]).astype(np.uint64)
    lut = O.hist_lut(hs, hr)
    wg = np.stack([lut[c][w[..., c]] for c in range(3)], -1)
    exp_rmse = np.sqrt(((wg[ov].astype(np.uint8).astype(np.int64) - np.array([110, 126, 66])) ** 2).sum() / (3 * ov.sum()))
    assert abs(rmse - exp_rmse) < 1e-12
    top = w[ov][:, 0].max()
    assert lut[0][top] == 255.0 and 109.0 <= lut[0][top - 1] < 110.0
    assert valid.sum() == (m1 | m2).sum()


def test_composite_identical_views_and_constant_gain():
    rng = np.random.default_rng(2)
    img = np.clip(nd.gaussian_filter(rng.random((80, 100, 3)) * 255, (3, 3, 0)), 0, 255).astype(np.uint8)
    # the same view over itself.  The identity warp's normalise/unnormalise
    # round trip lands some samples just below an integer, which truncates
    # (as the reference's grid_sample does), so the gain is the closed form
    # of exact sums over the warped image, and the blend stays within 2 LSB
    pano, valid, seam, gain, rmse, vert = O.composite(img, img, np.eye(3), 100, 80, 0.0, 0.0)
    w, m = O.grid_warp(img, np.eye(3), 100, 80, 0.0, 0.0)
    assert valid.all() and m.all()
    n = float(w.shape[0] * w.shape[1])
    m1 = (w.reshape(-1, 3).astype(np.int64).sum(0) / n).astype(np.float32)
    m2 = (img.reshape(-1, 3).astype(np.int64).sum(0) / n).astype(np.float32)
    g = np.clip((m2 + np.float32(1e-3)) / (m1 + np.float32(1e-3)), np.float32(0.5), np.float32(2.0))
    np.testing.assert_array_equal(gain, g)
    wg = np.clip(w.astype(np.float32) * g, 0, 255).astype(np.uint8).astype(np.int64)
    assert abs(rmse - np.sqrt(((wg - img) ** 2).sum() / (3 * n))) < 1e-12
    assert np.abs(pano.astype(int) - img).max() <= 2
    # two flat views shifted by 40 px: the gain is (c2 + 1e-3) / (m1 + 1e-3)
    # with m1 the exact overlap mean of the warped first view; a tall overlap
    # gives a vertical seam and the blend stays within 2 LSB of the second colour
    a = np.empty((60, 80, 3), np.uint8); a[:] = (100, 120, 60)
    b = np.empty((60, 80, 3), np.uint8); b[:] = (110, 126, 66)
    Ht = np.array([[1.0, 0, 40.0], [0, 1.0, 0], [0, 0, 1.0]])
    cw, ch, ox, oy = O.canvas(60, 80, 60, 80, Ht)
    pano, valid, seam, gain, rmse, vert = O.composite(a, b, Ht, cw, ch, ox, oy)
    w, m1 = O.grid_warp(a, Ht, cw, ch, ox, oy)
    m2 = np.zeros((ch, cw), bool); m2[:60, :80] = True
    ov = m1 & m2
    mean1 = (w[ov].astype(np.int64).sum(0) / ov.sum()).astype(np.float32)
    exp_gain = (np.float32([110, 126, 66]) + np.float32(1e-3)) / (mean1 + np.float32(1e-3))
    np.testing.assert_array_equal(gain, exp_gain)
    assert vert == 1 and valid.sum() == (m1 | m2).sum()
    assert (np.abs(pano[valid].astype(int) - np.array([110, 126, 66])) <= 2).all()
    wg = np.clip(w[ov].astype(np.float32) * gain, 0, 255).astype(np.uint8).astype(np.int64)
    assert abs(rmse - np.sqrt(((wg - np.array([110, 126, 66])) ** 2).sum() / (3 * ov.sum()))) < 1e-12


# ---- CLAHE (core.py:2562-2565, cv2.createCLAHE(...).apply) ----

def test_clahe_constant_image_known_answers():
    # 64x64, 8x8 tiles: tile area 64, limit max(1, (int)(2 * 64 / 256)) = 1.
    # Constant 100: clipped 63 -> batch 0, residual 63, step 256 // 63 = 4, so
    # bins 0, 4, ..., 248 get one each; cumsum[100] = 25 (multiples of 4
    # below 100) + 1 (limit) + 1 (100 is one of them) = 27 -> round(27 * 255 / 64) = 108.
    assert (O.clahe(np.full((64, 64), 100, np.uint8), 2.0, (8, 8)) == 108).all()
    # constant 90 (not a multiple of 4): 23 + 1 = 24 -> 95.625 -> 96
    assert (O.clahe(np.full((64, 64), 90, np.uint8), 2.0, (8, 8)) == 96).all()
    # one 64x64 tile: area 4096, limit 32; clipped 4064 -> batch 15,
    # residual 224, step 1 (bins 0..223 get one more).
    # v = 100: 100 * 16 + (32 + 15 + 1) = 1648 -> 102.598 -> 103
    # v = 230: 230 * 15 + 224 + (32 + 15) = 3721 -> 231.65 -> 232
    assert (O.clahe(np.full((64, 64), 100, np.uint8), 2.0, (1, 1)) == 103).all()
    assert (O.clahe(np.full((64, 64), 230, np.uint8), 2.0, (1, 1)) == 232).all()


def test_clahe_two_level_tile_known_answer():
    # one 64x64 tile, half 10 and half 240: each bin holds 2048 > limit 32,
    # clipped 2 * 2016 = 4032 -> batch 15, residual 192, step 1.
    # LUT[10]  = 10 * 16 + (32 + 15 + 1) = 208 -> 12.95 -> 13
    # LUT[240] = 240 * 15 + 192 + (48 - 16) + (32 + 15) = 3871 -> 240.99 -> 241
    img = np.full((64, 64), 10, np.uint8)
    img[:, 32:] = 240
    out = O.clahe(img, 2.0, (1, 1))
    assert (out[:, :32] == 13).all() and (out[:, 32:] == 241).all()


def test_clahe_one_tile_no_clip_is_histogram_equalisation():
    # clip 0 and a single tile reduce to textbook equalisation,
    # round_even(cdf[v] * 255 / N) (float32 product as OpenCV's LUT)
    rng = np.random.default_rng(5)
    img = np.clip(rng.normal(90, 30, (57, 83)), 0, 255).astype(np.uint8)
    cdf = np.cumsum(np.bincount(img.ravel(), minlength=256))
    lut = np.rint(cdf.astype(np.float32) * np.float32(255.0 / img.size)).astype(np.uint8)
    np.testing.assert_array_equal(O.clahe(img, 0.0, (1, 1)), lut[img])
    # with clipping the one-tile map stays monotone in the input value
    out = O.clahe(img, 2.0, (1, 1))
    order = np.argsort(img.ravel(), kind="stable")
    assert (np.diff(out.ravel()[order].astype(int)) >= 0).all()


def test_clahe_bilinear_ramp_between_tile_centres():
    # two 32-wide tiles, left all 50, right all 200, no clipping:
    # LUT_left[v] = 255 for v >= 50 else 0, LUT_right[v] = 255 for v >= 200.
    # Right pixels map to 255 through both; left pixels get 255 times the
    # left tile's weight: 1 up to its centre (x = 15.5), then 1.5 - x / 32.
    img = np.full((8, 64), 50, np.uint8)
    img[:, 32:] = 200
    x = np.arange(64)
    exp = np.where((x >= 16) & (x < 32), np.rint(255.0 * (1.5 - x / 32.0)), 255).astype(np.uint8)
    out = O.clahe(img, 0.0, (2, 1))
    np.testing.assert_array_equal(out, np.broadcast_to(exp, (8, 64)))
    # the same along y (tiles (1, 2) on the transposed image)
    np.testing.assert_array_equal(O.clahe(np.ascontiguousarray(img.T), 0.0, (1, 2)), np.broadcast_to(exp, (8, 64)).T)
    # right tile: 200 pixels of 30 (columns 32..56) and 56 of 50, so
    # LUT_right[30] = round(200 * 255 / 256) = 199 and LUT_left[30] = 0; the
    # 30s left of the right centre blend to 199 (x / 32 - 0.5), fractions
    # above one half (rounded, not truncated) and a tie 99.5 -> 100 at x = 32
    img = np.full((8, 64), 50, np.uint8)
    img[:, 32:57] = 30
    exp = np.full(64, 255.0)
    exp[32:48] = np.rint(199.0 * (x[32:48] / 32.0 - 0.5))
    exp[48:57] = 199
    assert exp[32] == 100 and exp[33] == 106
    np.testing.assert_array_equal(O.clahe(img, 0.0, (2, 1)), np.broadcast_to(exp.astype(np.uint8), (8, 64)))


def test_clahe_ragged_image_tiles_reflect101_padding():
    # H = 61, W = 60, 4 x 4 tiles: H is ragged, so the image is padded by 3
    # rows AND by 4 - 60 % 4 = 4 columns (OpenCV pads both axes), giving
    # 16 x 16 tiles of the 64 x 64 BORDER_REFLECT_101 extension (numpy's
    # "reflect").  At a tile centre (16 m + 8, 16 k + 8) the bilinear weights
    # are exactly (1, 0), so the output is that tile's own equalisation LUT.
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (61, 60)).astype(np.uint8)
    img[:30, :30] //= 3
    ext = np.pad(img, ((0, 3), (0, 4)), mode="reflect")
    out = O.clahe(img, 0.0, (4, 4))
    for m in range(4):
        for k in range(4):
            tile = ext[16 * m:16 * m + 16, 16 * k:16 * k + 16]
            cdf = np.cumsum(np.bincount(tile.ravel(), minlength=256))
            lut = np.rint(cdf * 255.0 / 256.0).astype(np.uint8)
            y, x = 16 * m + 8, 16 * k + 8
            assert out[y, x] == lut[img[y, x]], (m, k)


def test_clahe_rejects_bad_arguments():
    with pytest.raises(ValueError):
        O.clahe(np.zeros((8, 8), np.uint8), 2.0, (0, 8))


def test_bgr_to_gray_pinned_by_bt601():
    """cv2.cvtColor(BGR2GRAY) (core.py:2548) is Y = 0.299 R + 0.587 G + 0.114 B
    rounded to nearest (OpenCV's documented BT.601 luma).  Pinned without the
    oracle's fixed-point constants: a gray pixel maps to itself, the primaries
    give OpenCV's documented 29 / 150 / 76 (blue, green, red: a swapped channel
    order fails), and over every 4th value of each channel the result is the
    nearest integer to the float luma wherever that is not within 0.01 of a
    half (the fixed-point coefficients differ from the decimals by < 2e-5)."""
    v = np.arange(256, dtype=np.uint8)
    np.testing.assert_array_equal(O.bgr_to_gray(np.stack([v, v, v], -1)[None])[0], v)
    prim = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]], np.uint8)
    np.testing.assert_array_equal(O.bgr_to_gray(prim)[0], [29, 150, 76, 255, 0])
    g = np.arange(0, 256, 4, dtype=np.int64)
    b, gg, r = np.meshgrid(g, g, g, indexing="ij")
    bgr = np.stack([b, gg, r], -1).reshape(-1, 1, 3).astype(np.uint8)
    got = O.bgr_to_gray(bgr)[:, 0].astype(np.int64)
    y = 0.114 * b.ravel() + 0.587 * gg.ravel() + 0.299 * r.ravel()
    assert np.abs(got - y).max() <= 0.5 + 0.006
    clear = np.abs((y - np.floor(y)) - 0.5) > 0.01
    assert clear.mean() > 0.97
    np.testing.assert_array_equal(got[clear], np.floor(y[clear] + 0.5).astype(np.int64))


def _feather_weight(n, p):
    """Distance of source pixel p (0..n-1) to the nearest invalid pixel of its
    own image along one axis: 1 at either edge pixel."""
    return np.minimum(p + 1, n - p)


@pytest.mark.parametrize("tx,ty,a,b", [(7, 3, 200, 40), (-5, 11, 13, 250), (30, -20, 255, 0)])
def test_feather_overlap_closed_form(tx, ty, a, b):
    """Two constant images, the first moved by an integer translation: inside
    the overlap the blend is trunc((a w1 + b w2) / (w1 + w2)) with w the
    per-image distance to the nearest invalid canvas pixel, here
    min(x + 1, w - x, y + 1, h - y) in each image's own pixels (for a
    translation the canvas distance to the warped image's invalid region, so
    the definition of the stitch's feather weights).  Outside the overlap
    each image comes through unchanged.  Swapped weights or a +-1 offset in
    either weight changes many overlap pixels by several levels."""
    h1, w1, h2, w2 = 40, 50, 36, 44
    img1 = np.full((h1, w1, 3), a, np.uint8)
    img2 = np.full((h2, w2, 3), b, np.uint8)
    T = np.array([[1.0, 0, tx], [0, 1.0, ty], [0, 0, 1.0]])
    cw, ch, ox, oy = O.canvas(h1, w1, h2, w2, T)
    assert ox == max(0.0, -tx) and oy == max(0.0, -ty)
    pano, valid = O.warp_feather(img1, img2, T, cw, ch, ox, oy)
    v, u = np.mgrid[0:ch, 0:cw]
    x1, y1 = u - ox - tx, v - oy - ty           # img1 source pixel
    x2, y2 = u - round(ox), v - round(oy)       # img2 pasted at (round(ox), round(oy))
    in1 = (x1 >= 0) & (x1 <= w1 - 1) & (y1 >= 0) & (y1 <= h1 - 1)
    in2 = (x2 >= 0) & (x2 < w2) & (y2 >= 0) & (y2 < h2)
    W1 = np.where(in1, np.minimum(_feather_weight(w1, x1), _feather_weight(h1, y1)), 0).astype(np.float64)
    W2 = np.where(in2, np.minimum(_feather_weight(w2, x2), _feather_weight(h2, y2)), 0).astype(np.float64)
    np.testing.assert_array_equal(valid, in1 | in2)
    both = in1 & in2
    assert both.sum() > 200
    exact = (a * W1 + b * W2) / np.maximum(W1 + W2, 1)
    got = pano[..., 0].astype(np.int64)
    for c in range(3):
        np.testing.assert_array_equal(pano[..., c], pano[..., 0])
    np.testing.assert_array_equal(got[in1 & ~in2], a)
    np.testing.assert_array_equal(got[in2 & ~in1], b)
    assert not pano[~(in1 | in2)].any()
    # float32 arithmetic: equal to trunc of the exact value away from integers
    frac = exact - np.floor(exact)
    sure = both & (frac > 1e-4) & (frac < 1 - 1e-4)
    assert sure.sum() > 0.5 * both.sum()
    np.testing.assert_array_equal(got[sure], np.floor(exact[sure]).astype(np.int64))
    assert np.abs(got[both] - exact[both]).max() < 1.0


def _largest_valid_rect_area(m):
    """Brute force: the largest all-valid axis-aligned rectangle of a small mask."""
    h, w = m.shape
    c = np.zeros((h + 1, w + 1), np.int64)
    c[1:, 1:] = np.cumsum(np.cumsum(m, 0), 1)
    best = 0
    for y0 in range(h):
        for y1 in range(y0 + 1, h + 1):
            for x0 in range(w):
                for x1 in range(x0 + 1, w + 1):
                    area = (y1 - y0) * (x1 - x0)
                    if area > best and c[y1, x1] - c[y0, x1] - c[y1, x0] + c[y0, x0] == area:
                        best = area
    return best


def _box_of(crop_of_index, shape):
    """Rectangle (y0, y1, x0, x1) of a crop of an index image."""
    ys, xs = np.divmod(crop_of_index, shape[1])
    return int(ys.min()), int(ys.max()) + 1, int(xs.min()), int(xs.max()) + 1


def test_autocrop_closed_forms():
    """_autocrop (core.py:3627-3684): a rectangle mask gives itself; the union
    of two overlapping rectangles (two views under a translation) gives their
    intersection; an empty mask returns the input; a mask too thin for an
    inner rectangle falls back to its bounding box."""
    idx = np.arange(30 * 40).reshape(30, 40)
    m = np.zeros((30, 40), bool); m[4:21, 6:33] = True
    assert _box_of(O.autocrop(idx, m), m.shape) == (4, 21, 6, 33)
    for (ay, ax, by, bx) in [(2, 3, 9, 15), (9, 15, 2, 3), (2, 15, 9, 3)]:
        m = np.zeros((30, 40), bool)
        m[ay:ay + 15, ax:ax + 20] = True
        m[by:by + 15, bx:bx + 20] = True
        y0, y1 = max(ay, by), min(ay, by) + 15
        x0, x1 = max(ax, bx), min(ax, bx) + 20
        assert _box_of(O.autocrop(idx, m), m.shape) == (y0, y1, x0, x1), (ay, ax, by, bx)
    assert O.autocrop(idx, np.zeros((30, 40), bool)) is idx
    m = np.zeros((30, 40), bool); m[10, 5:25] = True          # one row: no inner rectangle
    assert _box_of(O.autocrop(idx, m), m.shape) == (10, 11, 5, 25)
    m = np.zeros((30, 40), bool); m[3:20, 7] = True
    assert _box_of(O.autocrop(idx, m), m.shape) == (3, 20, 7, 8)


def _view_footprint(rng, h, w):
    """Canvas footprint of two views: an axis-aligned rectangle (the pasted
    view) and a rotated, scaled rectangle (the warped view, rotation within
    15 degrees), rasterised."""
    m = np.zeros((h, w), bool)
    y0, x0 = rng.integers(0, h // 3), rng.integers(0, w // 3)
    m[y0:y0 + rng.integers(h // 2, h - y0 + 1), x0:x0 + rng.integers(w // 2, w - x0 + 1)] = True
    cx, cy = rng.uniform(0.4, 0.6) * w, rng.uniform(0.4, 0.6) * h
    th = np.radians(rng.uniform(-15, 15))
    a, b = rng.uniform(0.25, 0.45) * w, rng.uniform(0.25, 0.45) * h
    v, u = np.mgrid[0:h, 0:w]
    du, dv = u - cx, v - cy
    m |= (np.abs(np.cos(th) * du + np.sin(th) * dv) <= a) & (np.abs(-np.sin(th) * du + np.cos(th) * dv) <= b)
    return m


def test_autocrop_row_column_convex_masks():
    """On masks whose every row and column is one run of valid pixels (the
    union of two overlapping view footprints), the rule's column range is exactly
    the set of columns valid in every row of the bounding box, and its row
    range the rows valid in every column: the crop is all valid, and no larger
    than the largest valid rectangle found by brute force.  A mask whose
    every-row column set or every-column row set is shorter than 2 falls back
    to its bounding box."""
    rng = np.random.default_rng(12)
    n_checked = 0
    for t in range(40):
        h, w = rng.integers(12, 24, 2)
        m = _view_footprint(rng, h, w)
        runs_r = [np.flatnonzero(np.diff(np.r_[0, r.astype(np.int8), 0])) for r in m if r.any()]
        runs_c = [np.flatnonzero(np.diff(np.r_[0, c.astype(np.int8), 0])) for c in m.T if c.any()]
        if not m.any() or any(len(x) != 2 for x in runs_r + runs_c):
            continue
        rows = np.flatnonzero(m.any(1)); cols = np.flatnonzero(m.any(0))
        sub = m[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1]
        full_c = np.flatnonzero(sub.all(0)); full_r = np.flatnonzero(sub.all(1))
        idx = np.arange(m.size).reshape(m.shape)
        y0, y1, x0, x1 = _box_of(O.autocrop(idx, m), m.shape)
        if len(full_c) >= 2 and len(full_r) >= 2:
            assert (y0, y1, x0, x1) == (rows[0] + full_r[0], rows[0] + full_r[-1] + 1,
                                        cols[0] + full_c[0], cols[0] + full_c[-1] + 1), t
            assert m[y0:y1, x0:x1].all()
            assert (y1 - y0) * (x1 - x0) <= _largest_valid_rect_area(m)
            n_checked += 1
        else:
            assert (y0, y1, x0, x1) == (rows[0], rows[-1] + 1, cols[0], cols[-1] + 1), t
    assert n_checked >= 10


def test_u8_unit_by_double_product_is_exact():
    # the device column prefilter forms (float)v / 255 as (float)(v * (1 / 255.0))
    # in double: equal to the correctly rounded float quotient for every byte
    v = np.arange(256)
    q = np.float32(v) / np.float32(255)
    np.testing.assert_array_equal((v.astype(np.float64) * (1.0 / 255.0)).astype(np.float32), q)
    # a plausible slip, a float reciprocal, differs for some bytes
    assert not np.array_equal(np.float32(v) * np.float32(1.0 / 255.0), q)
