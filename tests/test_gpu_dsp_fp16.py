"""DSP-SIFT descriptors (core.py:845-997) and the binary16 pyramid numerics of
the reference's low-VRAM mode (core.py:1207, 1244, 1274, 1306-1308): GPU
bit-exact against the CPU oracle, single-image and batched, including one of
the reference's own fixture images in the mode its fixtures were made in."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "imc")
FIXTURE_MODE = dict(dsp=True, dsp_n_scales=5, fp16_pyramid=True, n_octaves=4, double_px_cap=4_194_304)


def _sift(**kw):
    from pysift_amd import PySIFT
    return PySIFT(**kw)


def _oracle_params(s):
    return O.params(n_octaves=s.n_octaves, n_scales=s.n_scales, sigma0=s.sigma0,
                    contrast_thresh=s.contrast_thresh, edge_thresh=s.edge_thresh,
                    double_image=s.double_image, rootsift=s.rootsift, double_px_cap=s.double_px_cap,
                    max_cands_per_octave=s.max_cands_per_octave, max_keypoints=s.max_keypoints,
                    dsp_n_scales=s.dsp_n_scales if s.dsp else 0, fp16_pyramid=s.fp16_pyramid)


def _compare(img, s, tag):
    kp, df, du = s.detect_raw(img)
    kp = kp.cpu().numpy(); df = df.cpu().numpy(); du = du.cpu().numpy()
    rk, rd, ru = O.detect_describe(img, _oracle_params(s))
    print(f"[case] {tag}: gpu {len(kp)} kp, oracle {len(rk)} kp", flush=True)
    assert kp.shape == rk.shape, (kp.shape, rk.shape)
    np.testing.assert_array_equal(kp, rk)
    np.testing.assert_array_equal(df, rd)
    np.testing.assert_array_equal(du, ru)
    return len(kp)


@pytest.mark.parametrize("H,W,kw", [
    (257, 333, FIXTURE_MODE),
    (300, 200, dict(dsp=True, dsp_n_scales=3)),
    (480, 640, dict(fp16_pyramid=True)),
    (150, 151, dict(fp16_pyramid=True, double_image=False, dsp=True, dsp_n_scales=4)),
    (900, 1100, dict(FIXTURE_MODE, double_px_cap=500_000)),        # above the cap: not doubled
])
def test_detect_bitwise_dsp_fp16(H, W, kw):
    img = syn.synth_gray(H, W, H * 3 + W)
    n = _compare(img, _sift(**kw), f"{H}x{W} {kw}")
    assert n > 20


def test_pyramid_fp16_levels():
    # the device keeps the float32 chain; the stored (binary16) values of
    # octaves >= 2 are what its DoG and samplers read
    import torch
    from pysift_amd import _lib
    H, W = 240, 320
    img = syn.synth_gray(H, W, 21)
    s = _sift(fp16_pyramid=True, n_octaves=4)
    s.detect_raw(img)
    p = _oracle_params(s)
    S = s.n_scales
    for o in range(4):
        print(f"[case] fp16 pyramid octave {o}", flush=True)
        g_ref, d_ref = O.pyramid_octave(img, p, o)
        Ho = ctypes.c_int(); Wo = ctypes.c_int(); g = ctypes.c_void_p(); d = ctypes.c_void_p()
        _lib.check(_lib.lib().msift_pyramid_view(s._h, H, W, _lib.ptr(s._ws), o, ctypes.byref(Ho),
                                                 ctypes.byref(Wo), ctypes.byref(g), ctypes.byref(d)))
        assert not d.value   # the DoG is formed by its readers, not stored
        base = s._ws.data_ptr()
        n = Ho.value * Wo.value
        gs = s._ws[g.value - base: g.value - base + 4 * n * (S + 3)].view(torch.float32)
        gs = gs.reshape(S + 3, Ho.value, Wo.value).cpu().numpy()
        if o >= 2:
            # the levels are stored in float32; their readers round them to binary16
            gs = O.f16_round(gs)
            assert np.array_equal(O.f16_round(g_ref), g_ref)
        ds = gs[1:] - gs[:-1]
        np.testing.assert_array_equal(gs, g_ref, err_msg=f"gauss octave {o}")
        np.testing.assert_array_equal(ds, d_ref, err_msg=f"dog octave {o}")


def test_batched_dsp_fp16_equals_single():
    from pysift_amd.batch import BatchPlan
    sizes = [(240, 320), (301, 417), (150, 151), (97, 203)]
    imgs = [syn.synth_gray(h, w, 60 + i) for i, (h, w) in enumerate(sizes)]
    s = _sift(**FIXTURE_MODE)
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.run(graph=False)
    res = plan.results()
    for i, img in enumerate(imgs):
        print(f"[case] batched fixture-mode image {i} {img.shape}", flush=True)
        kp, df, _ = s.detect_raw(img)
        np.testing.assert_array_equal(res[i][0].cpu().numpy(), kp.cpu().numpy())
        np.testing.assert_array_equal(res[i][1].cpu().numpy(), df.cpu().numpy())


def test_reference_fixture_image_matches_oracle():
    # one of the reference's IMC images in its fixture mode: GPU == oracle,
    # and the oracle is held to the reference's RTX 3050 output on CPU
    from PIL import Image
    im = Image.open(os.path.join(GOLDEN, "imc_british_museum.jpg"))
    im.draft("L", im.size)
    g = np.asarray(im.convert("L"))
    n = _compare(g, _sift(**FIXTURE_MODE), "imc_british_museum DSP-5 fp16")
    assert n > 3000


def test_grad_rejects_dsp_and_fp16():
    import torch
    from pysift_amd.autograd import sift_descriptors
    x = torch.zeros((1, 64, 64), device="cuda")
    k = torch.zeros((1, 1, 5), device="cuda")
    for kw in (dict(dsp=True), dict(fp16_pyramid=True)):
        with pytest.raises(ValueError):
            sift_descriptors(x, k, _sift(double_image=False, **kw))
