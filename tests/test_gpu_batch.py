"""Batched detectAndCompute: every image bit-identical to the single-image path
(itself bit-exact vs the oracle), with HIP-graph replay and every blur tile."""
import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _images():
    sizes = [(240, 320), (301, 417), (150, 151), (480, 640), (97, 203), (200, 300)]
    imgs = [syn.synth_gray(h, w, 40 + i) for i, (h, w) in enumerate(sizes)]
    imgs[4] = np.full((97, 203), 128, np.uint8)          # flat: 0 keypoints, retry path
    return imgs


def test_batch_matches_single_and_oracle():
    from pysift_amd import PySIFT
    from pysift_amd.batch import BatchPlan
    imgs = _images()
    s = PySIFT()
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.run(graph=False)
    res = plan.results()
    res_u8 = plan.results(u8=True)
    for i, img in enumerate(imgs):
        print(f"[case] batch image {i} {img.shape}", flush=True)
        kp, df, du = s.detect_raw(img)
        np.testing.assert_array_equal(res[i][0].cpu().numpy(), kp.cpu().numpy())
        np.testing.assert_array_equal(res[i][1].cpu().numpy(), df.cpu().numpy())
        np.testing.assert_array_equal(res_u8[i][1].cpu().numpy(), du.cpu().numpy())
    rk, rd, ru = O.detect_describe(imgs[0], O.params())
    np.testing.assert_array_equal(res[0][0].cpu().numpy(), rk)
    assert res[4][0].shape[0] == 0


def test_batch_graph_replay_and_tiles_identical():
    import torch
    from pysift_amd import PySIFT
    from pysift_amd.batch import BatchPlan
    imgs = _images()
    s = PySIFT()
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.run(graph=False)
    ref = [(k.cpu().numpy().copy(), d.cpu().numpy().copy()) for k, d in plan.results()]
    for th in (16, 32, 64):
        plan.set_blur_tile(th)
        for rep in range(3):
            plan.kp.zero_()
            plan.run(graph=True)
            torch.cuda.synchronize()
            got = plan.results()
            for (k, d), (rk, rd) in zip(got, ref):
                np.testing.assert_array_equal(k.cpu().numpy(), rk)
                np.testing.assert_array_equal(d.cpu().numpy(), rd)


def test_batch_non_doubled_and_caps():
    from pysift_amd import PySIFT
    from pysift_amd.batch import BatchPlan
    imgs = [syn.synth_gray(700, 900, 7), syn.synth_gray(300, 330, 8)]
    s = PySIFT(double_px_cap=500_000, max_keypoints=150, n_octaves=4)
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.run(graph=False)
    for (k, d), img in zip(plan.results(), imgs):
        kp, df, _ = s.detect_raw(img)
        np.testing.assert_array_equal(k.cpu().numpy(), kp.cpu().numpy())
        np.testing.assert_array_equal(d.cpu().numpy(), df.cpu().numpy())
