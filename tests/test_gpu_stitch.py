"""GPU warp/feather and the full two-image stitch against the CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 1])
def test_warp_feather_bitexact(seed):
    import torch
    from pysift_amd import GPUPyStitch
    from pysift_amd.stitch import compute_canvas_size
    rng = np.random.default_rng(seed)
    img1 = rng.integers(0, 256, (120, 160, 3)).astype(np.uint8)
    img2 = rng.integers(0, 256, (100, 150, 3)).astype(np.uint8)
    H = np.array([[0.9, 0.1, 60.3], [-0.07, 1.05, -20.6], [2e-4, -1e-4, 1.0]])
    cw, ch, ox, oy = compute_canvas_size(120, 160, 100, 150, H)
    assert (cw, ch, ox, oy) == O.canvas(120, 160, 100, 150, H)
    st = GPUPyStitch(verbose=False, save_output=False)
    p, v = st.warp_feather(torch.from_numpy(img1).cuda(), torch.from_numpy(img2).cuda(), H, cw, ch, ox, oy)
    rp, rv = O.warp_feather(img1, img2, H, cw, ch, ox, oy)
    np.testing.assert_array_equal(p.cpu().numpy(), rp)
    np.testing.assert_array_equal(v.cpu().numpy().astype(bool), rv)


def test_bgr_to_gray_bitexact():
    import torch
    from pysift_amd import GPUPyStitch
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (77, 91, 3)).astype(np.uint8)
    st = GPUPyStitch(verbose=False, save_output=False)
    g = st._gray(torch.from_numpy(img).cuda()).cpu().numpy()
    np.testing.assert_array_equal(g, O.bgr_to_gray(img))


def _oracle_stitch(va, vb, thr, n_hyp, max_kp=8000):
    p = O.params(n_octaves=4)
    kps, ds = [], []
    for v in (va, vb):
        k, _, du = O.detect_describe(O.clahe(O.bgr_to_gray(v), 2.0, (8, 8)), p)
        if len(k) > max_kp:
            order = np.argsort(-k[:, 4], kind="stable")[:max_kp]
            k, du = k[order], du[order]
        kps.append(k); ds.append(du)
    _, _, pairs = O.match_u8(ds[0], ds[1], 0.75)
    p1 = kps[0][pairs[:, 0], :2].astype(np.float64)
    p2 = kps[1][pairs[:, 1], :2].astype(np.float64)
    H, mask = O.ransac(p1, p2, seed=42, n_hyp=n_hyp, thr=thr, method="magsac", lo=True)
    return kps, pairs, H, mask


def test_stitch_pair_matches_oracle_and_ground_truth():
    from pysift_amd import GPUPyStitch
    from pysift_amd.stitch import compute_canvas_size
    va, vb, Hgt = syn.synth_pair(seed=1000, view_hw=(384, 512), canvas_hw=(768, 1024))
    st = GPUPyStitch(inlier_thresh=3.0, ransac_iters=2000, verbose=False, save_output=False)
    pano = st.stitch(va, vb)
    kps, pairs, rH, rmask = _oracle_stitch(va, vb, 3.0, 2000)
    print(f"[case] stitch 512x384: {len(pairs)} oracle matches, {int(rmask.sum())} inliers", flush=True)
    assert st.last_metrics["matches"] == len(pairs)
    assert st.last_metrics["inliers"] == int(rmask.sum())
    np.testing.assert_array_equal(st.last_homography, rH)
    # sanity bound only (parity is the bit-exact homography above): this pair
    # overlaps 30-40 % and gives ~40 matches after CLAHE; the oracle's own
    # corner error is 3.8 px, a wrong model is off by hundreds of pixels
    err = syn.corner_error(st.last_homography, Hgt, 512, 384)
    assert err < 6.0, err
    cw, ch, ox, oy = compute_canvas_size(384, 512, 384, 512, rH)
    rp, rv = O.warp_feather(va, vb, rH, cw, ch, ox, oy)
    np.testing.assert_array_equal(pano, O.autocrop(rp, rv))


def test_stitch_errors():
    from pysift_amd import GPUPyStitch
    st = GPUPyStitch(verbose=False, save_output=False)
    flat = np.full((200, 300, 3), 90, np.uint8)
    with pytest.raises(RuntimeError):
        st.stitch(flat, flat)
    with pytest.raises(ValueError):
        st.stitch(flat)


def test_autocrop_box_matches_oracle_autocrop():
    # the device rectangle against the oracle's restatement of _autocrop
    # (core.py:3627-3684) on masks with ragged edges, holes, one valid row or
    # column, an all-valid and an empty mask
    import torch
    from pysift_amd.stitch import autocrop_box
    rng = np.random.default_rng(7)
    masks = []
    for _ in range(6):
        h, w = rng.integers(20, 90, 2)
        m = np.zeros((h, w), bool)
        y0, x0 = rng.integers(0, h // 3), rng.integers(0, w // 3)
        y1, x1 = rng.integers(2 * h // 3, h + 1), rng.integers(2 * w // 3, w + 1)
        m[y0:y1, x0:x1] = True
        for r in range(y0, y1):                       # ragged left and right edges
            m[r, x0:x0 + rng.integers(0, 4)] = False
            m[r, max(x1 - rng.integers(0, 4), x0):x1] = False
        m[rng.integers(0, h, 5), rng.integers(0, w, 5)] = False   # holes
        masks.append(m)
    one_row = np.zeros((30, 40), bool); one_row[12, 5:30] = True
    one_col = np.zeros((30, 40), bool); one_col[3:25, 17] = True
    masks += [one_row, one_col, np.ones((25, 35), bool), np.zeros((10, 12), bool)]
    for m in masks:
        pano = np.arange(m.size * 3, dtype=np.int64).reshape(m.shape + (3,))
        box = autocrop_box(torch.from_numpy(m.astype(np.uint8)).cuda())
        got = pano if box is None else pano[box[0]:box[1], box[2]:box[3]]
        np.testing.assert_array_equal(got, O.autocrop(pano, m))
