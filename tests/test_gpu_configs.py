"""The benchmark workloads at their full sizes, checked against the CPU oracle.

* 2 x 1024x768 BGR stitch (CLAHE, detect, exact match, 2000-hypothesis
  RANSAC, feather): matches, inliers, homography and panorama bit-exact; the
  homography within 2 px (mean corner error) of the ground truth.
* 64 uneven images up to 1920x1080: the batched graph-replayed pipeline
  equals the per-image path on every image and the oracle on a fixed sample;
  the 32 warped pairs (2k-8k keypoints per image) matched in one launch and
  verified by batched RANSAC (1024 hypotheses per pair) equal the per-pair
  calls on every pair and the oracle on a fixed sample.
* one 4000x3000 pair with 20k-40k keypoints each: detection equal to the
  oracle, the exact matcher (both methods) equal to the integer oracle on the
  full keypoint sets and on 40000 x 40000 random descriptors, 16384-hypothesis
  RANSAC and the feather panorama equal to the oracle.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    O.set_threads(min(os.cpu_count() or 8, 16))


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
    return pairs, H, mask


def test_stitch_1024x768_bgr_pair():
    from pysift_amd import GPUPyStitch
    va, vb, Hgt = syn.synth_pair(seed=1000)
    assert va.shape == (768, 1024, 3)
    st = GPUPyStitch(inlier_thresh=3.0, ransac_iters=2000, verbose=False, save_output=False)
    pano = st.stitch(va, vb)
    pairs, rH, rmask = _oracle_stitch(va, vb, 3.0, 2000)
    print(f"[case] stitch 1024x768: {len(pairs)} matches, {int(rmask.sum())} inliers", flush=True)
    assert st.last_metrics["matches"] == len(pairs)
    assert st.last_metrics["inliers"] == int(rmask.sum())
    np.testing.assert_array_equal(st.last_homography, rH)
    err = syn.corner_error(st.last_homography, Hgt, 1024, 768)
    print(f"[case] stitch 1024x768 corner error {err:.3f} px", flush=True)
    assert err < 2.0
    cw, ch, ox, oy = O.canvas(768, 1024, 768, 1024, rH)
    rp, rv = O.warp_feather(va, vb, rH, cw, ch, ox, oy)
    np.testing.assert_array_equal(pano, O.autocrop(rp, rv))


def _sample(sizes, k=6):
    """Indices of a fixed sample: the smallest and largest image and evenly spaced others."""
    area = np.array([h * w for h, w in sizes])
    idx = {int(area.argmin()), int(area.argmax())}
    idx.update(int(i) for i in np.linspace(0, len(sizes) - 1, k - 2).round())
    return sorted(idx)


def test_batch64_uneven_equals_per_image_and_oracle():
    import torch
    from pysift_amd import PySIFT
    from pysift_amd.batch import BatchPlan
    imgs = syn.synth_batch(64, 2000)
    s = PySIFT()
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.capture()
    plan.run()
    torch.cuda.synchronize()
    res = [(k.cpu().numpy(), d.cpu().numpy()) for k, d in plan.results()]
    res_u8 = [d.cpu().numpy() for _, d in plan.results(u8=True)]
    for i, img in enumerate(imgs):
        kp, df, du = s.detect_raw(img)
        np.testing.assert_array_equal(res[i][0], kp.cpu().numpy(), err_msg=f"image {i} {img.shape}")
        np.testing.assert_array_equal(res[i][1], df.cpu().numpy(), err_msg=f"image {i}")
        np.testing.assert_array_equal(res_u8[i], du.cpu().numpy(), err_msg=f"image {i}")
    for i in _sample([im.shape for im in imgs]):
        rk, rd, ru = O.detect_describe(imgs[i], O.params())
        print(f"[case] batch64 image {i} {imgs[i].shape}: oracle {len(rk)} keypoints", flush=True)
        np.testing.assert_array_equal(res[i][0], rk)
        np.testing.assert_array_equal(res[i][1], rd)
        np.testing.assert_array_equal(res_u8[i], ru)


def test_pairs32_batched_match_ransac_equals_per_pair_and_oracle():
    import torch
    from pysift_amd import PySIFT
    from pysift_amd.batch import BatchPlan
    from pysift_amd.matcher import MatchBatchPlan, match_u8
    from pysift_amd.ransac import RansacBatchPlan, ransac_homography
    n_pairs, n_hyp = 32, 1024
    pairs = syn.synth_pairs_batch(n_pairs, 2100)
    imgs = [im for a, b, _ in pairs for im in (a, b)]
    s = PySIFT()
    plan = BatchPlan(s, [i.shape for i in imgs])
    plan.pack(imgs)
    plan.run(graph=False)
    res = plan.results(u8=True)
    counts = [k.shape[0] for k, _ in res]
    print(f"[case] pairs32: keypoints per image {min(counts)}..{max(counts)}, median {int(np.median(counts))}",
          flush=True)
    assert 2000 <= min(counts) and max(counts) <= 8000
    A, B = res[0::2], res[1::2]
    mp = MatchBatchPlan([k.shape[0] for k, _ in A], [k.shape[0] for k, _ in B])
    DA = mp.pack([d for _, d in A], mp.Ns)
    DB = mp.pack([d for _, d in B], mp.Ms)
    KA = torch.cat([k for k, _ in A]).contiguous()
    KB = torch.cat([k for k, _ in B]).contiguous()
    rp = RansacBatchPlan(mp.Ns, n_hyp)
    mp.run(DA, DB, 0.75, "shift")
    p1, p2 = mp.gather_points(KA, KB)
    H, mask, ninl, _ = rp.run(p1, p2, mp.npairs, seed=42, thr=3.0)
    torch.cuda.synchronize()
    Hn, maskn, npn = H.cpu().numpy(), mask.cpu().numpy(), mp.npairs.cpu().numpy()
    sample = _sample([a.shape for a, _, _ in pairs])
    errs = []
    for p in range(n_pairs):
        (ka, da), (kb, db) = A[p], B[p]
        _, _, pr = match_u8(da, db, 0.75, "shift")
        assert pr.shape[0] == npn[p], p
        q1 = ka[pr[:, 0].long(), :2].double()
        q2 = kb[pr[:, 1].long(), :2].double()
        Hs, ms = ransac_homography(q1, q2, seed=42 + p, n_hyp=n_hyp, thr=3.0)
        np.testing.assert_array_equal(Hn[p], Hs, err_msg=f"pair {p}")
        off = rp.offsets[p]
        np.testing.assert_array_equal(maskn[off:off + len(ms)].astype(bool), ms, err_msg=f"pair {p}")
        errs.append(syn.corner_error(Hs, pairs[p][2], pairs[p][0].shape[1], pairs[p][0].shape[0]))
        if p in sample:
            an, bn = da.cpu().numpy(), db.cpu().numpy()
            _, _, rpairs = O.match_u8(an, bn, 0.75)
            np.testing.assert_array_equal(pr.cpu().numpy(), rpairs, err_msg=f"pair {p}")
            kan, kbn = ka.cpu().numpy(), kb.cpu().numpy()
            rH, rm = O.ransac(kan[rpairs[:, 0], :2].astype(np.float64), kbn[rpairs[:, 1], :2].astype(np.float64),
                              seed=42 + p, n_hyp=n_hyp, thr=3.0)
            np.testing.assert_array_equal(Hs, rH, err_msg=f"pair {p}")
            np.testing.assert_array_equal(ms, rm, err_msg=f"pair {p}")
            print(f"[case] pair {p} {pairs[p][0].shape}: {len(an)} x {len(bn)} descriptors, {len(rpairs)} matches, "
                  f"{int(rm.sum())} inliers, corner error {errs[-1]:.3f} px", flush=True)
    assert np.median(errs) < 1.0


def _rand_desc(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 128, (n, 128)).astype(np.uint8)
    hot = rng.random((n, 128)) < 0.1
    d[hot] = rng.integers(128, 256, hot.sum()).astype(np.uint8)
    return d


def test_12mp_pair_detect_match_ransac_panorama():
    import torch
    from pysift_amd import GPUPyStitch, PySIFT
    from pysift_amd.matcher import match_u8
    from pysift_amd.ransac import ransac_homography
    a, b, Hgt = syn.synth_12mp_pair()
    H, W = a.shape
    s = PySIFT(max_keypoints=40000, max_cands_per_octave=200000)
    p = O.params(max_keypoints=40000, max_cands_per_octave=200000)
    out = []
    for name, img in (("a", a), ("b", b)):
        kp, df, du = s.detect_raw(img)
        rk, rd, ru = O.detect_describe(img, p)
        locs = len(np.unique(rk[:, :2], axis=0))
        print(f"[case] 12 MP image {name}: {len(rk)} keypoints at {locs} locations", flush=True)
        np.testing.assert_array_equal(kp.cpu().numpy(), rk)
        np.testing.assert_array_equal(df.cpu().numpy(), rd)
        np.testing.assert_array_equal(du.cpu().numpy(), ru)
        assert 20000 <= len(rk) <= 40000 and locs >= 20000
        out.append((kp, du))
    (kpa, dua), (kpb, dub) = out
    rf, rb, rpairs = O.match_u8(dua.cpu().numpy(), dub.cpu().numpy(), 0.75)
    for method in ("shift", "hilo"):
        f, bw, pr = match_u8(dua, dub, 0.75, method)
        np.testing.assert_array_equal(f.cpu().numpy(), rf, err_msg=method)
        np.testing.assert_array_equal(bw.cpu().numpy(), rb, err_msg=method)
        np.testing.assert_array_equal(pr.cpu().numpy(), rpairs, err_msg=method)
    print(f"[case] 12 MP match: {len(rpairs)} mutual pairs", flush=True)
    p1 = kpa[torch.from_numpy(rpairs[:, 0]).cuda().long(), :2].double()
    p2 = kpb[torch.from_numpy(rpairs[:, 1]).cuda().long(), :2].double()
    Hs, ms = ransac_homography(p1, p2, seed=42, n_hyp=16384, thr=3.0)
    rH, rm = O.ransac(p1.cpu().numpy(), p2.cpu().numpy(), seed=42, n_hyp=16384, thr=3.0)
    np.testing.assert_array_equal(Hs, rH)
    np.testing.assert_array_equal(ms, rm)
    err = syn.corner_error(Hs, Hgt, W, H)
    print(f"[case] 12 MP RANSAC: {int(rm.sum())} inliers, corner error {err:.3f} px", flush=True)
    assert err < 1.0
    st = GPUPyStitch(save_output=False, verbose=False)
    b1, b2 = np.repeat(a[..., None], 3, 2), np.repeat(b[..., None], 3, 2)
    cw, ch, ox, oy = O.canvas(H, W, H, W, Hs)
    pano, valid = st.warp_feather(torch.from_numpy(b1).cuda(), torch.from_numpy(b2).cuda(), Hs, cw, ch, ox, oy)
    rp, rv = O.warp_feather(b1, b2, Hs, cw, ch, ox, oy)
    np.testing.assert_array_equal(pano.cpu().numpy(), rp)
    np.testing.assert_array_equal(valid.cpu().numpy().astype(bool), rv)


def test_match_40000x40000_exact():
    import torch
    from pysift_amd.matcher import match_u8
    n = 40000
    a = _rand_desc(n, 401)
    b = _rand_desc(n, 402)
    k = n // 2
    b[:k] = np.clip(a[:k].astype(int) + np.random.default_rng(403).integers(-3, 4, (k, 128)), 0, 255).astype(np.uint8)
    rf, rb, rp = O.match_u8(a, b, 0.75)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    for method in ("shift", "hilo"):
        print(f"[case] match 40000x40000 {method}", flush=True)
        f, bw, pr = match_u8(ta, tb, 0.75, method)
        np.testing.assert_array_equal(f.cpu().numpy(), rf, err_msg=method)
        np.testing.assert_array_equal(bw.cpu().numpy(), rb, err_msg=method)
        np.testing.assert_array_equal(pr.cpu().numpy(), rp, err_msg=method)
    assert len(rp) >= k - 100
