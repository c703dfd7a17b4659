"""Batched matcher and batched RANSAC: every pair's result equals the
single-pair call (and so the integer / FP64 CPU oracle) bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PAIR_SIZES = [(0, 50), (50, 0), (1, 1), (2, 3), (300, 700), (513, 257), (64, 64), (1500, 1100), (0, 0), (257, 513)]


def _rand_desc(n, seed, hi_frac=0.1):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 128, (n, 128)).astype(np.uint8)
    hot = rng.random((n, 128)) < hi_frac
    d[hot] = rng.integers(128, 256, hot.sum()).astype(np.uint8)
    return d


def _pair(N, M, seed):
    a = _rand_desc(N, seed)
    b = _rand_desc(M, seed + 1)
    k = min(N, M) // 2
    if k:
        b[:k] = np.clip(a[:k].astype(int) + np.random.default_rng(seed).integers(-3, 4, (k, 128)), 0, 255).astype(np.uint8)
    return a, b


@pytest.mark.parametrize("method", ["hilo", "shift", "twopass"])
def test_batched_match_equals_single_and_oracle(method):
    import torch
    from pysift_amd.matcher import match_u8, match_u8_batched
    print(f"[case] batched match {method} {len(PAIR_SIZES)} pairs", flush=True)
    data = [_pair(N, M, 100 + 7 * p) for p, (N, M) in enumerate(PAIR_SIZES)]
    ta = [torch.from_numpy(a).cuda() for a, _ in data]
    tb = [torch.from_numpy(b).cuda() for _, b in data]
    res = match_u8_batched(ta, tb, 0.75, method)
    for p, ((a, b), (f, bw, pr)) in enumerate(zip(data, res)):
        if a.shape[0] and b.shape[0]:
            sf, sb, sp = match_u8(ta[p], tb[p], 0.75, method)
            np.testing.assert_array_equal(f.cpu().numpy(), sf.cpu().numpy())
            np.testing.assert_array_equal(bw.cpu().numpy(), sb.cpu().numpy())
            np.testing.assert_array_equal(pr.cpu().numpy(), sp.cpu().numpy())
        rf, rb, rp = O.match_u8(a, b, 0.75)
        np.testing.assert_array_equal(f.cpu().numpy(), rf.reshape(-1, 3)[:a.shape[0]])
        np.testing.assert_array_equal(bw.cpu().numpy(), rb.reshape(-1, 3)[:b.shape[0]])
        np.testing.assert_array_equal(pr.cpu().numpy(), rp.reshape(-1, 2))


def _corr(N, seed, out_frac=0.3, noise=0.5):
    rng = np.random.default_rng(seed)
    th = rng.uniform(-0.2, 0.2)
    Hgt = np.array([[np.cos(th), -np.sin(th), rng.uniform(-100, 100)],
                    [np.sin(th), np.cos(th), rng.uniform(-60, 60)], [rng.uniform(-1e-4, 1e-4), 5e-5, 1.0]])
    p1 = rng.uniform([0, 0], [1024, 768], (N, 2))
    q = np.c_[p1, np.ones(N)] @ Hgt.T
    p2 = q[:, :2] / q[:, 2:] + rng.normal(0, noise, (N, 2))
    nout = int(out_frac * N)
    p2[:nout] = rng.uniform([0, 0], [1024, 768], (nout, 2))
    return p1, p2


@pytest.mark.parametrize("method,lo", [("magsac", True), ("classic", True), ("magsac", False)])
def test_batched_ransac_equals_single_per_pair(method, lo):
    import torch
    from pysift_amd.ransac import RansacBatchPlan, ransac_homography
    sizes = [4, 37, 3, 500, 0, 1200, 60]
    print(f"[case] batched ransac {method} lo={lo} sizes={sizes}", flush=True)
    pts = [_corr(n, 900 + n) if n else (np.zeros((0, 2)), np.zeros((0, 2))) for n in sizes]
    caps = [n + 5 for n in sizes]                # room beyond the count: only npts points are used
    plan = RansacBatchPlan(caps, n_hyp=256)
    tot = sum(caps)
    p1 = torch.zeros((tot, 2), dtype=torch.float64, device="cuda")
    p2 = torch.full((tot, 2), 7.0, dtype=torch.float64, device="cuda")
    for (a, b), off, n in zip(pts, plan.offsets, sizes):
        if n:
            p1[off:off + n] = torch.from_numpy(a).cuda()
            p2[off:off + n] = torch.from_numpy(b).cuda()
    npts = torch.tensor(sizes, dtype=torch.int32, device="cuda")
    H, mask, ninl, best = plan.run(p1, p2, npts, seed=42, thr=3.0, method=method, lo=lo)
    H = H.cpu().numpy(); mask = mask.cpu().numpy(); ninl = ninl.cpu().numpy(); best = best.cpu().numpy()
    for p, ((a, b), off, n) in enumerate(zip(pts, plan.offsets, sizes)):
        if n < 4:
            assert np.all(H[p] == 0) and ninl[p] == 0 and best[p] == -1
            continue
        sH, smask, sninl, sbest, _ = ransac_homography(a, b, seed=42 + p, n_hyp=256, thr=3.0, method=method, lo=lo,
                                                       return_device=True)
        np.testing.assert_array_equal(H[p], sH.cpu().numpy())
        np.testing.assert_array_equal(mask[off:off + n], smask.cpu().numpy())
        assert ninl[p] == int(sninl.item()) and best[p] == int(sbest.item())
        rH, rmask = O.ransac(a, b, seed=42 + p, n_hyp=256, thr=3.0, method=method, lo=lo)
        np.testing.assert_array_equal(H[p], rH)


def test_gather_points_follows_pairs():
    import torch
    from pysift_amd.matcher import MatchBatchPlan
    sizes = [(40, 60), (0, 10), (90, 30)]
    data = [_pair(N, M, 300 + p) for p, (N, M) in enumerate(sizes)]
    plan = MatchBatchPlan([N for N, _ in sizes], [M for _, M in sizes])
    A = plan.pack([torch.from_numpy(a).cuda() for a, _ in data], plan.Ns)
    B = plan.pack([torch.from_numpy(b).cuda() for _, b in data], plan.Ms)
    plan.run(A, B)
    rng = np.random.default_rng(5)
    kpa = torch.from_numpy(rng.random((sum(plan.Ns), 6)).astype(np.float32)).cuda()
    kpb = torch.from_numpy(rng.random((sum(plan.Ms), 6)).astype(np.float32)).cuda()
    p1, p2 = plan.gather_points(kpa, kpb)
    res = plan.results()
    ka, kb, q1, q2 = kpa.cpu().numpy(), kpb.cpu().numpy(), p1.cpu().numpy(), p2.cpu().numpy()
    for p, (_, _, pr) in enumerate(res):
        pr = pr.cpu().numpy()
        ao, bo = plan.a_off[p], plan.b_off[p]
        for k, (i, j) in enumerate(pr):
            np.testing.assert_array_equal(q1[ao + k], ka[ao + i, :2].astype(np.float64))
            np.testing.assert_array_equal(q2[ao + k], kb[bo + j, :2].astype(np.float64))
