"""GPU RANSAC vs the CPU oracle: winning hypothesis, every score, H and the
inlier mask are bit-exact; H recovers a known homography."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _corr(N, seed, out_frac=0.3, noise=0.5):
    rng = np.random.default_rng(seed)
    Hgt = np.array([[0.95, 0.08, 120.0], [-0.05, 1.02, -30.0], [1e-4, -5e-5, 1.0]])
    p1 = rng.uniform([0, 0], [1024, 768], (N, 2))
    q = np.c_[p1, np.ones(N)] @ Hgt.T
    p2 = q[:, :2] / q[:, 2:]
    p2 += rng.normal(0, noise, p2.shape)
    nout = int(out_frac * N)
    p2[:nout] = rng.uniform([0, 0], [1024, 768], (nout, 2))
    return p1, p2, Hgt


@pytest.mark.parametrize("N,B,method,lo", [(4, 16, "magsac", True), (37, 100, "magsac", True),
                                           (500, 2000, "magsac", True), (500, 2000, "classic", True),
                                           (1200, 1024, "magsac", False), (3000, 333, "classic", True)])
def test_ransac_bitexact(N, B, method, lo):
    import torch
    from pysift_amd.ransac import ransac_homography
    print(f"[case] ransac N={N} B={B} {method} lo={lo}", flush=True)
    p1, p2, Hgt = _corr(N, N + B)
    H, mask, ninl, best, ws = ransac_homography(p1, p2, seed=42, n_hyp=B, thr=3.0, method=method, lo=lo,
                                                return_device=True)
    rH, rmask, rbest, rsc, rHs = O.ransac(p1, p2, seed=42, n_hyp=B, thr=3.0, method=method, lo=lo, return_all=True)
    scores = ws[:8 * B].view(torch.float64).cpu().numpy()
    np.testing.assert_array_equal(scores, rsc)
    assert int(best.item()) == rbest
    np.testing.assert_array_equal(H.cpu().numpy(), rH)
    np.testing.assert_array_equal(mask.cpu().numpy().astype(bool), rmask)
    assert int(ninl.item()) == int(rmask.sum())
    if N >= 500:
        c = np.array([[0, 0, 1], [1024, 0, 1], [1024, 768, 1], [0, 768, 1]], float).T
        a = H.cpu().numpy() @ c; a = a[:2] / a[2]
        g = Hgt @ c; g = g[:2] / g[2]
        assert np.abs(a - g).max() < 2.0


def test_ransac_too_few_points_raises():
    from pysift_amd.ransac import ransac_homography
    with pytest.raises(ValueError):
        ransac_homography(np.zeros((3, 2)), np.zeros((3, 2)))
