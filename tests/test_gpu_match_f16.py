"""FP16-MFMA float-descriptor matcher: squared distances against FP64, and
the nearest neighbours against an FP64 brute force."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _unit_desc(n, seed):
    # RootSIFT-like: non-negative, unit L2, a few dominant bins
    rng = np.random.default_rng(seed)
    d = rng.gamma(0.6, 1.0, (n, 128))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


@pytest.mark.parametrize("N,M", [(1, 1), (37, 300), (513, 700), (2000, 1500)])
def test_f16_distances_and_neighbours_against_fp64(N, M):
    import torch
    from pysift_amd.matcher import match_f16
    print(f"[case] match_f16 {N}x{M}", flush=True)
    a = _unit_desc(N, N + 7)
    b = _unit_desc(M, M + 11)
    k = min(N, M) // 2
    b[:k] = a[:k] + np.random.default_rng(3).normal(0, 0.01, (k, 128)).astype(np.float32)
    b[:k] /= np.linalg.norm(b[:k], axis=1, keepdims=True)
    fwd, bwd, pairs = match_f16(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    f = fwd.cpu().numpy()
    j = f[:, 0]
    d1 = f[:, 1].view(np.float32)
    A, B = a.astype(np.float64), b.astype(np.float64)
    D = (A * A).sum(1)[:, None] + (B * B).sum(1)[None, :] - 2 * A @ B.T
    # distance of the returned neighbour against FP64 (fp16 rounding of unit inputs)
    err = np.abs(d1 - D[np.arange(N), j])
    assert err.max() < 2e-3, err.max()
    # the returned neighbour is the FP64 nearest one up to that error
    assert np.all(D[np.arange(N), j] <= D.min(1) + 4e-3)
    if N >= 37:
        assert np.mean(j == D.argmin(1)) > 0.97
    p = pairs.cpu().numpy()
    if k >= 10:
        assert np.mean(p[:, 0] == p[:, 1]) > 0.9 and len(p) >= 0.8 * k
