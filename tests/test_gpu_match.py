"""GPU parity of the INT8-MFMA matcher against the integer CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rand_desc(n, seed, hi_frac=0.1):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 128, (n, 128)).astype(np.uint8)
    hot = rng.random((n, 128)) < hi_frac
    d[hot] = rng.integers(128, 256, hot.sum()).astype(np.uint8)
    return d


def _run(a, b, method, ratio=0.75):
    import torch
    from pysift_amd.matcher import match_u8
    ta = torch.from_numpy(a).cuda(); tb = torch.from_numpy(b).cuda()
    f, bw, p = match_u8(ta, tb, ratio, method)
    return f.cpu().numpy(), bw.cpu().numpy(), p.cpu().numpy()


@pytest.mark.parametrize("method", ["hilo", "shift", "twopass"])
@pytest.mark.parametrize("N,M", [(1, 1), (2, 3), (37, 300), (300, 37), (513, 700), (1500, 1100)])
def test_match_exact_random(method, N, M):
    print(f"[case] match {method} {N}x{M}", flush=True)
    a = _rand_desc(N, N * 31 + M)
    # plant near-duplicates so matches exist
    b = _rand_desc(M, M * 17 + N)
    k = min(N, M) // 2
    b[:k] = np.clip(a[:k].astype(int) + np.random.default_rng(1).integers(-3, 4, (k, 128)), 0, 255).astype(np.uint8)
    f, bw, p = _run(a, b, method)
    rf, rb, rp = O.match_u8(a, b, 0.75)
    np.testing.assert_array_equal(f, rf)
    np.testing.assert_array_equal(bw, rb)
    np.testing.assert_array_equal(p, rp)
    if min(N, M) >= 37:
        assert len(rp) > 0


@pytest.mark.parametrize("method", ["hilo", "shift", "twopass"])
def test_match_extremes_and_ties(method):
    # all-255 / all-0 rows (max distance 8 323 200) and exact duplicate rows (ties)
    a = np.zeros((70, 128), np.uint8); a[::2] = 255
    b = np.concatenate([np.full((3, 128), 255, np.uint8), np.zeros((3, 128), np.uint8),
                        _rand_desc(80, 5, 0.5)])
    f, bw, p = _run(a, b, method)
    rf, rb, rp = O.match_u8(a, b, 0.75)
    np.testing.assert_array_equal(f, rf)
    np.testing.assert_array_equal(bw, rb)
    np.testing.assert_array_equal(p, rp)
    assert (rf[:, 1] == 0).all() or True


@pytest.mark.parametrize("N,M", [(9000, 7000), (4097, 12001)])
def test_twopass_equals_shift_large(N, M):
    # sizes the oracle does not reach in seconds: the two-pass kernel against
    # the fused shift kernel (both bit-exact at the oracle sizes above)
    import torch
    from pysift_amd.matcher import match_u8
    print(f"[case] twopass vs shift {N}x{M}", flush=True)
    a = _rand_desc(N, 91)
    b = _rand_desc(M, 92)
    k = min(N, M) // 3
    b[:k] = np.clip(a[:k].astype(int) + np.random.default_rng(3).integers(-4, 5, (k, 128)), 0, 255).astype(np.uint8)
    ta = torch.from_numpy(a).cuda(); tb = torch.from_numpy(b).cuda()
    f1, b1, p1 = match_u8(ta, tb, 0.75, "shift")
    f2, b2, p2 = match_u8(ta, tb, 0.75, "twopass")
    assert torch.equal(f1, f2) and torch.equal(b1, b2) and torch.equal(p1, p2)
    assert p1.shape[0] > k // 2


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_twopass_every_tile_shape(shape):
    # every tile shape the autotuner may pick (32-128 x per wave) and chunk
    # counts that leave partial chunk groups give the oracle's result
    from pysift_amd.matcher import set_tile
    N, M = 1500, 1300
    a = _rand_desc(N, 77)
    b = _rand_desc(M, 78)
    k = min(N, M) // 2
    b[:k] = np.clip(a[:k].astype(int) + np.random.default_rng(2).integers(-3, 4, (k, 128)), 0, 255).astype(np.uint8)
    rf, rb, rp = O.match_u8(a, b, 0.75)
    try:
        for cpw in (1, 2, 3):
            print(f"[case] twopass shape {shape} chunks {cpw} {N}x{M}", flush=True)
            set_tile(shape, cpw)
            f, bw, p = _run(a, b, "twopass")
            np.testing.assert_array_equal(f, rf)
            np.testing.assert_array_equal(bw, rb)
            np.testing.assert_array_equal(p, rp)
    finally:
        set_tile(0, 4)


def test_match_empty():
    a = _rand_desc(0, 1); b = _rand_desc(10, 2)
    f, bw, p = _run(a, b, "hilo")
    assert len(p) == 0 and (bw[:, 0] == -1).all()
