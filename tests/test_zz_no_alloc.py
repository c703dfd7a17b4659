"""msift_match_u8_noalloc (static device scratch, no allocation) against
msift_match_u8 (caller workspace), bit for bit, through the C-ABI: two calls
in a row on one GPU, the first at the size limit, the second smaller, so the
second call runs over the scratch contents the first one left."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_ROWS = 16384


def _desc(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 128, (n, 128)).astype(np.uint8)
    hot = rng.random((n, 128)) < 0.1
    d[hot] = rng.integers(128, 256, hot.sum()).astype(np.uint8)
    return d


def _pair(N, M, seed):
    a = _desc(N, seed)
    b = _desc(M, seed + 1)
    k = min(N, M) // 2
    noise = np.random.default_rng(seed + 2).integers(-3, 4, (k, 128))
    b[:k] = np.clip(a[:k].astype(int) + noise, 0, 255).astype(np.uint8)
    return a, b


def _outputs(torch, N, M):
    return (torch.full((max(N, 1), 3), -7, dtype=torch.int32, device="cuda"),
            torch.full((max(M, 1), 3), -7, dtype=torch.int32, device="cuda"),
            torch.full((max(min(N, M), 1), 2), -7, dtype=torch.int32, device="cuda"),
            torch.full((1,), -7, dtype=torch.int32, device="cuda"))


def _call(torch, a, b, method, noalloc):
    from pysift_amd import _lib
    L = _lib.lib()
    N, M = a.shape[0], b.shape[0]
    fwd, bwd, pairs, npairs = _outputs(torch, N, M)
    s = _lib.stream_ptr(a.device)
    if noalloc:
        rc = L.msift_match_u8_noalloc(_lib.ptr(a), N, _lib.ptr(b), M, 0.75, method, _lib.ptr(fwd), _lib.ptr(bwd),
                                      _lib.ptr(pairs), _lib.ptr(npairs), s)
    else:
        n = ctypes.c_size_t()
        _lib.check(L.msift_match_workspace_bytes(N, M, ctypes.byref(n)))
        ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
        rc = L.msift_match_u8(_lib.ptr(a), N, _lib.ptr(b), M, 0.75, method, _lib.ptr(ws), ws.numel(),
                              _lib.ptr(fwd), _lib.ptr(bwd), _lib.ptr(pairs), _lib.ptr(npairs), s)
    _lib.check(rc, "noalloc" if noalloc else "match_u8")
    torch.cuda.synchronize()
    k = int(npairs.item())
    return fwd[:N].cpu().numpy(), bwd[:M].cpu().numpy(), pairs[:k].cpu().numpy(), k


@pytest.mark.parametrize("method", [0, 1, 2])
def test_noalloc_equals_workspace_entry_two_calls(method):
    import torch
    for N, M, seed in ((MAX_ROWS, MAX_ROWS, 11), (700, 900, 12)):
        print(f"[case] noalloc method {method} {N}x{M}", flush=True)
        a_np, b_np = _pair(N, M, seed)
        a = torch.from_numpy(a_np).cuda()
        b = torch.from_numpy(b_np).cuda()
        ref = _call(torch, a, b, method, noalloc=False)
        got = _call(torch, a, b, method, noalloc=True)
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
        assert ref[3] > 0
        if N < 2000:
            rf, rb, rp = O.match_u8(a_np, b_np, 0.75)
            np.testing.assert_array_equal(got[0], rf)
            np.testing.assert_array_equal(got[1], rb)
            np.testing.assert_array_equal(got[2], rp)


def test_noalloc_rejects_sizes_beyond_scratch():
    import torch
    from pysift_amd import _lib
    L = _lib.lib()
    a = torch.zeros((MAX_ROWS + 1, 128), dtype=torch.uint8, device="cuda")
    b = torch.zeros((8, 128), dtype=torch.uint8, device="cuda")
    fwd, bwd, pairs, npairs = _outputs(torch, MAX_ROWS + 1, 8)
    rc = L.msift_match_u8_noalloc(_lib.ptr(a), MAX_ROWS + 1, _lib.ptr(b), 8, 0.75, 1, _lib.ptr(fwd), _lib.ptr(bwd),
                                  _lib.ptr(pairs), _lib.ptr(npairs), _lib.stream_ptr(a.device))
    assert rc == _lib.MSIFT_EINVAL
    assert "16384" in _lib.last_error()
    torch.cuda.synchronize()
    assert int(npairs.item()) == -7      # nothing ran
