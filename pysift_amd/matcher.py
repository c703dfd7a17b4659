"""Exact mutual ratio-test matcher on uint8 descriptors (INT8 MFMA).

Restates GPUPyStitch._match_features / _match_symmetric (reference
pysift/core.py:2840-2940, 2652-2701) on the exact integer distance
d = sum (a - b)^2 of uint8 descriptors: ratio test d1 < ratio^2 d2 in both
directions plus the mutual check, matches ordered by query index.
"""
import ctypes

from . import _lib

METHODS = {"hilo": 0, "shift": 1}


def match_workspace_bytes(N, M):
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().msift_match_workspace_bytes(int(N), int(M), ctypes.byref(n)), "match workspace")
    return n.value


def match_u8(a, b, ratio=0.75, method="hilo", sync=True, ws=None):
    """a: (N,128) uint8 CUDA tensor, b: (M,128).  Returns (fwd (N,3), bwd (M,3),
    pairs (K,2)) int32 device tensors; fwd[i] = (best j, d1, d2)."""
    import torch
    if a.dtype != torch.uint8 or b.dtype != torch.uint8 or a.dim() != 2 or b.dim() != 2 \
            or a.shape[1] != 128 or b.shape[1] != 128:
        raise ValueError("descriptors must be (N,128) uint8")
    a = a.contiguous(); b = b.contiguous()
    dev = a.device
    N, M = a.shape[0], b.shape[0]
    if ws is None or ws.numel() < match_workspace_bytes(N, M):
        ws = torch.empty(max(match_workspace_bytes(N, M), 1), dtype=torch.uint8, device=dev)
    fwd = torch.empty((max(N, 1), 3), dtype=torch.int32, device=dev)
    bwd = torch.empty((max(M, 1), 3), dtype=torch.int32, device=dev)
    pairs = torch.empty((max(min(N, M), 1), 2), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = _lib.lib().msift_match_u8(_lib.ptr(a), N, _lib.ptr(b), M, float(ratio), METHODS[method],
                                   _lib.ptr(ws), ws.numel(), _lib.ptr(fwd), _lib.ptr(bwd), _lib.ptr(pairs),
                                   _lib.ptr(npairs), _lib.stream_ptr(dev))
    _lib.check(rc, "match_u8")
    if not sync:
        return fwd[:N], bwd[:M], pairs, npairs
    k = int(npairs.item())
    return fwd[:N], bwd[:M], pairs[:k]
