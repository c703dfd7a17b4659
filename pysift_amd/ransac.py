"""GPU RANSAC homography with LO refit (reference GPUPyStitch._ransac,
pysift/core.py:3011-3097): Philox4x32-10 hypotheses, FP64 DLT, MAGSAC soft
score (default) or inlier count, symmetric reprojection error."""
import ctypes

import numpy as np

from . import _lib


def ransac_workspace_bytes(N, n_hyp):
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().msift_ransac_workspace_bytes(int(N), int(n_hyp), ctypes.byref(n)), "ransac workspace")
    return n.value


def ransac_homography(pts1, pts2, seed=42, n_hyp=1500, thr=5.0, sigma_max=None, method="magsac", lo=True,
                      min_inliers=10, device=None, return_device=False, ws=None):
    """pts1, pts2: (N,2) float64 (NumPy or CUDA tensors).  Returns
    (H (3,3) float64, inlier mask (N,) bool) as NumPy arrays, or device tensors
    (H, mask, ninl, best, ws) with return_device=True."""
    import torch
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    p1 = torch.as_tensor(pts1, dtype=torch.float64).to(dev).contiguous()
    p2 = torch.as_tensor(pts2, dtype=torch.float64).to(dev).contiguous()
    N = p1.shape[0]
    if N < 4:
        raise ValueError(f"Need >= 4 correspondences, got {N}")
    if p1.shape != (N, 2) or p2.shape != (N, 2):
        raise ValueError("points must be (N,2)")
    sm = thr if sigma_max is None else sigma_max
    need = ransac_workspace_bytes(N, n_hyp)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=dev)
    H = torch.empty(9, dtype=torch.float64, device=dev)
    mask = torch.empty(N, dtype=torch.uint8, device=dev)
    ninl = torch.zeros(1, dtype=torch.int32, device=dev)
    best = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = _lib.lib().msift_ransac_homography(
        _lib.ptr(p1), _lib.ptr(p2), N, int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_hyp), float(thr), float(sm),
        0 if method == "magsac" else 1, int(bool(lo)), int(min_inliers), _lib.ptr(ws), ws.numel(),
        _lib.ptr(H), _lib.ptr(mask), _lib.ptr(ninl), _lib.ptr(best), _lib.stream_ptr(dev))
    _lib.check(rc, "ransac")
    if return_device:
        return H.view(3, 3), mask, ninl, best, ws
    return H.view(3, 3).cpu().numpy(), mask.cpu().numpy().astype(bool)
