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


class RansacBatchPlan:
    """Batched RANSAC over pairs of point sets (pair p has room for caps[p]
    points at point offset `offsets[p]`); counts are read on the device.  Pair
    p uses key seed + p, so it equals ransac_homography(seed=seed + p)."""

    def __init__(self, caps, n_hyp=1024, device=None):
        import torch
        self.caps = [int(c) for c in caps]
        self.np = len(self.caps)
        self.n_hyp = int(n_hyp)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        L = _lib.lib()
        arr = (ctypes.c_int * self.np)(*self.caps)
        p = ctypes.c_void_p()
        _lib.check(L.msift_ransac_batch_create(self.np, arr, self.n_hyp, ctypes.byref(p)), "ransac batch")
        self._p = p
        ws = ctypes.c_size_t()
        offs = (ctypes.c_longlong * self.np)()
        _lib.check(L.msift_ransac_batch_sizes(p, ctypes.byref(ws), offs), "ransac batch sizes")
        self.offsets = list(offs)
        dev = self.device
        self.ws = torch.empty(max(ws.value, 1), dtype=torch.uint8, device=dev)
        self.H = torch.empty((self.np, 3, 3), dtype=torch.float64, device=dev)
        self.mask = torch.zeros(max(sum(self.caps), 1), dtype=torch.uint8, device=dev)
        self.ninl = torch.zeros(self.np, dtype=torch.int32, device=dev)
        self.best = torch.zeros(self.np, dtype=torch.int32, device=dev)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and _lib._lib is not None:
            _lib._lib.msift_ransac_batch_destroy(p)
            self._p = None

    def run(self, p1, p2, npts, seed=42, thr=5.0, sigma_max=None, method="magsac", lo=True, min_inliers=10):
        """p1, p2: (sum caps, 2) float64 device tensors; npts: (n_pairs,) int32 device."""
        import torch
        need = max((o + c for o, c in zip(self.offsets, self.caps)), default=0)
        for name, t in (("p1", p1), ("p2", p2)):
            if t.dtype != torch.float64 or t.dim() != 2 or t.shape[1] != 2 or t.shape[0] < need:
                raise ValueError(f"{name} must be a ({need}, 2) float64 tensor (got {tuple(t.shape)} {t.dtype})")
            if t.device != self.device:
                raise ValueError(f"{name} is on {t.device}, the plan on {self.device}")
        if npts.dtype != torch.int32 or npts.dim() != 1 or npts.numel() != self.np or npts.device != self.device:
            raise ValueError(f"npts must be a ({self.np},) int32 tensor on {self.device}")
        if lo and min_inliers < 4:
            raise ValueError("min_inliers must be >= 4 when lo is set")
        sm = thr if sigma_max is None else sigma_max
        rc = _lib.lib().msift_ransac_homography_batched(
            self._p, _lib.ptr(p1.contiguous()), _lib.ptr(p2.contiguous()), _lib.ptr(npts), int(seed) & 0xFFFFFFFFFFFFFFFF,
            float(thr), float(sm), 0 if method == "magsac" else 1, int(bool(lo)), int(min_inliers), _lib.ptr(self.ws),
            self.ws.numel(), _lib.ptr(self.H), _lib.ptr(self.mask), _lib.ptr(self.ninl), _lib.ptr(self.best),
            _lib.stream_ptr(self.device))
        _lib.check(rc, "ransac_homography_batched")
        return self.H, self.mask, self.ninl, self.best


def ransac_homography_batched(point_pairs, seed=42, n_hyp=1024, thr=5.0, sigma_max=None, method="magsac", lo=True,
                              min_inliers=10, device=None):
    """point_pairs: list of (pts1 (N,2), pts2 (N,2)).  Returns per pair
    (H (3,3) float64, mask (N,) bool) NumPy arrays; pair p is seeded seed + p."""
    import torch
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    caps = [int(np.asarray(a).shape[0]) if not hasattr(a, "shape") else int(a.shape[0]) for a, _ in point_pairs]
    plan = RansacBatchPlan(caps, n_hyp, dev)
    tot = max(sum(caps), 1)
    p1 = torch.zeros((tot, 2), dtype=torch.float64, device=dev)
    p2 = torch.zeros((tot, 2), dtype=torch.float64, device=dev)
    for (a, b), off, c in zip(point_pairs, plan.offsets, caps):
        if c:
            p1[off:off + c] = torch.as_tensor(a, dtype=torch.float64).to(dev)
            p2[off:off + c] = torch.as_tensor(b, dtype=torch.float64).to(dev)
    npts = torch.tensor(caps, dtype=torch.int32, device=dev)
    H, mask, _, _ = plan.run(p1, p2, npts, seed, thr, sigma_max, method, lo, min_inliers)
    Hn = H.cpu().numpy()
    m = mask.cpu().numpy().astype(bool)
    return [(Hn[p], m[off:off + c]) for p, (off, c) in enumerate(zip(plan.offsets, caps))]
