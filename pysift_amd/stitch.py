"""GPUPyStitch drop-in: two-image stitch on MI355X.

Mirrors pysift.core.GPUPyStitch (reference pysift/core.py:2400-4092):
constructor keywords, stitch(*imgs | [imgs]) -> uint8 (H, W, 3) panorama,
RuntimeError for fewer than 4 matches / inliers (core.py:3745-3746,
3764-3765), ValueError for a wrong image count (core.py:4003-4004).  The
stages are the library's HIP kernels: BGR->gray, detectAndCompute, the exact
INT8 matcher, Philox RANSAC with LO refit, and warp + feather blend (the
blend that replaces the reference's gain/seam/multiband compositor).
"""
import os
import time
from datetime import datetime

import numpy as np

from . import _lib
from .clahe import CLAHE
from .detector import DMatch, PySIFT
from .matcher import match_float, match_u8
from .ransac import ransac_homography


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pysift_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def compute_canvas_size(h1, w1, h2, w2, H):
    """_compute_canvas_size (core.py:3171-3187): bounding box of H(img1 corners)
    and img2; returns (cw, ch, ox, oy)."""
    c1 = np.array([[0, 0, 1], [w1, 0, 1], [w1, h1, 1], [0, h1, 1]], np.float64).T
    wc = H @ c1
    wc = wc / wc[2]
    ax = np.concatenate([wc[0], [0, w2, w2, 0]])
    ay = np.concatenate([wc[1], [0, 0, h2, h2]])
    ox = max(0.0, -ax.min())
    oy = max(0.0, -ay.min())
    return int(np.ceil(ax.max() + ox)) + 1, int(np.ceil(ay.max() + oy)) + 1, ox, oy


def autocrop_box(both_d):
    """autocrop's rectangle from a device validity mask, computed on the
    device (the same first/last rules: argmax gives the first maximum):
    returns (y0, y1, x0, x1), half-open, or None when no pixel is valid."""
    torch = _torch()
    both_d = both_d != 0
    rows, cols = both_d.any(1), both_d.any(0)
    ri, ci = torch.nonzero(rows).flatten(), torch.nonzero(cols).flatten()
    if ri.numel() == 0:
        return None
    r0, r1, c0, c1 = (int(x) for x in torch.stack([ri[0], ri[-1], ci[0], ci[-1]]).cpu())
    sub = both_d[r0:r1 + 1, c0:c1 + 1].to(torch.uint8)
    nr, nc = sub.shape
    left, right = sub.argmax(1), nc - 1 - sub.flip(1).argmax(1)
    top, bot = sub.argmax(0), nr - 1 - sub.flip(0).argmax(0)
    ic0, ic1, ir0, ir1 = (int(x) for x in torch.stack([left.max(), right.min(), top.max(), bot.min()]).cpu())
    if ic0 >= ic1 or ir0 >= ir1:
        return r0, r1 + 1, c0, c1 + 1
    return r0 + ir0, r0 + ir1 + 1, c0 + ic0, c0 + ic1 + 1


class GPUPyStitch:
    """Two-image panorama stitcher (reference: core.py:2400-4092)."""

    def __init__(self, n_depth_layers=4, n_blend_levels=6, ransac_iters=1500, inlier_thresh=5.0,
                 ransac_method="magsac", sigma_max=None, match_ratio=0.75, match_batch=512, max_keypoints=8000,
                 pca_dims=64, int8_matching=False, orientation="histogram", descriptor="sift", matcher="ratio",
                 color_match="gain", dsp_sift=False, contrast_thresh=0.04, n_octaves=4, n_scales=3, sigma0=1.6,
                 edge_thresh=10.0, double_image=True, max_dimension=0, denoise=False, sharpen=False,
                 clahe_clip=2.0, clahe_tile=8, output_max_dim=0, post_sharpen=False, output_format="png",
                 output_quality=90, *, match_method="twopass", ransac_seed=42, save_output=True, verbose=True,
                 blend="feather", device=None):
        if matcher != "ratio":
            raise NotImplementedError("pysift_amd implements only the ratio-test matcher")
        if max_dimension or denoise or sharpen or output_max_dim or post_sharpen:
            raise NotImplementedError("OpenCV pre/post-processing is not implemented in pysift_amd")
        if ransac_method not in ("magsac", "classic"):
            raise ValueError(f"ransac_method must be 'magsac' or 'classic' (got {ransac_method!r})")
        self.n_depth_layers = n_depth_layers   # MiDaS is unavailable: global homography (core.py:3788-3792)
        if blend not in ("feather", "multiband"):
            raise ValueError(f"blend must be 'feather' or 'multiband' (got {blend!r})")
        if color_match not in ("gain", "hist"):
            raise ValueError(f"color_match must be 'gain' or 'hist' (got {color_match!r})")
        self.blend = blend                     # "multiband": the reference's gain + seam + pyramid blend
        self.n_blend_levels = int(n_blend_levels)
        self.ransac_iters = int(ransac_iters)
        self.inlier_thresh = float(inlier_thresh)
        self.ransac_method = ransac_method
        self.sigma_max = float(sigma_max) if sigma_max is not None else float(inlier_thresh)
        self.match_ratio = float(match_ratio)
        self.match_batch = match_batch
        self.max_keypoints = int(max_keypoints)
        self.pca_dims = pca_dims
        self.int8_matching = int8_matching
        self.matcher = matcher
        self.descriptor = descriptor
        self.color_match = color_match
        self.clahe_clip = clahe_clip
        self.clahe_tile = clahe_tile
        self._clahe = CLAHE(clipLimit=clahe_clip, tileGridSize=(clahe_tile, clahe_tile))
        self.output_format = output_format.lower()
        self.output_quality = output_quality
        if match_method not in ("hilo", "shift", "twopass", "float"):
            raise ValueError(f"match_method must be 'hilo', 'shift', 'twopass' or 'float' (got {match_method!r})")
        # int8_matching (core.py:2894-2910) asks for INT8 GEMM matching.  The
        # integer matchers ('twopass', 'shift', 'hilo') are INT8 MFMA GEMMs and exact, so
        # the flag selects them; the reference's per-row-scaled approximate
        # INT8 similarity on the PCA/float path is not reproduced.
        if int8_matching and match_method == "float":
            raise NotImplementedError("int8_matching with match_method='float' (the approximate per-row-scaled INT8 "
                                      "similarity) is not implemented; the default match_method='twopass' is exact INT8 MFMA")
        self.match_method = match_method
        self._pca = None   # (mean, proj) of the float path, fitted on the first pair (core.py:2729-2736)
        self.ransac_seed = int(ransac_seed)
        self.save_output = bool(save_output)
        self.verbose = bool(verbose)
        self._device = device
        self._detector = PySIFT(n_octaves=n_octaves, n_scales=n_scales, sigma0=sigma0,
                                contrast_thresh=contrast_thresh, edge_thresh=edge_thresh, dsp=dsp_sift,
                                orientation=orientation, descriptor=descriptor, double_image=double_image,
                                max_keypoints=max_keypoints, device=device)
        self.last_timings = None
        self.last_metrics = None
        self.last_homography = None

    # ------------------------------------------------------------------
    def _dev(self):
        torch = _torch()
        return torch.device(self._device) if self._device is not None else torch.device("cuda", torch.cuda.current_device())

    def _upload(self, img):
        torch = _torch()
        if isinstance(img, np.ndarray):
            if img.dtype != np.uint8 or img.ndim != 3 or img.shape[2] != 3:
                raise ValueError("images must be (H, W, 3) uint8 BGR")
            return torch.from_numpy(np.ascontiguousarray(img)).to(self._dev())
        t = img if isinstance(img, torch.Tensor) else torch.from_dlpack(img)
        if t.dtype != torch.uint8 or t.dim() != 3 or t.shape[2] != 3:
            raise ValueError("images must be (H, W, 3) uint8 BGR")
        return t.to(self._dev()).contiguous()

    def _gray(self, bgr):
        torch = _torch()
        H, W = bgr.shape[:2]
        g = torch.empty((H, W), dtype=torch.uint8, device=bgr.device)
        _lib.check(_lib.lib().msift_bgr_to_gray(_lib.ptr(bgr), H, W, _lib.ptr(g), _lib.stream_ptr(bgr.device)),
                   "bgr_to_gray")
        return g

    def _extract_features(self, bgr):
        """_extract_features (core.py:2535-2572): gray -> CLAHE(clahe_clip,
        clahe_tile x clahe_tile) in place -> detectAndCompute -> keep the
        max_keypoints strongest (stable).  Descriptors are uint8 for the exact
        matcher, float32 for match_method "float"."""
        torch = _torch()
        gray = self._gray(bgr)
        self._clahe.apply(gray, dst=gray)
        kp, df, du = self._detector.detect_raw(gray)
        d = df if self.match_method == "float" else du
        if kp.shape[0] > self.max_keypoints:
            order = torch.sort(kp[:, 4], descending=True, stable=True).indices[:self.max_keypoints]
            kp, d = kp[order], d[order]
        return kp, d, tuple(bgr.shape[:2])

    def _match_float(self, d1, d2):
        """The reference's float path (core.py:2875-2939): PCA to pca_dims for
        SIFT descriptors (fitted on the first pair, then cached), binary16
        similarity, symmetric ratio test."""
        dims = int(self.pca_dims) if (self.descriptor == "sift" and self.pca_dims) else 0
        fwd, bwd, pairs, pca = match_float(d1, d2, self.match_ratio, dims, self._pca)
        if pca is not None:
            self._pca = pca
        return fwd, bwd, pairs

    def _match_features(self, d1, d2):
        """Mutual ratio test; returns (N,2) int32 device pairs (i, j)."""
        if d1.shape[0] == 0 or d2.shape[0] < 2:
            torch = _torch()
            return torch.zeros((0, 2), dtype=torch.int32, device=d1.device)
        if self.match_method == "float":
            return self._match_float(d1, d2)[2]
        _, _, pairs = match_u8(d1, d2, self.match_ratio, self.match_method)
        return pairs

    def match(self, d1, d2):
        """cv2.DMatch list (distance = sqrt of the squared distance: integer
        for the exact matcher, 2 - 2 s for the float path)."""
        if self.match_method == "float":
            fwd, _, pairs = self._match_float(d1, d2)
            p = pairs.cpu().numpy()
            f = fwd.cpu().numpy()[:, 1].view(np.float32)
            return [DMatch(int(i), int(j), float(np.sqrt(f[i]))) for i, j in p]
        fwd, _, pairs = match_u8(d1, d2, self.match_ratio, self.match_method)
        p = pairs.cpu().numpy()
        f = fwd.cpu().numpy()
        return [DMatch(int(i), int(j), float(np.sqrt(f[i, 1]))) for i, j in p]

    def _ransac(self, pts1, pts2):
        return ransac_homography(pts1, pts2, seed=self.ransac_seed, n_hyp=self.ransac_iters,
                                 thr=self.inlier_thresh, sigma_max=self.sigma_max,
                                 method=self.ransac_method, lo=True, min_inliers=10,
                                 device=pts1.device if hasattr(pts1, "device") else None, return_device=True)

    def _stitch_pair(self, img1, img2):
        """_stitch_pair (core.py:3720-3818): img2 is the reference frame."""
        torch = _torch()
        timings, metrics = {}, {}
        t0 = time.perf_counter()
        b1, b2 = self._upload(img1), self._upload(img2)
        kp1, d1, _ = self._extract_features(b1)
        kp2, d2, _ = self._extract_features(b2)
        pairs = self._match_features(d1, d2)
        torch.cuda.current_stream(b1.device).synchronize()
        timings["st1"] = time.perf_counter() - t0
        metrics.update(kp1=int(kp1.shape[0]), kp2=int(kp2.shape[0]), matches=int(pairs.shape[0]))
        if pairs.shape[0] < 4:
            raise RuntimeError(f"Insufficient matches ({pairs.shape[0]}), need >= 4")
        pts1 = kp1[pairs[:, 0].long(), :2].double().contiguous()
        pts2 = kp2[pairs[:, 1].long(), :2].double().contiguous()
        timings["st2"] = 0.0
        t0 = time.perf_counter()
        Hd, mask, ninl, best, _ = self._ransac(pts1, pts2)
        n_inl = int(ninl.item())
        timings["st3"] = time.perf_counter() - t0
        metrics.update(inliers=n_inl, inlier_ratio=n_inl / max(int(pairs.shape[0]), 1))
        if n_inl < 4:
            raise RuntimeError(f"Only {n_inl} inliers after RANSAC, need >= 4")
        H = Hd.cpu().numpy()
        self.last_homography = H
        timings["st4"] = 0.0
        metrics.update(depth_cv=None, bands=0)
        t0 = time.perf_counter()
        h1, w1 = b1.shape[:2]
        h2, w2 = b2.shape[:2]
        cw, ch, ox, oy = compute_canvas_size(h1, w1, h2, w2, H)
        seam_rmse = float("nan")
        if self.blend == "multiband":
            pano_d, valid_d, _, stats = self.composite_multiband(b1, b2, H, cw, ch, ox, oy)
            seam_rmse = stats["rmse"]
        else:
            pano_d, valid_d = self.warp_feather(b1, b2, H, cw, ch, ox, oy)
        # autocrop on the device; only the cropped panorama crosses to the host
        box = autocrop_box(valid_d)
        n_valid = int(valid_d.sum().item())
        pano = (pano_d if box is None else pano_d[box[0]:box[1], box[2]:box[3]]).cpu().numpy()
        timings["st5"] = time.perf_counter() - t0
        metrics.update(seam_rmse=seam_rmse, coverage=float(n_valid) / float(ch * cw) * 100.0,
                       canvas_w=cw, canvas_h=ch)
        return pano, timings, metrics

    def warp_feather(self, b1, b2, H, cw, ch, ox, oy):
        torch = _torch()
        pano = torch.empty((ch, cw, 3), dtype=torch.uint8, device=b1.device)
        valid = torch.empty((ch, cw), dtype=torch.uint8, device=b1.device)
        Hc = np.ascontiguousarray(H, np.float64)
        import ctypes
        rc = _lib.lib().msift_warp_feather(_lib.ptr(b1), b1.shape[0], b1.shape[1], _lib.ptr(b2), b2.shape[0],
                                           b2.shape[1], Hc.ctypes.data_as(ctypes.c_void_p), cw, ch, float(ox),
                                           float(oy), _lib.ptr(pano), _lib.ptr(valid), _lib.stream_ptr(b1.device))
        _lib.check(rc, "warp_feather")
        return pano, valid

    def composite_multiband(self, b1, b2, H, cw, ch, ox, oy, n_levels=None, seam_mask=False):
        """_blend_pair (self.color_match: 'gain' or 'hist') plus the warp and paste of
        _stitch_pair (core.py:3686-3714, 3783-3802) on the device: returns
        (pano (ch, cw, 3) u8, valid (ch, cw) u8, seam mask float32 or None,
        stats dict with gain, seam orientation and seam RMSE)."""
        torch = _torch()
        import ctypes
        L = n_levels if n_levels is not None else self.n_blend_levels
        dev = b1.device
        n = ctypes.c_size_t()
        _lib.check(_lib.lib().msift_composite_workspace_bytes(cw, ch, L, ctypes.byref(n)), "composite")
        ws = torch.empty(n.value, dtype=torch.uint8, device=dev)
        pano = torch.empty((ch, cw, 3), dtype=torch.uint8, device=dev)
        valid = torch.empty((ch, cw), dtype=torch.uint8, device=dev)
        sm = torch.empty((ch, cw), dtype=torch.float32, device=dev) if seam_mask else None
        st = torch.zeros(13, dtype=torch.int64, device=dev)   # msift_composite_stats, 104 bytes
        Hc = np.ascontiguousarray(H, np.float64)
        _lib.check(_lib.lib().msift_composite_multiband(
            _lib.ptr(b1), b1.shape[0], b1.shape[1], _lib.ptr(b2), b2.shape[0], b2.shape[1],
            Hc.ctypes.data_as(ctypes.c_void_p), cw, ch, float(ox), float(oy), L, 1 if self.color_match == "hist" else 0,
            _lib.ptr(ws), ws.numel(),
            _lib.ptr(pano), _lib.ptr(valid), _lib.ptr(sm), _lib.ptr(st), _lib.stream_ptr(dev)), "composite")
        raw = st.cpu().numpy().tobytes()
        gain = np.frombuffer(raw[72:84], np.float32).copy()
        vertical, first, has = np.frombuffer(raw[84:96], np.int32)
        rmse = float(np.frombuffer(raw[96:104], np.float64)[0])
        stats = {"gain": gain, "vertical": int(vertical) if has else -1, "img1_first": int(first),
                 "n_overlap": int(np.frombuffer(raw[48:56], np.uint64)[0]), "rmse": rmse}
        return pano, valid, sm, stats

    def stitch(self, *args):
        """Stitch two images (BGR uint8); returns the uint8 (H, W, 3) panorama."""
        images = list(args[0]) if len(args) == 1 and isinstance(args[0], (list, tuple)) else list(args)
        if len(images) == 3:
            raise NotImplementedError("3-image stitching is not implemented in pysift_amd")
        if len(images) != 2:
            raise ValueError(f"stitch() accepts 2 or 3 images, got {len(images)}")
        pano, timings, metrics = self._stitch_pair(images[0], images[1])
        self.last_timings, self.last_metrics = timings, metrics
        if self.verbose:
            self._print_report(timings, metrics)
        if self.save_output:
            self._save_output(pano)
        return pano

    def _print_report(self, timings, metrics):
        print("=" * 42)
        print("  GPUPyStitch Report (pysift_amd)")
        print("=" * 42)
        print(f"Stage 1 - Feature Extraction:     {timings['st1']:>6.3f}s  (img1: {metrics['kp1']} kp, "
              f"img2: {metrics['kp2']} kp, matches: {metrics['matches']})")
        print(f"Stage 3 - Robust Fitting (RANSAC): {timings['st3']:>6.3f}s  (inliers: {metrics['inliers']})")
        print(f"Stage 5 - Composition:            {timings['st5']:>6.3f}s  (coverage: {metrics['coverage']:.1f}%)")
        print("-" * 44)
        print(f"Total:                             {sum(timings.values()):>6.3f}s")

    def _save_output(self, pano):
        """Writes panorama_<timestamp>.<ext> to the working directory (core.py:4077-4092)."""
        from PIL import Image
        ts = datetime.now().strftime("%Y%m%d_%H%M%S")
        ext = "jpg" if self.output_format in ("jpg", "jpeg") else "png"
        fname = f"panorama_{ts}.{ext}"
        img = Image.fromarray(pano[:, :, ::-1].copy())
        if ext == "jpg":
            img.save(fname, quality=self.output_quality)
        else:
            img.save(fname)
        if self.verbose:
            print(f"Saved: {fname}  ({pano.shape[1]}x{pano.shape[0]} px, {os.path.getsize(fname) / 2**20:.1f} MB)")
