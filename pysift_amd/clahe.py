"""Contrast-limited adaptive histogram equalisation on the device.

GPUPyStitch._extract_features runs cv2.createCLAHE(clipLimit=clahe_clip,
tileGridSize=(clahe_tile, clahe_tile)).apply(gray) ahead of detection
(core.py:2562-2565; cv2 is the reference's path whenever kornia is absent,
core.py:2551-2561).  createCLAHE / CLAHE.apply here take the same arguments
and keep OpenCV's 8-bit semantics, computed by msift_clahe (three HIP
kernels: tile histograms, clipped-CDF LUTs, bilinear LUT blend).
"""
import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class CLAHE:
    """cv2.CLAHE counterpart: clipLimit (<= 0: no clipping) and tileGridSize
    = (tiles_x, tiles_y), each 1..128."""

    def __init__(self, clipLimit=40.0, tileGridSize=(8, 8)):
        self.setClipLimit(clipLimit)
        self.setTilesGridSize(tileGridSize)
        self._ws = {}

    def getClipLimit(self):
        return self._clip

    def setClipLimit(self, clipLimit):
        clip = float(clipLimit)
        if clip != clip:
            raise ValueError("clipLimit must be a number")
        self._clip = max(clip, 0.0)

    def getTilesGridSize(self):
        return self._tiles

    def setTilesGridSize(self, tileGridSize):
        tx, ty = (int(v) for v in tileGridSize)
        if not (1 <= tx <= 128 and 1 <= ty <= 128):
            raise ValueError("tileGridSize must lie in [1, 128] on each axis")
        self._tiles = (tx, ty)
        self._ws = {}

    def collectGarbage(self):
        self._ws = {}

    def _workspace(self, device):
        torch = _torch()
        ws = self._ws.get(device)
        if ws is None:
            n = ctypes.c_size_t()
            _lib.check(_lib.lib().msift_clahe_workspace_bytes(self._tiles[0], self._tiles[1], ctypes.byref(n)),
                       "clahe")
            ws = torch.empty(n.value, dtype=torch.uint8, device=device)
            self._ws[device] = ws
        return ws

    def apply(self, src, dst=None):
        """(H, W) uint8 image (device tensor, DLPack object or numpy array)
        -> (H, W) uint8 device tensor.  src may have a row pitch (stride(0) >=
        W, stride(1) == 1); dst, when given, is a dense (H, W) uint8 tensor on
        the same device and may be src itself."""
        torch = _torch()
        if isinstance(src, np.ndarray):
            src = torch.from_numpy(np.ascontiguousarray(src)).cuda()
        elif not isinstance(src, torch.Tensor):
            src = torch.from_dlpack(src)
        if src.dtype != torch.uint8 or src.dim() != 2:
            raise ValueError("CLAHE.apply expects a (H, W) uint8 image")
        if not src.is_cuda:
            src = src.cuda()
        if src.stride(1) != 1 or src.stride(0) < src.shape[1]:
            src = src.contiguous()
        H, W = src.shape
        if dst is None:
            dst = torch.empty((H, W), dtype=torch.uint8, device=src.device)
        elif (dst.dtype != torch.uint8 or tuple(dst.shape) != (H, W) or not dst.is_contiguous()
              or dst.device != src.device):
            raise ValueError("dst must be a dense (H, W) uint8 tensor on the source's device")
        if H == 0 or W == 0:
            return dst
        ws = self._workspace(src.device)
        _lib.check(_lib.lib().msift_clahe(_lib.ptr(src), H, W, src.stride(0), self._clip, self._tiles[0],
                                          self._tiles[1], _lib.ptr(ws), ws.numel(), _lib.ptr(dst),
                                          _lib.stream_ptr(src.device)), "clahe")
        return dst


def createCLAHE(clipLimit=40.0, tileGridSize=(8, 8)):
    """cv2.createCLAHE (OpenCV's defaults: clipLimit 40, 8 x 8 tiles)."""
    return CLAHE(clipLimit, tileGridSize)
