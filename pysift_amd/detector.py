"""PySIFT drop-in: detectAndCompute on MI355X through the C-ABI.

Mirrors pysift.core.PySIFT (reference pysift/core.py:1151-2172): constructor
arguments, defaults of the >=12 GB device branch (core.py:1191, 1207, 1270),
return types (a cv2.KeyPoint-compatible list plus an (N,128) float32 array,
or device tensors with gpu_output=True) and the ValueError for a bad
dsp_n_scales (core.py:1203-1204).
"""
import ctypes
import math

import numpy as np

from . import _lib


class KeyPoint:
    """cv2.KeyPoint duck type (pt, size, angle, response, octave, class_id)."""

    __slots__ = ("pt", "size", "angle", "response", "octave", "class_id")

    def __init__(self, x, y, size, angle=-1.0, response=0.0, octave=0, class_id=-1):
        self.pt = (float(x), float(y))
        self.size = float(size)
        self.angle = float(angle)
        self.response = float(response)
        self.octave = int(octave)
        self.class_id = int(class_id)

    def __repr__(self):
        return (f"KeyPoint(pt=({self.pt[0]:.3f}, {self.pt[1]:.3f}), size={self.size:.3f}, "
                f"angle={self.angle:.3f}, response={self.response:.5f}, octave={self.octave})")


class DMatch:
    """cv2.DMatch duck type (queryIdx, trainIdx, imgIdx, distance)."""

    __slots__ = ("queryIdx", "trainIdx", "imgIdx", "distance")

    def __init__(self, _queryIdx=-1, _trainIdx=-1, _distance=0.0, _imgIdx=0):
        self.queryIdx = int(_queryIdx)
        self.trainIdx = int(_trainIdx)
        self.imgIdx = int(_imgIdx)
        self.distance = float(_distance)

    def __repr__(self):
        return f"DMatch({self.queryIdx} -> {self.trainIdx}, d={self.distance:.4f})"


_DSP_PRESETS = {
    3: [1.0 / math.sqrt(2), 1.0, math.sqrt(2)],
    4: [0.5, 1.0 / math.sqrt(2), 1.0, math.sqrt(2)],
    5: [0.5, 1.0 / math.sqrt(2), 1.0, math.sqrt(2), 2.0],
}

PHASE_KEYS = ("gauss_pyr_ms", "dog_pyr_ms", "extrema_ms", "stack_ms", "orient_ms", "descriptor_ms",
              "cv_kpt_ms", "total_ms")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pysift_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class PySIFT:
    """GPU SIFT detector/descriptor (reference: core.py:1151-2172)."""

    def __init__(self, n_octaves=None, n_scales=3, sigma0=1.6, contrast_thresh=0.04, edge_thresh=10.0,
                 dsp=False, dsp_n_scales=3, fp16_pyramid=None, orientation="histogram", descriptor="sift",
                 double_image=True, rootsift=True, *, max_keypoints=8000, max_cands_per_octave=25000,
                 double_px_cap=16_000_000, device=None):
        if dsp_n_scales not in _DSP_PRESETS:
            raise ValueError(f"dsp_n_scales must be 3, 4, or 5 (got {dsp_n_scales})")
        if orientation != "histogram" or descriptor != "sift":
            raise NotImplementedError("learned orientation/descriptor backends are not implemented in pysift_amd")
        self.n_octaves = 5 if n_octaves is None else int(n_octaves)
        self.n_scales = int(n_scales)
        self.sigma0 = float(sigma0)
        self.contrast_thresh = float(contrast_thresh)
        self.edge_thresh = float(edge_thresh)
        self.dsp = bool(dsp)
        self.dsp_n_scales = dsp_n_scales
        self._dsp_scales_np = np.array(_DSP_PRESETS[dsp_n_scales], np.float32)
        # an MI355X is the reference's high-VRAM device (core.py:1207): float32
        # storage unless fp16_pyramid=True asks for the low-VRAM numerics
        self.fp16_pyramid = bool(fp16_pyramid) if fp16_pyramid is not None else False
        self.orientation = orientation
        self.descriptor = descriptor
        self.double_image = bool(double_image)
        self.rootsift = bool(rootsift)
        self.max_keypoints = int(max_keypoints)
        self.max_cands_per_octave = int(max_cands_per_octave)
        self.double_px_cap = int(double_px_cap)
        self._high_vram = True
        self._phase_timings = None
        self._n_calls = 0
        self._device = device
        self._params = _lib.MsiftParams(self.n_octaves, self.n_scales, self.sigma0, self.contrast_thresh,
                                        self.edge_thresh, int(self.double_image), int(self.rootsift),
                                        self.double_px_cap, self.max_cands_per_octave, self.max_keypoints,
                                        5, 100, 1, self.dsp_n_scales if self.dsp else 0, int(self.fp16_pyramid))
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.msift_create(ctypes.byref(self._params), ctypes.byref(h)), "PySIFT")
        self._h = h
        self._ws = None
        self._ws_shape = None
        self._out = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.msift_destroy(h)
            self._h = None

    # ------------------------------------------------------------------
    @property
    def capacity(self):
        """Largest keypoint count one call can return (<= 4 angles per keypoint)."""
        return 4 * self.max_keypoints

    def _device_of(self):
        torch = _torch()
        return torch.device(self._device) if self._device is not None else torch.device("cuda", torch.cuda.current_device())

    def _to_device_u8(self, gray):
        torch = _torch()
        dev = self._device_of()
        if isinstance(gray, np.ndarray):
            if gray.dtype != np.uint8 or gray.ndim != 2:
                raise ValueError("gray must be a 2-D uint8 array")
            arr = np.ascontiguousarray(gray)
            if not arr.flags.writeable:   # torch.from_numpy needs a writable buffer
                arr = arr.copy()
            return torch.from_numpy(arr).to(dev, non_blocking=False)
        if not isinstance(gray, torch.Tensor):
            gray = torch.from_dlpack(gray)
        if gray.dtype != torch.uint8 or gray.dim() != 2:
            raise ValueError("gray must be a 2-D uint8 tensor")
        return gray.to(dev).contiguous()

    def workspace_bytes(self, H, W):
        n = ctypes.c_size_t()
        _lib.check(_lib.lib().msift_workspace_bytes(self._h, H, W, ctypes.byref(n)), "workspace")
        return n.value

    def _ensure_buffers(self, H, W, dev):
        torch = _torch()
        if self._ws_shape != (H, W) or self._ws is None or self._ws.device != dev:
            self._ws = torch.empty(self.workspace_bytes(H, W), dtype=torch.uint8, device=dev)
            self._ws_shape = (H, W)
        if self._out is None or self._out[0].device != dev:
            cap = self.capacity
            self._out = (torch.empty((cap, 6), dtype=torch.float32, device=dev),
                         torch.empty((cap, 128), dtype=torch.float32, device=dev),
                         torch.empty((cap, 128), dtype=torch.uint8, device=dev),
                         torch.zeros(1, dtype=torch.int32, device=dev))

    def launch(self, img, out=None, profile=False):
        """Enqueue the device pipeline on the current stream without waiting.
        img: (H, W) uint8 CUDA tensor.  out: optional (kp6, f32, u8, count)
        buffers with room for `capacity` keypoints.  Returns those buffers;
        the count stays on the device."""
        torch = _torch()
        if img.dtype != torch.uint8 or img.dim() != 2 or not img.is_cuda:
            raise ValueError("img must be a 2-D uint8 CUDA tensor")
        if img.stride(1) != 1:
            raise ValueError("img rows must be contiguous (stride(1) == 1); pass img.contiguous()")
        H, W = img.shape
        dev = img.device
        self._ensure_buffers(H, W, dev)
        kp, df, du, cnt = out if out is not None else self._out
        phase = (ctypes.c_float * 8)() if profile else None
        rc = _lib.lib().msift_detect_describe(
            self._h, _lib.ptr(img), H, W, img.stride(0), _lib.ptr(self._ws), self._ws.numel(), self.capacity,
            _lib.ptr(kp), _lib.ptr(df), _lib.ptr(du), _lib.ptr(cnt),
            ctypes.cast(phase, ctypes.c_void_p) if profile else None, _lib.stream_ptr(dev))
        _lib.check(rc, "detectAndCompute")
        if profile:
            self._phase_timings = {k: float(phase[i]) for i, k in enumerate(PHASE_KEYS)}
        self._n_calls += 1
        return kp, df, du, cnt

    def new_outputs(self, device=None):
        torch = _torch()
        dev = device or self._device_of()
        cap = self.capacity
        return (torch.empty((cap, 6), dtype=torch.float32, device=dev),
                torch.empty((cap, 128), dtype=torch.float32, device=dev),
                torch.empty((cap, 128), dtype=torch.uint8, device=dev),
                torch.zeros(1, dtype=torch.int32, device=dev))

    def detect_raw(self, gray, profile=False):
        """Run the device pipeline; returns (kp6, desc_f32, desc_u8) views of
        length N on the device: kp6 columns x, y, size, angle, response, octave."""
        torch = _torch()
        img = self._to_device_u8(gray)
        with torch.cuda.device(img.device):
            kp, df, du, cnt = self.launch(img, profile=profile)
            n = min(int(cnt.item()), self.capacity)
        # fresh tensors: the launch buffers are reused by the next call
        return kp[:n].clone(), df[:n].clone(), du[:n].clone()

    def detectAndCompute(self, gray, mask=None, profile=False, gpu_output=False, u8=False):
        """Detect keypoints and compute descriptors (core.py:2021-2172).

        gray: (H, W) uint8 NumPy array, torch tensor or DLPack capsule owner.
        mask is accepted and ignored, as in the reference.
        Returns (list[KeyPoint], np.ndarray (N,128) float32), or with
        gpu_output=True device tensors ((N,4) [x,y,size,angle], (N,128)).
        u8=True returns the uint8 descriptors instead of float32.
        """
        kp, df, du = self.detect_raw(gray, profile=profile)
        desc = du if u8 else df
        if gpu_output:
            return kp[:, :4].clone(), desc.clone()
        kp_np = kp.cpu().numpy()
        d_np = desc.cpu().numpy()
        kps = [KeyPoint(r[0], r[1], r[2], r[3], r[4], int(r[5])) for r in kp_np]
        return kps, d_np
