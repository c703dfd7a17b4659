"""Differentiable SIFT descriptors: a torch.autograd.Function over the HIP
forward and hand-written backward (libpysift_amd msift_desc_forward /
msift_desc_backward).

The forward builds the Gaussian pyramid of float images in [0, 1] (single-
size base, or the x2 spline base with double_image) and describes fixed,
detached keypoints; the backward recomputes the pyramid (no level is kept
between the two calls), differentiates the descriptor histogram and its
normalisation on one wavefront per keypoint, accumulates the per-pixel
scatters in fixed point (bitwise reproducible) and runs the blur and
decimation adjoints down to the pixels.  No reference counterpart: the
reference has no gradient path.
"""
import ctypes

from . import _lib
from .detector import PySIFT


def _ws(sift, B, H, W, device):
    import torch
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().msift_grad_workspace_bytes(sift._h, B, H, W, ctypes.byref(n)), "grad workspace")
    return torch.empty(max(n.value, 1), dtype=torch.uint8, device=device)


def _check(images, kps):
    import torch
    if not isinstance(images, torch.Tensor) or images.dim() != 3 or images.dtype != torch.float32 or not images.is_cuda:
        raise ValueError("images must be a (B, H, W) float32 device tensor")
    if not isinstance(kps, torch.Tensor) or kps.dim() != 3 or kps.shape[0] != images.shape[0] or kps.shape[2] != 5:
        raise ValueError("keypoints must be a (B, K, 5) tensor of (octave, si, row, col, angle)")


class SiftDescriptorFn:
    """apply(images (B, H, W) f32, keypoints (B, K, 5) f32, sift) -> (B, K, 128)."""

    _fn = None

    @classmethod
    def apply(cls, images, keypoints, sift):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(images, keypoints, sift)


def _make_function():
    import torch

    class _SiftDescriptor(torch.autograd.Function):
        @staticmethod
        def forward(ctx, images, keypoints, sift):
            _check(images, keypoints)
            imgs = images.detach().contiguous()
            kps = keypoints.detach().to(device=imgs.device, dtype=torch.float32).contiguous()
            B, H, W = imgs.shape
            K = kps.shape[1]
            ws = _ws(sift, B, H, W, imgs.device)
            desc = torch.empty((B, K, 128), dtype=torch.float32, device=imgs.device)
            rc = _lib.lib().msift_desc_forward(sift._h, _lib.ptr(imgs), B, H, W, _lib.ptr(kps), K, _lib.ptr(ws),
                                               ws.numel(), _lib.ptr(desc), _lib.stream_ptr(imgs.device))
            _lib.check(rc, "descriptor forward")
            ctx.save_for_backward(imgs, kps)
            ctx.sift = sift
            return desc

        @staticmethod
        def backward(ctx, gdesc):
            imgs, kps = ctx.saved_tensors
            sift = ctx.sift
            B, H, W = imgs.shape
            K = kps.shape[1]
            g = gdesc.detach().to(torch.float32).contiguous()
            ws = _ws(sift, B, H, W, imgs.device)
            gimg = torch.empty_like(imgs)
            rc = _lib.lib().msift_desc_backward(sift._h, _lib.ptr(imgs), B, H, W, _lib.ptr(kps), K, _lib.ptr(g),
                                                _lib.ptr(ws), ws.numel(), _lib.ptr(gimg), _lib.stream_ptr(imgs.device))
            _lib.check(rc, "descriptor backward")
            return gimg, None, None

    return _SiftDescriptor


_default = {}


def sift_descriptors(images, keypoints, sift=None):
    """Descriptors (B, K, 128) of float images (B, H, W) in [0, 1] at fixed
    keypoints (B, K, 5) = (octave, si, row, col, angle deg); differentiable
    with respect to `images`.  `sift` is a PySIFT in its plain mode (no DSP,
    float32 storage); with double_image=True (and H * W below its
    double_px_cap) octave 0 is the x2 spline upsample and keypoint rows and
    columns are in its pixels (default: PySIFT(double_image=False))."""
    if sift is None:
        if "s" not in _default:
            _default["s"] = PySIFT(double_image=False)
        sift = _default["s"]
    if sift.dsp or sift.fp16_pyramid:
        raise ValueError("the descriptor gradient needs the plain descriptor (dsp=False, fp16_pyramid=False)")
    return SiftDescriptorFn.apply(images, keypoints.detach(), sift)
