"""Batched detectAndCompute over images of uneven size.

One plan per list of image sizes: every pipeline stage runs as one launch over
all images (per-image tile tables in device memory), the whole call is
captured once in a HIP graph and replayed, and the blur tile height is chosen
by a timing autotuner (results do not depend on it).  Each image's output is
bit-identical to PySIFT.detectAndCompute on that image alone.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib

_TUNE_FILE = os.path.join(os.path.expanduser("~"), ".cache", "pysift_amd", "autotune.json")
BLUR_TILES = (16, 32, 64)


def _load_tune():
    try:
        with open(_TUNE_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_tune(d):
    try:
        os.makedirs(os.path.dirname(_TUNE_FILE), exist_ok=True)
        with open(_TUNE_FILE, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass


class BatchPlan:
    """Device plan for a fixed list of (H, W) image sizes."""

    def __init__(self, sift, sizes, device=None):
        import torch
        self.sift = sift
        self.sizes = [(int(h), int(w)) for h, w in sizes]
        self.n = len(self.sizes)
        dev = torch.device(device) if device is not None else sift._device_of()
        self.device = dev
        Hs = (ctypes.c_int * self.n)(*[h for h, _ in self.sizes])
        Ws = (ctypes.c_int * self.n)(*[w for _, w in self.sizes])
        L = _lib.lib()
        p = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(L.msift_batch_create(sift._h, self.n, Hs, Ws, ctypes.byref(p)), "batch plan")
        self._p = p
        ws = ctypes.c_size_t(); inb = ctypes.c_size_t()
        offs = (ctypes.c_size_t * self.n)()
        _lib.check(L.msift_batch_sizes(p, ctypes.byref(ws), ctypes.byref(inb), offs), "batch sizes")
        self.input_offsets = [int(o) for o in offs]
        self.cap = self.n * sift.capacity
        self.ws = torch.empty(ws.value, dtype=torch.uint8, device=dev)
        self.packed = torch.zeros(max(inb.value, 1), dtype=torch.uint8, device=dev)
        self.kp = torch.empty((self.cap, 6), dtype=torch.float32, device=dev)
        self.desc = torch.empty((self.cap, 128), dtype=torch.float32, device=dev)
        self.desc_u8 = torch.empty((self.cap, 128), dtype=torch.uint8, device=dev)
        self.counts = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.offsets = torch.zeros(self.n + 1, dtype=torch.int32, device=dev)
        self.graph = None
        self.blur_tile = 0

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and _lib._lib is not None:
            _lib._lib.msift_batch_destroy(p)
            self._p = None

    def set_blur_tile(self, th):
        _lib.check(_lib.lib().msift_batch_set_blur_tile(self._p, int(th)), "blur tile")
        self.blur_tile = int(th)
        self.graph = None

    def pack(self, images):
        """Copy the images (uint8 (H, W), NumPy or tensors) into the packed input."""
        import torch
        if len(images) != self.n:
            raise ValueError(f"plan holds {self.n} images, got {len(images)}")
        for img, off, (h, w) in zip(images, self.input_offsets, self.sizes):
            t = torch.from_numpy(np.ascontiguousarray(img)) if isinstance(img, np.ndarray) else img
            if tuple(t.shape) != (h, w) or t.dtype != torch.uint8:
                raise ValueError(f"image of shape {tuple(t.shape)} does not match plan size {(h, w)}")
            self.packed[off:off + h * w].copy_(t.reshape(-1), non_blocking=True)

    def launch(self):
        """Enqueue the batched pipeline on the current stream (capturable)."""
        rc = _lib.lib().msift_detect_describe_batched(
            self._p, _lib.ptr(self.packed), _lib.ptr(self.ws), self.ws.numel(), self.cap, _lib.ptr(self.kp),
            _lib.ptr(self.desc), _lib.ptr(self.desc_u8), _lib.ptr(self.counts), _lib.ptr(self.offsets),
            _lib.stream_ptr(self.device))
        _lib.check(rc, "detect_describe_batched")

    def capture(self):
        """Record the whole batched call as one HIP graph (after a warm-up run)."""
        import torch
        with torch.cuda.device(self.device):
            self.launch()
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.launch()
            torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def run(self, graph=True):
        if graph:
            if self.graph is None:
                self.capture()
            self.graph.replay()
        else:
            self.launch()

    def results(self, u8=False):
        """Per-image (kp6, desc) device tensors (synchronises)."""
        offs = self.offsets.cpu().numpy()
        cnt = self.counts.cpu().numpy()
        d = self.desc_u8 if u8 else self.desc
        out = []
        for i in range(self.n):
            a, c = int(offs[i]), int(min(cnt[i], max(self.cap - offs[i], 0)))
            out.append((self.kp[a:a + c], d[a:a + c]))
        return out

    def autotune(self, iters=3):
        """Time the batched call for each blur tile height, keep the fastest
        (cached per batch signature in ~/.cache/pysift_amd/autotune.json)."""
        import torch
        key = f"blur:{self.n}:{sum(h * w for h, w in self.sizes)}"
        cache = _load_tune()
        if key in cache:
            self.set_blur_tile(cache[key])
            return cache[key], {}
        times = {}
        with torch.cuda.device(self.device):
            for th in BLUR_TILES:
                self.set_blur_tile(th)
                self.launch()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(iters):
                    self.launch()
                e1.record()
                torch.cuda.synchronize()
                times[th] = e0.elapsed_time(e1) / iters
        best = min(times, key=times.get)
        self.set_blur_tile(best)
        cache[key] = best
        _save_tune(cache)
        return best, times


def detect_batch(sift, images, graph=True, u8=False):
    """Convenience: plan + pack + run + per-image results."""
    plan = BatchPlan(sift, [np.asarray(i).shape[:2] if isinstance(i, np.ndarray) else tuple(i.shape[:2])
                            for i in images])
    plan.pack(images)
    plan.run(graph=graph)
    return plan.results(u8=u8)
