"""The descriptor gradient at the differentiation workload's full size: 16
float32 512x512 images in [0, 1], 1024 fixed keypoints each, loss
sum(W * descriptors) with W ~ N(0, 1), checked on 256 random pixels against
FP64 central differences at eps 1e-3.

Checked pixels: the loss is piecewise analytic (orientation bins, the clip at
0.2, zero gradient magnitudes and RootSIFT's empty bins switch it between
pieces), and a central difference whose stencil straddles a switch bounds no
derivative.  A pixel is checked when the FP64 forward's discrete state (every
sample's orientation bin and zero/nonzero magnitude, every clip, every empty
bin) is the same at img - eps e_p and img + eps e_p for every keypoint the
pixel reaches; the other pixels are reported, not asserted.
"""
import concurrent.futures as cf
import multiprocessing as mp
import os

import numpy as np
import pytest

from oracle import desc_torch64 as T
from pysift_amd import synthetic as syn

pytestmark = pytest.mark.gpu

B, H, W, K, N_OCT, EPS = 16, 512, 512, 1024, 5, 1e-3


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("rootsift", [True, False], ids=["rootsift", "l2sift"])
def test_gradient_16x512_vs_fp64_central_differences(rootsift):
    import torch
    from pysift_amd import PySIFT
    from pysift_amd.autograd import sift_descriptors
    imgs = syn.synth_float_batch(B, H, W, seed=3000)
    kps = syn.random_keypoints(B, K, H, W, n_oct=N_OCT, seed=3100)
    Wt = np.random.default_rng(3200).standard_normal((B, K, 128)).astype(np.float32)
    sift = PySIFT(double_image=False, rootsift=rootsift)
    x = torch.from_numpy(imgs).cuda().requires_grad_(True)
    d = sift_descriptors(x, torch.from_numpy(kps).cuda(), sift)
    (d * torch.from_numpy(Wt).cuda()).sum().backward()
    torch.cuda.synchronize()
    g = x.grad.cpu().numpy()
    rng = np.random.default_rng(3300)
    pix = [(b, int(y), int(xx)) for b in range(B) for y, xx in zip(rng.integers(0, H, 16), rng.integers(0, W, 16))]
    jobs = [(imgs[b].astype(np.float64), kps[b], Wt[b], [(y, xx)], EPS, rootsift, N_OCT, 120) for b, y, xx in pix]
    workers = max(1, min(os.cpu_count() or 4, 16) - 1)
    print(f"[case] gradient 16x512^2 rootsift={rootsift}: central differences on {len(pix)} pixels, "
          f"{workers} workers", flush=True)
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(T.fd_job, jobs))
    fd = np.concatenate([r[0] for r in res])
    ck = np.concatenate([r[1] for r in res])
    gp = np.array([g[b, y, xx] for b, y, xx in pix])
    e_all, e_ck = _rel(gp, fd), _rel(gp[ck], fd[ck])
    print(f"[case] checked pixels {int(ck.sum())}/{len(ck)}; GPU vs FD: checked {e_ck:.3e}, all {e_all:.3e}",
          flush=True)
    # the FP64 autograd gradient on the checked pixels of the first 4 images:
    # the rule leaves the central differences within 1e-4 of the exact derivative
    sel = [i for i, (b, _, _) in enumerate(pix) if b < 4 and ck[i]]
    refs = {b: T.loss_and_grad(imgs[b].astype(np.float64), kps[b], Wt[b], n_oct=N_OCT, rootsift=rootsift)[1].numpy()
            for b in range(4)}
    rp = np.array([refs[pix[i][0]][pix[i][1], pix[i][2]] for i in sel])
    e_ref = _rel(rp, fd[sel])
    print(f"[case] FP64 autograd vs FD on {len(sel)} checked pixels: {e_ref:.3e}", flush=True)
    assert ck.sum() >= 40
    assert e_ref <= 1e-4
    assert e_ck <= 1e-3
