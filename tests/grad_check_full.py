"""Full-size accuracy check of the descriptor gradient (run by hand on a GPU
host; pytest does not collect it): 16 float32 512x512 images, 1024 fixed
keypoints each, loss sum(W * descriptors), W ~ N(0, 1).

Prints one JSON line with
* the GPU image gradient against the FP64 torch-autograd gradient (all pixels
  of every image, norm-wise relative error);
* the GPU gradient and the FP64 autograd gradient against FP64 central
  differences at 256 random pixels, eps 1e-3 and eps 1e-5;
* the stage checks: the descriptor stage's level gradient against FP64
  autograd on the GPU's own levels, and the blur / decimation adjoint against
  the FP64 vector-Jacobian product of the pyramid (images 0 and 1).

    python tests/grad_check_full.py [--workers 12]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

B, H, W, K, N_OCT = 16, 512, 512, 1024, 5


def _inputs():
    from pysift_amd import synthetic as syn
    imgs = syn.synth_float_batch(B, H, W, seed=3000)
    kps = syn.random_keypoints(B, K, H, W, n_oct=N_OCT, seed=3100)
    Wt = np.random.default_rng(3200).standard_normal((B, K, 128)).astype(np.float32)
    return imgs, kps, Wt


def _fd_job(job):
    import torch
    torch.set_num_threads(1)
    from oracle import desc_torch64 as T
    b, pix, eps, rootsift, Wb = job
    imgs, kps, _ = _inputs()
    return T.central_difference(imgs[b].astype(np.float64), kps[b], Wb, pix, eps=eps, n_oct=N_OCT, radius=120,
                                rootsift=rootsift)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _gpu_backward(torch, L, _lib, sift, ti, tk, tw, ws):
    gimg = torch.empty_like(ti)
    _lib.check(L.msift_desc_backward(sift._h, _lib.ptr(ti), B, H, W, _lib.ptr(tk), K, _lib.ptr(tw), _lib.ptr(ws),
                                     ws.numel(), _lib.ptr(gimg), _lib.stream_ptr(ti.device)))
    torch.cuda.synchronize()
    return gimg.cpu().numpy()


def _stage_checks(torch, T, L, _lib, sift, ws, imgs, kps, Wt, g, rootsift):
    ph = T.patch_half_sizes()
    stage_desc, stage_adj = [], []
    for b in (0, 1):
        x = torch.tensor(imgs[b].astype(np.float64), requires_grad=True)
        octs = T.pyramid(x, N_OCT)
        acc = torch.zeros((), dtype=torch.float64)
        num = den = 0.0
        for o in range(N_OCT):
            Ho, Wo = ctypes.c_int(), ctypes.c_int()
            pg, pd = ctypes.c_void_p(), ctypes.c_void_p()
            _lib.check(L.msift_grad_level_view(sift._h, B, H, W, _lib.ptr(ws), b, o, ctypes.byref(Ho),
                                               ctypes.byref(Wo), ctypes.byref(pg), ctypes.byref(pd)))
            m = Ho.value * Wo.value
            base = ws.data_ptr()
            g32 = ws[pg.value - base:pg.value - base + 16 * m].view(torch.float32).view(4, Ho.value, Wo.value).cpu()
            dg = ws[pd.value - base:pd.value - base + 24 * m].view(torch.int64).view(3, Ho.value, Wo.value).cpu()
            dg = dg.numpy().astype(np.float64) / 2.0 ** 32
            acc = acc + (octs[o][1:] * torch.from_numpy(dg)).sum()
            Gt = g32.double().requires_grad_(True)
            loss = torch.zeros((), dtype=torch.float64)
            for si in (1, 2, 3):
                sel = (kps[b, :, 0].astype(int) == o) & (kps[b, :, 1].astype(int) == si)
                if sel.any():
                    d = T.describe_group(Gt[si], ph[si], kps[b, sel, 2], kps[b, sel, 3], kps[b, sel, 4], rootsift)
                    loss = loss + (torch.from_numpy(Wt[b, sel].astype(np.float64)) * d).sum()
            if loss.requires_grad:
                (ref,) = torch.autograd.grad(loss, Gt)
                num += float(np.sum((dg - ref[1:].numpy()) ** 2))
                den += float(np.sum(ref[1:].numpy() ** 2))
        (ref,) = torch.autograd.grad(acc, x)
        stage_adj.append(_rel(g[b], ref.numpy()))
        stage_desc.append(float(np.sqrt(num / den)))
    return stage_desc, stage_adj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=12)
    args = ap.parse_args()
    import concurrent.futures as cf
    import multiprocessing as mp
    import torch
    from oracle import desc_torch64 as T
    from pysift_amd import PySIFT, _lib
    imgs, kps, Wt0 = _inputs()
    dev = torch.device("cuda", 0)
    L = _lib.lib()
    ti, tk = (torch.from_numpy(a).to(dev) for a in (imgs, kps))
    rng = np.random.default_rng(3300)
    pix = {b: [(int(y), int(x)) for y, x in zip(rng.integers(0, H, 16), rng.integers(0, W, 16))] for b in range(B)}
    result = {"workload": f"{B}x{H}x{W} float32, {K} fixed keypoints/image, loss sum(W*desc), W~N(0,1)",
              "fd_pixels": 256}
    ex = cf.ProcessPoolExecutor(max_workers=args.workers, mp_context=mp.get_context("spawn"))
    for case in ("rootsift", "rootsift_regular_kps", "l2_sift"):
        rootsift = case != "l2_sift"
        sift = PySIFT(double_image=False, rootsift=rootsift, device=dev)
        Wt = Wt0.copy()
        if case == "rootsift_regular_kps":
            # keypoints whose FP64 RootSIFT descriptor has a nearly empty bin
            # (0 < q < 1e-6, where d sqrt(q)/dq > 500) get weight 0
            n_sing = 0
            for b in range(B):
                d = T.descriptors(torch.from_numpy(imgs[b].astype(np.float64)), kps[b], N_OCT).numpy()
                sing = ((d > 0) & (d * d < 1e-6)).any(1)
                Wt[b, sing] = 0.0
                n_sing += int(sing.sum())
            result["singular_keypoints"] = n_sing
        tw = torch.from_numpy(Wt).to(dev)
        n = ctypes.c_size_t()
        _lib.check(L.msift_grad_workspace_bytes(sift._h, B, H, W, ctypes.byref(n)))
        ws = torch.empty(n.value, dtype=torch.uint8, device=dev)
        g = _gpu_backward(torch, L, _lib, sift, ti, tk, tw, ws)
        sd, sa = _stage_checks(torch, T, L, _lib, sift, ws, imgs, kps, Wt, g, rootsift)
        refs = np.stack([T.loss_and_grad(imgs[b].astype(np.float64), kps[b], Wt[b], n_oct=N_OCT,
                                         rootsift=rootsift)[1].numpy() for b in range(B)])
        e2e = [_rel(g[b], refs[b]) for b in range(B)]
        fdres = {}
        for eps in (1e-3, 1e-5):
            jobs = [(b, [p], eps, rootsift, Wt[b]) for b in range(B) for p in pix[b]]
            fd = np.concatenate(list(ex.map(_fd_job, jobs)))
            gp = np.array([g[b, y, x] for b in range(B) for (y, x) in pix[b]])
            rp = np.array([refs[b, y, x] for b in range(B) for (y, x) in pix[b]])
            fdres[f"{eps:g}"] = {"gpu_vs_fd": _rel(gp, fd), "fp64_autograd_vs_fd": _rel(rp, fd)}
        result[case] = {"gpu_vs_fp64_autograd_rel": {"max": max(e2e), "median": float(np.median(e2e)),
                                                     "all_pixels": _rel(g, refs)},
                        "central_differences": fdres,
                        "stage_desc_vs_fp64_autograd_rel": sd, "stage_adjoint_vs_fp64_vjp_rel": sa}
        print(f"{case} done", file=sys.stderr, flush=True)
    ex.shutdown()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
