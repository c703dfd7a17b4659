"""Benchmark: detectAndCompute ms/MP on the two-image stitch workload
(two 1024x768 views of one seeded synthetic canvas), plus the exact INT8
matcher and the full stitch when available.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = detectAndCompute on both views (gray, device resident), enqueued
back to back; K steps are timed between two synchronisations.  Prints ONE
JSON line.  The pyramid stage is timed separately with HIP events on the
launch stream to report its HBM roofline; the CPU oracle is timed on a
bounded sample of the same workload as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0
REF_MS_PER_MP = 72.8 / (768 * 1024 / 1e6)   # reference RTX 3050, 768x1024


# HBM bytes the pyramid kernels move per 1024x768 view (FETCH_SIZE + WRITE_SIZE
# counter runs, scaled by a 256 MiB copy kernel): profiles/pyramid_1024x768_traffic.md
PYRAMID_TRAFFIC_BYTES = 262_225_042


def octave_pixels(H, W, n_oct, doubled):
    h, w = (2 * H, 2 * W) if doubled else (H, W)
    tot = 0
    for _ in range(n_oct):
        tot += h * w
        h, w = (h + 1) // 2, (w + 1) // 2
    return tot


def pyramid_bytes(H, W, n_oct, S=3, doubled=True):
    """Algorithmic pyramid bytes: 1 B per input pixel + 4 B x the S+3
    Gaussian levels per octave pixel that the pyramid hands to its readers
    (the extrema form the DoG from all of them; orientation and descriptors
    read levels 1..S)."""
    return H * W + 4 * (S + 3) * octave_pixels(H, W, n_oct, doubled)


def cpu_baseline(views, seconds=12.0):
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    O.set_threads(threads)
    p = O.params()
    mp = 0.0
    t0 = time.perf_counter()
    n = 0
    while True:
        for v in views:
            O.detect_describe(v, p)
            mp += v.size / 1e6
            n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    # the same restatement on one thread (a shorter sample: whole views until ~6 s)
    O.set_threads(1)
    mp1, n1, t1 = 0.0, 0, time.perf_counter()
    while time.perf_counter() - t1 < 6.0:
        v = views[n1 % len(views)]
        O.detect_describe(v, p)
        mp1 += v.size / 1e6
        n1 += 1
    dt1 = time.perf_counter() - t1
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(dt * 1e3 / mp, 3), "unit": "ms/MP", "cores": threads, "kind": "port",
            "sample": f"oracle detect_describe on {n} x 1024x768 views ({dt:.1f} s, {threads} OpenMP threads)",
            "single_thread": {"value": round(dt1 * 1e3 / mp1, 3), "unit": "ms/MP",
                              "sample": f"{n1} views on 1 thread ({dt1:.1f} s)"},
            "cpu_model": model, "host_cpus": os.cpu_count()}


def _timed(fn, min_s=0.0):
    """(seconds per call, calls) of fn() repeated until min_s has passed (once at least)."""
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        if time.perf_counter() - t0 >= min_s:
            return (time.perf_counter() - t0) / n, n


def cpu_legs(threads):
    """The CPU oracle (the restated classic pipeline) on the other workloads,
    at `threads` OpenMP threads and at one thread, each on a bounded sample:
    * exact 40000 x 40000 uint8 matching: a slice of the query rows against
      all 40000 (the oracle's cost is linear in the rows), scaled to 40000;
    * 16384-hypothesis FP64 RANSAC on 5000 synthetic correspondences;
    * the 1024x768 BGR stitch (CLAHE, detect x2, exact match, 2000-hypothesis
      RANSAC, feather warp);
    * detectAndCompute over a fixed 8-image sample of the 64-image batch,
      scaled to images per second."""
    import numpy as np
    from oracle import oracle as O
    from pysift_amd import synthetic as syn
    out = {}
    rng = np.random.default_rng(9)
    n = 40000
    a = rng.integers(0, 256, (n, 128)).astype(np.uint8)
    b = a.copy()
    b[: n // 2] = np.clip(a[: n // 2].astype(int) + rng.integers(-4, 5, (n // 2, 128)), 0, 255).astype(np.uint8)
    b[n // 2:] = rng.integers(0, 256, (n - n // 2, 128)).astype(np.uint8)
    pts1 = rng.uniform(0, 1000, (5000, 2))
    Hs = np.array([[1.02, 0.05, 30.0], [-0.04, 0.98, 12.0], [1e-5, -2e-5, 1.0]])
    q = np.c_[pts1, np.ones(5000)] @ Hs.T
    pts2 = q[:, :2] / q[:, 2:] + rng.normal(0, 0.5, (5000, 2))
    pts2[:1000] = rng.uniform(0, 1000, (1000, 2))
    va, vb, _ = syn.synth_pair(seed=1000)
    imgs = syn.synth_batch(64, 2000)
    sample = imgs[::8]
    for tag, thr, rows in (("threads", threads, 4000), ("single_thread", 1, 500)):
        O.set_threads(thr)
        sec, _ = _timed(lambda: O.match_u8(a[:rows], b, 0.75))
        ms40k = sec * 1e3 * n / rows
        leg = {"threads": thr,
               "match_40000x40000": {"ms": round(ms40k, 1), "tops": round(2.0 * n * n * 128 / (ms40k * 1e-3) / 1e12, 4),
                                     "sample": f"{rows} x {n} rows, scaled to {n} x {n}"}}
        sec, _ = _timed(lambda: O.ransac(pts1, pts2, seed=42, n_hyp=16384, thr=3.0))
        leg["ransac_16384"] = {"ms": round(sec * 1e3, 1), "sample": "5000 correspondences, 20% outliers"}

        def stitch():
            p4 = O.params(n_octaves=4)
            kd = []
            for v in (va, vb):
                k, _, du = O.detect_describe(O.clahe(O.bgr_to_gray(v), 2.0, (8, 8)), p4)
                kd.append((k, du))
            _, _, pr = O.match_u8(kd[0][1], kd[1][1], 0.75)
            H, _ = O.ransac(kd[0][0][pr[:, 0], :2].astype(np.float64), kd[1][0][pr[:, 1], :2].astype(np.float64),
                            seed=42, n_hyp=2000, thr=3.0)
            cw, ch, ox, oy = O.canvas(768, 1024, 768, 1024, H)
            O.warp_feather(va, vb, H, cw, ch, ox, oy)
        sec, _ = _timed(stitch)
        leg["stitch_1024x768"] = {"ms": round(sec * 1e3, 1)}
        samp = sample if thr > 1 else sample[:2]
        t0 = time.perf_counter()
        for im in samp:
            O.detect_describe(im, O.params())
        dt = time.perf_counter() - t0
        leg["batch64_detect"] = {"img_s": round(len(samp) / dt, 2), "mpix_s": round(sum(i.size for i in samp) / 1e6 / dt, 3),
                                 "sample": f"{len(samp)} of the 64 images ({sum(i.size for i in samp) / 1e6:.2f} MP)"}
        out[tag] = leg
    O.set_threads(threads)
    return out


def bench_match(torch, dev, n, method, iters=5):
    """Exact INT8-MFMA matching of two n x 128 uint8 descriptor sets; returns
    (TOPS counted once as 2 n^2 128 / t, ms, mutual pairs).  The two-pass
    method's tile shape is autotuned first (untimed, cached per size class)."""
    from pysift_amd.matcher import autotune, match_u8, match_workspace_bytes
    g = torch.Generator(device="cpu").manual_seed(4)
    a = torch.randint(0, 256, (n, 128), dtype=torch.uint8, generator=g)
    b = a.clone()
    b[: n // 2] = torch.clamp(a[: n // 2].int() + torch.randint(-4, 5, (n // 2, 128), generator=g), 0, 255).to(torch.uint8)
    b[n // 2:] = torch.randint(0, 256, (n - n // 2, 128), dtype=torch.uint8, generator=g)
    a, b = a.to(dev), b.to(dev)
    if method == "twopass":
        autotune(n, n, dev)
    ws = torch.empty(match_workspace_bytes(n, n), dtype=torch.uint8, device=dev)
    out = match_u8(a, b, 0.75, method, sync=False, ws=ws)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        out = match_u8(a, b, 0.75, method, sync=False, ws=ws)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    return 2.0 * n * n * 128 / (ms * 1e-3) / 1e12, ms, int(out[3].item())


def bench_match_f16(torch, dev, n, iters=5, sample=256):
    """FP16-MFMA / FP32-accumulate matching of two n x 128 unit float
    descriptor sets; TOPS counted once (2 n^2 128 / t) and the squared-distance
    error of the returned neighbours against FP64 on `sample` rows."""
    import numpy as np
    from pysift_amd.matcher import match_f16, match_workspace_bytes
    g = torch.Generator(device="cpu").manual_seed(5)
    a = torch.rand((n, 128), generator=g) ** 3
    a = a / a.norm(dim=1, keepdim=True)
    b = a.clone()
    b[: n // 2] = a[: n // 2] + 0.01 * torch.randn((n // 2, 128), generator=g)
    b[n // 2:] = torch.rand((n - n // 2, 128), generator=g) ** 3
    b = b / b.norm(dim=1, keepdim=True)
    ta, tb = a.to(dev), b.to(dev)
    ws = torch.empty(match_workspace_bytes(n, n), dtype=torch.uint8, device=dev)
    fwd, _, pairs = match_f16(ta, tb, 0.75, ws=ws)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        fwd, _, pairs = match_f16(ta, tb, 0.75, ws=ws)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    rows = np.random.default_rng(6).choice(n, sample, replace=False)
    A = a[rows].double()
    D = (A * A).sum(1)[:, None] + (b.double() ** 2).sum(1)[None, :] - 2 * A @ b.double().T
    f = fwd.cpu().numpy()[rows]
    d1 = f[:, 1].view(np.float32).astype(np.float64)
    Dn = D.numpy()
    err = np.abs(d1 - Dn[np.arange(sample), f[:, 0]])
    same = float(np.mean(f[:, 0] == Dn.argmin(1)))
    return {"n": n, "tops": round(2.0 * n * n * 128 / (ms * 1e-3) / 1e12, 2), "ms": round(ms, 4),
            "pairs": int(pairs.shape[0]), "d2_abs_err_vs_fp64": {"max": float(err.max()), "mean": float(err.mean())},
            "nn_equal_fp64": same, "sample_rows": sample}


def bench_stitch(va, vb, iters=5):
    from pysift_amd import GPUPyStitch
    st = GPUPyStitch(inlier_thresh=3.0, ransac_iters=2000, verbose=False, save_output=False)
    st.stitch(va, vb)
    t0 = time.perf_counter()
    for _ in range(iters):
        st.stitch(va, vb)
    return (time.perf_counter() - t0) * 1e3 / iters, st.last_metrics


def bench_batch(torch, dev, n=64, iters=5):
    """64 uneven images (320x240 .. 1920x1080): batched graph-replayed pipeline
    vs the plain per-image path, inputs resident in HBM."""
    from pysift_amd import PySIFT
    from pysift_amd import synthetic as syn
    from pysift_amd.batch import BatchPlan
    imgs = syn.synth_batch(n, 2000)
    s = PySIFT(device=dev)
    plan = BatchPlan(s, [i.shape for i in imgs], device=dev)
    plan.pack(imgs)
    tile, tune = plan.autotune()
    plan.capture()
    plan.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        plan.run()
    e1.record()
    torch.cuda.synchronize()
    t_batch = e0.elapsed_time(e1) / iters
    dimgs = [torch.from_numpy(i).to(dev) for i in imgs]
    outs = s.new_outputs(dev)
    for img in dimgs:
        s.launch(img, outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        for img in dimgs:
            s.launch(img, outs)
    torch.cuda.synchronize()
    t_single = (time.perf_counter() - t0) * 1e3 / iters
    mp = sum(i.size for i in imgs) / 1e6
    # identical outputs: every image's keypoints and descriptors from the
    # batched call against the per-image path
    res = plan.results()
    identical = True
    for (kb, db), img in zip(res, dimgs):
        kp, df, _, cnt = s.launch(img, outs)
        k = int(cnt.item())
        identical = identical and kb.shape[0] == k and torch.equal(kb, kp[:k]) and torch.equal(db, df[:k])
    return {"images": n, "mpix": round(mp, 2), "identical": bool(identical),
            "batched_img_s": round(n / (t_batch * 1e-3), 1),
            "batched_ms": round(t_batch, 3), "per_image_path_img_s": round(n / (t_single * 1e-3), 1),
            "per_image_path_ms": round(t_single, 3), "speedup": round(t_single / t_batch, 2),
            "blur_tile": tile, "tune_ms": {str(k): round(v, 3) for k, v in tune.items()},
            "keypoints": int(plan.counts.sum().item())}


def bench_pairs(torch, dev, n_pairs=32, n_hyp=1024, iters=5):
    """32 homography-warped pairs at the uneven batch sizes: the 64 images are
    detected by one batched graph replay, then every pair is matched in one
    launch (exact INT8 MFMA, two-pass method: a - 128 as int8) and verified by batched RANSAC
    (1024 Philox hypotheses per pair, 3 px), against the per-pair loop
    (match_u8 + RANSAC per pair) on the same descriptors."""
    import numpy as np
    from pysift_amd import PySIFT
    from pysift_amd import synthetic as syn
    from pysift_amd.batch import BatchPlan
    from pysift_amd.matcher import MatchBatchPlan, match_u8
    from pysift_amd.ransac import RansacBatchPlan, ransac_homography
    pairs = syn.synth_pairs_batch(n_pairs, 2100)
    imgs = [im for a, b, _ in pairs for im in (a, b)]
    s = PySIFT(device=dev)
    plan = BatchPlan(s, [i.shape for i in imgs], device=dev)
    plan.pack(imgs)
    plan.autotune()
    plan.capture()
    plan.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        plan.run()
    e1.record()
    torch.cuda.synchronize()
    t_detect = e0.elapsed_time(e1) / iters
    res = plan.results(u8=True)
    A, B = res[0::2], res[1::2]
    mp = MatchBatchPlan([k.shape[0] for k, _ in A], [k.shape[0] for k, _ in B], device=dev)
    DA = mp.pack([d for _, d in A], mp.Ns)
    DB = mp.pack([d for _, d in B], mp.Ms)
    KA = torch.cat([k for k, _ in A]).contiguous()
    KB = torch.cat([k for k, _ in B]).contiguous()
    rp = RansacBatchPlan(mp.Ns, n_hyp, dev)

    def batched():
        mp.run(DA, DB, 0.75, "twopass")
        p1, p2 = mp.gather_points(KA, KB)
        return rp.run(p1, p2, mp.npairs, seed=42, thr=3.0)

    batched()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        H, mask, ninl, _ = batched()
    e1.record()
    torch.cuda.synchronize()
    t_batched = e0.elapsed_time(e1) / iters

    def per_pair():
        out = []
        for p in range(n_pairs):
            (ka, da), (kb, db) = A[p], B[p]
            _, _, pr = match_u8(da, db, 0.75, "twopass")
            if pr.shape[0] < 4:
                out.append(None)
                continue
            p1 = ka[pr[:, 0].long(), :2].double()
            p2 = kb[pr[:, 1].long(), :2].double()
            out.append(ransac_homography(p1, p2, seed=42 + p, n_hyp=n_hyp, thr=3.0, return_device=True)[0])
        return out

    per_pair()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        single = per_pair()
    torch.cuda.synchronize()
    t_single = (time.perf_counter() - t0) * 1e3 / iters
    Hn = H.cpu().numpy()
    same = all(sh is not None and np.array_equal(Hn[p], sh.cpu().numpy()) for p, sh in enumerate(single))
    errs = [syn.corner_error(Hn[p], pairs[p][2], pairs[p][0].shape[1], pairs[p][0].shape[0]) for p in range(n_pairs)]
    kp = [k.shape[0] for k, _ in res]
    return {"pairs": n_pairs, "hypotheses_per_pair": n_hyp, "detect_64_images_ms": round(t_detect, 3),
            "keypoints_per_image": {"min": min(kp), "median": int(np.median(kp)), "max": max(kp)},
            "matches": int(mp.npairs.sum().item()), "inliers": int(ninl.sum().item()),
            "match_ransac_batched_ms": round(t_batched, 3), "match_ransac_per_pair_ms": round(t_single, 3),
            "speedup": round(t_single / t_batched, 2), "batched_equals_per_pair": bool(same),
            "corner_error_px": {"median": round(float(np.median(errs)), 3), "max": round(float(np.max(errs)), 3),
                                "pairs_below_2px": int(np.sum(np.array(errs) < 2.0))}}


def bench_config4(torch, dev, iters=3, check=True):
    """4000x3000 grayscale pair (the second warped by a known homography):
    detect with lifted caps (40 000 keypoints, 200 000 candidates per octave),
    exact INT8 match (checked against the integer CPU oracle), the FP16 float
    path with its error against FP64, 16 384-hypothesis FP64 RANSAC and the
    feather-blended panorama."""
    import numpy as np
    from pysift_amd import GPUPyStitch, PySIFT
    from pysift_amd import synthetic as syn
    from pysift_amd.matcher import autotune, match_f16, match_u8, match_workspace_bytes
    from pysift_amd.ransac import ransac_homography
    from pysift_amd.stitch import compute_canvas_size
    a, b, Hgt = syn.synth_12mp_pair()
    H, W = a.shape
    s = PySIFT(max_keypoints=40000, max_cands_per_octave=200000, device=dev)
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    outs = s.new_outputs(dev)
    s.launch(ta, outs, profile=True)
    torch.cuda.synchronize()
    pyr = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        s.launch(ta, outs)
    e1.record()
    torch.cuda.synchronize()
    t_det = e0.elapsed_time(e1) / iters
    for _ in range(iters):
        s.launch(ta, outs, profile=True)
        pyr.append(s._phase_timings["gauss_pyr_ms"])
    t_pyr = float(np.median(pyr))
    kpa, dfa, dua = s.detect_raw(ta)
    kpb, dfb, dub = s.detect_raw(tb)
    N, M = kpa.shape[0], kpb.shape[0]
    ws = torch.empty(match_workspace_bytes(N, M), dtype=torch.uint8, device=dev)
    res = {"image": [H, W], "keypoints": [N, M], "detect_ms": round(t_det, 3), "ms_per_mp": round(t_det / (H * W / 1e6), 4),
           "pyramid_ms": round(t_pyr, 4)}
    Bp = pyramid_bytes(H, W, 5)
    res["pyramid_roofline"] = {"algorithmic_bytes": Bp, "achieved_gbs": round(Bp / (t_pyr * 1e-3) / 1e9, 1),
                               "frac": round(Bp / (t_pyr * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    autotune(N, M, dev)
    for method in ("hilo", "shift", "twopass"):
        match_u8(dua, dub, 0.75, method, sync=False, ws=ws)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            fwd, bwd, pairs, npairs = match_u8(dua, dub, 0.75, method, sync=False, ws=ws)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        res[f"match_{method}"] = {"ms": round(ms, 3), "tops": round(2.0 * N * M * 128 / (ms * 1e-3) / 1e12, 2),
                                  "pairs": int(npairs.item())}
    fwd, bwd, pairs = match_u8(dua, dub, 0.75, "twopass", ws=ws)
    same = True
    for method in ("hilo", "shift"):
        f2, b2, p2 = match_u8(dua, dub, 0.75, method, ws=ws)
        same = same and bool(torch.equal(f2, fwd) and torch.equal(b2, bwd) and torch.equal(p2, pairs))
    res["methods_identical"] = same
    if check:
        from oracle import oracle as O
        O.set_threads(min(os.cpu_count() or 8, 16))
        t0 = time.perf_counter()
        rf, rb, rp = O.match_u8(dua.cpu().numpy(), dub.cpu().numpy(), 0.75)
        res["match_exact_vs_cpu_oracle"] = bool(np.array_equal(fwd.cpu().numpy(), rf) and
                                                np.array_equal(bwd.cpu().numpy(), rb) and
                                                np.array_equal(pairs.cpu().numpy(), rp))
        res["cpu_oracle_match_s"] = round(time.perf_counter() - t0, 2)
    f16 = match_f16(dfa, dfb, 0.75, ws=ws)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        f16 = match_f16(dfa, dfb, 0.75, ws=ws)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    rows = np.random.default_rng(8).choice(N, min(256, N), replace=False)
    A64 = dfa[rows].double().cpu()
    B64 = dfb.double().cpu()
    D = ((A64 * A64).sum(1)[:, None] + (B64 * B64).sum(1)[None, :] - 2 * A64 @ B64.T).numpy()
    ff = f16[0].cpu().numpy()[rows]
    err = np.abs(ff[:, 1].view(np.float32) - D[np.arange(len(rows)), ff[:, 0]])
    res["match_f16"] = {"ms": round(ms, 3), "tops": round(2.0 * N * M * 128 / (ms * 1e-3) / 1e12, 2),
                        "pairs": int(f16[2].shape[0]), "d2_abs_err_vs_fp64_max": float(err.max()),
                        "nn_equal_fp64": float(np.mean(ff[:, 0] == D.argmin(1)))}
    p1 = kpa[pairs[:, 0].long(), :2].double().contiguous()
    p2 = kpb[pairs[:, 1].long(), :2].double().contiguous()
    ransac_homography(p1, p2, seed=42, n_hyp=16384, thr=3.0, return_device=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        Hd, mask, ninl, best, _ = ransac_homography(p1, p2, seed=42, n_hyp=16384, thr=3.0, return_device=True)
    e1.record()
    torch.cuda.synchronize()
    Hn = Hd.cpu().numpy()
    res["ransac_16384"] = {"ms": round(e0.elapsed_time(e1) / iters, 3), "matches": int(pairs.shape[0]),
                           "inliers": int(ninl.item()), "corner_error_px": round(syn.corner_error(Hn, Hgt, W, H), 3)}
    st = GPUPyStitch(save_output=False, verbose=False)
    b1 = torch.from_numpy(np.repeat(a[..., None], 3, 2)).to(dev)
    b2 = torch.from_numpy(np.repeat(b[..., None], 3, 2)).to(dev)
    cw, ch, ox, oy = compute_canvas_size(H, W, H, W, Hn)
    st.warp_feather(b1, b2, Hn, cw, ch, ox, oy)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        pano, valid = st.warp_feather(b1, b2, Hn, cw, ch, ox, oy)
    e1.record()
    torch.cuda.synchronize()
    res["panorama"] = {"ms": round(e0.elapsed_time(e1) / iters, 3), "canvas": [ch, cw]}
    return res


def bench_match_float(torch, dev, sizes=(8000, 40000), iters=5):
    """The reference's float matching path (PCA-64 + binary16 similarity +
    symmetric ratio test) on two n x 128 unit descriptor sets: the PCA fit
    (once per stitcher in the reference), the projection of both sets, and
    the similarity match; TOPS of the match counted once (2 n^2 64 / t)."""
    from pysift_amd.matcher import (l2_normalize, match_sim_f16, match_workspace_bytes, pca_fit,
                                    pca_project)
    res = {}
    for n in sizes:
        g = torch.Generator(device="cpu").manual_seed(5)
        a = torch.rand((n, 128), generator=g) ** 3
        b = a.clone()
        b[: n // 2] = a[: n // 2] + 0.01 * torch.randn((n // 2, 128), generator=g)
        b[n // 2:] = torch.rand((n - n // 2, 128), generator=g) ** 3
        ta, tb = a.to(dev), b.to(dev)
        ws = torch.empty(match_workspace_bytes(n, n), dtype=torch.uint8, device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        mean, proj, _, sweeps = pca_fit(ta, tb, 64)
        torch.cuda.synchronize()
        e0.record()
        mean, proj, _, sweeps = pca_fit(ta, tb, 64)
        e1.record()
        torch.cuda.synchronize()
        t_fit = e0.elapsed_time(e1)
        na, nb = l2_normalize(ta), l2_normalize(tb)
        e0.record()
        for _ in range(iters):
            pa, pb = pca_project(na, mean, proj), pca_project(nb, mean, proj)
        e1.record()
        torch.cuda.synchronize()
        t_proj = e0.elapsed_time(e1) / iters
        _, _, pairs = match_sim_f16(pa, pb, 0.75, ws=ws)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            _, _, pairs = match_sim_f16(pa, pb, 0.75, ws=ws)
        e1.record()
        torch.cuda.synchronize()
        t_match = e0.elapsed_time(e1) / iters
        res[str(n)] = {"pca_fit_ms": round(t_fit, 3), "jacobi_sweeps": int(sweeps.item()),
                       "project_both_ms": round(t_proj, 4), "match_ms": round(t_match, 4),
                       "match_tops": round(2.0 * n * n * 64 / (t_match * 1e-3) / 1e12, 2),
                       "fp16_mfma_peak_tops": 2500.0, "pairs": int(pairs.shape[0])}
    return res


def bench_composite(torch, dev, va, vb, Hgt, iters=5):
    """The reference's composition (gain + DP seam + 6-level multiband blend)
    on the stitch pair's canvas and on two 4000x3000 views overlapping by a
    third (a 6667x3000 canvas), beside the feather blend on the same canvas."""
    import numpy as np
    from pysift_amd import GPUPyStitch
    from pysift_amd import synthetic as syn
    from pysift_amd.stitch import compute_canvas_size
    st = GPUPyStitch(save_output=False, verbose=False, blend="multiband")
    g = syn.synth_gray(3000, 6667, 9)
    big = np.repeat(g[..., None], 3, 2)
    Hb = np.array([[1.0, 0.0, -2667.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    res = {}
    for name, a, b, H in (("stitch pair", va, vb, Hgt), ("2x 4000x3000", big[:, :4000].copy(), big[:, 2667:].copy(), Hb)):
        b1, b2 = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
        cw, ch, ox, oy = compute_canvas_size(a.shape[0], a.shape[1], b.shape[0], b.shape[1], H)
        st.composite_multiband(b1, b2, H, cw, ch, ox, oy)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            _, _, _, stats = st.composite_multiband(b1, b2, H, cw, ch, ox, oy)
        e1.record()
        torch.cuda.synchronize()
        t_mb = e0.elapsed_time(e1) / iters
        e0.record()
        for _ in range(iters):
            st.warp_feather(b1, b2, H, cw, ch, ox, oy)
        e1.record()
        torch.cuda.synchronize()
        res[name] = {"canvas": [ch, cw], "multiband_ms": round(t_mb, 3),
                     "feather_ms": round(e0.elapsed_time(e1) / iters, 3),
                     "seam": "vertical" if stats["vertical"] == 1 else "horizontal", "seam_rmse": round(stats["rmse"], 3)}
    return res


def bench_clahe(torch, dev, gray, iters=50):
    """CLAHE(2.0, 8 x 8) as the stitcher runs it ahead of detection, on one
    1024x768 view and on a 4000x3000 image, timed with HIP events on the
    launch stream.  Algorithmic HBM bytes: 3 per pixel (the histogram pass and
    the apply pass each read the image once, the apply pass writes it)."""
    from pysift_amd import synthetic as syn
    from pysift_amd.clahe import CLAHE
    c = CLAHE(2.0, (8, 8))
    res = {}
    for name, arr in (("1024x768", gray), ("4000x3000", syn.synth_gray(3000, 4000, 5))):
        img = torch.from_numpy(arr).to(dev)
        out = torch.empty_like(img)
        c.apply(img, dst=out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            c.apply(img, dst=out)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        res[name] = {"ms": round(ms, 4), "achieved_gbs": round(3 * img.numel() / (ms * 1e-3) / 1e9, 1)}
    return res


def bench_grad(torch, dev, B=16, H=512, W=512, K=1024, iters=5):
    """16 float 512x512 images, 1024 fixed keypoints each, loss sum(W * desc):
    HIP forward and hand-written backward (pyramid recomputed), and one
    torch.autograd step through them."""
    import numpy as np
    from pysift_amd import PySIFT
    from pysift_amd import synthetic as syn
    from pysift_amd.autograd import sift_descriptors
    s = PySIFT(double_image=False, device=dev)
    x = torch.from_numpy(syn.synth_float_batch(B, H, W, seed=3000)).to(dev)
    kps = torch.from_numpy(syn.random_keypoints(B, K, H, W, n_oct=5, seed=3100)).to(dev)
    w = torch.from_numpy(np.random.default_rng(3200).standard_normal((B, K, 128)).astype(np.float32)).to(dev)

    def step():
        xx = x.clone().requires_grad_(True)
        (sift_descriptors(xx, kps, s) * w).sum().backward()
        return xx.grad

    step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        d = sift_descriptors(x, kps, s)
    e1.record()
    torch.cuda.synchronize()
    t_fwd = e0.elapsed_time(e1) / iters
    e0.record()
    for _ in range(iters):
        step()
    e1.record()
    torch.cuda.synchronize()
    t_step = e0.elapsed_time(e1) / iters
    return {"images": B, "hw": [H, W], "keypoints_per_image": K, "forward_ms": round(t_fwd, 3),
            "forward_backward_ms": round(t_step, 3), "backward_ms": round(t_step - t_fwd, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--pyramid-only", action="store_true", help="(profiling) run only detect on view a")
    ap.add_argument("--calib-copy", action="store_true",
                    help="(counter calibration) one 256 MiB device-to-device copy, then exit")
    ap.add_argument("--match-n", type=int, default=40000, help="descriptors per side for the match TOPS leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the match / stitch legs")
    ap.add_argument("--only", choices=["match", "batch", "stitch", "grad", "pairs", "config4", "clahe", "match_float", "composite"], default=None,
                    help="(profiling) run a single extra leg and print its numbers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: pysift_amd has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pysift_amd import PySIFT
    from pysift_amd import synthetic as syn

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": src.numel() * 4}))
        return
    va, vb, Hgt = syn.synth_pair(seed=1000)
    if args.only:
        if args.only == "match":
            out = {m: bench_match(torch, dev, args.match_n, m) for m in ("hilo", "shift", "twopass")}
            out["f16"] = bench_match_f16(torch, dev, args.match_n)
            print(json.dumps(out))
        elif args.only == "batch":
            print(json.dumps(bench_batch(torch, dev)))
        elif args.only == "grad":
            print(json.dumps(bench_grad(torch, dev)))
        elif args.only == "pairs":
            print(json.dumps(bench_pairs(torch, dev)))
        elif args.only == "config4":
            print(json.dumps(bench_config4(torch, dev)))
        elif args.only == "composite":
            print(json.dumps(bench_composite(torch, dev, va, vb, Hgt)))
        elif args.only == "match_float":
            print(json.dumps(bench_match_float(torch, dev)))
        elif args.only == "clahe":
            g = ((va[..., 0].astype(np.int32) * 1868 + va[..., 1].astype(np.int32) * 9617
                  + va[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8)
            print(json.dumps(bench_clahe(torch, dev, g)))
        else:
            print(json.dumps(bench_stitch(va, vb)))
        return
    # BGR -> gray, fixed-point BT.601 as cv2.cvtColor (the stitcher's first kernel)
    gray = [((v[..., 0].astype(np.int32) * 1868 + v[..., 1].astype(np.int32) * 9617
              + v[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8) for v in (va, vb)]
    imgs = [torch.from_numpy(g).to(dev) for g in gray]
    H, W = gray[0].shape
    s = PySIFT(device=dev)
    outs = [s.new_outputs(dev) for _ in imgs]
    views = imgs[:1] if args.pyramid_only else imgs

    def step():
        for img, out in zip(views, outs):
            s.launch(img, out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    mp = sum(v.numel() for v in views) / 1e6
    ms_per_mp = dt * 1e3 / mp
    counts = [int(o[3].item()) for o in outs[:len(views)]]

    # pyramid stage timed with HIP events on the launch stream (median of 20)
    pyr = []
    for _ in range(20):
        s.launch(imgs[0], outs[0], profile=True)
        pyr.append(s._phase_timings["gauss_pyr_ms"])
    t_pyr = float(np.median(pyr)) * 1e-3
    B = pyramid_bytes(H, W, s.n_octaves)
    achieved = B / t_pyr / 1e9

    line = {
        "metric": "ms/MP detectAndCompute",
        "value": round(ms_per_mp, 4),
        "unit": "ms/MP",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(ms_per_mp / REF_MS_PER_MP, 6),
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "stitch pair views 2x1024x768 u8 gray: detectAndCompute each "
                               "(5 octaves, S=3, doubled, 8000 kp cap)", "images_per_step": len(views),
                   "keypoints": counts},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": PYRAMID_TRAFFIC_BYTES,
                     "kernel": "pyramid (k_ups_* + 26 x k_blur) per 1024x768 view",
                     "algorithmic_bytes": B,
                     "bytes_model": "1 B per input pixel + 4 B x (S+3) Gaussian levels per octave pixel",
                     "avg_ms": round(t_pyr * 1e3, 4)},
        "phase_ms": {k: round(v, 4) for k, v in s._phase_timings.items()},
    }
    if not args.no_extra and not args.pyramid_only:
        extra = {}
        for method in ("hilo", "shift", "twopass"):
            tops, ms, npairs = bench_match(torch, dev, args.match_n, method)
            extra[f"match_{method}"] = {"n": args.match_n, "tops": round(tops, 2), "ms": round(ms, 4),
                                        "mfma_int8_peak_tops": PEAK_INT8_TOPS, "pairs": npairs}
        extra["match_f16"] = bench_match_f16(torch, dev, args.match_n)
        extra["batch64"] = bench_batch(torch, dev)
        extra["pairs32"] = bench_pairs(torch, dev)
        extra["grad16x512"] = bench_grad(torch, dev)
        extra["clahe"] = bench_clahe(torch, dev, gray[0])
        extra["match_float_pca64"] = bench_match_float(torch, dev)
        extra["composite_multiband"] = bench_composite(torch, dev, va, vb, Hgt)
        extra["config4_12mp_pair"] = bench_config4(torch, dev)
        st_ms, met = bench_stitch(va, vb)
        extra["stitch_pair_ms"] = round(st_ms, 3)
        extra["stitch_metrics"] = {k: met[k] for k in ("kp1", "kp2", "matches", "inliers")}
        line["extra"] = extra
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(gray)
        if not args.no_extra:
            line["cpu_baseline"]["other_workloads"] = cpu_legs(line["cpu_baseline"]["cores"])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
