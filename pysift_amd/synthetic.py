"""Seeded synthetic inputs for the benchmark workloads (no datasets offline).

Content recipe (same for every workload): background 128, random Gaussian
blobs (sigma 2-20 px, amplitude 30-120, random sign), random axis-aligned and
rotated rectangles of random intensity, one 8-px checker patch, N(0,3) noise,
clipped to 0..255 and rounded.  Shape counts scale with image area so that
large images keep the 640x480 feature density.

Workloads that name a keypoint density (2k-8k keypoints per image for the
32 matched pairs, 20k-40k per 12 MP image) add fine texture: small Gaussian
blobs (pairs_fine, FINE_12MP), drawn from their own random stream after the
rest of the content, so images without it are unchanged.  The 64-image
detection batch keeps the plain recipe.
"""
import math

import numpy as np

_BASE_AREA = 640 * 480

# fine texture (blobs per MP, sigma range px, amplitude range) of the workloads
# that name a keypoint density
FINE_12MP = (2000, (1.0, 3.0), (15.0, 60.0))


def pairs_fine(H, W):
    """Fine texture of a pair image of H x W: 9100 - 1250 * MP blobs (sigma
    0.8-2 px, amplitude 15-60), which puts the returned keypoints of the
    320x240 .. 1920x1080 sizes in 2k-8k."""
    mp = H * W / 1e6
    return ((9100.0 - 1250.0 * mp) / mp, (0.8, 2.0), (15.0, 60.0))


def _fine_texture(img, seed, per_mp, sigma, amp, n_sigma=6):
    """Adds per_mp * area/1e6 small Gaussian blobs (peak amplitude ~U(amp),
    random sign, sigma ~U(sigma) quantised to n_sigma values) to img in place:
    each blob is an impulse of mass amplitude * 2 pi sigma^2 split bilinearly
    over its four nearest pixels, and each sigma group is one Gaussian filter
    of its impulse image."""
    from scipy.ndimage import gaussian_filter
    H, W = img.shape
    n = int(round(per_mp * H * W / 1e6))
    if n <= 0:
        return img
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), 0xF17E]))
    cy = rng.uniform(0, H - 1, n)
    cx = rng.uniform(0, W - 1, n)
    levels = np.linspace(sigma[0], sigma[1], n_sigma)
    g = rng.integers(0, n_sigma, n)
    am = rng.uniform(amp[0], amp[1], n) * rng.choice([-1.0, 1.0], n)
    y0 = np.minimum(np.floor(cy).astype(np.int64), H - 2)
    x0 = np.minimum(np.floor(cx).astype(np.int64), W - 2)
    fy, fx = cy - y0, cx - x0
    for k, sg in enumerate(levels):
        sel = g == k
        if not sel.any():
            continue
        m = am[sel] * (2.0 * math.pi * sg * sg)
        imp = np.zeros((H, W), np.float64)
        for dy, wy in ((0, 1.0 - fy[sel]), (1, fy[sel])):
            for dx, wx in ((0, 1.0 - fx[sel]), (1, fx[sel])):
                np.add.at(imp, (y0[sel] + dy, x0[sel] + dx), m * wy * wx)
        img += gaussian_filter(imp, sg, mode="constant", truncate=4.0)
    return img


def synth_gray(H, W, seed, n_blobs=None, n_rects=None, checker=128, as_float=False, fine=None):
    """uint8 (H, W) image (or float64 before quantisation when as_float).
    fine: optional (blobs per MP, sigma range, amplitude range) of fine
    texture (FINE_BATCH, FINE_12MP)."""
    rng = np.random.default_rng(seed)
    scale = (H * W) / _BASE_AREA
    nb = int(round(200 * scale)) if n_blobs is None else n_blobs
    nr = int(round(50 * scale)) if n_rects is None else n_rects
    img = np.full((H, W), 128.0, np.float64)
    # Gaussian blobs, evaluated on a +-4 sigma window
    cy = rng.uniform(0, H, nb); cx = rng.uniform(0, W, nb)
    sg = rng.uniform(2, 20, nb); am = rng.uniform(30, 120, nb) * rng.choice([-1.0, 1.0], nb)
    for i in range(nb):
        r = int(math.ceil(4 * sg[i]))
        y0, y1 = max(0, int(cy[i]) - r), min(H, int(cy[i]) + r + 1)
        x0, x1 = max(0, int(cx[i]) - r), min(W, int(cx[i]) + r + 1)
        if y0 >= y1 or x0 >= x1:
            continue
        yy = np.arange(y0, y1)[:, None] - cy[i]
        xx = np.arange(x0, x1)[None, :] - cx[i]
        img[y0:y1, x0:x1] += am[i] * np.exp(-(yy * yy + xx * xx) / (2 * sg[i] ** 2))
    # rectangles (painted over)
    ry = rng.uniform(0, H, nr); rx = rng.uniform(0, W, nr)
    ra = rng.uniform(8, 160, nr); rb = rng.uniform(8, 160, nr)
    rot = rng.random(nr) < 0.5
    th = rng.uniform(0, math.pi, nr); val = rng.uniform(0, 255, nr)
    for i in range(nr):
        t = th[i] if rot[i] else 0.0
        c, s = math.cos(t), math.sin(t)
        ext = 0.5 * (abs(c) * ra[i] + abs(s) * rb[i]) + 1, 0.5 * (abs(s) * ra[i] + abs(c) * rb[i]) + 1
        y0, y1 = max(0, int(ry[i] - ext[1])), min(H, int(ry[i] + ext[1]) + 1)
        x0, x1 = max(0, int(rx[i] - ext[0])), min(W, int(rx[i] + ext[0]) + 1)
        if y0 >= y1 or x0 >= x1:
            continue
        yy = np.arange(y0, y1)[:, None] - ry[i]
        xx = np.arange(x0, x1)[None, :] - rx[i]
        u = c * xx + s * yy
        v = -s * xx + c * yy
        m = (np.abs(u) <= ra[i] / 2) & (np.abs(v) <= rb[i] / 2)
        img[y0:y1, x0:x1][m] = val[i]
    # one 8-px checker patch
    if checker and H > 16 and W > 16:
        cs = min(checker, H // 2, W // 2)
        py = int(rng.integers(0, H - cs + 1)); px = int(rng.integers(0, W - cs + 1))
        yy, xx = np.mgrid[0:cs, 0:cs]
        img[py:py + cs, px:px + cs] = np.where(((yy // 8) + (xx // 8)) % 2 == 0, 40.0, 215.0)
    if fine is not None and H > 1 and W > 1:
        _fine_texture(img, seed, *fine)
    img += rng.normal(0.0, 3.0, (H, W))
    img = np.clip(img, 0, 255)
    if as_float:
        return img
    return np.round(img).astype(np.uint8)


def _bilinear(img, sx, sy):
    """Bilinear sample of a (H, W, C) float image at (sx, sy); zeros outside."""
    H, W = img.shape[:2]
    x0 = np.floor(sx).astype(np.int64); y0 = np.floor(sy).astype(np.int64)
    ax = (sx - x0)[..., None]; ay = (sy - y0)[..., None]
    out = np.zeros(sx.shape + img.shape[2:], np.float64)
    for dy, wy in ((0, 1 - ay), (1, ay)):
        for dx, wx in ((0, 1 - ax), (1, ax)):
            xi = x0 + dx; yi = y0 + dy
            ok = (xi >= 0) & (xi < W) & (yi >= 0) & (yi < H)
            v = np.zeros(sx.shape + img.shape[2:], np.float64)
            v[ok] = img[yi[ok], xi[ok]]
            out += wx * wy * v
    return out


def synth_pair(seed=1000, view_hw=(768, 1024), canvas_hw=(1536, 2048), overlap=(0.30, 0.40)):
    """Two BGR uint8 views of one tinted canvas; the second is warped by a
    random homography (rotation +-15 deg, scale 0.8-1.25, perspective
    ~U(-2e-4, 2e-4)).  Returns (view_a, view_b, H_gt) with H_gt mapping
    view-a pixel coordinates to view-b pixel coordinates."""
    rng = np.random.default_rng(seed)
    Hc, Wc = canvas_hw
    h, w = view_hw
    gray = synth_gray(Hc, Wc, seed + 1, as_float=True)
    tint = rng.uniform(0.8, 1.2, 3)
    canvas = np.clip(gray[..., None] * tint[None, None, :], 0, 255)
    y0, x0 = (Hc - h) // 2, max(0, (Wc - 2 * w) // 2)
    view_a = np.round(canvas[y0:y0 + h, x0:x0 + w]).astype(np.uint8)
    th = math.radians(rng.uniform(-15, 15))
    sc = rng.uniform(0.8, 1.25)
    px, py = rng.uniform(-2e-4, 2e-4, 2)
    target = rng.uniform(*overlap)
    cxv, cyv = w / 2.0, h / 2.0
    best = None
    for shift in np.linspace(0.2, 0.9, 71):
        # G maps view-b pixels to view-a pixels: centre, perspective, rotate+scale, shift
        T0 = np.array([[1, 0, -cxv], [0, 1, -cyv], [0, 0, 1]], float)
        P = np.array([[1, 0, 0], [0, 1, 0], [px, py, 1]], float)
        R = np.array([[sc * math.cos(th), -sc * math.sin(th), 0], [sc * math.sin(th), sc * math.cos(th), 0], [0, 0, 1]])
        T1 = np.array([[1, 0, cxv + shift * w], [0, 1, cyv], [0, 0, 1]], float)
        G = T1 @ R @ P @ T0
        Hgt = np.linalg.inv(G)
        # overlap = fraction of view-a pixels that land inside view b
        ys, xs = np.mgrid[0:h:8, 0:w:8]
        pts = np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)])
        q = Hgt @ pts
        qx, qy = q[0] / q[2], q[1] / q[2]
        ov = float(np.mean((q[2] > 0) & (qx >= 0) & (qx < w) & (qy >= 0) & (qy < h)))
        if best is None or abs(ov - target) < abs(best[0] - target):
            best = (ov, G, Hgt)
    _, G, Hgt = best
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    q = G @ np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)])
    sx = (q[0] / q[2]).reshape(h, w) + x0
    sy = (q[1] / q[2]).reshape(h, w) + y0
    view_b = np.round(np.clip(_bilinear(canvas, sx, sy), 0, 255)).astype(np.uint8)
    return view_a, view_b, Hgt / Hgt[2, 2]


def synth_batch_sizes(n=64, seed=2000, lo=(240, 320), hi=(1080, 1920)):
    rng = np.random.default_rng(seed)
    Hs = rng.integers(lo[0], hi[0] + 1, n)
    Ws = rng.integers(lo[1], hi[1] + 1, n)
    return [(int(a), int(b)) for a, b in zip(Hs, Ws)]


def synth_batch(n=64, seed=2000):
    """Uneven batch: sizes uniform in 320x240 .. 1920x1080."""
    return [synth_gray(h, w, seed + 1 + i) for i, (h, w) in enumerate(synth_batch_sizes(n, seed))]


def corner_error(H_est, H_gt, w, h):
    """Mean corner transfer error (px) between two homographies."""
    c = np.array([[0, 0, 1], [w, 0, 1], [w, h, 1], [0, h, 1]], float).T
    a = H_est @ c; a = a[:2] / a[2]
    b = H_gt @ c; b = b[:2] / b[2]
    return float(np.mean(np.linalg.norm(a - b, axis=0)))


def synth_float_batch(B=16, H=512, W=512, seed=3000):
    """(B, H, W) float32 images in [0, 1]: the synthetic content of synth_gray
    (seed + image index), quantised, then scaled by 1/255."""
    return np.stack([synth_gray(H, W, seed + b).astype(np.float32) / np.float32(255.0) for b in range(B)])


def random_keypoints(B, K, H, W, n_oct=5, S=3, seed=3100, margin=2, max_octave=None):
    """(B, K, 5) float32 fixed keypoints (octave, si, row, col, angle): octave
    uniform in [0, min(n_oct, max_octave + 1)), si in [1, S], position uniform
    over the octave image inside `margin`, angle uniform in [0, 360)."""
    rng = np.random.default_rng(seed)
    top = n_oct if max_octave is None else min(n_oct, max_octave + 1)
    out = np.empty((B, K, 5), np.float32)
    o = rng.integers(0, top, (B, K))
    Ho = np.ceil(H / 2.0 ** o)
    Wo = np.ceil(W / 2.0 ** o)
    out[..., 0] = o
    out[..., 1] = rng.integers(1, S + 1, (B, K))
    out[..., 2] = margin + rng.random((B, K)) * np.maximum(Ho - 2 * margin - 1, 1)
    out[..., 3] = margin + rng.random((B, K)) * np.maximum(Wo - 2 * margin - 1, 1)
    out[..., 4] = rng.random((B, K)) * 360.0
    return out


def synth_warped_pair(h, w, seed, content_seed=None, fine=None):
    """(A, B, H_gt) uint8 grayscale: A = synth_gray(h, w), B = A warped about
    its centre by a random homography (rotation +-15 deg, scale 0.8-1.25,
    perspective ~U(-2e-4, 2e-4) scaled to the image size, shift ~U(-10 %, 10 %)),
    bilinear, zeros outside; H_gt maps A pixel coordinates to B's."""
    rng = np.random.default_rng(seed)
    a = synth_gray(h, w, seed + 1000 if content_seed is None else content_seed, as_float=True, fine=fine)
    th = math.radians(rng.uniform(-15, 15))
    sc = rng.uniform(0.8, 1.25)
    px, py = rng.uniform(-2e-4, 2e-4, 2) * (1024.0 / max(h, w))
    tx, ty = rng.uniform(-0.1, 0.1) * w, rng.uniform(-0.1, 0.1) * h
    T0 = np.array([[1, 0, -w / 2.0], [0, 1, -h / 2.0], [0, 0, 1]], float)
    P = np.array([[1, 0, 0], [0, 1, 0], [px, py, 1]], float)
    R = np.array([[sc * math.cos(th), -sc * math.sin(th), 0], [sc * math.sin(th), sc * math.cos(th), 0], [0, 0, 1]])
    T1 = np.array([[1, 0, w / 2.0 + tx], [0, 1, h / 2.0 + ty], [0, 0, 1]], float)
    G = T1 @ R @ P @ T0                       # B pixel -> A pixel
    b = np.empty((h, w), np.float64)
    for y0 in range(0, h, 512):               # row bands keep the temporaries small
        y1 = min(h, y0 + 512)
        ys, xs = np.mgrid[y0:y1, 0:w].astype(np.float64)
        q = G @ np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)])
        sx = (q[0] / q[2]).reshape(y1 - y0, w)
        sy = (q[1] / q[2]).reshape(y1 - y0, w)
        b[y0:y1] = np.clip(_bilinear(a[..., None], sx, sy)[..., 0], 0, 255)
    Hgt = np.linalg.inv(G)
    return np.round(a).astype(np.uint8), np.round(b).astype(np.uint8), Hgt / Hgt[2, 2]


def synth_pairs_batch(n_pairs=32, seed=2100, dense=True):
    """n_pairs warped pairs (synth_warped_pair) at the uneven batch sizes,
    with the fine texture of pairs_fine (dense=False: the plain recipe)."""
    sizes = synth_batch_sizes(2 * n_pairs, seed)
    return [synth_warped_pair(sizes[2 * p][0], sizes[2 * p][1], seed + 7 * p, seed + 1000 + p,
                              fine=pairs_fine(*sizes[2 * p]) if dense else None)
            for p in range(n_pairs)]


def synth_12mp_pair(seed=4000, fine=FINE_12MP):
    """The 4000x3000 pair: synth_warped_pair with the fine texture that gives
    20k-40k keypoints per image."""
    return synth_warped_pair(3000, 4000, seed, fine=fine)
