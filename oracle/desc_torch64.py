"""FP64 PyTorch restatement of the differentiable SIFT descriptor.

Test infrastructure only (the oracle of the gradient path): the device
library never imports it.  It restates, in float64 with torch autograd,

* the Gaussian pyramid of a float image in [0, 1] with a single-size base
  or the x2 cubic B-spline base (pysift/core.py:1234-1352, 1270-1274): base
  blur sqrt(sigma0^2 - 0.5^2) (doubled: sqrt(sigma0^2 - 1)), levels
  1..S by the incremental sigmas, next octave = level S[::2, ::2]; each blur is
  the separable H-then-V correlation with normalised exp(-x^2 / 2 sigma^2) taps,
  radius int(4 sigma + 0.5), half-sample reflect boundary (core.py:475-608);
* the 4x4x8 descriptor at fixed keypoints (core.py:693-835) and RootSIFT
  (core.py:1870-1873).

The sampling geometry that only depends on the keypoint (rotated coordinates,
window membership, spatial bins and weights, Gaussian window) is computed in
float32 with the same operation sequence as the device kernel, so both select
the same pixels and bins; everything that depends on the image (gradients,
magnitude, orientation, histogram, normalisation) is float64 and
differentiable.  Used for the autograd reference gradient and for float64
central finite differences.
"""
import math

import numpy as np
import torch

F64 = torch.float64


def sigma_schedule(sigma0=1.6, S=3, doubled=False):
    """(base sigma, [incremental sigmas of levels 1..S+2]) as core.py:1240-1284."""
    k = 2.0 ** (1.0 / S)
    sig = [sigma0 * k ** s for s in range(S + 3)]
    b = 1.0 if doubled else 0.5
    base = math.sqrt(max(sigma0 * sigma0 - b * b, 1e-6))
    inc = [math.sqrt(max(sig[s] ** 2 - sig[s - 1] ** 2, 1e-6)) for s in range(1, S + 3)]
    return base, inc


def gauss_taps(sigma):
    r = int(4.0 * sigma + 0.5)
    x = torch.arange(-r, r + 1, dtype=F64)
    v = torch.exp(-0.5 * (x / sigma) ** 2)
    return v / v.sum(), r


def reflect_index(n, r):
    """scipy 'reflect' (half-sample symmetric, period 2n) indices of -r .. n+r-1."""
    i = np.arange(-r, n + r) % (2 * n)
    i = np.where(i >= n, 2 * n - 1 - i, i)
    return torch.from_numpy(i.astype(np.int64))


def _corr_last(x, w, r):
    n = x.shape[-1]
    xp = x[..., reflect_index(n, r)]
    return xp.unfold(-1, 2 * r + 1, 1) @ w


def blur(x, sigma):
    """H pass then V pass on the last two dims (float64)."""
    if sigma < 0.01:
        return x
    w, r = gauss_taps(sigma)
    y = _corr_last(x, w, r)
    return _corr_last(y.transpose(-1, -2), w, r).transpose(-1, -2)


# cubic B-spline prefilter taps sqrt(3) z^|k|, z = sqrt(3) - 2, |k| <= 30: the
# FIR form of the mirror-boundary recursive prefilter of scipy.ndimage.zoom
# (order 3), as the detector uses for the x2 base (core.py:1274)
_UPS_K = 30
_UPS_H0 = float.fromhex("0x1.bb67ae8584caap+0")
_UPS_Z = float.fromhex("-0x1.126145e9ecd58p-2")


def _mirror(i, n):
    """scipy 'mirror' boundary (period 2n - 2) of index i."""
    if n == 1:
        return 0
    p = 2 * n - 2
    i = abs(i) % p
    return p - i if i >= n else i


def _bspline_w(t):
    s = 1.0 - t
    return [s ** 3 / 6.0, (2.0 / 3.0 - t * t) + 0.5 * t ** 3, (2.0 / 3.0 - s * s) + 0.5 * s ** 3, t ** 3 / 6.0]


def spline_matrices(n):
    """(prefilter (n, n), interpolation (2n, n)) float64 matrices of the x2
    cubic B-spline upsample along one axis of n samples: output o sits at
    o (n - 1) / (2n - 1) in input coordinates (zoom's grid_mode=False)."""
    taps = [_UPS_H0]
    for _ in range(_UPS_K):
        taps.append(taps[-1] * _UPS_Z)
    P = np.zeros((n, n))
    for y in range(n):
        for k in range(-_UPS_K, _UPS_K + 1):
            P[y, _mirror(y + k, n)] += taps[abs(k)]
    no = 2 * n
    z = (n - 1) / (no - 1)
    M = np.zeros((no, n))
    for o in range(no):
        t = o * z
        f = math.floor(t)
        for b, w in enumerate(_bspline_w(t - f)):
            M[o, _mirror(f - 1 + b, n)] += w
    return torch.from_numpy(P), torch.from_numpy(M)


def upsample2x(img):
    """x2 cubic B-spline upsample of (H, W) float64 (differentiable)."""
    H, W = img.shape
    Ph, Mh = spline_matrices(H)
    Pw, Mw = spline_matrices(W)
    return Mh @ (Ph @ img @ Pw.T) @ Mw.T


def pyramid(img, n_oct, S=3, sigma0=1.6, doubled=False):
    """img (H, W) float64 -> list over octaves of (S+1, Ho, Wo) level stacks;
    doubled: the base is the x2 spline upsample, blurred to sigma0 from an
    assumed blur of 1.0 (core.py:1270-1284)."""
    base, inc = sigma_schedule(sigma0, S, doubled=doubled)
    g0 = blur(upsample2x(img) if doubled else img, base)
    octs = []
    for _ in range(n_oct):
        levels = [g0]
        for lv in range(S):
            levels.append(blur(levels[-1], inc[lv]))
        octs.append(torch.stack(levels))
        g0 = levels[S][::2, ::2]
    return octs


def patch_half_sizes(S=3, sigma0=1.6):
    """ph per scale index: max(4, round(3 sigma sqrt2 2.5)) in float32 (core.py:1806-1810)."""
    k = np.float32(2.0 ** (1.0 / S))
    out = []
    for si in range(S + 1):
        pw = np.float32(math.pow(float(k), float(si)))
        t0 = np.float32(np.float32(3.0 * sigma0) * pw)
        t0 = np.float32(t0 * np.float32(1.41421354))
        t0 = np.float32(t0 * np.float32(2.5))
        ph = float(np.rint(t0))
        out.append(4 if ph < 4.0 else int(ph))
    return out


def _geometry(ph, ri, ci, angle_deg, H, W):
    """Float32 sampling geometry of a group of keypoints with one patch size,
    as the device kernel: arrays over the valid (keypoint, pixel) pairs."""
    f = np.float32
    ri = np.atleast_1d(np.asarray(ri, f))
    ci = np.atleast_1d(np.asarray(ci, f))
    ang_rad = (np.atleast_1d(np.asarray(angle_deg, f)) * f(0.0174532924)).astype(f)
    cos_a = np.cos(ang_rad.astype(np.float64)).astype(f)[:, None]
    sin_a = np.sin(ang_rad.astype(np.float64)).astype(f)[:, None]
    scale_sp = f(f(4.0) / f(f(2.0) * f(ph)))
    sigma_d = f(f(ph) * f(0.5657))
    inv_2s2 = f(f(1.0) / f(f(f(2.0) * sigma_d) * sigma_d + f(1e-12)))
    side = 2 * ph + 1
    r_off, c_off = np.divmod(np.arange(side * side), side)
    r_off = (r_off - ph).astype(np.int64)[None, :]
    c_off = (c_off - ph).astype(np.int64)[None, :]
    rr = ri.astype(np.int64)[:, None] + r_off
    cc = ci.astype(np.int64)[:, None] + c_off
    xr = (cos_a * c_off.astype(f) - sin_a * r_off.astype(f)).astype(f)
    yr = (sin_a * c_off.astype(f) + cos_a * r_off.astype(f)).astype(f)
    xb = (xr * scale_sp + f(1.5)).astype(f)
    yb = (yr * scale_sp + f(1.5)).astype(f)
    ok = (rr >= 1) & (rr < H - 1) & (cc >= 1) & (cc < W - 1) & (xb >= -0.5) & (xb < 3.5) & (yb >= -0.5) & (yb < 3.5)
    kidx = np.broadcast_to(np.arange(len(ri))[:, None], ok.shape)[ok]
    xb, yb, xr, yr, rr, cc = xb[ok], yb[ok], xr[ok], yr[ok], rr[ok], cc[ok]
    gw = np.exp(-(xr.astype(np.float64) ** 2 + yr.astype(np.float64) ** 2) * float(inv_2s2))
    xi0 = np.floor(xb).astype(np.int64)
    yi0 = np.floor(yb).astype(np.int64)
    dxb = (xb - xi0.astype(f)).astype(np.float64)
    dyb = (yb - yi0.astype(f)).astype(np.float64)
    return dict(k=kidx, rr=rr, cc=cc, gw=gw, xi0=xi0, yi0=yi0, dxb=dxb, dyb=dyb,
                ang_rad=ang_rad.astype(np.float64)[kidx])


def _normalize(hist, rootsift):
    t = hist / torch.sqrt((hist * hist).sum(-1, keepdim=True) + 1e-14)
    u = torch.clamp(t, max=0.2)
    v = u / torch.sqrt((u * u).sum(-1, keepdim=True) + 1e-14)
    if rootsift:
        q = v / (v.abs().sum(-1, keepdim=True) + 1e-7)
        pos = q > 0
        v = torch.where(pos, torch.sqrt(torch.where(pos, q, torch.ones_like(q))), torch.zeros_like(q))
    return v


def _state(oi0, m2, hist, rootsift):
    """Discrete decisions of the descriptor: orientation bin of every sample,
    zero / nonzero gradient magnitude of every sample, the clip at 0.2 and
    (RootSIFT) the zero / positive pattern of every bin.  Between two inputs
    with the same state the descriptor is an analytic function of the pixels."""
    t = hist / torch.sqrt((hist * hist).sum(-1, keepdim=True) + 1e-14)
    st = [oi0.numpy(), (m2 > 0).numpy(), (t > 0.2).numpy()]
    if rootsift:
        st.append((hist > 0).numpy())
    return st


def describe_group(G, ph, ri, ci, angle_deg, rootsift=True, return_state=False):
    """Descriptors (n, 128) float64 of n keypoints with patch half-size ph on
    one level image G (Ho, Wo) float64 (and their discrete state)."""
    H, W = G.shape
    n = len(np.atleast_1d(ri))
    g = _geometry(ph, ri, ci, angle_deg, H, W)
    rr = torch.from_numpy(g["rr"])
    cc = torch.from_numpy(g["cc"])
    dx = G[rr, cc + 1] - G[rr, cc - 1]
    dy = G[rr + 1, cc] - G[rr - 1, cc]
    m2 = dx * dx + dy * dy
    nz = m2 > 0
    mag = torch.where(nz, torch.sqrt(torch.where(nz, m2, torch.ones_like(m2))), torch.zeros_like(m2))
    theta = torch.atan2(dy, dx)
    ori = torch.remainder((theta - torch.from_numpy(g["ang_rad"])) * 57.29577951 + 360.0, 360.0)
    ob = ori * 0.02222222
    oi0 = torch.floor(ob.detach()).to(torch.int64)
    df = ob - oi0.to(F64)
    wmag = mag * torch.from_numpy(g["gw"])
    xi0 = torch.from_numpy(g["xi0"])
    yi0 = torch.from_numpy(g["yi0"])
    dxb = torch.from_numpy(g["dxb"])
    dyb = torch.from_numpy(g["dyb"])
    base = torch.from_numpy(g["k"]) * 128
    hist = torch.zeros(n * 128, dtype=F64)
    o0, o1 = oi0 % 8, (oi0 + 1) % 8
    for yy, wy in ((yi0, 1.0 - dyb), (yi0 + 1, dyb)):
        for xx, wx in ((xi0, 1.0 - dxb), (xi0 + 1, dxb)):
            cell = base + yy.clamp(0, 3) * 32 + xx.clamp(0, 3) * 8
            hist = hist.index_add(0, cell + o0, wy * wx * (1.0 - df) * wmag)
            hist = hist.index_add(0, cell + o1, wy * wx * df * wmag)
    if return_state:
        return _normalize(hist.view(n, 128), rootsift), _state(oi0, m2, hist.view(n, 128), rootsift)
    return _normalize(hist.view(n, 128), rootsift)


def describe(G, ph, ri, ci, angle_deg, rootsift=True):
    """Descriptor (128,) float64 of one keypoint on level image G (Ho, Wo) float64."""
    return describe_group(G, ph, [ri], [ci], [angle_deg], rootsift)[0]


def descriptors_from_pyramid(octs, kps, S=3, sigma0=1.6, rootsift=True, state=None):
    """kps (K, 5) rows (octave, si, ri, ci, angle) -> (K, 128); rows with an
    octave outside the pyramid or si outside [1, S] give zeros.  A list
    passed as `state` receives the discrete state of every keypoint group."""
    kps = np.asarray(kps, dtype=np.float32).reshape(-1, 5)
    n_oct = len(octs)
    ph = patch_half_sizes(S, sigma0)
    out = torch.zeros((len(kps), 128), dtype=F64)
    o = kps[:, 0]
    si = kps[:, 1].astype(np.int64)
    valid = (o >= 0) & (o.astype(np.int64) < n_oct) & (si >= 1) & (si <= S)
    oi = o.astype(np.int64)
    rows = []
    parts = []
    for oo in range(n_oct):
        for ss in range(1, S + 1):
            sel = np.nonzero(valid & (oi == oo) & (si == ss))[0]
            if len(sel) == 0:
                continue
            if state is None:
                parts.append(describe_group(octs[oo][ss], ph[ss], kps[sel, 2], kps[sel, 3], kps[sel, 4], rootsift))
            else:
                d, st = describe_group(octs[oo][ss], ph[ss], kps[sel, 2], kps[sel, 3], kps[sel, 4], rootsift, True)
                parts.append(d)
                state.extend(st)
            rows.append(torch.from_numpy(sel))
    if parts:
        out = out.index_copy(0, torch.cat(rows), torch.cat(parts))
    return out


def descriptors(img, kps, n_oct, S=3, sigma0=1.6, rootsift=True, doubled=False):
    """img (H, W) float64, kps (K, 5) rows (octave, si, ri, ci, angle) -> (K, 128)."""
    return descriptors_from_pyramid(pyramid(img, n_oct, S, sigma0, doubled), kps, S, sigma0, rootsift)


def loss_and_grad(img, kps, Wt, n_oct, S=3, sigma0=1.6, rootsift=True, doubled=False):
    """sum(Wt * descriptors) and its autograd gradient with respect to img."""
    x = torch.as_tensor(img, dtype=F64).clone().requires_grad_(True)
    d = descriptors(x, kps, n_oct, S, sigma0, rootsift, doubled)
    loss = (torch.as_tensor(Wt, dtype=F64).reshape(d.shape) * d).sum()
    (g,) = torch.autograd.grad(loss, x)
    return float(loss.detach()), g


def central_difference(img, kps, Wt, pixels, eps=1e-3, n_oct=5, S=3, sigma0=1.6, rootsift=True, radius=None,
                       return_checked=False, doubled=False):
    """FP64 central differences (L(img + eps e_p) - L(img - eps e_p)) / (2 eps)
    of the loss at pixels [(y, x), ...].  The pyramid is linear in the image,
    so it is built once for img and once for each impulse e_p and the
    perturbed pyramids are img's +- eps times the impulse's.  With `radius`
    (octave pixels), only keypoints whose centre lies within `radius` of the
    pixel in their own octave are re-evaluated (the others do not change);
    None re-evaluates all.  With a doubled base, octave o's pixels are
    2^(o - 1) image pixels.

    return_checked: also return, per pixel, whether the descriptor's discrete
    state (every sample's orientation bin and zero/nonzero magnitude, the clip
    at 0.2, RootSIFT's zero/positive bins) is the same at img + eps e_p and
    img - eps e_p for every re-evaluated keypoint.  At such a pixel the loss
    has no kink on the stencil, so the central difference approximates the
    derivative to O(eps^2); elsewhere it straddles a point where the loss is
    not differentiable and bounds no derivative."""
    img = np.asarray(img, dtype=np.float64)
    kps = np.asarray(kps, dtype=np.float32).reshape(-1, 5)
    Wt = torch.as_tensor(np.asarray(Wt, dtype=np.float64).reshape(-1, 128))
    H, W = img.shape
    with torch.no_grad():
        base = pyramid(torch.from_numpy(img), n_oct, S, sigma0, doubled)
        out, checked = [], []
        for (y, x) in pixels:
            sel = np.ones(len(kps), dtype=bool)
            if radius is not None:
                sc = 2.0 ** (np.maximum(kps[:, 0], 0) - (1 if doubled else 0))
                sel = (kps[:, 0] >= 0) & (np.abs(kps[:, 2] - y / sc) <= radius) & \
                      (np.abs(kps[:, 3] - x / sc) <= radius)
            if not sel.any():
                out.append(0.0)
                checked.append(True)
                continue
            e = torch.zeros((H, W), dtype=F64)
            e[y, x] = 1.0
            imp = pyramid(e, n_oct, S, sigma0, doubled)
            ls, states = [], []
            for sgn in (1.0, -1.0):
                octs = [b + (sgn * eps) * d for b, d in zip(base, imp)]
                st = [] if return_checked else None
                ls.append(float((Wt[sel] * descriptors_from_pyramid(octs, kps[sel], S, sigma0, rootsift, st)).sum()))
                states.append(st)
            out.append((ls[0] - ls[1]) / (2 * eps))
            if return_checked:
                checked.append(all(np.array_equal(a, b) for a, b in zip(states[0], states[1])))
    if return_checked:
        return np.array(out), np.array(checked, dtype=bool)
    return np.array(out)


def fd_job(job):
    """One worker task of central_difference for a process pool:
    job = (img, kps, Wt, pixels, eps, rootsift, n_oct, radius[, doubled]) -> (fd, checked)."""
    img, kps, Wt, pixels, eps, rootsift, n_oct, radius = job[:8]
    doubled = job[8] if len(job) > 8 else False
    torch.set_num_threads(1)
    return central_difference(img, kps, Wt, pixels, eps=eps, n_oct=n_oct, rootsift=rootsift, radius=radius,
                              return_checked=True, doubled=doubled)
