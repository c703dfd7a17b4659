"""ctypes wrapper of the CPU oracle (oracle/libsift_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results; the pysift_amd package never
imports it.  Each wrapper names the reference function (pysift/core.py:<line>)
that the C routine restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsift_oracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Params(ctypes.Structure):
    _fields_ = [
        ("n_octaves", ctypes.c_int), ("n_scales", ctypes.c_int), ("sigma0", ctypes.c_double),
        ("contrast_thresh", ctypes.c_double), ("edge_thresh", ctypes.c_double),
        ("double_image", ctypes.c_int), ("rootsift", ctypes.c_int), ("double_px_cap", ctypes.c_int),
        ("max_cands_per_octave", ctypes.c_int), ("max_keypoints", ctypes.c_int), ("border", ctypes.c_int),
        ("retry_below", ctypes.c_int), ("adaptive_thresh", ctypes.c_int),
        ("dsp_n_scales", ctypes.c_int), ("fp16_pyramid", ctypes.c_int),
    ]


def params(n_octaves=5, n_scales=3, sigma0=1.6, contrast_thresh=0.04, edge_thresh=10.0,
           double_image=True, rootsift=True, double_px_cap=16_000_000, max_cands_per_octave=25000,
           max_keypoints=8000, border=5, retry_below=100, adaptive_thresh=True, dsp_n_scales=0,
           fp16_pyramid=False):
    """Defaults of PySIFT.__init__ on a >=12 GB device (core.py:1184-1270);
    dsp_n_scales 3/4/5 selects DSP-SIFT (0: plain), fp16_pyramid the
    binary16 storage of the low-VRAM mode."""
    return Params(n_octaves, n_scales, sigma0, contrast_thresh, edge_thresh, int(double_image),
                  int(rootsift), double_px_cap, max_cands_per_octave, max_keypoints, border,
                  retry_below, int(adaptive_thresh), int(dsp_n_scales), int(fp16_pyramid))


def params_low_vram(**kw):
    """The reference's defaults below 12 GB of device memory (core.py:1191,
    1207, 1270): 4 octaves, binary16 pyramid storage, doubling below 4.19 MP."""
    d = dict(n_octaves=4, double_px_cap=4_194_304, fp16_pyramid=True)
    d.update(kw)
    return params(**d)


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _declare(L):
    L.or_exp.restype = _D; L.or_exp.argtypes = [_D]
    L.or_log2.restype = _D; L.or_log2.argtypes = [_D]
    L.or_atan2.restype = _D; L.or_atan2.argtypes = [_D, _D]
    L.or_exp_f.restype = ctypes.c_float; L.or_exp_f.argtypes = [ctypes.c_float]
    L.or_atan2_f.restype = ctypes.c_float; L.or_atan2_f.argtypes = [ctypes.c_float, ctypes.c_float]
    L.or_exp_f_arr.restype = None; L.or_exp_f_arr.argtypes = [_P, _P, ctypes.c_int]
    L.or_f16_round_arr.restype = None; L.or_f16_round_arr.argtypes = [_P, _P, ctypes.c_int]
    L.or_clahe.restype = _I; L.or_clahe.argtypes = [_P, _I, _I, _I, _D, _I, _I, _P]
    L.or_l2n_rows.restype = None; L.or_l2n_rows.argtypes = [_P, _I, _P]
    L.or_grid_warp.restype = None; L.or_grid_warp.argtypes = [_P, _I, _I, _P, _I, _I, _D, _D, _P, _P]
    L.or_gauss_weights_sp.restype = None; L.or_gauss_weights_sp.argtypes = [_P]
    L.or_gauss_weights_cv.restype = None; L.or_gauss_weights_cv.argtypes = [_P]
    L.or_gauss_down.restype = None; L.or_gauss_down.argtypes = [_P, _I, _I, _P, _P]
    L.or_zoom_up.restype = None; L.or_zoom_up.argtypes = [_P, _I, _I, _I, _I, _P]
    L.or_sobel.restype = None; L.or_sobel.argtypes = [_P, _I, _I, _I, _P]
    L.or_cv_blur25.restype = None; L.or_cv_blur25.argtypes = [_P, _I, _I, _P, _P]
    L.or_dp_seam.restype = None; L.or_dp_seam.argtypes = [_P, _P, _I, _I, _I, _P]
    L.or_composite.restype = _I
    L.or_composite.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I, _D, _D, _I, _I, _P, _P, _P, _P, _P, _P]
    L.or_hist_lut.restype = None; L.or_hist_lut.argtypes = [_P, _P, _P]
    L.or_pca_fit.restype = _I; L.or_pca_fit.argtypes = [_P, _I, _P, _I, _I, _P, _P, _P, _P]
    L.or_pca_project.restype = None; L.or_pca_project.argtypes = [_P, _I, _P, _P, _I, _P]
    L.or_sim16.restype = None; L.or_sim16.argtypes = [_P, _I, _P, _I, _I, _P]
    L.or_match_sim16.restype = None; L.or_match_sim16.argtypes = [_P, _I, _P, _I, _I, _D, _P, _P, _P, _P]
    L.or_descriptor_dsp.restype = None
    L.or_descriptor_dsp.argtypes = [_P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I, _I, _I, _P, _P]
    L.or_atan2_f_arr.restype = None; L.or_atan2_f_arr.argtypes = [_P, _P, _P, ctypes.c_int]
    L.or_sincos.restype = None; L.or_sincos.argtypes = [_D, _P, _P]
    L.or_radius.restype = _I; L.or_radius.argtypes = [_D]
    L.or_taps.restype = None; L.or_taps.argtypes = [_D, _I, _P]
    L.or_blur.restype = None; L.or_blur.argtypes = [_P, _P, _I, _I, _D]
    L.or_upsample2x.restype = None; L.or_upsample2x.argtypes = [_P, _I, _I, _P]
    L.or_adaptive_thresh.restype = _D; L.or_adaptive_thresh.argtypes = [_P, _I, _I, _I, _D]
    L.or_detect_describe.restype = _I
    L.or_detect_describe.argtypes = [_P, _I, _I, _I, ctypes.POINTER(Params), _I, _P, _P, _P]
    L.or_pyramid_octave.restype = _I
    L.or_pyramid_octave.argtypes = [_P, _I, _I, _I, ctypes.POINTER(Params), _I, _P, _P, _P, _P]
    L.or_match_u8.restype = None
    L.or_match_u8.argtypes = [_P, _I, _P, _I, _D, _P, _P, _P, _P]
    L.or_philox4x32_10.restype = None; L.or_philox4x32_10.argtypes = [_P, _P, _P]
    L.or_ransac_sample.restype = None; L.or_ransac_sample.argtypes = [_I, _I, ctypes.c_uint64, _P]
    L.or_dlt_ls.restype = _I; L.or_dlt_ls.argtypes = [_P, _P, _I, _P]
    L.or_ransac.restype = _I
    L.or_ransac.argtypes = [_P, _P, _I, ctypes.c_uint64, _I, _D, _D, _I, _I, _I, _P, _P, _P, _P]
    L.or_canvas.restype = None; L.or_canvas.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P]
    L.or_warp_feather.restype = None
    L.or_warp_feather.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I, _D, _D, _P, _P]
    L.or_bgr_to_gray.restype = None; L.or_bgr_to_gray.argtypes = [_P, _I, _I, _P]
    L.or_set_threads.restype = _I; L.or_set_threads.argtypes = [_I]
    L.or_extrema_octave.restype = _I
    L.or_extrema_octave.argtypes = [_P, _I, _I, _I, ctypes.c_float, _I, _I, _P]
    L.or_refine_octave.restype = _I
    L.or_refine_octave.argtypes = [_P, _I, _I, _I, _I, ctypes.c_float, ctypes.c_float, _P, _I, _P, _P]
    L.or_orientation.restype = _I
    L.or_orientation.argtypes = [_P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]
    L.or_descriptor.restype = None
    L.or_descriptor.argtypes = [_P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I, _I, _P, _P]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    return lib().or_set_threads(int(n))


# ---- elementary functions -------------------------------------------------
def exp(x):
    return lib().or_exp(float(x))


def log2(x):
    return lib().or_log2(float(x))


def atan2(y, x):
    return lib().or_atan2(float(y), float(x))


def exp_f(x):
    """Single-precision exp of the orientation / descriptor weights (array)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().or_exp_f_arr(_ptr(x), _ptr(y), x.size)
    return y


def atan2_f(y, x):
    """Single-precision atan2 of the orientation / descriptor gradients (arrays)."""
    y = np.ascontiguousarray(y, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    a = np.empty_like(y)
    lib().or_atan2_f_arr(_ptr(y), _ptr(x), _ptr(a), y.size)
    return a


def sincos(x):
    s = ctypes.c_double(); c = ctypes.c_double()
    lib().or_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


# ---- pyramid ----------------------------------------------------------------
def radius(sigma):
    """core.py:561: int(truncate*sigma + 0.5), truncate=4 (core.py:1288)."""
    return lib().or_radius(float(sigma))


def taps(sigma):
    r = radius(sigma)
    t = np.empty(2 * r + 1, np.float32)
    lib().or_taps(float(sigma), r, _ptr(t))
    return t


def blur(x, sigma):
    """_fast_gaussian (core.py:549-608)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().or_blur(_ptr(x), _ptr(out), x.shape[0], x.shape[1], float(sigma))
    return out


def upsample2x(x):
    """cpnd.zoom(x, 2.0, order=3) (core.py:1274)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((2 * x.shape[0], 2 * x.shape[1]), np.float32)
    lib().or_upsample2x(_ptr(x), x.shape[0], x.shape[1], _ptr(out))
    return out


def adaptive_thresh(img, base=0.04):
    """_adaptive_contrast_thresh (core.py:1129-1148)."""
    img = np.ascontiguousarray(img, np.uint8)
    return lib().or_adaptive_thresh(_ptr(img), img.shape[0], img.shape[1], img.shape[1], float(base))


def pyramid_octave(img, p, which):
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    Ho = ctypes.c_int(); Wo = ctypes.c_int()
    if lib().or_pyramid_octave(_ptr(img), H, W, W, ctypes.byref(p), which, ctypes.byref(Ho),
                               ctypes.byref(Wo), None, None) != 0:
        raise ValueError("octave out of range")
    S = p.n_scales
    g = np.empty((S + 3, Ho.value, Wo.value), np.float32)
    d = np.empty((S + 2, Ho.value, Wo.value), np.float32)
    lib().or_pyramid_octave(_ptr(img), H, W, W, ctypes.byref(p), which, ctypes.byref(Ho),
                            ctypes.byref(Wo), _ptr(g), _ptr(d))
    return g, d


def detect_describe(img, p=None, cap=200000):
    """PySIFT.detectAndCompute (core.py:2021-2172).  Returns
    kp (N,6) [x,y,size,angle,response,octave], desc f32 (N,128), desc u8 (N,128)."""
    p = p or params()
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    kp = np.empty((cap, 6), np.float32)
    d = np.empty((cap, 128), np.float32)
    du = np.empty((cap, 128), np.uint8)
    n = lib().or_detect_describe(_ptr(img), H, W, W, ctypes.byref(p), cap, _ptr(kp), _ptr(d), _ptr(du))
    if n < 0:
        raise ValueError("oracle detect failed")
    return kp[:n].copy(), d[:n].copy(), du[:n].copy()


# ---- matcher ----------------------------------------------------------------
def match_u8(a, b, ratio=0.75):
    """Exact integer restatement of _match_features (core.py:2840-2940)."""
    a = np.ascontiguousarray(a, np.uint8); b = np.ascontiguousarray(b, np.uint8)
    N, M = a.shape[0], b.shape[0]
    fwd = np.empty((max(N, 1), 3), np.int32); bwd = np.empty((max(M, 1), 3), np.int32)
    pairs = np.empty((max(min(N, M), 1), 2), np.int32)
    npairs = ctypes.c_int()
    lib().or_match_u8(_ptr(a), N, _ptr(b), M, float(ratio), _ptr(fwd), _ptr(bwd), _ptr(pairs),
                      ctypes.byref(npairs))
    return fwd[:N], bwd[:M], pairs[:npairs.value].copy()


# ---- RANSAC -----------------------------------------------------------------
def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32); k = np.ascontiguousarray(key, np.uint32)
    o = np.empty(4, np.uint32)
    lib().or_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def ransac_sample(B, N, seed):
    idx = np.empty((B, 4), np.int32)
    lib().or_ransac_sample(B, N, seed, _ptr(idx))
    return idx


def dlt_ls(p1, p2):
    """_compute_homography (core.py:2959-2981)."""
    p1 = np.ascontiguousarray(p1, np.float64); p2 = np.ascontiguousarray(p2, np.float64)
    H = np.empty(9, np.float64)
    if lib().or_dlt_ls(_ptr(p1), _ptr(p2), p1.shape[0], _ptr(H)) != 0:
        raise ValueError("need >= 4 points")
    return H.reshape(3, 3)


def ransac(p1, p2, seed=42, n_hyp=2000, thr=3.0, sigma_max=None, method="magsac", lo=True,
           min_inliers=10, return_all=False):
    """_ransac (core.py:3011-3097) with Philox hypotheses."""
    p1 = np.ascontiguousarray(p1, np.float64); p2 = np.ascontiguousarray(p2, np.float64)
    N = p1.shape[0]
    if N < 4:
        raise ValueError(f"Need >= 4 correspondences, got {N}")
    sm = thr if sigma_max is None else sigma_max
    H = np.empty(9, np.float64); mask = np.empty(N, np.uint8)
    sc = np.empty(n_hyp, np.float64); Hs = np.empty((n_hyp, 9), np.float64)
    best = lib().or_ransac(_ptr(p1), _ptr(p2), N, seed, n_hyp, float(thr), float(sm),
                           0 if method == "magsac" else 1, int(lo), int(min_inliers),
                           _ptr(H), _ptr(mask), _ptr(sc), _ptr(Hs))
    if return_all:
        return H.reshape(3, 3), mask.astype(bool), best, sc, Hs.reshape(-1, 3, 3)
    return H.reshape(3, 3), mask.astype(bool)


# ---- warp / feather ---------------------------------------------------------
def canvas(h1, w1, h2, w2, H):
    """_compute_canvas_size (core.py:3171-3187)."""
    H = np.ascontiguousarray(H, np.float64)
    cw = ctypes.c_int(); ch = ctypes.c_int(); ox = ctypes.c_double(); oy = ctypes.c_double()
    lib().or_canvas(h1, w1, h2, w2, _ptr(H), ctypes.byref(cw), ctypes.byref(ch), ctypes.byref(ox),
                    ctypes.byref(oy))
    return cw.value, ch.value, ox.value, oy.value


def warp_feather(img1, img2, H, cw, ch, ox, oy):
    img1 = np.ascontiguousarray(img1, np.uint8); img2 = np.ascontiguousarray(img2, np.uint8)
    H = np.ascontiguousarray(H, np.float64)
    pano = np.empty((ch, cw, 3), np.uint8); valid = np.empty((ch, cw), np.uint8)
    lib().or_warp_feather(_ptr(img1), img1.shape[0], img1.shape[1], _ptr(img2), img2.shape[0],
                          img2.shape[1], _ptr(H), cw, ch, float(ox), float(oy), _ptr(pano), _ptr(valid))
    return pano, valid.astype(bool)


def autocrop(pano, both):
    """_autocrop (core.py:3627-3684): the bounding box of the valid pixels,
    then inside it the column range [max first-valid, min last-valid] over the
    rows and the row range [max first-valid, min last-valid] over the columns
    (an all-invalid row or column counts as first 0 and last n-1, as NumPy's
    argmax of an all-False line gives 0); a range that is empty or one wide
    falls back to the bounding box; no valid pixel returns pano unchanged."""
    both = np.asarray(both, bool)
    rows = np.flatnonzero(both.any(1))
    cols = np.flatnonzero(both.any(0))
    if rows.size == 0:
        return pano
    r0, r1, c0, c1 = int(rows[0]), int(rows[-1]), int(cols[0]), int(cols[-1])
    sub = both[r0:r1 + 1, c0:c1 + 1]
    nr, nc = sub.shape
    ic0 = int(np.argmax(sub, axis=1).max())
    ic1 = int((nc - 1 - np.argmax(sub[:, ::-1], axis=1)).min())
    ir0 = int(np.argmax(sub, axis=0).max())
    ir1 = int((nr - 1 - np.argmax(sub[::-1, :], axis=0)).min())
    if ic0 >= ic1 or ir0 >= ir1:
        return pano[r0:r1 + 1, c0:c1 + 1]
    return pano[r0 + ir0:r0 + ir1 + 1, c0 + ic0:c0 + ic1 + 1]


def bgr_to_gray(bgr):
    """cv2.cvtColor(BGR2GRAY) (core.py:2548): OpenCV's fixed-point BT.601."""
    bgr = np.ascontiguousarray(bgr, np.uint8)
    g = np.empty(bgr.shape[:2], np.uint8)
    lib().or_bgr_to_gray(_ptr(bgr), bgr.shape[0], bgr.shape[1], _ptr(g))
    return g


# ---- single-stage hooks -----------------------------------------------------
def extrema_octave(D, thr, border=5, cap=1 << 20):
    """find_extrema_fused on one DoG stack (core.py:176-217): (n,3) s,r,c."""
    D = np.ascontiguousarray(D, np.float32)
    S2, H, W = D.shape
    out = np.empty((cap, 3), np.int32)
    n = lib().or_extrema_octave(_ptr(D), S2 - 2, H, W, float(thr), border, cap, _ptr(out))
    return out[:n].copy()


def refine_octave(D, cands, thr, edge=12.1, border=5):
    """refine_keypoints on one DoG stack (core.py:343-463), FP64 solve.
    Returns valid (n,) bool and out (n,4) ri, ci, si, response."""
    D = np.ascontiguousarray(D, np.float32)
    c = np.ascontiguousarray(cands, np.int32).reshape(-1, 3)
    S2, H, W = D.shape
    valid = np.zeros(len(c), np.int32); out = np.zeros((len(c), 4), np.float32)
    lib().or_refine_octave(_ptr(D), S2 - 2, H, W, border, float(thr), float(edge), _ptr(c), len(c),
                           _ptr(valid), _ptr(out))
    return valid.astype(bool), out


def orientation(G, ri, ci, sigma):
    """orientation_hist for one keypoint on one Gaussian level (core.py:228-333)."""
    G = np.ascontiguousarray(G, np.float32)
    a = np.empty(4, np.float32)
    n = lib().or_orientation(_ptr(G), G.shape[0], G.shape[1], float(ri), float(ci), float(sigma), _ptr(a))
    return a[:n].copy()


def descriptor(G, ri, ci, ang_rad, ph, rootsift=True, dsp_n_scales=0):
    """sift_descriptor_coop (or, with dsp_n_scales 3/4/5, sift_descriptor_dsp_coop)
    + RootSIFT for one keypoint (core.py:693-835, 845-997, 1870-1873)."""
    G = np.ascontiguousarray(G, np.float32)
    d = np.empty(128, np.float32); u = np.empty(128, np.uint8)
    lib().or_descriptor_dsp(_ptr(G), G.shape[0], G.shape[1], float(ri), float(ci), float(ang_rad), int(ph),
                            int(rootsift), int(dsp_n_scales), _ptr(d), _ptr(u))
    return d, u


def f16_round(x):
    """x (float32) rounded to binary16 and back, as a float16 array stores it."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().or_f16_round_arr(_ptr(x), _ptr(y), x.size)
    return y


def clahe(img, clip=2.0, tiles=(8, 8)):
    """CLAHE with OpenCV semantics on a u8 image (core.py:2562-2565); tiles = (x, y)."""
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    out = np.empty_like(img)
    if lib().or_clahe(_ptr(img), H, W, W, float(clip), int(tiles[0]), int(tiles[1]), _ptr(out)) != 0:
        raise ValueError("bad CLAHE arguments")
    return out


# ---- the reference's float matcher (core.py:2574-2743, 2840-2940) ------------
def l2n_rows(x):
    """x / (||x|| + 1e-7) per row (core.py:2879-2880), float32, (n, 128)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().or_l2n_rows(_ptr(x), x.shape[0], _ptr(y))
    return y


def pca_fit(xa, xb, dims=64):
    """PCA of the stacked L2-normalised rows (core.py:2728-2736): returns
    (mean (128,) f32, proj (128, dims) f32, eigenvalues descending (128,) f64,
    Jacobi sweeps)."""
    xa = np.ascontiguousarray(xa, np.float32); xb = np.ascontiguousarray(xb, np.float32)
    mean = np.empty(128, np.float32); proj = np.empty((128, dims), np.float32); ev = np.empty(128, np.float64)
    sw = ctypes.c_int()
    if lib().or_pca_fit(_ptr(xa), xa.shape[0], _ptr(xb), xb.shape[0], dims, _ptr(mean), _ptr(proj), _ptr(ev),
                        ctypes.byref(sw)) != 0:
        raise ValueError("bad PCA arguments")
    return mean, proj, ev, sw.value


def pca_project(x, mean, proj):
    """(x - mean) @ proj, rows re-normalised with +1e-7 (core.py:2737-2742)."""
    x = np.ascontiguousarray(x, np.float32)
    mean = np.ascontiguousarray(mean, np.float32); proj = np.ascontiguousarray(proj, np.float32)
    dims = proj.shape[1]
    y = np.empty((x.shape[0], dims), np.float32)
    lib().or_pca_project(_ptr(x), x.shape[0], _ptr(mean), _ptr(proj), dims, _ptr(y))
    return y


def sim16(a, b):
    """binary16 similarity matrix (the fp16-autocast mm output, core.py:2668-2670)."""
    a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
    out = np.empty((a.shape[0], b.shape[0]), np.float32)
    lib().or_sim16(_ptr(a), a.shape[0], _ptr(b), b.shape[0], a.shape[1], _ptr(out))
    return out


def match_sim16(a, b, ratio=0.75):
    """_match_symmetric on unit rows a (N, D), b (M, D): returns fwd (N, 3),
    bwd (M, 3) (index, d1, d2 with the squared distances as float32 bits) and
    the mutual pairs."""
    a = np.ascontiguousarray(a, np.float32); b = np.ascontiguousarray(b, np.float32)
    N, M = a.shape[0], b.shape[0]
    fwd = np.empty((max(N, 1), 3), np.int32); bwd = np.empty((max(M, 1), 3), np.int32)
    pairs = np.empty((max(min(N, M), 1), 2), np.int32)
    npairs = ctypes.c_int()
    lib().or_match_sim16(_ptr(a), N, _ptr(b), M, a.shape[1], float(ratio), _ptr(fwd), _ptr(bwd), _ptr(pairs),
                         ctypes.byref(npairs))
    return fwd[:N], bwd[:M], pairs[:npairs.value].copy()


def match_float(da, db, ratio=0.75, dims=64, pca=None):
    """The reference's float path end to end (core.py:2875-2939 with
    int8_matching off): normalise, PCA to dims (fitted on this pair unless
    `pca` = (mean, proj) is given; dims >= 128 or 0 skips it), fp16 similarity,
    symmetric ratio test, mutual pairs.  Returns (fwd, bwd, pairs, pca)."""
    a, b = l2n_rows(da), l2n_rows(db)
    if 0 < dims < 128:
        if pca is None:
            m, p, _, _ = pca_fit(a, b, dims)
            pca = (m, p)
        a, b = pca_project(a, *pca), pca_project(b, *pca)
    fwd, bwd, pairs = match_sim16(a, b, ratio)
    return fwd, bwd, pairs, pca


# ---- compositing: gain + DP seam + multiband (core.py:3211-3239, 3287-3553, 3686-3714) ----
def grid_warp(img, H, cw, ch, ox, oy):
    """_warp_image (core.py:3211-3239): (warped uint8 (ch, cw, 3), mask bool)."""
    img = np.ascontiguousarray(img, np.uint8); H = np.ascontiguousarray(H, np.float64)
    out = np.empty((ch, cw, 3), np.uint8); m = np.empty((ch, cw), np.uint8)
    lib().or_grid_warp(_ptr(img), img.shape[0], img.shape[1], _ptr(H), cw, ch, float(ox), float(oy), _ptr(out), _ptr(m))
    return out, m.astype(bool)


def gauss_weights_sp():
    w = np.empty(9, np.float32); lib().or_gauss_weights_sp(_ptr(w)); return w


def gauss_weights_cv():
    w = np.empty(25, np.float32); lib().or_gauss_weights_cv(_ptr(w)); return w


def gauss_down(x):
    x = np.ascontiguousarray(x, np.float32)
    H, W = x.shape
    out = np.empty(((H + 1) // 2, (W + 1) // 2), np.float32)
    lib().or_gauss_down(_ptr(x), H, W, _ptr(gauss_weights_sp()), _ptr(out))
    return out


def zoom_up(x, th, tw):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((th, tw), np.float32)
    lib().or_zoom_up(_ptr(x), x.shape[0], x.shape[1], th, tw, _ptr(out))
    return out


def sobel(x, axis):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().or_sobel(_ptr(x), x.shape[0], x.shape[1], int(axis), _ptr(out))
    return out


def cv_blur25(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().or_cv_blur25(_ptr(x), x.shape[0], x.shape[1], _ptr(gauss_weights_cv()), _ptr(out))
    return out


def dp_seam(cost, ov, vertical=True):
    cost = np.ascontiguousarray(cost, np.float32); ov = np.ascontiguousarray(ov, np.uint8)
    H, W = cost.shape
    seam = np.empty(H if vertical else W, np.int32)
    lib().or_dp_seam(_ptr(cost), _ptr(ov), H, W, int(bool(vertical)), _ptr(seam))
    return seam


def hist_lut(hs, hr):
    """The per-channel LUT of _hist_match (core.py:3591-3626) from the overlap
    histograms hs, hr (3, 256) of the image to correct and of the reference."""
    hs = np.ascontiguousarray(hs, np.uint64); hr = np.ascontiguousarray(hr, np.uint64)
    lut = np.empty((3, 256), np.float32)
    lib().or_hist_lut(_ptr(hs), _ptr(hr), _ptr(lut))
    return lut


def composite(img1, img2, H, cw, ch, ox, oy, n_levels=6, color_match="gain"):
    """_blend_pair (color_match 'gain' or 'hist') plus the warp and paste of
    _stitch_pair (core.py:3686-3714, 3783-3802): returns (pano uint8 (ch, cw,
    3), valid bool, seam mask float32, gain (3,), seam RMSE, vertical)."""
    img1 = np.ascontiguousarray(img1, np.uint8); img2 = np.ascontiguousarray(img2, np.uint8)
    H = np.ascontiguousarray(H, np.float64)
    pano = np.empty((ch, cw, 3), np.uint8); valid = np.empty((ch, cw), np.uint8)
    seam = np.empty((ch, cw), np.float32); gain = np.empty(3, np.float32)
    rmse = ctypes.c_double(); vert = ctypes.c_int()
    if lib().or_composite(_ptr(img1), img1.shape[0], img1.shape[1], _ptr(img2), img2.shape[0], img2.shape[1],
                          _ptr(H), cw, ch, float(ox), float(oy), int(n_levels), 1 if color_match == "hist" else 0,
                          _ptr(pano), _ptr(valid), _ptr(seam),
                          _ptr(gain), ctypes.byref(rmse), ctypes.byref(vert)) != 0:
        raise ValueError("bad composite arguments")
    return pano, valid.astype(bool), seam, gain, rmse.value, vert.value
