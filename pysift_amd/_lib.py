"""ctypes binding of libpysift_amd.so (the C-ABI declared in include/pysift_amd.h).

Device buffers are torch tensors passed by data_ptr(); the stream is the
current torch stream.  A missing or unloadable library raises ImportError:
there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpysift_amd.so")

MSIFT_OK = 0
MSIFT_EINVAL = -1
MSIFT_EHIP = -2
MSIFT_EWORKSPACE = -3
MSIFT_ESTATE = -4

EXPORTS = [
    "msift_abi_version", "msift_last_error", "msift_create", "msift_destroy", "msift_workspace_bytes",
    "msift_detect_describe", "msift_pyramid_view", "msift_match_workspace_bytes", "msift_match_u8",
    "msift_match_f16", "msift_ransac_workspace_bytes", "msift_ransac_homography", "msift_warp_feather",
    "msift_bgr_to_gray", "msift_batch_create", "msift_batch_destroy", "msift_batch_sizes",
    "msift_batch_set_blur_tile", "msift_set_blur_tile", "msift_detect_describe_batched", "msift_grad_workspace_bytes",
    "msift_desc_forward", "msift_desc_backward", "msift_grad_level_view", "msift_match_batch_create",
    "msift_match_batch_destroy", "msift_match_batch_sizes", "msift_match_u8_batched",
    "msift_match_batch_gather_points", "msift_ransac_batch_create", "msift_ransac_batch_destroy",
    "msift_ransac_batch_sizes", "msift_ransac_homography_batched", "msift_clahe_workspace_bytes", "msift_clahe",
    "msift_l2_normalize", "msift_pca_workspace_bytes", "msift_pca_fit", "msift_pca_project", "msift_match_sim_f16",
    "msift_composite_workspace_bytes", "msift_composite_multiband", "msift_match_u8_noalloc", "msift_match_set_tile",
]


class MsiftParams(ctypes.Structure):
    _fields_ = [
        ("n_octaves", ctypes.c_int), ("n_scales", ctypes.c_int), ("sigma0", ctypes.c_double),
        ("contrast_thresh", ctypes.c_double), ("edge_thresh", ctypes.c_double),
        ("double_image", ctypes.c_int), ("rootsift", ctypes.c_int), ("double_px_cap", ctypes.c_int),
        ("max_cands_per_octave", ctypes.c_int), ("max_keypoints", ctypes.c_int), ("border", ctypes.c_int),
        ("retry_below", ctypes.c_int), ("adaptive_thresh", ctypes.c_int),
        ("dsp_n_scales", ctypes.c_int), ("fp16_pyramid", ctypes.c_int),
    ]


_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t


def _declare(L):
    L.msift_abi_version.restype = _I; L.msift_abi_version.argtypes = []
    L.msift_last_error.restype = _I; L.msift_last_error.argtypes = [ctypes.c_char_p, _SZ]
    L.msift_create.restype = _I; L.msift_create.argtypes = [ctypes.POINTER(MsiftParams), ctypes.POINTER(_P)]
    L.msift_destroy.restype = _I; L.msift_destroy.argtypes = [_P]
    L.msift_workspace_bytes.restype = _I
    L.msift_workspace_bytes.argtypes = [_P, _I, _I, ctypes.POINTER(_SZ)]
    L.msift_detect_describe.restype = _I
    L.msift_detect_describe.argtypes = [_P, _P, _I, _I, _I, _P, _SZ, _I, _P, _P, _P, _P, _P, _P]
    L.msift_pyramid_view.restype = _I
    L.msift_pyramid_view.argtypes = [_P, _I, _I, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                     ctypes.POINTER(_P), ctypes.POINTER(_P)]
    L.msift_match_workspace_bytes.restype = _I
    L.msift_match_workspace_bytes.argtypes = [_I, _I, ctypes.POINTER(_SZ)]
    L.msift_match_u8.restype = _I
    L.msift_match_u8.argtypes = [_P, _I, _P, _I, _D, _I, _P, _SZ, _P, _P, _P, _P, _P]
    L.msift_match_set_tile.restype = _I
    L.msift_match_set_tile.argtypes = [_I, _I]
    L.msift_match_u8_noalloc.restype = _I
    L.msift_match_u8_noalloc.argtypes = [_P, _I, _P, _I, _D, _I, _P, _P, _P, _P, _P]
    L.msift_match_f16.restype = _I
    L.msift_match_f16.argtypes = [_P, _I, _P, _I, _D, _P, _SZ, _P, _P, _P, _P, _P]
    L.msift_ransac_workspace_bytes.restype = _I
    L.msift_ransac_workspace_bytes.argtypes = [_I, _I, ctypes.POINTER(_SZ)]
    L.msift_ransac_homography.restype = _I
    L.msift_ransac_homography.argtypes = [_P, _P, _I, ctypes.c_uint64, _I, _D, _D, _I, _I, _I, _P, _SZ,
                                          _P, _P, _P, _P, _P]
    L.msift_warp_feather.restype = _I
    L.msift_warp_feather.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I, _D, _D, _P, _P, _P]
    L.msift_bgr_to_gray.restype = _I
    L.msift_bgr_to_gray.argtypes = [_P, _I, _I, _P, _P]
    L.msift_clahe_workspace_bytes.restype = _I
    L.msift_clahe_workspace_bytes.argtypes = [_I, _I, ctypes.POINTER(_SZ)]
    L.msift_clahe.restype = _I
    L.msift_clahe.argtypes = [_P, _I, _I, _I, _D, _I, _I, _P, _SZ, _P, _P]
    L.msift_l2_normalize.restype = _I; L.msift_l2_normalize.argtypes = [_P, _I, _P, _P]
    L.msift_pca_workspace_bytes.restype = _I
    L.msift_pca_workspace_bytes.argtypes = [_I, _I, ctypes.POINTER(_SZ)]
    L.msift_pca_fit.restype = _I
    L.msift_pca_fit.argtypes = [_P, _I, _P, _I, _I, _P, _SZ, _P, _P, _P, _P, _P]
    L.msift_pca_project.restype = _I; L.msift_pca_project.argtypes = [_P, _I, _P, _P, _I, _P, _P]
    L.msift_composite_workspace_bytes.restype = _I
    L.msift_composite_workspace_bytes.argtypes = [_I, _I, _I, ctypes.POINTER(_SZ)]
    L.msift_composite_multiband.restype = _I
    L.msift_composite_multiband.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I, _D, _D, _I, _I, _P, _SZ, _P, _P, _P, _P,
                                            _P]
    L.msift_match_sim_f16.restype = _I
    L.msift_match_sim_f16.argtypes = [_P, _I, _P, _I, _I, _D, _P, _SZ, _P, _P, _P, _P, _P]
    L.msift_batch_create.restype = _I
    L.msift_batch_create.argtypes = [_P, _I, _P, _P, ctypes.POINTER(_P)]
    L.msift_batch_destroy.restype = _I; L.msift_batch_destroy.argtypes = [_P]
    L.msift_batch_sizes.restype = _I; L.msift_batch_sizes.argtypes = [_P, _P, _P, _P]
    L.msift_batch_set_blur_tile.restype = _I; L.msift_batch_set_blur_tile.argtypes = [_P, _I]
    L.msift_set_blur_tile.restype = _I; L.msift_set_blur_tile.argtypes = [_P, _I]
    L.msift_detect_describe_batched.restype = _I
    L.msift_detect_describe_batched.argtypes = [_P, _P, _P, _SZ, _I, _P, _P, _P, _P, _P, _P]
    L.msift_grad_workspace_bytes.restype = _I
    L.msift_grad_workspace_bytes.argtypes = [_P, _I, _I, _I, ctypes.POINTER(_SZ)]
    L.msift_desc_forward.restype = _I
    L.msift_desc_forward.argtypes = [_P, _P, _I, _I, _I, _P, _I, _P, _SZ, _P, _P]
    L.msift_desc_backward.restype = _I
    L.msift_desc_backward.argtypes = [_P, _P, _I, _I, _I, _P, _I, _P, _P, _SZ, _P, _P]
    L.msift_match_batch_create.restype = _I
    L.msift_match_batch_create.argtypes = [_I, _P, _P, ctypes.POINTER(_P)]
    L.msift_match_batch_destroy.restype = _I; L.msift_match_batch_destroy.argtypes = [_P]
    L.msift_match_batch_sizes.restype = _I; L.msift_match_batch_sizes.argtypes = [_P, _P, _P, _P]
    L.msift_match_u8_batched.restype = _I
    L.msift_match_u8_batched.argtypes = [_P, _P, _P, _D, _I, _P, _SZ, _P, _P, _P, _P, _P]
    L.msift_match_batch_gather_points.restype = _I
    L.msift_match_batch_gather_points.argtypes = [_P, _P, _I, _P, _I, _P, _P, _P, _P, _P]
    L.msift_ransac_batch_create.restype = _I
    L.msift_ransac_batch_create.argtypes = [_I, _P, _I, ctypes.POINTER(_P)]
    L.msift_ransac_batch_destroy.restype = _I; L.msift_ransac_batch_destroy.argtypes = [_P]
    L.msift_ransac_batch_sizes.restype = _I; L.msift_ransac_batch_sizes.argtypes = [_P, _P, _P]
    L.msift_ransac_homography_batched.restype = _I
    L.msift_ransac_homography_batched.argtypes = [_P, _P, _P, _P, ctypes.c_uint64, _D, _D, _I, _I, _I, _P, _SZ,
                                                  _P, _P, _P, _P, _P]
    L.msift_grad_level_view.restype = _I
    L.msift_grad_level_view.argtypes = [_P, _I, _I, _I, _P, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                        ctypes.POINTER(_P), ctypes.POINTER(_P)]


def lib():
    """Load the in-tree library (raises ImportError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run pysift_amd/build.py (hipcc, gfx950)")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise ImportError(f"cannot load {LIB_PATH}: {e}") from e
        _declare(L)
        if L.msift_abi_version() != 2:
            raise ImportError("libpysift_amd ABI version mismatch")
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().msift_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    """Map a C status code to the reference's Python exception types."""
    if rc == MSIFT_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == MSIFT_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def stream_ptr(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Raw device pointer of a torch tensor (None -> NULL)."""
    return ctypes.c_void_p(0 if t is None else t.data_ptr())
