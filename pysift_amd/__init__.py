"""pysift_amd: MI355X-native drop-in for PySIFT's detect -> describe -> match
-> RANSAC -> warp/feather path (reference SivaIITM/PySIFT v0.1.4).

Public names follow the reference package facade (pysift/__init__.py:1-7):
PySIFT and GPUPyStitch, plus cv2-compatible KeyPoint/DMatch duck types.
"""
from .detector import DMatch, KeyPoint, PySIFT  # noqa: F401

__version__ = "0.1.4"


def __getattr__(name):
    if name == "GPUPyStitch":
        from .stitch import GPUPyStitch
        return GPUPyStitch
    if name == "SiftDescriptorFn":
        from .autograd import SiftDescriptorFn
        return SiftDescriptorFn
    raise AttributeError(name)
