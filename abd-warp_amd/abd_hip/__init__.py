"""abd_hip — MI355X (gfx950) implementation of the abd-warp per-Newton-iteration
contact path: culling, IPC barrier energy / gradient / PSD-projected Hessian,
cubic CCD, and the block-sparse Newton solve, behind the reference's Python
call shapes.  All computation runs in libabd_hip.so (hand-written HIP); torch
tensors are only the device buffers.
"""
from . import _lib
from .const_params import ContactParams
from .scenes import SceneSpec

__all__ = ["ContactParams", "SceneSpec", "load", "AffineBodySimulator"]


def load():
    """Loads libabd_hip.so (raises ImportError when it has not been built)."""
    return _lib.load()


def __getattr__(name):
    if name == "AffineBodySimulator":
        from .abd import AffineBodySimulator
        return AffineBodySimulator
    raise AttributeError(name)
