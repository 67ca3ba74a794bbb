"""flyhip: MI355X-native whole-brain leaky integrate-and-fire engine.

Drop-in for the reference's ``Brain`` (flyzero/brain.py) and ``BatchBrain`` (flyzero/gpu.py):
Python host code over hand-written HIP kernels (``libflyhip.so``, C ABI in include/flyhip.h).
"""

from .params import LIFParams

__all__ = ["LIFParams", "Brain", "BatchBrain"]


def __getattr__(name):
    # the engine classes load torch and the native library on first use
    if name == "Brain":
        from .brain import Brain
        return Brain
    if name == "BatchBrain":
        from .gpu import BatchBrain
        return BatchBrain
    raise AttributeError(name)
