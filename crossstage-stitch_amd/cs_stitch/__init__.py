"""MI355X-native per-pair stitching path (S1-S6) behind the reference's stage
and harness interfaces. The compute runs in libcsstitch.so (HIP kernels for
gfx950); this package only binds it."""
from . import _ffi
from ._ffi import load
from .stages import StageError

__all__ = ["load", "StageError", "_ffi"]
