"""benchmarks_amd -- MI355X-native matrix-free Laplacian (CEED BK3/BK5, BP3/BP5)
behind the reference's operator interface.  The compute path is libmfx.so
(hand-written gfx950 HIP kernels behind the C-ABI in include/mfx.h)."""
from ._lib import LIB_PATH, MfxError, lib  # noqa: F401

__all__ = ["LIB_PATH", "MfxError", "lib"]
