"""libcpab_amd: CPA (continuous piecewise-affine) transformations with the
integration, its gradient and the interpolation that follows it running as
hand-written HIP kernels on AMD Instinct MI355X (gfx950).

The public API mirrors SkafteNicki/libcpab (libcpab/__init__.py:9-14).
"""
from .cpab import Cpab
from .alignment import CpabAligner
from .sequential import CpabSequential

__all__ = ["Cpab", "CpabAligner", "CpabSequential"]
