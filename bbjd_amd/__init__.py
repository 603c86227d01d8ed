"""bbjd_amd -- Jacobi-Davidson QZ eigensolver for A x = lambda B x on one
AMD Instinct MI355X (gfx950).

``jdqz`` is the scipy-eigs-style entry (the reference's bbjd.jdqz surface);
``HAVE_C_EXT`` tells whether libbbjd.so is built (reference __init__.py:9-12).
"""
from ._lib import LIB_PATH
import os as _os

HAVE_C_EXT = _os.path.exists(LIB_PATH)

from .bbjd import jdqz  # noqa: E402
from .pencil import BandedTauPencil, DeviceOperator  # noqa: E402
from .sparse import SparseDeviceOperator  # noqa: E402

__all__ = ["jdqz", "HAVE_C_EXT", "BandedTauPencil", "DeviceOperator", "SparseDeviceOperator"]
