"""quarry: GPU BVH build and spatial queries on MI355X (gfx950).

Drop-in for the reference's Python module (bindings/python/quarry_py.cpp):
describe(), demo_mesh(), demo_gaussians() and the Scene / PointCloud /
GaussianCloud classes live in quarry.api; the low-level ctypes binding of the
C ABI is quarry._abi.
"""
from . import _abi  # noqa: F401
from ._abi import QuarryError  # noqa: F401

__version__ = "1.1.0"


def describe() -> str:
    """Name of the HIP device, or raises QuarryError when none is visible."""
    d = _abi.describe()
    if d is None:
        raise QuarryError(_abi.lib().qr_last_error().decode())
    return d
