"""quarry: GPU BVH build and spatial queries on MI355X (gfx950).

Drop-in for the reference's Python module (bindings/python/quarry_py.cpp:887-1050):
describe(), demo_mesh(), demo_gaussians() and the Scene / PointCloud /
GaussianCloud classes, all running on the HIP library libquarry_hip.so through
its C ABI (quarry._abi).  The library is loaded on first use and a missing or
unloadable library raises immediately: there is no CPU fallback.
"""
from . import _abi  # noqa: F401
from ._abi import QuarryError  # noqa: F401
from .api import GaussianCloud, PointCloud, Scene  # noqa: F401
from .demo import demo_gaussians, demo_mesh  # noqa: F401

__version__ = "1.1.0"


def describe() -> str:
    """HIP device string, or an empty string when no GPU is available
    (quarry_py.cpp:891-897)."""
    d = _abi.describe()
    return d if d is not None else ""
