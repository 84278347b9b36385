"""opentimer_amd -- MI355X-native RC-tree Elmore timing and levelization.

The compute lives in libotg.so (hand-written gfx950 HIP kernels behind the C
ABI of include/otg.h); this package is its host-side mirror.
"""
from . import native  # noqa: F401  (raises ImportError when libotg.so is missing)

__all__ = ["native"]
__version__ = "0.1.0"
