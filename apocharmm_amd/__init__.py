"""apocharmm_amd: MI355X-native CHARMM nonbonded + bonded force path.

The force evaluation lives in libapocharmm_amd.so (HIP, gfx950) behind the C
ABI declared in include/apocharmm_amd.h; this package is its host-side mirror
of the reference's ForceManager surface plus seeded synthetic workloads.
"""
from . import _lib, systems  # noqa: F401
from ._lib import ENERGY_NAMES, ApoError, ApoInvalidArgument  # noqa: F401


def ForceManager(*args, **kwargs):
    from .force_manager import ForceManager as _FM
    return _FM(*args, **kwargs)


__all__ = ["ForceManager", "systems", "ENERGY_NAMES", "ApoError", "ApoInvalidArgument"]
