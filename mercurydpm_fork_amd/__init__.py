"""MI355X-native DEM time-step engine for MercuryDPM's hot path.

The compute path is the HIP library ``lib/libmercurydpm_dem.so`` behind the C
ABI of ``include/mercurydpm_dem.h``; ``dem.Engine`` binds it with ctypes.
"""
from . import dem, species  # noqa: F401

__all__ = ["dem", "species"]
