"""MI355X-native (gfx950) relational hot path of FlexPushdownDB.

Hash-join build/probe, blocked and legacy bloom filters (predicate transfer)
and hash group-by, as hand-written HIP kernels behind a C ABI
(include/fpdb_amd.h, libfpdb_amd.so), with the reference's PhysicalOp layer
rebuilt in C++ on top (flexpushdowndb_amd/pop, libfpdb_pop.so).

There is no CPU fallback: the operators raise if the HIP library is missing.
"""
from . import _capi  # noqa: F401
from .device import Context, DeviceArray, DeviceColumn  # noqa: F401

__all__ = ["Context", "DeviceArray", "DeviceColumn"]
