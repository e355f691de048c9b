"""MI355X-native EPREM energetic-particle update.

The compute path is the C-ABI library built from eprem_amd/csrc
(libeprem_hip.so, entry points in include/eprem_hip.h); this package holds
the host-side mirror of the reference's configuration, the ctypes binding
and the output digest tool.
"""
from . import abi, config, constants, digest  # noqa: F401
