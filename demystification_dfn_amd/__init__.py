"""MI355X-native per-RK-stage hot path of DemystiFicatioN.

The compute lives in the HIP library ``csrc/libdfn.so`` (C ABI, include/dfn.h);
this package only loads it, describes problems and mirrors the reference's
interfaces (see :mod:`.host`).
"""
from . import _abi  # noqa: F401
from .problems import Problem, fault_network, network_80k, network_small, single_fault  # noqa: F401
