"""Loader of the gfx950 device library (csrc/libdfn.so) and the device simulation.

There is no fallback: if the library is missing, or no gfx950 device is present,
every entry point raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from .host import Simulation

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libdfn.so")
_lib = None
_fns = None


def load():
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        _lib = C.CDLL(LIB_PATH)
        _fns = _abi.bind(_lib, "dfn_")
    return _fns


def device_info():
    buf = C.create_string_buffer(512)
    load()["device_info"](buf, 512)
    return buf.value.decode()


def bench_stream(nbytes=1 << 30, reps=20):
    """GB/s (read + write) of a plain copy kernel over nbytes on the device."""
    g = np.zeros(1)
    _abi.check(load()["bench_stream"](int(nbytes), int(reps), _abi.dptr(g)), "dfn_bench_stream")
    return float(g[0])


class DeviceSimulation(Simulation):
    """A problem resident on the MI355X: operators, state and solver loop on the device."""

    def __init__(self, problem):
        super().__init__(load(), problem, tag="dfn")

    def bench_kernel(self, reps=20):
        ms = np.zeros(4)
        _abi.check(self._fns["bench_kernel"](self._h, int(reps), _abi.dptr(ms)), "dfn_bench_kernel")
        return dict(ms_kernel=float(ms[0]), ms_dominant=float(ms[1]), alg_bytes=float(ms[2]),
                    dominant_bytes=float(ms[3]))
