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


class DeviceEnsemble:
    """Independent networks advanced together on the MI355X (dfn_ensemble_*): one
    block-diagonal operator layout, shared stage launches, per-network clocks,
    step sizes and accept/reject decisions."""

    def __init__(self, problems):
        self._fns = load()
        self.problems = list(problems)
        self.nnet = len(self.problems)
        arr = (_abi.DfnProblem * self.nnet)()
        self._keep = []
        for i, p in enumerate(self.problems):
            cp, keep = p.to_c()
            arr[i] = cp
            self._keep.append(keep)
        h = C.c_void_p()
        rc = self._fns["ensemble_create"](arr, self.nnet, C.byref(h))
        self._h = h
        _abi.check(rc, "dfn_ensemble_create")

    def close(self):
        if self._h is not None and self._h.value:
            self._fns["ensemble_destroy"](self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initiate_hmat(self):
        _abi.check(self._fns["ensemble_build_hmat"](self._h), "dfn_ensemble_build_hmat")

    def set_precision(self, prec):
        _abi.check(self._fns["ensemble_set_precision"](self._h, int(prec)), "dfn_ensemble_set_precision")

    def set_graph(self, on):
        _abi.check(self._fns["ensemble_set_graph"](self._h, 1 if on else 0), "dfn_ensemble_set_graph")

    def solve(self, nsteps):
        _abi.check(self._fns["ensemble_solve"](self._h, int(nsteps)), "dfn_ensemble_solve")

    def get_state(self, net):
        N = self.problems[net].N
        Y, P, Pd, t = np.zeros(3 * N), np.zeros(N), np.zeros(N), C.c_double()
        _abi.check(self._fns["ensemble_get_state"](self._h, int(net), _abi.dptr(Y), _abi.dptr(P), _abi.dptr(Pd),
                                                   C.byref(t)), "dfn_ensemble_get_state")
        return Y, P, Pd, t.value

    def stats(self, net):
        s = _abi.DfnRunStats()
        _abi.check(self._fns["ensemble_get_stats"](self._h, int(net), C.byref(s)), "dfn_ensemble_get_stats")
        return s.as_dict()

    def info(self):
        v = np.zeros(6, dtype=np.int64)
        _abi.check(self._fns["ensemble_info"](self._h, _abi.i64ptr(v)), "dfn_ensemble_info")
        return dict(zip(("networks", "elements", "max_elements", "S_entries", "alg_bytes", "graph"),
                        (int(x) for x in v)))
