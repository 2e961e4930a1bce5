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

    def tuning(self):
        """The autotuner's launch shapes for the active operator layout (dfn_tuning)."""
        c, ms = np.zeros(8, np.int32), np.zeros(2)
        _abi.check(self._fns["tuning"](self._h, _abi.i32ptr(c), _abi.dptr(ms)), "dfn_tuning")
        return dict(vtx_waves=int(c[0]), vtx_order=int(c[1]), panel_rows=int(c[2]), panel_order=int(c[3]),
                    panel_split=int(c[4]), fused_near=int(c[5]), rank_ceiling=int(c[6]), vtx_pairs=int(c[7]),
                    ms_default=float(ms[0]), ms_tuned=float(ms[1]))

    def kernel_timing(self, on=True):
        """Start (clearing the sums) or stop timing the dominant kernel of every
        stress evaluation (dfn_kernel_timing)."""
        _abi.check(self._fns["kernel_timing"](self._h, 1 if on else 0), "dfn_kernel_timing")

    def kernel_timing_read(self):
        """(mean ms per launch, launches timed) since the last start (dfn_kernel_timing_read)."""
        ms, n = np.zeros(1), np.zeros(1, np.int64)
        _abi.check(self._fns["kernel_timing_read"](self._h, _abi.dptr(ms), _abi.i64ptr(n)), "dfn_kernel_timing_read")
        return float(ms[0]), int(n[0])

    def set_formulation(self, form):
        """dfn_set_formulation: _abi.FORM_REFERENCE or _abi.FORM_NEWTON_V."""
        _abi.check(self._fns["set_formulation"](self._h, int(form)), "dfn_set_formulation")

    # -- solver output files (dfn_set_output, dfn_flush_output) ----------------
    def set_output(self, directory, stride_time, freq_writing_file):
        """Solver output files (dfn_set_output); directory None stops it."""
        d = os.fsencode(directory) if directory is not None else None
        _abi.check(self._fns["set_output"](self._h, d, int(stride_time), int(freq_writing_file)), "dfn_set_output")

    def flush_output(self):
        _abi.check(self._fns["flush_output"](self._h), "dfn_flush_output")

    # -- event monitor (dfn_monitor, dfn_monitor_take, dfn_moment) --------------
    def monitor(self, on=True):
        _abi.check(self._fns["monitor"](self._h, 1 if on else 0), "dfn_monitor")

    def monitor_take(self):
        """(times, maxV[records, faults]) recorded since the last call."""
        n = C.c_int64()
        _abi.check(self._fns["monitor_take"](self._h, 0, None, None, C.byref(n)), "dfn_monitor_take")
        nf = self.problem.nb_fault
        t = np.zeros(n.value)
        v = np.zeros(n.value * nf)
        if n.value:
            _abi.check(self._fns["monitor_take"](self._h, n.value, _abi.dptr(t), _abi.dptr(v), C.byref(n)),
                       "dfn_monitor_take")
        return t, v.reshape(-1, nf)

    def moment(self):
        """(M0 per fault, slip per element)."""
        m = np.zeros(self.problem.nb_fault)
        s = np.zeros(self.N)
        _abi.check(self._fns["moment"](self._h, _abi.dptr(m), _abi.dptr(s)), "dfn_moment")
        return m, s

    def kernel_multi(self, X):
        """Stress rates of the rows of X (nrhs x N slip-rate vectors) with the FP64
        H-matrix in one multi-RHS application (dfn_kernel_multi): (sT, sN) of the
        same shape (sN None in Mode III)."""
        X = np.ascontiguousarray(X, np.float64)
        if X.ndim != 2 or X.shape[1] != self.N:
            raise ValueError(f"kernel_multi: X must be (nrhs, {self.N})")
        sT = np.zeros_like(X)
        sN = np.zeros_like(X) if self.mode == _abi.MODE_II else None
        _abi.check(self._fns["kernel_multi"](self._h, X.shape[0], _abi.dptr(X), _abi.dptr(sT),
                                             _abi.dptr(sN) if sN is not None else None), "dfn_kernel_multi")
        return sT, sN

    def bench_kernel_multi(self, X, reps=5):
        """dfn_bench_kernel_multi on the rows of X."""
        X = np.ascontiguousarray(X, np.float64)
        out = np.zeros(4)
        _abi.check(self._fns["bench_kernel_multi"](self._h, X.shape[0], _abi.dptr(X), int(reps), _abi.dptr(out)),
                   "dfn_bench_kernel_multi")
        return dict(ms_apply=float(out[0]), ms_panel=float(out[1]), panel_flops=float(out[2]),
                    alg_bytes=float(out[3]))

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

    def solve_until(self, t_end, max_steps=1_000_000):
        _abi.check(self._fns["ensemble_solve_until"](self._h, float(t_end), int(max_steps)),
                   "dfn_ensemble_solve_until")

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

    def set_controller(self, mode, mlp=None):
        h = mlp._h if mlp is not None else None
        _abi.check(self._fns["ensemble_set_controller"](self._h, int(mode), h), "dfn_ensemble_set_controller")
        self._mlp = mlp  # the network must outlive its use here

    def set_recording(self, on):
        _abi.check(self._fns["ensemble_set_recording"](self._h, 1 if on else 0), "dfn_ensemble_set_recording")

    def take_records(self):
        n = C.c_int64()
        _abi.check(self._fns["ensemble_take_records"](self._h, None, None, 0, C.byref(n)), "dfn_ensemble_take_records")
        X = np.zeros((n.value, _abi.MLP_INPUTS), np.float32)
        T = np.zeros(n.value, np.float32)
        if n.value:
            _abi.check(self._fns["ensemble_take_records"](self._h, _abi.fptr(X), _abi.fptr(T), n.value, C.byref(n)),
                       "dfn_ensemble_take_records")
        return X[:n.value], T[:n.value]

    def info(self):
        v = np.zeros(9, dtype=np.int64)
        _abi.check(self._fns["ensemble_info"](self._h, _abi.i64ptr(v)), "dfn_ensemble_info")
        return dict(zip(("networks", "elements", "max_elements", "S_entries", "alg_bytes", "graph", "fused_near",
                         "rank_ceiling", "vtx_pairs"),
                        (int(x) for x in v)))

    def kernel_timing(self, on=True):
        """Live per-launch timing of the dominant kernel (dfn_ensemble_kernel_timing)."""
        _abi.check(self._fns["ensemble_kernel_timing"](self._h, 1 if on else 0), "dfn_ensemble_kernel_timing")

    def kernel_timing_read(self):
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        _abi.check(self._fns["ensemble_kernel_timing_read"](self._h, C.byref(ms), C.byref(n), C.byref(b)),
                   "dfn_ensemble_kernel_timing_read")
        return ms.value, n.value, b.value


class DeviceMlp:
    """The step-size proposal network on the device (dfn_mlp_*): 8 -> 64 -> 64 -> 1,
    GELU, BF16 weights, FP32 accumulation, Adam training."""

    def __init__(self):
        self._fns = load()
        h = C.c_void_p()
        _abi.check(self._fns["mlp_create"](C.byref(h)), "dfn_mlp_create")
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            self._fns["mlp_destroy"](self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, norm=None):
        p = np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        _abi.check(self._fns["mlp_set_params"](self._h, _abi.fptr(p), _abi.fptr(nm)), "dfn_mlp_set_params")

    def get_params(self):
        p = np.zeros(_abi.MLP_PARAMS, np.float32)
        nm = np.zeros(2 * _abi.MLP_INPUTS + 2, np.float32)
        _abi.check(self._fns["mlp_get_params"](self._h, _abi.fptr(p), _abi.fptr(nm)), "dfn_mlp_get_params")
        return p, nm

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros(X.shape[0], np.float32)
        _abi.check(self._fns["mlp_forward"](self._h, _abi.fptr(X), X.shape[0], _abi.fptr(Y)), "dfn_mlp_forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(_abi.MLP_PARAMS, np.float32)
        loss = C.c_double()
        _abi.check(self._fns["mlp_loss_grad"](self._h, _abi.fptr(X), _abi.fptr(T), X.shape[0], C.byref(loss),
                                              _abi.fptr(g)), "dfn_mlp_loss_grad")
        return loss.value, g

    def train(self, X, T, epochs=10, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = C.c_double()
        _abi.check(self._fns["mlp_train"](self._h, _abi.fptr(X), _abi.fptr(T), X.shape[0], int(epochs), int(batch),
                                          float(lr), int(seed), C.byref(loss)), "dfn_mlp_train")
        return loss.value
