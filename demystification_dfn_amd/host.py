"""Host-side mirror of the reference's hot-path interfaces over the C ABI.

Method names follow the reference: ``kernel`` (the kernel procedure pointer,
src/variables.f90:17-26), ``friction_law`` (src/variables.f90:9-14),
``compute_P`` (src/fluid_diffusion.f90:24), ``rkqs`` (src/ode_solver/ode_bs.f90:125),
``Hgemv_right`` (src/hmatrix/calculation_hmat.f90:3), ``initiate_hmat``
(src/hmatrix/initialisation_hmat.f90:13) and ``solve`` (src/solver.f90:42).
Errors the reference signals with ``ier`` stay in the return value; errors it
handles with STOP raise :class:`DfnError`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import check, dptr, i32ptr, i64ptr


class Simulation:
    """One problem on one library (device library or, in tests, the CPU oracle)."""

    def __init__(self, fns, problem, tag="dfn"):
        self._fns = fns
        self.tag = tag
        self.problem = problem
        self.N = problem.N
        self.mode = problem.fracture_mode
        cp, self._keep = problem.to_c()
        h = C.c_void_p()
        rc = fns["create"](C.byref(cp), C.byref(h))
        self._h = h
        if rc != _abi.DFN_OK:
            if h.value:
                fns["destroy"](h)
                self._h = C.c_void_p()
            check(rc, f"{tag}_create")

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if self._h is not None and self._h.value:
            self._fns["destroy"](self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def initiate_hmat(self):
        check(self._fns["build_hmat"](self._h), f"{self.tag}_build_hmat")

    def set_precision(self, prec):
        check(self._fns["set_precision"](self._h, prec), f"{self.tag}_set_precision")

    def hmat_stats(self, op=_abi.OP_T):
        s = np.zeros(10, dtype=np.int64)
        check(self._fns["hmat_stats"](self._h, op, i64ptr(s)), f"{self.tag}_hmat_stats")
        keys = ("nb_dense", "nb_lowrank", "S_dense", "S_lowrank_U", "S_lowrank_V", "max_rank",
                "nbclus", "nbblclus", "n", "sum_rank")
        return dict(zip(keys, (int(v) for v in s)))

    def hmat_export(self, op=_abi.OP_T):
        st = self.hmat_stats(op)
        nd, nl, n = st["nb_dense"], st["nb_lowrank"], st["n"]
        d = dict(n=n, perm=np.zeros(n, np.int32),
                 dense_row0=np.zeros(nd, np.int32), dense_col0=np.zeros(nd, np.int32),
                 dense_m=np.zeros(nd, np.int32), dense_n=np.zeros(nd, np.int32),
                 dense_off=np.zeros(nd, np.int64), dense_data=np.zeros(st["S_dense"]),
                 lr_row0=np.zeros(nl, np.int32), lr_col0=np.zeros(nl, np.int32),
                 lr_m=np.zeros(nl, np.int32), lr_n=np.zeros(nl, np.int32), lr_k=np.zeros(nl, np.int32),
                 lr_uoff=np.zeros(nl, np.int64), lr_voff=np.zeros(nl, np.int64),
                 lr_u=np.zeros(st["S_lowrank_U"]), lr_v=np.zeros(st["S_lowrank_V"]))
        check(self._fns["hmat_export"](
            self._h, op, i32ptr(d["perm"]), i32ptr(d["dense_row0"]), i32ptr(d["dense_col0"]),
            i32ptr(d["dense_m"]), i32ptr(d["dense_n"]), i64ptr(d["dense_off"]), dptr(d["dense_data"]),
            i32ptr(d["lr_row0"]), i32ptr(d["lr_col0"]), i32ptr(d["lr_m"]), i32ptr(d["lr_n"]),
            i32ptr(d["lr_k"]), i64ptr(d["lr_uoff"]), i64ptr(d["lr_voff"]), dptr(d["lr_u"]),
            dptr(d["lr_v"])), f"{self.tag}_hmat_export")
        return d

    def hmat_import(self, op, d):
        check(self._fns["hmat_import"](
            self._h, op, int(d["n"]), i32ptr(d["perm"]), len(d["dense_m"]), i32ptr(d["dense_row0"]),
            i32ptr(d["dense_col0"]), i32ptr(d["dense_m"]), i32ptr(d["dense_n"]), i64ptr(d["dense_off"]),
            dptr(d["dense_data"]), len(d["lr_m"]), i32ptr(d["lr_row0"]), i32ptr(d["lr_col0"]),
            i32ptr(d["lr_m"]), i32ptr(d["lr_n"]), i32ptr(d["lr_k"]), i64ptr(d["lr_uoff"]),
            i64ptr(d["lr_voff"]), dptr(d["lr_u"]), dptr(d["lr_v"])), f"{self.tag}_hmat_import")

    # -- state --------------------------------------------------------------
    def set_state(self, Y=None, P=None, time=0.0):
        Y = None if Y is None else np.ascontiguousarray(Y, np.float64)
        P = None if P is None else np.ascontiguousarray(P, np.float64)
        check(self._fns["set_state"](self._h, dptr(Y), dptr(P), float(time)), f"{self.tag}_set_state")

    def get_state(self):
        Y = np.zeros(3 * self.N)
        P = np.zeros(self.N)
        Pd = np.zeros(self.N)
        t = C.c_double()
        check(self._fns["get_state"](self._h, dptr(Y), dptr(P), dptr(Pd), C.byref(t)), f"{self.tag}_get_state")
        return Y, P, Pd, t.value

    def stats(self):
        s = _abi.DfnRunStats()
        check(self._fns["get_stats"](self._h, C.byref(s)), f"{self.tag}_get_stats")
        return s.as_dict()

    # -- hot path -------------------------------------------------------------
    def kernel(self, slip):
        slip = np.ascontiguousarray(slip, np.float64)
        sT = np.zeros(self.N)
        sN = np.zeros(self.N)
        check(self._fns["kernel"](self._h, dptr(slip), dptr(sT), dptr(sN)), f"{self.tag}_kernel")
        return sT, sN

    def Hgemv_right(self, x, op=_abi.OP_T, alpha=1.0, beta=0.0, y=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.N) if y is None else np.array(y, np.float64, copy=True)
        check(self._fns["hgemv"](self._h, op, alpha, beta, dptr(x), dptr(y)), f"{self.tag}_hgemv")
        return y

    def friction_law(self, x, Y):
        Y = np.ascontiguousarray(Y, np.float64)
        Yd = np.zeros(3 * self.N)
        check(self._fns["friction_law"](self._h, float(x), dptr(Y), dptr(Yd)), f"{self.tag}_friction_law")
        return Yd

    def compute_P(self, dt):
        Pt = np.zeros(self.N)
        Pd = np.zeros(self.N)
        ier = C.c_int32(0)
        check(self._fns["compute_P"](self._h, float(dt), dptr(Pt), dptr(Pd), C.byref(ier)), f"{self.tag}_compute_P")
        return Pt, Pd, int(ier.value)

    def rkqs(self, y, dydx, x, htry, eps, yscal):
        y = np.array(y, np.float64, copy=True)
        dydx = np.ascontiguousarray(dydx, np.float64)
        yscal = np.ascontiguousarray(yscal, np.float64)
        xx = C.c_double(x)
        hdid = C.c_double()
        hnext = C.c_double()
        check(self._fns["rkqs"](self._h, dptr(y), dptr(dydx), C.byref(xx), float(htry), float(eps),
                                dptr(yscal), C.byref(hdid), C.byref(hnext)), f"{self.tag}_rkqs")
        return y, xx.value, hdid.value, hnext.value

    def solve(self, nsteps, t_final=0.0):
        s = _abi.DfnRunStats()
        rc = self._fns["solve"](self._h, int(nsteps), float(t_final), C.byref(s))
        check(rc, f"{self.tag}_solve")
        return s.as_dict()
