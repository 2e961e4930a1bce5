"""Problems from the reference's HDF5 configuration files.

`read_config` reads a ``problems/<name>/config`` file with the library's HDF5
reader (dfn_h5_*, no libhdf5) and returns a :class:`Problem` with the content
src/configuration.f90:70-379 reads: hyperparameters, material and loading, and per
fault its element count, nodes, friction, state, pressure, permeability data and
injection sources.  The H-matrix parameters are not in the file; the reference
takes them from src/variables.f90:122-124 (nleaf 100, eta 3, eps_ACA 1e-8), the
Problem defaults.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from .problems import Problem, concatenate_faults

PER_FAULT = ("a", "b", "Dc", "f0", "V0", "normal_loading_dot", "shear_loading_dot", "V", "theta", "sigmaN", "P",
             "permeability", "kmax", "kmin", "Tk", "Lk", "Snk")
MATERIAL = ("Sigma11_dot", "Sigma12_dot", "Sigma22_dot", "Sigma31_dot", "Sigma32_dot", "cp", "cs", "mu", "porosity",
            "fluid_comp", "rock_comp", "fluid_density", "dyn_viscosity")


class H5File:
    """Read-only view of an HDF5 file through the library's reader."""

    def __init__(self, path):
        from .device import load
        self._f = load()
        h = C.c_void_p()
        _abi.check(self._f["h5_open"](os.fsencode(path), C.byref(h)), f"dfn_h5_open({path})")
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            self._f["h5_close"](self._h)
        self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self, name):
        cls, rank = C.c_int32(), C.c_int32()
        dims = np.zeros(8, np.int64)
        _abi.check(self._f["h5_info"](self._h, name.encode(), C.byref(cls), C.byref(rank), _abi.i64ptr(dims)),
                   f"dfn_h5_info({name})")
        return cls.value, tuple(int(x) for x in dims[:rank.value])

    def has(self, name):
        cls, rank = C.c_int32(), C.c_int32()
        return self._f["h5_info"](self._h, name.encode(), C.byref(cls), C.byref(rank), None) == _abi.DFN_OK

    def f64(self, name):
        _, shape = self.info(name)
        n = int(np.prod(shape)) if shape else 1
        a = np.zeros(n)
        _abi.check(self._f["h5_read_f64"](self._h, name.encode(), _abi.dptr(a), n), f"dfn_h5_read_f64({name})")
        return a.reshape(shape) if shape else a[0]

    def i64(self, name):
        _, shape = self.info(name)
        n = int(np.prod(shape)) if shape else 1
        a = np.zeros(n, np.int64)
        _abi.check(self._f["h5_read_i64"](self._h, name.encode(), _abi.i64ptr(a), n), f"dfn_h5_read_i64({name})")
        return a.reshape(shape) if shape else int(a[0])

    def string(self, name):
        buf = C.create_string_buffer(4096)
        _abi.check(self._f["h5_read_string"](self._h, name.encode(), buf, 4096), f"dfn_h5_read_string({name})")
        return buf.value.decode()


def write_h5(path, datasets):
    """Writes {name: FP64 array of rank 1 or 2} as the root group's datasets of a
    new HDF5 file (dfn_h5_write_f64, at most 8 datasets)."""
    from .device import load
    f = load()
    names = list(datasets)
    arrs = [np.ascontiguousarray(datasets[k], np.float64) for k in names]
    n = len(names)
    cnames = (C.c_char_p * n)(*[k.encode() for k in names])
    ranks = np.array([a.ndim for a in arrs], np.int32)
    dims = np.zeros(2 * n, np.int64)
    for i, a in enumerate(arrs):
        k = min(a.ndim, 2)  # the library rejects other ranks
        dims[2 * i:2 * i + k] = a.shape[:k]
    ptrs = (C.POINTER(C.c_double) * n)(*[_abi.dptr(a) for a in arrs])
    _abi.check(f["h5_write_f64"](os.fsencode(path), n, cnames, _abi.i32ptr(ranks), _abi.i64ptr(dims), ptrs),
               f"dfn_h5_write_f64({path})")


OUTPUT_FIELDS = ("V", "theta", "P", "tractionTel", "tractionNel", "permeability", "time")


def read_output(path):
    """{name: array} of a solver output file outputNNNNNNN (src/solver.f90:227-317):
    the element fields have shape (N, saves), time has shape (saves,)."""
    with H5File(path) as f:
        return {k: f.f64(k) for k in OUTPUT_FIELDS}


def read_config(path, **hmatrix):
    """Problem from a reference configuration file; hmatrix may override eta,
    eps_aca, nleaf and max_rank."""
    with H5File(path) as f:
        nb_fault = f.i64("hyperparameter/nb_fault")
        mode = f.string("hyperparameter/fracture_mode")
        kernel = f.string("hyperparameter/static_kernel")
        faults, sources = [], []
        for i in range(1, nb_fault + 1):
            g = f"fault{i}/"
            nel = f.i64(g + "nb_element")
            node = np.asarray(f.f64(g + "node"), np.float64).reshape(nel + 1, 2)
            fd = dict(node=node)
            for nm in PER_FAULT:
                fd[nm] = np.asarray(f.f64(g + nm), np.float64).reshape(nel)
            faults.append(fd)
            ns = f.i64(g + "nb_source")
            if ns > 0:
                q = np.atleast_1d(f.f64(g + "Q"))
                tb = np.atleast_1d(f.f64(g + "t_injection_beg"))
                te = np.atleast_1d(f.f64(g + "t_injection_end"))
                ix = np.atleast_1d(f.i64(g + "index_injection"))
                for s in range(ns):
                    sources.append((i, int(ix[s]), float(q[s]), float(tb[s]), float(te[s])))
        nels, node, arrays = concatenate_faults(faults)
        mat = {nm: float(f.f64("material_loading/" + nm)) for nm in MATERIAL}
        prob = Problem(name=f.string("hyperparameter/simulation_name"),
                       fracture_mode=_abi.MODE_II if mode.strip().lower() == "modeii" else _abi.MODE_III,
                       static_kernel=_abi.KERNEL_HMATRIX if kernel.strip().lower() == "hmatrix"
                       else _abi.KERNEL_DIRECT,
                       fault_nb_element=nels, node=node, arrays=arrays, tol_solver=float(f.f64("hyperparameter/tol_solver")),
                       permeability_coupling=int(f.i64("hyperparameter/permeability_coupling")), sources=sources,
                       meta=dict(max_it=f.i64("hyperparameter/max_it"), stride_time=f.i64("hyperparameter/stride_time"),
                                 friction_law=f.string("hyperparameter/friction_law"), config=str(path)),
                       **mat, **hmatrix)
    return prob
