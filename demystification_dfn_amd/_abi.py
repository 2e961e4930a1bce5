"""ctypes mirror of include/dfn.h.

The same declarations are bound against two libraries: the device library
(``libdfn.so``, prefix ``dfn_``) and, in the tests only, the CPU oracle
(prefix ``dfno_``).  Nothing here computes anything.
"""
import ctypes as C

import numpy as np

DFN_OK = 0
DFN_ERR_ARG = 1
DFN_ERR_HIP = 2
DFN_ERR_CG = 3
DFN_ERR_UNDERFLOW = 4
DFN_ERR_TENSILE = 5
DFN_ERR_STATE = 6
DFN_ERR_NOMEM = 7
DFN_ERR_UNSUPPORTED = 8
DFN_ERR_CG_TIMEOUT = 9

ERROR_NAMES = {
    DFN_OK: "ok",
    DFN_ERR_ARG: "bad argument",
    DFN_ERR_HIP: "HIP runtime error",
    DFN_ERR_CG: "compute_P did not converge",
    DFN_ERR_UNDERFLOW: "stepsize underflow in rkqs",
    DFN_ERR_TENSILE: "normal traction became extensional",
    DFN_ERR_STATE: "call out of order",
    DFN_ERR_NOMEM: "allocation failed",
    DFN_ERR_UNSUPPORTED: "unsupported option",
    DFN_ERR_CG_TIMEOUT: "compute_P workgroups lost contact",
}

MODE_II, MODE_III = 2, 3
KERNEL_DIRECT, KERNEL_HMATRIX = 0, 1
OP_T, OP_N = 0, 1
PREC_FP64, PREC_FP32, PREC_MIXED, PREC_FP32_STORE, PREC_MIXED_SPLIT = 0, 1, 2, 3, 4
STEP_RKQS, STEP_PI, STEP_LEARNED = 0, 1, 2

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class DfnProblem(C.Structure):
    _fields_ = [
        ("fracture_mode", C.c_int32),
        ("static_kernel", C.c_int32),
        ("nb_fault", C.c_int32),
        ("nb_element", C.c_int32),
        ("fault_nb_element", _i64p),
        ("node", _dp),
        ("a", _dp), ("b", _dp), ("Dc", _dp), ("f0", _dp), ("V0", _dp),
        ("normal_loading_dot", _dp), ("shear_loading_dot", _dp),
        ("V", _dp), ("theta", _dp), ("sigmaN", _dp), ("P", _dp),
        ("permeability", _dp), ("kmin", _dp), ("kmax", _dp), ("Snk", _dp), ("Lk", _dp), ("Tk", _dp),
        ("mu", C.c_double), ("cs", C.c_double), ("cp", C.c_double),
        ("Sigma11_dot", C.c_double), ("Sigma12_dot", C.c_double), ("Sigma22_dot", C.c_double),
        ("Sigma31_dot", C.c_double), ("Sigma32_dot", C.c_double),
        ("porosity", C.c_double), ("fluid_comp", C.c_double), ("rock_comp", C.c_double),
        ("dyn_viscosity", C.c_double), ("fluid_density", C.c_double),
        ("tol_solver", C.c_double),
        ("permeability_coupling", C.c_int32),
        ("nb_source", C.c_int32),
        ("source_fault", _i64p), ("source_index", _i64p),
        ("source_Q", _dp), ("source_t_beg", _dp), ("source_t_end", _dp),
        ("eta", C.c_double), ("eps_aca", C.c_double),
        ("nleaf", C.c_int32), ("max_rank", C.c_int32),
    ]


class DfnRunStats(C.Structure):
    _fields_ = [
        ("accepted_steps", C.c_int64),
        ("rejected_attempts", C.c_int64),
        ("rhs_evaluations", C.c_int64),
        ("cg_solves", C.c_int64),
        ("cg_iterations", C.c_int64),
        ("cg_halvings", C.c_int64),
        ("time", C.c_double),
        ("t_try", C.c_double),
        ("max_V", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_H = C.c_void_p
# name -> argtypes (restype is always c_int)
SIGNATURES = {
    "create": [C.POINTER(DfnProblem), C.POINTER(C.c_void_p)],
    "destroy": [_H],
    "build_hmat": [_H],
    "set_precision": [_H, C.c_int32],
    "hmat_stats": [_H, C.c_int32, _i64p],
    "hmat_export": [_H, C.c_int32, _i32p, _i32p, _i32p, _i32p, _i32p, _i64p, _dp,
                    _i32p, _i32p, _i32p, _i32p, _i32p, _i64p, _i64p, _dp, _dp],
    "hmat_import": [_H, C.c_int32, C.c_int32, _i32p, C.c_int32, _i32p, _i32p, _i32p, _i32p, _i64p, _dp,
                    C.c_int32, _i32p, _i32p, _i32p, _i32p, _i32p, _i64p, _i64p, _dp, _dp],
    "set_state": [_H, _dp, _dp, C.c_double],
    "get_state": [_H, _dp, _dp, _dp, C.POINTER(C.c_double)],
    "get_stats": [_H, C.POINTER(DfnRunStats)],
    "kernel": [_H, _dp, _dp, _dp],
    # device pointers slip, stressT, stressN and a hipStream_t (NULL: default stream)
    "kernel_noalloc": [_H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "kernel_multi": [_H, C.c_int32, _dp, _dp, _dp],
    "bench_kernel_multi": [_H, C.c_int32, _dp, C.c_int32, _dp],
    "hgemv": [_H, C.c_int32, C.c_double, C.c_double, _dp, _dp],
    "friction_law": [_H, C.c_double, _dp, _dp],
    "compute_P": [_H, C.c_double, _dp, _dp, _i32p],
    "rkqs": [_H, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp],
    "solve": [_H, C.c_int64, C.c_double, C.POINTER(DfnRunStats)],
    "set_formulation": [_H, C.c_int32],
    "tuning": [_H, _i32p, _dp],
    "kernel_timing": [_H, C.c_int32],
    "kernel_timing_read": [_H, _dp, _i64p],
    "monitor": [_H, C.c_int32],
    "monitor_take": [_H, C.c_int64, _dp, _dp, _i64p],
    "moment": [_H, _dp, _dp],
    "h5_open": [C.c_char_p, C.POINTER(C.c_void_p)],
    "h5_close": [_H],
    "h5_info": [_H, C.c_char_p, _i32p, _i32p, _i64p],
    "h5_read_f64": [_H, C.c_char_p, _dp, C.c_int64],
    "h5_read_i64": [_H, C.c_char_p, _i64p, C.c_int64],
    "h5_read_string": [_H, C.c_char_p, C.c_char_p, C.c_int32],
    "h5_write_f64": [C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), _i32p, _i64p, C.POINTER(_dp)],
    "set_output": [_H, C.c_char_p, C.c_int64, C.c_int64],
    "flush_output": [_H],
    "bench_kernel": [_H, C.c_int32, _dp],
    "bench_stream": [C.c_int64, C.c_int32, _dp],
    "device_info": [C.c_char_p, C.c_int32],
    # step-size proposal network
    "mlp_create": [C.POINTER(C.c_void_p)],
    "mlp_destroy": [_H],
    "mlp_set_params": [_H, _fp, _fp],
    "mlp_get_params": [_H, _fp, _fp],
    "mlp_forward": [_H, _fp, C.c_int64, _fp],
    "mlp_loss_grad": [_H, _fp, _fp, C.c_int64, _dp, _fp],
    "mlp_train": [_H, _fp, _fp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _dp],
    # ensembles of independent networks
    "ensemble_create": [C.POINTER(DfnProblem), C.c_int32, C.POINTER(C.c_void_p)],
    "ensemble_destroy": [_H],
    "ensemble_build_hmat": [_H],
    "ensemble_set_precision": [_H, C.c_int32],
    "ensemble_set_graph": [_H, C.c_int32],
    "ensemble_solve": [_H, C.c_int64],
    "ensemble_solve_until": [_H, C.c_double, C.c_int64],
    "ensemble_get_state": [_H, C.c_int32, _dp, _dp, _dp, C.POINTER(C.c_double)],
    "ensemble_get_stats": [_H, C.c_int32, C.POINTER(DfnRunStats)],
    "ensemble_info": [_H, _i64p],
    "ensemble_kernel_timing": [_H, C.c_int32],
    "ensemble_kernel_timing_read": [_H, _dp, _i64p, _dp],
    "ensemble_set_controller": [_H, C.c_int32, _H],
    "ensemble_set_recording": [_H, C.c_int32],
    "ensemble_take_records": [_H, _fp, _fp, C.c_int64, _i64p],
}

# entry points every build of the device library must export (checked by the CPU tests)
REQUIRED = list(SIGNATURES)


def bind(lib, prefix, names=None):
    """Set argtypes/restype for `prefix + name` on lib; returns {name: fn}."""
    out = {}
    for name in names or SIGNATURES:
        fn = getattr(lib, prefix + name, None)
        if fn is None:
            continue
        fn.argtypes = SIGNATURES[name]
        fn.restype = C.c_int
        out[name] = fn
    return out


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i64p)


def fptr(a):
    """Pointer to a C-contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


FORM_REFERENCE = 0
FORM_NEWTON_V = 1

MLP_INPUTS = 8
MLP_PARAMS = 4801


def i32ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


class DfnError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {ERROR_NAMES.get(code, 'error')} (status {code})")


def check(code, where):
    if code != DFN_OK:
        raise DfnError(code, where)
