"""ORACLE -- test infrastructure only.

ctypes loader for the CPU restatement (oracle/liboracle.so, prefix dfno_) and,
when it has been built, the reference's own Fortran hot-path modules
(oracle/_ref/libdfn_ref.so).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use this module, and only as the checker.
"""
import ctypes as C
import os

from demystification_dfn_amd import _abi
from demystification_dfn_amd.host import Simulation

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdfn_ref.so")

_lib = None
_fns = None


def load():
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIB)
        _fns = _abi.bind(_lib, "dfno_")
        _lib.dfno_handle_set_threads.argtypes = [C.c_void_p, C.c_int]
        _lib.dfno_handle_set_threads.restype = C.c_int
    return _fns


class OracleSimulation(Simulation):
    def __init__(self, problem, threads=1):
        super().__init__(load(), problem, tag="dfno")
        _lib.dfno_handle_set_threads(self._h, int(threads))

    def set_threads(self, n):
        _lib.dfno_handle_set_threads(self._h, int(n))

    def set_cg_variant(self, extra=0, reverse=False):
        """Test hook: every pressure solve runs `extra` CG iterations past its
        stopping test (src/fluid_diffusion.f90:97-131) and, with `reverse`, sums
        its dot products from the last element; (0, False) is the reference."""
        _lib.dfno_debug_set_cg_variant.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.dfno_debug_set_cg_variant.restype = C.c_int
        assert _lib.dfno_debug_set_cg_variant(self._h, int(extra), 1 if reverse else 0) == 0

    def set_t_try(self, t_try):
        """Trial step of the solver loop (t_try, src/solver.f90:31)."""
        _lib.dfno_set_t_try.argtypes = [C.c_void_p, C.c_double]
        _lib.dfno_set_t_try.restype = C.c_int
        assert _lib.dfno_set_t_try(self._h, float(t_try)) == 0


# --------------------------------------------------------------------------
# Introspection of the oracle and the reference build
# --------------------------------------------------------------------------
import numpy as np  # noqa: E402

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _bind_debug():
    load()
    _lib.dfno_debug_geometry.argtypes = [C.c_void_p] + [_dp] * 6
    _lib.dfno_debug_crossings.argtypes = [C.c_void_p, _i64p, C.c_int32, _i32p]
    _lib.dfno_eval_matrix.argtypes = [C.c_void_p, C.c_int32, _dp]
    _lib.dfno_debug_stress_rate.argtypes = [C.c_void_p, _dp]
    _lib.dfno_hmat_tree.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _i32p, C.c_int32, _i32p, _i32p,
                                    _i64p, C.c_int32]


def geometry(sim):
    _bind_debug()
    N = sim.N
    out = dict(elem=np.zeros(2 * N), ds=np.zeros(N), tang=np.zeros(2 * N), nrm=np.zeros(2 * N),
               sld=np.zeros(N), nld=np.zeros(N))
    rc = _lib.dfno_debug_geometry(sim._h, *(_p(out[k]) for k in ("elem", "ds", "tang", "nrm", "sld", "nld")))
    assert rc == 0
    return out


def stress_rate(sim):
    """The shear stress rate of the oracle's last friction-law evaluation."""
    _bind_debug()
    out = np.zeros(sim.N)
    assert _lib.dfno_debug_stress_rate(sim._h, _p(out)) == 0
    return out


def crossings(sim):
    _bind_debug()
    n = C.c_int32()
    _lib.dfno_debug_crossings(sim._h, None, 0, C.byref(n))
    cr = np.zeros(6 * max(n.value, 1), np.int64)
    _lib.dfno_debug_crossings(sim._h, _p(cr, _i64p), n.value, C.byref(n))
    return cr[:6 * n.value].reshape(n.value, 6)


def eval_matrix(sim, op=0):
    _bind_debug()
    M = np.zeros(sim.N * sim.N)
    assert _lib.dfno_eval_matrix(sim._h, op, _p(M)) == 0
    return M.reshape(sim.N, sim.N, order="F")


def hmat_tree(sim, op=0):
    _bind_debug()
    N = sim.N
    mc = np.zeros(N, np.int32)
    nbc, nbb = C.c_int32(), C.c_int32()
    rc = _lib.dfno_hmat_tree(sim._h, op, _p(mc, _i32p), C.byref(nbc), None, 0, C.byref(nbb), None, None, 0)
    assert rc == 0, rc
    ind = np.zeros(2 * nbc.value, np.int32)
    blt = np.zeros(3 * nbb.value, np.int32)
    inf = np.zeros(6 * nbb.value, np.int64)
    _lib.dfno_hmat_tree(sim._h, op, _p(mc, _i32p), C.byref(nbc), _p(ind, _i32p), nbc.value, C.byref(nbb),
                        _p(blt, _i32p), _p(inf, _i64p), nbb.value)
    return dict(mat_corr=mc, ind_clusters=ind, bltree=blt.reshape(-1, 3), info=inf.reshape(-1, 6))


class RefSimulation:
    """The reference's own Fortran modules (oracle/_ref/libdfn_ref.so) on a problem.

    The reference keeps its state in module variables, so only one instance is
    live at a time.  Geometry, loading projections and crossings are taken from
    the oracle (they are pinned by their own closed-form tests)."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_LIB)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = C.CDLL(REF_LIB)
            lib.ref_init.argtypes = ([C.c_int, C.c_int, _i64p] + [_dp] * 17 + [_dp, _i32p, C.c_int, _i64p, _i64p] +
                                     [_dp] * 3 + [C.c_int, _i64p])
            lib.ref_set_state.argtypes = [_dp, _dp, C.c_double]
            lib.ref_get_state.argtypes = [_dp, _dp, _dp, _dp]
            lib.ref_kernel.argtypes = [_dp, _dp, _dp]
            lib.ref_eval_matrix.argtypes = [C.c_int, C.c_int, _dp]
            lib.ref_friction_law.argtypes = [C.c_double, _dp, _dp]
            lib.ref_compute_P.argtypes = [C.c_double, _dp, _dp, _i32p]
            lib.ref_rkqs.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
            lib.ref_solve.argtypes = [C.c_int64, C.c_double, _i64p]
            lib.ref_set_t_try.argtypes = [C.c_double]
            lib.ref_trees.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int, _i32p, _i32p, _i32p,
                                      _i32p, _i32p, _i64p]
            cls._lib = lib
        return cls._lib

    def __init__(self, problem):
        lib = self.lib()
        osim = OracleSimulation(problem)
        g = geometry(osim)
        cr = crossings(osim)
        Y, P, _, _ = osim.get_state()
        osim.close()
        self.N = N = problem.N
        self.problem = problem
        A = problem.arrays
        keep = []

        def arr(a, dt=np.float64):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        src = np.array(problem.sources, np.float64).reshape(-1, 5)
        scal = arr([problem.mu, problem.cs, problem.cp, problem.porosity, problem.fluid_comp, problem.rock_comp,
                    problem.dyn_viscosity, problem.tol_solver])
        ints = arr([problem.fracture_mode, problem.static_kernel, problem.permeability_coupling], np.int32)
        crf = arr(cr.reshape(-1) if len(cr) else np.zeros(6), np.int64)
        lib.ref_init(N, problem.nb_fault, _p(arr(problem.fault_nb_element, np.int64), _i64p),
                     _p(arr(problem.node.reshape(-1))), _p(g["elem"]), _p(g["ds"]), _p(g["tang"]), _p(g["nrm"]),
                     _p(arr(A["a"])), _p(arr(A["b"])), _p(arr(A["Dc"])), _p(arr(A["f0"])), _p(arr(A["V0"])),
                     _p(g["nld"]), _p(g["sld"]), _p(arr(A["permeability"])), _p(arr(A["kmin"])),
                     _p(arr(A["Snk"])), _p(arr(Y)), _p(arr(P)), _p(scal), _p(ints, _i32p), len(src),
                     _p(arr(src[:, 0].astype(np.int64), np.int64), _i64p),
                     _p(arr(src[:, 1].astype(np.int64), np.int64), _i64p),
                     _p(arr(src[:, 2])), _p(arr(src[:, 3])), _p(arr(src[:, 4])), len(cr), _p(crf, _i64p))
        self._keep = keep

    def set_state(self, Y, P, time=0.0):
        self.lib().ref_set_state(_p(np.ascontiguousarray(Y, np.float64)), _p(np.ascontiguousarray(P, np.float64)),
                                 float(time))

    def get_state(self):
        Y, P, Pd, t = np.zeros(3 * self.N), np.zeros(self.N), np.zeros(self.N), np.zeros(1)
        self.lib().ref_get_state(_p(Y), _p(P), _p(Pd), _p(t))
        return Y, P, Pd, float(t[0])

    def kernel(self, slip):
        s = np.array(slip, np.float64)
        sT, sN = np.zeros(self.N), np.zeros(self.N)
        self.lib().ref_kernel(_p(s), _p(sT), _p(sN))
        return sT, sN

    def eval_matrix(self, op=0):
        M = np.zeros(self.N * self.N)
        self.lib().ref_eval_matrix(op, 1 if self.problem.fracture_mode == 3 else 0, _p(M))
        return M.reshape(self.N, self.N, order="F")

    def friction_law(self, x, Y):
        Yd = np.zeros(3 * self.N)
        self.lib().ref_friction_law(float(x), _p(np.ascontiguousarray(Y, np.float64)), _p(Yd))
        return Yd

    def compute_P(self, dt):
        Pt, Pd = np.zeros(self.N), np.zeros(self.N)
        ier = C.c_int32()
        self.lib().ref_compute_P(float(dt), _p(Pt), _p(Pd), C.byref(ier))
        return Pt, Pd, int(ier.value)

    def rkqs(self, y, dydx, x, htry, eps, yscal):
        y = np.array(y, np.float64)
        xx = np.array([x], np.float64)
        hd, hn = np.zeros(1), np.zeros(1)
        self.lib().ref_rkqs(_p(y), _p(np.ascontiguousarray(dydx, np.float64)), _p(xx), float(htry), float(eps),
                            _p(np.ascontiguousarray(yscal, np.float64)), _p(hd), _p(hn))
        return y, float(xx[0]), float(hd[0]), float(hn[0])

    def set_t_try(self, t):
        self.lib().ref_set_t_try(float(t))

    def solve(self, nsteps, t_final=0.0):
        acc = C.c_int64()
        self.lib().ref_solve(int(nsteps), float(t_final), C.byref(acc))
        return acc.value

    @classmethod
    def trees(cls, elem, nleaf, eta):
        """Reference cluster + block cluster trees on element midpoints (2,n)."""
        lib = cls.lib()
        n = elem.size // 2
        cn = np.ascontiguousarray(elem, np.float64).copy()
        mc = np.zeros(n, np.int32)
        nbc, nbb = C.c_int32(), C.c_int32()
        cap_cl, cap_bl = 4 * n + 8, 64 * n + 64
        ind = np.zeros(2 * cap_cl, np.int32)
        blt = np.zeros(3 * cap_bl, np.int32)
        inf = np.zeros(6 * cap_bl, np.int64)
        lib.ref_trees(n, _p(cn), int(nleaf), float(eta), cap_cl, cap_bl, _p(mc, _i32p), C.byref(nbc),
                      _p(ind, _i32p), C.byref(nbb), _p(blt, _i32p), _p(inf, _i64p))
        assert nbc.value <= cap_cl and nbb.value <= cap_bl
        return dict(mat_corr=mc, ind_clusters=ind[:2 * nbc.value], bltree=blt[:3 * nbb.value].reshape(-1, 3),
                    info=inf[:6 * nbb.value].reshape(-1, 6))


def compress(A, B, eps):
    """compression_low_rank on factors A (m x k), B (n x k): (A', B', rank)."""
    load()
    fn = _lib.dfno_debug_compress
    fn.argtypes = [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, C.c_double, _i32p]
    m, k = A.shape
    n = B.shape[0]
    a = np.asfortranarray(A, np.float64).copy(order="F")
    b = np.asfortranarray(B, np.float64).copy(order="F")
    r = C.c_int32()
    assert fn(m, n, k, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), float(eps), C.byref(r)) == 0
    return a[:, :r.value], b[:, :r.value], r.value


def full_aca(M, eps, max_rank):
    """The full-pivot ACA of compute_full_ACA on the block M: (A, B, rank)."""
    load()
    fn = _lib.dfno_debug_full_aca
    fn.argtypes = [C.c_int32, C.c_int32, _dp, C.c_double, C.c_int32, _dp, _dp, _i32p]
    m, n = M.shape
    R = np.asfortranarray(M, np.float64).copy(order="F")
    A = np.zeros((m, max_rank), order="F")
    B = np.zeros((n, max_rank), order="F")
    r = C.c_int32()
    assert fn(m, n, R.ctypes.data_as(_dp), float(eps), int(max_rank), A.ctypes.data_as(_dp), B.ctypes.data_as(_dp),
              C.byref(r)) == 0
    return A[:, :r.value], B[:, :r.value], r.value
