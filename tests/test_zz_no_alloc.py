"""dfn_kernel_noalloc, the dense stress evaluation that allocates nothing (its
partial sums sit in a fixed device array per GPU), against dfn_kernel on the same
handle and slip: bit for bit, on successive calls whose sizes shrink, so that the
second call runs on a scratch array still holding the first call's partials."""
import ctypes as C

import numpy as np
import pytest

import cases
from demystification_dfn_amd import _abi, problems

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """The HIP runtime the device library already loaded (device memory for the test)."""

    def __init__(self):
        from demystification_dfn_amd import device
        device.load()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), 8 * max(n, 1)) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = self.alloc(a.size)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), 8 * a.size, _H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8 * n, _D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc_matches_kernel(hip, p, seed):
    from demystification_dfn_amd.device import DeviceSimulation
    print(f"[gpu] noalloc {p.name} N={p.N} mode={p.fracture_mode}", flush=True)
    d = DeviceSimulation(p)
    N = p.N
    slip = 10.0 ** np.random.default_rng(seed).uniform(-12, 0, N)
    refT, refN = d.kernel(slip)
    ds, dT, dN = hip.put(slip), hip.alloc(N), hip.alloc(N)
    fn = d._fns["kernel_noalloc"]
    _abi.check(fn(d._h, ds, dT, dN if p.fracture_mode == 2 else None, None), "dfn_kernel_noalloc")
    np.testing.assert_array_equal(hip.get(dT, N), refT)
    if p.fracture_mode == 2:
        np.testing.assert_array_equal(hip.get(dN, N), refN)
    d.close()


def test_noalloc_kernel_matches_dfn_kernel_on_shrinking_inputs():
    hip = _Hip()
    try:
        # 4000-element single fault (Mode III), then a 1029-element Mode II network,
        # then a 300-element Mode III fault: each call smaller than the one before
        _noalloc_matches_kernel(hip, problems.single_fault(n=4000), 0)
        _noalloc_matches_kernel(hip, cases.reference_network(), 1)
        _noalloc_matches_kernel(hip, cases.single_fault_small(), 2)
    finally:
        hip.free()


def test_noalloc_kernel_rejects_what_it_cannot_hold():
    from demystification_dfn_amd.device import DeviceSimulation
    hip = _Hip()
    try:
        # Mode II needs 2 * ceil((N+1)/64) * N doubles: N = 8192 is one too many
        big = problems.single_fault(n=8192, length=81920.0, mode=_abi.MODE_II, name="noalloc_too_large")
        print(f"[gpu] noalloc size check N={big.N}", flush=True)
        d = DeviceSimulation(big)
        ds, dT, dN = hip.alloc(big.N), hip.alloc(big.N), hip.alloc(big.N)
        assert d._fns["kernel_noalloc"](d._h, ds, dT, dN, None) == _abi.DFN_ERR_ARG
        d.close()
        # H-matrix handles have no no-allocation path
        hm = problems.network_small(n_faults=6, seed=2, static_kernel=_abi.KERNEL_HMATRIX)
        d = DeviceSimulation(hm)
        assert d._fns["kernel_noalloc"](d._h, ds, dT, dN, None) == _abi.DFN_ERR_UNSUPPORTED
        d.close()
        # the largest Mode II size still runs, and matches dfn_kernel
        ok = problems.single_fault(n=8191, length=81910.0, mode=_abi.MODE_II, name="noalloc_largest")
        _noalloc_matches_kernel(hip, ok, 3)
    finally:
        hip.free()
