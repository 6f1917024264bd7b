"""ddh_chi2_device, the entry point that allocates nothing, against ddh_chi2:
two calls in a row on one GPU, the second on a smaller problem (so the fixed
scratch array holds the first call's partial sums when it starts), for the
diagonal and the full-Jones solver. Device memory for the inputs and the
output comes from the HIP runtime the library links (libamdhip64.so.7, so the
test and the library share one runtime); the library call itself allocates
nothing."""
import ctypes as C

import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from dp3_amd import warmstart as WS

pytestmark = pytest.mark.gpu


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_rt():
    H.lib()  # load the library first: the runtime below is the one it links
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DeviceBuffer:
    def __init__(self, host):
        self.rt = hip_rt()
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.rt.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.rt.hipFree(self.ptr) == 0


def device_chi2(solver, dp, solutions):
    sol = DeviceBuffer(np.ascontiguousarray(solutions).view(np.float64).ravel())
    out = DeviceBuffer(np.full(1, np.nan))
    rc = H.lib().ddh_chi2_device(solver.handle, dp.handle, sol.ptr, out.ptr, None)
    assert rc == 0, H.last_error()
    assert sol.rt.hipDeviceSynchronize() == 0
    value = float(out.read()[0])
    sol.free()
    out.free()
    return value


def interval(w, seed, **kw):
    q = WS.IntervalSequence(w, 1, seed=seed, solver_kwargs=kw)
    g = q.unit()
    q.solver.solve_problem(q.problems[0], g)
    return q, g


def test_chi2_device_equals_chi2_two_calls_second_smaller():
    print("case: no-alloc chi2, 62 then 12 stations", flush=True)
    big = S.Workload("noalloc-big", 62, 5, 2, 1, 48, 6, 4, 20, 1e-5)
    small = S.Workload("noalloc-small", 12, 3, 1, 1, 16, 2, 3, 20, 1e-5)
    for w, seed in ((big, 3), (small, 4)):
        q, g = interval(w, seed)
        ref = H.chi2(q.solver, q.problems[0], g)
        got = device_chi2(q.solver, q.problems[0], g)
        print(f"{w.name}: chi2 {got:.10g} (ddh_chi2 {ref:.10g})")
        assert got == ref
        # unit gains give another value through the same scratch
        u = q.unit()
        assert device_chi2(q.solver, q.problems[0], u) == H.chi2(q.solver, q.problems[0], u)
        q.release(0)


def test_chi2_device_full_jones_and_scalar_rejected():
    print("case: no-alloc chi2, full-Jones; scalar rejected", flush=True)
    w = S.Workload("noalloc-fj", 10, 2, 1, 1, 12, 2, 2, 20, 1e-5)
    rng = np.random.default_rng(5)
    n_bl = w.n_baselines
    a1 = np.array([a for a in range(w.n_stations) for b in range(a + 1, w.n_stations)], np.int32)
    a2 = np.array([b for a in range(w.n_stations) for b in range(a + 1, w.n_stations)], np.int32)
    ch = np.diff(S.channel_begin(w.n_channels, w.n_channel_blocks)).astype(np.int32)
    sol_index = np.zeros((w.n_directions, w.solint), np.uint32)
    sol_index[1] = 1
    fj = H.Solver(w.n_stations, [1, 1], w.n_channel_blocks, n_correlations=4,
                  n_solution_polarizations=4)
    dp = H.DeviceProblem(fj, w.solint, a1, a2, ch, sol_index)
    shape = (w.solint, n_bl, w.n_channels, 4)
    data = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    model = (rng.normal(size=(2,) + shape) + 1j * rng.normal(size=(2,) + shape)).astype(np.complex64)
    dp.assign_and_weight(data, model, a1, a2)
    g = (np.eye(2)[None, None] + 0.1 * (rng.normal(size=(w.n_channel_blocks, w.n_stations * 2, 2, 2))
                                        + 1j * rng.normal(size=(w.n_channel_blocks,
                                                                w.n_stations * 2, 2, 2))))
    g = np.ascontiguousarray(g.reshape(w.n_channel_blocks, -1), dtype=np.complex128)
    ref = H.chi2(fj, dp, g)
    got = device_chi2(fj, dp, g)
    print(f"full-Jones: chi2 {got:.10g} (ddh_chi2 {ref:.10g})")
    assert got == ref
    sc = H.Solver(w.n_stations, [1, 1], w.n_channel_blocks, n_solution_polarizations=1)
    dps = H.DeviceProblem(sc, w.solint, a1, a2, ch, sol_index)
    buf = DeviceBuffer(np.zeros(64))
    rc = H.lib().ddh_chi2_device(sc.handle, dps.handle, buf.ptr, buf.ptr, None)
    assert rc != 0 and "scalar" in H.last_error()
    buf.free()
    for x in (dp, dps, fj, sc):
        x.close()
