"""ctypes bindings of libddecal_hip.so (include/ddecal_hip.h).

The library is the product: there is no CPU fallback. Loading fails loudly
when the shared library has not been built (python -m dp3_amd.build).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddecal_hip.so")

MODEL_C64 = 0
MODEL_F16 = 1
ITERATION_STREAM = 0
ITERATION_GRAM = 1

_LIB = None


class Config(C.Structure):
    _fields_ = [
        ("n_antennas", C.c_int32),
        ("n_directions", C.c_int32),
        ("n_channel_blocks", C.c_int32),
        ("n_correlations", C.c_int32),
        ("n_solutions_per_direction", C.POINTER(C.c_int32)),
        ("max_iterations", C.c_int32),
        ("min_iterations", C.c_int32),
        ("accuracy", C.c_double),
        ("constraint_accuracy", C.c_double),
        ("step_size", C.c_double),
        ("detect_stalling", C.c_int32),
        ("step_diff_sigma", C.c_double),
        ("phase_only", C.c_int32),
        ("model_storage", C.c_int32),
        ("iteration", C.c_int32),
        ("n_solution_polarizations", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("constraint_iterations", C.c_int64),
        ("has_converged", C.c_int32),
        ("n_step_magnitudes", C.c_int32),
        ("last_step_magnitude", C.c_double),
    ]


class CbView(C.Structure):
    _fields_ = [
        ("n_visibilities", C.c_int64),
        ("data", C.c_void_p),
        ("model", C.c_void_p),
        ("antenna_pairs", C.c_void_p),
        ("solution_map", C.c_void_p),
    ]


class Component(C.Structure):
    _fields_ = [
        ("ra", C.c_double), ("dec", C.c_double),
        ("I", C.c_double), ("Q", C.c_double), ("U", C.c_double), ("V", C.c_double),
        ("ref_freq", C.c_double),
        ("n_spectral_terms", C.c_int32),
        ("log_si", C.c_int32),
        ("spectral_terms", C.c_double * 8),
        ("has_rm", C.c_int32),
        ("is_gaussian", C.c_int32),
        ("pol_fraction", C.c_double), ("pol_angle", C.c_double), ("rm", C.c_double),
        ("major_axis", C.c_double), ("minor_axis", C.c_double),
        ("position_angle", C.c_double),
        ("pa_is_absolute", C.c_int32),
        ("direction", C.c_int32),
    ]


def components_from_sky(sky):
    """ddh_component array from a dp3_amd.synth.Sky (Stokes I, flat spectra)."""
    n = len(sky.direction)
    arr = (Component * n)()
    for i in range(n):
        c = arr[i]
        c.ra, c.dec, c.I = float(sky.ra[i]), float(sky.dec[i]), float(sky.flux[i])
        c.is_gaussian = int(sky.is_gaussian[i])
        c.major_axis, c.minor_axis = float(sky.major[i]), float(sky.minor[i])
        c.position_angle = float(sky.pa[i])
        c.direction = int(sky.direction[i])
    return arr


# Every exported symbol and its signature; tests check they are all present.
SIGNATURES = {
    "ddh_last_error": ([], C.c_char_p),
    "ddh_device_count": ([], C.c_int),
    "ddh_config_defaults": ([C.POINTER(Config)], None),
    "ddh_solver_create": ([C.POINTER(Config)], C.c_void_p),
    "ddh_solver_destroy": ([C.c_void_p], None),
    "ddh_solve": ([C.c_void_p, C.POINTER(CbView), C.c_int32, C.c_void_p, C.c_double,
                   C.POINTER(Result)], C.c_int),
    "ddh_problem_create": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p], C.c_void_p),
    "ddh_problem_destroy": ([C.c_void_p], None),
    "ddh_problem_upload": ([C.c_void_p, C.POINTER(CbView)], C.c_int),
    "ddh_problem_baseline_stride": ([C.c_void_p], C.c_int),
    "ddh_problem_data_ptr": ([C.c_void_p, C.c_int32], C.c_void_p),
    "ddh_problem_model_ptr": ([C.c_void_p, C.c_int32, C.c_int32], C.c_void_p),
    "ddh_problem_set_model_scale": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_problem_download_plane": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "ddh_solve_problem": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Result)], C.c_int),
    "ddh_solver_last_timing": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                C.POINTER(C.c_double)], C.c_int),
    "ddh_set_schedule": ([C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32], C.c_int),
    "ddh_set_profiling": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_set_gram_fusion": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_set_gram_solve_groups": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_set_antenna_solve_schedule": ([C.c_void_p, C.c_int32, C.c_int32], C.c_int),
    "ddh_antenna_solve_schedule": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
                                   C.c_int),
    "ddh_batch_set_antenna_solve_schedule": ([C.c_void_p, C.c_int32, C.c_int32], C.c_int),
    "ddh_batch_antenna_solve_schedule": ([C.c_void_p, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)], C.c_int),
    "ddh_batch_set_gram_solve_groups": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_batch_gram_solve_groups": ([C.c_void_p], C.c_int),
    "ddh_predict": ([C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_void_p,
                     C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p],
                    C.c_int),
    "ddh_problem_synthesize_data": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                     C.c_float, C.c_uint32], C.c_int),
    "ddh_calibrate_bandwidth": ([C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)], C.c_int),
    "ddh_batch_create": ([C.c_void_p, C.c_void_p, C.c_int32], C.c_void_p),
    "ddh_batch_destroy": ([C.c_void_p], None),
    "ddh_batch_set_schedule": ([C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32],
                               C.c_int),
    "ddh_batch_solve": ([C.c_void_p, C.c_void_p, C.POINTER(Result)], C.c_int),
    "ddh_batch_rebind": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_solver_set_smoothness": ([C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                   C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
    "ddh_solver_set_constraint_weights": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_solver_set_sub_solution_weights": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_problem_assign_and_weight_device": ([C.c_void_p, C.c_int32] + [C.c_void_p] * 6
                                             + [C.c_int32, C.c_void_p], C.c_int),
    "ddh_problem_assign_and_weight": ([C.c_void_p, C.c_int32] + [C.c_void_p] * 6
                                      + [C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "ddh_correct_and_subtract_device": ([C.c_int32] * 7 + [C.c_void_p] * 6 + [C.c_int32] * 2
                                        + [C.c_void_p], C.c_int),
    "ddh_correct_and_subtract": ([C.c_int32] * 7 + [C.c_void_p] * 6 + [C.c_int32] * 2
                                 + [C.POINTER(C.c_double)], C.c_int),
    "ddh_batch_solve_warm": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(Result)],
                             C.c_int),
    "ddh_batch_last_warm_time": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "ddh_batch_last_time": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "ddh_batch_set_profiling": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_batch_last_timing": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_int64)],
                              C.c_int),
    "ddh_autotune": ([C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                      C.POINTER(C.c_double)], C.c_int),
    "ddh_batch_autotune": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                            C.POINTER(C.c_double)], C.c_int),
    "ddh_one_step": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ddh_chi2": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "ddh_chi2_device": ([C.c_void_p] * 5, C.c_int),
    "ddh_mlp_create": ([C.c_int32, C.c_int32, C.c_int32, C.c_uint32], C.c_void_p),
    "ddh_mlp_destroy": ([C.c_void_p], None),
    "ddh_mlp_num_params": ([C.c_void_p], C.c_int64),
    "ddh_mlp_get_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_mlp_set_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "ddh_mlp_forward": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "ddh_mlp_set_fused_forward": ([C.c_void_p, C.c_int32], C.c_int),
    "ddh_mlp_loss_and_grads": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                C.POINTER(C.c_float), C.c_void_p], C.c_int),
    "ddh_mlp_train": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                       C.c_float, C.c_uint32, C.POINTER(C.c_float)], C.c_int),
    "ddh_gemm_bf16": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                       C.c_int32, C.c_int32], C.c_int),
    "ddh_warmstart_features": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int32, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "ddh_warmstart_apply": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "ddh_warmstart_features_2x2": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "ddh_warmstart_apply_2x2": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
}

N_WARM_FEATURES = 16
N_WARM_OUTPUTS = 4


def _rows_ptrs(a):
    return (C.c_void_p * a.shape[0])(*[a[i].ctypes.data for i in range(a.shape[0])])


def one_step(solver, dproblem, solutions):
    """Undamped solver update from `solutions` and its denominators (two per
    solution: a diagonal pair's, a scalar solution's summed denominator
    twice, or the diagonal of a 2x2 denominator)."""
    sol = np.ascontiguousarray(solutions, dtype=np.complex128)
    nxt = np.zeros_like(sol)
    den = np.zeros((sol.shape[0], solver.n_antennas * solver.n_sub * 2), np.float64)
    _check(lib().ddh_one_step(solver.handle, dproblem.handle, C.cast(_rows_ptrs(sol), C.c_void_p),
                              C.cast(_rows_ptrs(nxt), C.c_void_p),
                              C.cast(_rows_ptrs(den), C.c_void_p)), "ddh_one_step")
    return nxt, den


def chi2(solver, dproblem, solutions):
    sol = np.ascontiguousarray(solutions, dtype=np.complex128)
    out = C.c_double()
    _check(lib().ddh_chi2(solver.handle, dproblem.handle, C.cast(_rows_ptrs(sol), C.c_void_p),
                          C.byref(out)), "ddh_chi2")
    return out.value


def gemm_bf16(a, b, trans_a=False, trans_b=False):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    M = a.shape[1] if trans_a else a.shape[0]
    K = a.shape[0] if trans_a else a.shape[1]
    N = b.shape[0] if trans_b else b.shape[1]
    c = np.zeros((M, N), np.float32)
    _check(lib().ddh_gemm_bf16(_ptr(a), _ptr(b), _ptr(c), M, N, K, int(trans_a), int(trans_b)),
           "ddh_gemm_bf16")
    return c


def warmstart_features(g0, next1, ga, gb, den, n_cb, n_ant, n_sub, full_jones=False):
    """Features per solution; full_jones: 2x2 solutions (4 values each)."""
    arrs = [np.ascontiguousarray(x, np.complex128) for x in (g0, next1, ga, gb)]
    ne = 4 if full_jones else 2
    for a in arrs:
        if a.size != n_cb * n_ant * n_sub * ne:
            raise ValueError("solution arrays do not match the shape")
    d = np.ascontiguousarray(den, np.float64)
    if d.size != n_cb * n_ant * n_sub * 2:
        raise ValueError("denominators do not match the shape")
    out = np.zeros((n_cb * n_ant * n_sub, N_WARM_FEATURES), np.float32)
    fn = lib().ddh_warmstart_features_2x2 if full_jones else lib().ddh_warmstart_features
    _check(fn(*[_ptr(a) for a in arrs], _ptr(d), n_cb, n_ant, n_sub, _ptr(out)),
           "ddh_warmstart_features")
    return out


def warmstart_apply(prev, pred, full_jones=False):
    """prev * exp(pred) per solution (full_jones: row p of each 2x2 matrix
    times exp(pred_p))."""
    p = np.ascontiguousarray(prev, np.complex128)
    q = np.ascontiguousarray(pred, np.float32)
    ne = 4 if full_jones else 2
    if p.size % ne or q.size != p.size // ne * N_WARM_OUTPUTS:
        raise ValueError("predictions do not match the solutions")
    out = np.zeros_like(p)
    fn = lib().ddh_warmstart_apply_2x2 if full_jones else lib().ddh_warmstart_apply
    _check(fn(_ptr(p), _ptr(q), p.size // ne, _ptr(out)), "ddh_warmstart_apply")
    return out


class MLP:
    """3-layer perceptron on the GPU (bf16 MFMA GEMMs, FP32 accumulation)."""

    def __init__(self, n_in, n_hidden, n_out, seed=0):
        h = lib().ddh_mlp_create(n_in, n_hidden, n_out, seed)
        if not h:
            raise HipError(f"ddh_mlp_create failed: {last_error()}")
        self.handle = h
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out

    def num_params(self):
        return int(lib().ddh_mlp_num_params(self.handle))

    def get_params(self):
        p = np.zeros(self.num_params(), np.float32)
        _check(lib().ddh_mlp_get_params(self.handle, _ptr(p)), "ddh_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        _check(lib().ddh_mlp_set_params(self.handle, _ptr(p)), "ddh_mlp_set_params")

    def set_fused_forward(self, enable=True):
        _check(lib().ddh_mlp_set_fused_forward(self.handle, int(enable)),
               "ddh_mlp_set_fused_forward")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        _check(lib().ddh_mlp_forward(self.handle, _ptr(x), x.shape[0], _ptr(y)), "ddh_mlp_forward")
        return y

    def loss_and_grads(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float()
        _check(lib().ddh_mlp_loss_and_grads(self.handle, _ptr(x), _ptr(t), x.shape[0],
                                            C.byref(loss), _ptr(g)), "ddh_mlp_loss_and_grads")
        return loss.value, g

    def train(self, x, t, epochs=5, batch=4096, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        _check(lib().ddh_mlp_train(self.handle, _ptr(x), _ptr(t), x.shape[0], epochs, batch, lr,
                                   seed, C.byref(loss)), "ddh_mlp_train")
        return loss.value

    def close(self):
        if getattr(self, "handle", None):
            lib().ddh_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def calibrate_bandwidth(nbytes=1 << 30, repeats=10):
    """(copy GB/s counting read+write, streaming-read GB/s) of plain kernels."""
    c, r = C.c_double(), C.c_double()
    _check(lib().ddh_calibrate_bandwidth(nbytes, repeats, C.byref(c), C.byref(r)),
           "ddh_calibrate_bandwidth")
    return c.value, r.value


class HipError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"{LIB_PATH} is missing: build it with `python -m dp3_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def last_error():
    msg = lib().ddh_last_error()
    return msg.decode() if msg else ""


def _check(rc, what):
    if rc != 0:
        raise HipError(f"{what} failed ({rc}): {last_error()}")


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


@dataclass
class SolveResult:
    iterations: int
    constraint_iterations: int
    has_converged: bool
    n_step_magnitudes: int
    last_step_magnitude: float


class Solver:
    """IterativeDiagonalSolver on the GPU (one instance = one HIP stream)."""

    def __init__(self, n_antennas, n_solutions_per_direction, n_channel_blocks, *,
                 n_correlations=2, max_iterations=100, min_iterations=0, accuracy=1e-5,
                 constraint_accuracy=1e-4, step_size=0.2, detect_stalling=True,
                 step_diff_sigma=0.1, phase_only=False, model_storage=MODEL_C64,
                 iteration=ITERATION_STREAM, n_solution_polarizations=2):
        L = lib()
        self._nsd = np.ascontiguousarray(n_solutions_per_direction, dtype=np.int32)
        cfg = Config()
        L.ddh_config_defaults(C.byref(cfg))
        cfg.n_antennas = n_antennas
        cfg.n_directions = len(self._nsd)
        cfg.n_channel_blocks = n_channel_blocks
        cfg.n_correlations = n_correlations
        cfg.n_solutions_per_direction = self._nsd.ctypes.data_as(C.POINTER(C.c_int32))
        cfg.max_iterations = max_iterations
        cfg.min_iterations = min_iterations
        cfg.accuracy = accuracy
        cfg.constraint_accuracy = constraint_accuracy
        cfg.step_size = step_size
        cfg.detect_stalling = int(detect_stalling)
        cfg.step_diff_sigma = step_diff_sigma
        cfg.phase_only = int(phase_only)
        cfg.model_storage = model_storage
        cfg.iteration = iteration
        cfg.n_solution_polarizations = n_solution_polarizations
        self.n_solution_polarizations = n_solution_polarizations
        self.config = cfg
        self.n_antennas = n_antennas
        self.n_channel_blocks = n_channel_blocks
        self.n_correlations = n_correlations
        self.n_sub = int(self._nsd.sum())
        self.model_storage = model_storage
        self.iteration = iteration
        h = L.ddh_solver_create(C.byref(cfg))
        if not h:
            raise HipError(f"ddh_solver_create failed: {last_error()}")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().ddh_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def set_schedule(self, waves_per_tile=4, rows_per_trip=1, use_graph=0, poll_interval=8):
        _check(lib().ddh_set_schedule(self.handle, waves_per_tile, rows_per_trip, use_graph,
                                      poll_interval), "ddh_set_schedule")

    def set_profiling(self, enable=True):
        _check(lib().ddh_set_profiling(self.handle, int(enable)), "ddh_set_profiling")

    def set_gram_fusion(self, enable=True):
        """Fused Gram iteration (apply + antenna solve in one kernel) where the
        problem allows it; on by default. True (1): each iteration's finishing
        step runs inside the next iteration's kernel; 2: a separate finishing
        kernel per iteration (the same values); False (0): off."""
        _check(lib().ddh_set_gram_fusion(self.handle, int(enable)), "ddh_set_gram_fusion")

    def set_gram_solve_groups(self, groups):
        _check(lib().ddh_set_gram_solve_groups(self.handle, int(groups)),
               "ddh_set_gram_solve_groups")

    def set_antenna_solve_schedule(self, split=0, threads=0):
        _check(lib().ddh_set_antenna_solve_schedule(self.handle, int(split), int(threads)),
               "ddh_set_antenna_solve_schedule")

    def antenna_solve_schedule(self):
        sp, th = C.c_int32(), C.c_int32()
        _check(lib().ddh_antenna_solve_schedule(self.handle, C.byref(sp), C.byref(th)),
               "ddh_antenna_solve_schedule")
        return sp.value, th.value

    def last_timing(self):
        ms = C.c_double()
        n = C.c_int64()
        total = C.c_double()
        _check(lib().ddh_solver_last_timing(self.handle, C.byref(ms), C.byref(n),
                                            C.byref(total)), "ddh_solver_last_timing")
        return ms.value, n.value, total.value

    @staticmethod
    def _sol_ptrs(solutions):
        assert solutions.dtype == np.complex128 and solutions.flags.c_contiguous
        n_cb = solutions.shape[0]
        arr = (C.c_void_p * n_cb)(*[solutions[cb].ctypes.data for cb in range(n_cb)])
        return arr

    def solve(self, problem, solutions, time=0.0):
        """Drop-in solve from a SolveData-layout problem (dp3_amd.synth.SolveProblem).
        `solutions` [n_cb, n_ant*n_sub*2] complex128 is updated in place."""
        views = (CbView * problem.n_channel_blocks)()
        keep = []
        n_dir = problem.n_directions
        for cb in range(problem.n_channel_blocks):
            v0, v1 = int(problem.vis_offsets[cb]), int(problem.vis_offsets[cb + 1])
            d = np.ascontiguousarray(problem.data[v0:v1])
            m = np.ascontiguousarray(problem.model[v0 * n_dir:v1 * n_dir])
            pairs = np.ascontiguousarray(np.stack([problem.ant1[v0:v1], problem.ant2[v0:v1]],
                                                  axis=1).astype(np.uint32))
            sm = np.ascontiguousarray(problem.sol_map[v0 * n_dir:v1 * n_dir])
            keep += [d, m, pairs, sm]
            views[cb] = CbView(v1 - v0, d.ctypes.data, m.ctypes.data, pairs.ctypes.data,
                               sm.ctypes.data)
        res = Result()
        ptrs = self._sol_ptrs(solutions)
        _check(lib().ddh_solve(self.handle, views, problem.n_channel_blocks,
                               C.cast(ptrs, C.c_void_p), time, C.byref(res)), "ddh_solve")
        return SolveResult(res.iterations, res.constraint_iterations, bool(res.has_converged),
                           res.n_step_magnitudes, res.last_step_magnitude)

    def set_smoothness(self, frequencies, bandwidth_hz, bandwidth_ref_frequency_hz=0.0,
                       spectral_exponent=0.0, kernel_truncation=True, antenna_factors=None,
                       dd_factors=None):
        """Smoothness constraint on the device (None frequencies: off)."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (frequencies, antenna_factors, dd_factors)]
        _check(lib().ddh_solver_set_smoothness(self.handle, _ptr(keep[0]), float(bandwidth_hz),
                                               float(bandwidth_ref_frequency_hz),
                                               float(spectral_exponent), int(kernel_truncation),
                                               _ptr(keep[1]), _ptr(keep[2])),
               "ddh_solver_set_smoothness")

    def set_constraint_weights(self, weights):
        """[n_antennas, n_channel_blocks] weights of the smoothness constraint."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        assert w.size == self.n_antennas * self.n_channel_blocks
        _check(lib().ddh_solver_set_constraint_weights(self.handle, _ptr(w)),
               "ddh_solver_set_constraint_weights")

    def set_sub_solution_weights(self, weights):
        """[n_sub, n_antennas, n_channel_blocks] weights of the smoothness constraint."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        assert w.size == int(self._nsd.sum()) * self.n_antennas * self.n_channel_blocks
        _check(lib().ddh_solver_set_sub_solution_weights(self.handle, _ptr(w)),
               "ddh_solver_set_sub_solution_weights")

    def solve_problem(self, dproblem, solutions):
        res = Result()
        ptrs = self._sol_ptrs(solutions)
        _check(lib().ddh_solve_problem(self.handle, dproblem.handle, C.cast(ptrs, C.c_void_p),
                                       C.byref(res)), "ddh_solve_problem")
        return SolveResult(res.iterations, res.constraint_iterations, bool(res.has_converged),
                           res.n_step_magnitudes, res.last_step_magnitude)


def autotune(solver, dproblem):
    """Picks the fastest correlate schedule for this problem; returns
    (waves_per_tile, rows_per_trip, correlate ms)."""
    w, r, ms = C.c_int32(), C.c_int32(), C.c_double()
    _check(lib().ddh_autotune(solver.handle, dproblem.handle, C.byref(w), C.byref(r),
                              C.byref(ms)), "ddh_autotune")
    return w.value, r.value, ms.value


class Batch:
    """Many solution intervals solved together (one launch per kernel per
    iteration over all intervals)."""

    def __init__(self, solvers, dproblems):
        self.solvers = list(solvers)
        self.dproblems = list(dproblems)
        n = len(self.solvers)
        sh = (C.c_void_p * n)(*[s.handle for s in self.solvers])
        ph = (C.c_void_p * n)(*[p.handle for p in self.dproblems])
        h = lib().ddh_batch_create(C.cast(sh, C.c_void_p), C.cast(ph, C.c_void_p), n)
        if not h:
            raise HipError(f"ddh_batch_create failed: {last_error()}")
        self.handle = h

    def rebind(self, dproblems):
        """Next intervals of the same shapes; keeps buffers, schedule and graph."""
        self.dproblems = list(dproblems)
        n = len(self.dproblems)
        ph = (C.c_void_p * n)(*[p.handle for p in self.dproblems])
        _check(lib().ddh_batch_rebind(self.handle, C.cast(ph, C.c_void_p)), "ddh_batch_rebind")

    def set_schedule(self, waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=8):
        _check(lib().ddh_batch_set_schedule(self.handle, waves_per_tile, rows_per_trip,
                                            use_graph, poll_interval), "ddh_batch_set_schedule")

    def autotune(self):
        w, r, ms = C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().ddh_batch_autotune(self.handle, C.byref(w), C.byref(r), C.byref(ms)),
               "ddh_batch_autotune")
        return w.value, r.value, ms.value

    def set_gram_solve_groups(self, groups):
        _check(lib().ddh_batch_set_gram_solve_groups(self.handle, int(groups)),
               "ddh_batch_set_gram_solve_groups")

    def gram_solve_groups(self):
        return int(lib().ddh_batch_gram_solve_groups(self.handle))

    def set_antenna_solve_schedule(self, split=0, threads=0):
        _check(lib().ddh_batch_set_antenna_solve_schedule(self.handle, int(split), int(threads)),
               "ddh_batch_set_antenna_solve_schedule")

    def antenna_solve_schedule(self):
        sp, th = C.c_int32(), C.c_int32()
        _check(lib().ddh_batch_antenna_solve_schedule(self.handle, C.byref(sp), C.byref(th)),
               "ddh_batch_antenna_solve_schedule")
        return sp.value, th.value

    def _pointers(self, solutions):
        n = len(self.solvers)
        assert len(solutions) == n
        inner = []
        for sol in solutions:
            assert sol.dtype == np.complex128 and sol.flags.c_contiguous
            inner.append((C.c_void_p * sol.shape[0])(*[sol[cb].ctypes.data
                                                       for cb in range(sol.shape[0])]))
        outer = (C.c_void_p * n)(*[C.cast(x, C.c_void_p) for x in inner])
        return inner, outer, (Result * n)()

    @staticmethod
    def _results(res):
        return [SolveResult(r.iterations, r.constraint_iterations, bool(r.has_converged),
                            r.n_step_magnitudes, r.last_step_magnitude) for r in res]

    def solve(self, solutions):
        """solutions: list of [n_cb, n_ant*n_sub*2] complex128 arrays, in place."""
        inner, outer, res = self._pointers(solutions)
        _check(lib().ddh_batch_solve(self.handle, C.cast(outer, C.c_void_p), res),
               "ddh_batch_solve")
        return self._results(res)

    def solve_warm(self, net, solutions, stretch=4):
        """Learned warm start on the device from the previous solutions held in
        `solutions` (replaced by the final solutions)."""
        inner, outer, res = self._pointers(solutions)
        _check(lib().ddh_batch_solve_warm(self.handle, net.handle, stretch,
                                          C.cast(outer, C.c_void_p), res),
               "ddh_batch_solve_warm")
        return self._results(res)

    def last_warm_ms(self):
        ms = C.c_double()
        _check(lib().ddh_batch_last_warm_time(self.handle, C.byref(ms)),
               "ddh_batch_last_warm_time")
        return ms.value

    def last_time_ms(self):
        ms = C.c_double()
        _check(lib().ddh_batch_last_time(self.handle, C.byref(ms)), "ddh_batch_last_time")
        return ms.value

    def set_profiling(self, enable=True):
        _check(lib().ddh_batch_set_profiling(self.handle, int(enable)), "ddh_batch_set_profiling")

    def last_timing(self):
        """(correlate ms, antenna solve + decision ms, iterations that did work)
        of the last profiled solve."""
        ms, sm, n = C.c_double(), C.c_double(), C.c_int64()
        _check(lib().ddh_batch_last_timing(self.handle, C.byref(ms), C.byref(sm), C.byref(n)),
               "ddh_batch_last_timing")
        return ms.value, sm.value, n.value

    def close(self):
        if getattr(self, "handle", None):
            lib().ddh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class DeviceProblem:
    """A solution interval resident in device memory (library layout)."""

    def __init__(self, solver: Solver, n_times, antenna1, antenna2, channels_per_block,
                 sol_index):
        a1 = np.ascontiguousarray(antenna1, dtype=np.int32)
        a2 = np.ascontiguousarray(antenna2, dtype=np.int32)
        ch = np.ascontiguousarray(channels_per_block, dtype=np.int32)
        si = np.ascontiguousarray(sol_index, dtype=np.uint32)
        h = lib().ddh_problem_create(solver.handle, n_times, len(a1), _ptr(a1), _ptr(a2),
                                     _ptr(ch), _ptr(si))
        if not h:
            raise HipError(f"ddh_problem_create failed: {last_error()}")
        self.handle = h
        self.solver = solver
        self.n_times = n_times
        self.n_baselines = len(a1)
        self.channels_per_block = ch
        self.bl_stride = lib().ddh_problem_baseline_stride(h)

    @classmethod
    def from_problem(cls, solver, problem):
        """Structure of a SolveData-layout problem (dp3_amd.synth.SolveProblem)."""
        n_t, n_bl = problem.n_times, problem.n_baselines
        ch = np.diff(problem.channel_begin).astype(np.int32)
        v0 = 0
        nv0 = int(problem.vis_offsets[1])
        a1 = problem.ant1[0:n_bl * ch[0]:ch[0]]
        a2 = problem.ant2[0:n_bl * ch[0]:ch[0]]
        n_dir = problem.n_directions
        sm = problem.sol_map[:nv0 * n_dir].reshape(n_dir, n_t, n_bl * ch[0])[:, :, 0]
        obj = cls(solver, n_t, a1, a2, ch, sm)
        del v0
        return obj

    def upload(self, problem):
        views = (CbView * problem.n_channel_blocks)()
        keep = []
        n_dir = problem.n_directions
        for cb in range(problem.n_channel_blocks):
            v0, v1 = int(problem.vis_offsets[cb]), int(problem.vis_offsets[cb + 1])
            d = np.ascontiguousarray(problem.data[v0:v1])
            m = np.ascontiguousarray(problem.model[v0 * n_dir:v1 * n_dir])
            keep += [d, m]
            views[cb] = CbView(v1 - v0, d.ctypes.data, m.ctypes.data, None, None)
        _check(lib().ddh_problem_upload(self.handle, views), "ddh_problem_upload")

    def predict(self, ref, station_uvw, freq, chan_width, components, *, stokes_i_only=True,
                correct_freq_smearing=False, scaled_ncp_uvw=None):
        """Fills every model plane. station_uvw [n_t, n_st, 3], freq/chan_width
        over all channels, components a ddh_component array; scaled_ncp_uvw
        (3 values) turns on the time-smearing correction."""
        uvw = np.ascontiguousarray(station_uvw, dtype=np.float64)
        f = np.ascontiguousarray(freq, dtype=np.float64)
        w = np.ascontiguousarray(chan_width, dtype=np.float64)
        ncp = None if scaled_ncp_uvw is None else np.ascontiguousarray(scaled_ncp_uvw, np.float64)
        assert ncp is None or ncp.shape == (3,)
        assert uvw.shape[0] == self.n_times
        _check(lib().ddh_predict(self.handle, ref[0], ref[1], _ptr(uvw), uvw.shape[1], _ptr(f),
                                 _ptr(w), C.cast(components, C.c_void_p), len(components),
                                 int(stokes_i_only), int(correct_freq_smearing),
                                 None if ncp is None else _ptr(ncp)), "ddh_predict")

    def synthesize_data(self, gains, sol_index, noise_sigma, seed=0):
        """data = sum_d G_a M_d G_b^H + noise, on the device. gains [n_cb, n_ant*n_sub*2]."""
        g = np.ascontiguousarray(gains, dtype=np.complex128)
        n_sub = g.shape[1] // (2 * self.solver.n_antennas)
        ptrs = (C.c_void_p * g.shape[0])(*[g[cb].ctypes.data for cb in range(g.shape[0])])
        si = np.ascontiguousarray(sol_index, dtype=np.uint32)
        _check(lib().ddh_problem_synthesize_data(self.handle, C.cast(ptrs, C.c_void_p), n_sub,
                                                 _ptr(si), float(noise_sigma), seed),
               "ddh_problem_synthesize_data")

    def assign_and_weight(self, data, model, antenna1, antenna2, flags=None, weights=None,
                          linear_weighting=False):
        """Fills the problem from an interval's raw buffers (AssignAndWeight +
        SolveData packing on the device): data [n_t, n_bl, n_ch, 4] and model
        [n_dir, n_t, n_bl, n_ch, 4] complex64, flags (bool) and weights
        (float32) [n_t, n_bl, n_ch, 4] or None, over every baseline of the
        buffers with antenna1/antenna2 [n_bl]. Returns the kernels' device
        time in ms."""
        d = np.ascontiguousarray(data, dtype=np.complex64)
        m = np.ascontiguousarray(model, dtype=np.complex64)
        a1 = np.ascontiguousarray(antenna1, dtype=np.int32)
        a2 = np.ascontiguousarray(antenna2, dtype=np.int32)
        n_t, n_bl, n_ch, n_c = d.shape
        assert n_t == self.n_times and n_c == 4 and m.shape[1:] == d.shape
        assert len(a1) == n_bl and len(a2) == n_bl
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        for x in (f, w):
            assert x is None or x.shape == d.shape
        ms = C.c_double()
        _check(lib().ddh_problem_assign_and_weight(
            self.handle, n_bl, _ptr(a1), _ptr(a2), _ptr(d), _ptr(f), _ptr(w), _ptr(m),
            int(linear_weighting), C.byref(ms)),
            "ddh_problem_assign_and_weight")
        return ms.value

    def data_ptr(self, cb):
        return lib().ddh_problem_data_ptr(self.handle, cb)

    def model_ptr(self, cb, d):
        return lib().ddh_problem_model_ptr(self.handle, cb, d)

    def set_model_scale(self, scale):
        s = np.ascontiguousarray(scale, dtype=np.float32)
        _check(lib().ddh_problem_set_model_scale(self.handle, _ptr(s)),
               "ddh_problem_set_model_scale")

    def download_plane(self, cb, plane, n_corr=2, f16=False):
        rows = self.n_times * int(self.channels_per_block[cb])
        if f16 and plane > 0:
            out = np.zeros((rows, self.bl_stride, n_corr, 2), np.float16)
        else:
            out = np.zeros((rows, self.bl_stride, n_corr), np.complex64)
        _check(lib().ddh_problem_download_plane(self.handle, cb, plane, _ptr(out)),
               "ddh_problem_download_plane")
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib().ddh_problem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def correct_and_subtract(data, model, solutions, antenna1, antenna2, channel_block_start,
                         n_antennas, subtract=True, keep_model=False):
    """DDECal::CorrectAndSubtractModels on the GPU for host arrays: data
    [n_t, n_bl, n_ch, 4] and model [n_dir, n_t, n_bl, n_ch, 4] complex64,
    solutions [n_cb, n_ant, n_dir, n_pol] complex128. Returns (data, model,
    kernel ms) with the outputs as new arrays."""
    d = np.ascontiguousarray(data, dtype=np.complex64).copy()
    m = np.ascontiguousarray(model, dtype=np.complex64).copy()
    s = np.ascontiguousarray(solutions, dtype=np.complex128)
    n_dir, n_t, n_bl, n_ch, _ = m.shape
    a1 = np.ascontiguousarray(antenna1, dtype=np.int32)
    a2 = np.ascontiguousarray(antenna2, dtype=np.int32)
    cbs = np.ascontiguousarray(channel_block_start, dtype=np.int32)
    ms = C.c_double()
    _check(lib().ddh_correct_and_subtract(n_t, n_bl, n_ch, n_dir, n_antennas, s.shape[0],
                                          s.shape[-1], _ptr(a1), _ptr(a2), _ptr(cbs), _ptr(s),
                                          _ptr(d), _ptr(m), int(subtract), int(keep_model),
                                          C.byref(ms)), "ddh_correct_and_subtract")
    return d, m, ms.value
