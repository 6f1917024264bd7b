"""Numerical envelope of the device solver at high dynamic range: the
CPU-runnable shape (24 stations, 4 directions of one point each, 32 channels
in 4 blocks, solution interval 4, noise 0.01 Jy on re and im) with one
direction raised to 10^3 or 10^4 Jy, the regime of direction-dependent
calibration with an A-team source.

Every form runs from unit gains to the 1e-6 tolerance (iteration limit 1000,
no stall detection) and is compared with the FP64 oracle's converged
solution, which reaches the thermal chi^2:
  * the oracle itself (the reference precision: FP32 visibilities and sums),
  * the streaming iteration on complex64 planes (per-visibility residual),
  * the Gram iteration on complex64 planes (per-baseline cross products
    C = sum D conj(M_d), G = sum M_e conj(M_d) made once, U = C - sum gg G
    each iteration: a difference of two FP32 sums of the bright direction's
    products),
  * FP16 model planes with the streaming and the Gram iteration.
The reference precision itself does not reach the tolerance here: its FP32
per-visibility residual D - sum g M loses the faint directions' signal
against the bright one (the oracle stops at the limit, chi^2 0.4 % above the
FP64 optimum at 10^3 Jy and 74 % at 10^4 Jy). The complex64 device forms
must do at least as well as the reference precision (gain error and chi^2
excess against the FP64 oracle within 1.5 times the FP32 oracle's, or within
the parity bars). FP16 model planes keep 11
significand bits relative to their direction's scale, so the bright
direction's stored model carries a quantization error of about
2^-12 |M_bright| per visibility that no gain can fit; it enters the faint
directions' solutions like extra noise, which here is far above the 0.01 Jy
thermal noise (about 0.14 Jy rms at 10^3 Jy): FP16 storage is outside its
envelope, and the test records that it is (chi^2 excess above 10 %), so that
a claim to the contrary fails. Such skies keep complex64 model planes."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import oracle_model, phase_ref_relerr

pytestmark = pytest.mark.gpu

LIMIT = 1000
TOLERANCE = 1e-6


def bright_problem(flux):
    w = S.SMALL
    rng = np.random.default_rng(77)
    geom = S.make_geometry(w, rng)
    sky = S.make_sky(w, geom, rng)
    sky.flux[sky.direction == 0] = flux
    model = oracle_model(w, geom, sky)
    gains = S.true_gains(w, rng)
    return w, S.assemble(w, geom, model, gains, rng)


def gpu(P, storage, iteration):
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 max_iterations=LIMIT, accuracy=TOLERANCE, detect_stalling=False,
                 model_storage=storage, iteration=iteration)
    sol = P.unit_solutions()
    r = s.solve(P, sol)
    s.close()
    return sol, r.iterations


@pytest.mark.parametrize("flux", [1e3, 1e4])
def test_bright_direction_against_fp64_oracle(flux):
    print(f"case: bright direction {flux:g} Jy", flush=True)
    w, P = bright_problem(flux)
    kw = dict(max_iterations=LIMIT, accuracy=TOLERANCE, detect_stalling=False, n_threads=16)
    ref64 = P.unit_solutions()
    r64 = O.solve(P, ref64, fp64=True, **kw)
    ref32 = P.unit_solutions()
    r32 = O.solve(P, ref32, **kw)
    forms = {"oracle fp32": (ref32, r32.iterations)}
    for name, storage, iteration in (("stream c64", H.MODEL_C64, H.ITERATION_STREAM),
                                     ("gram c64", H.MODEL_C64, H.ITERATION_GRAM),
                                     ("stream fp16", H.MODEL_F16, H.ITERATION_STREAM),
                                     ("gram fp16", H.MODEL_F16, H.ITERATION_GRAM)):
        forms[name] = gpu(P, storage, iteration)
    c64 = O.chi2(P, ref64)
    err = {k: phase_ref_relerr(g, ref64, P.n_antennas, P.n_sub) for k, (g, _) in forms.items()}
    chi = {k: O.chi2(P, g) / c64 - 1.0 for k, (g, _) in forms.items()}
    print(f"FP64 oracle: {r64.iterations} iterations, chi2 {c64:.6g}")
    for k, (_, it) in forms.items():
        print(f"{k:12s}: {it} iterations, gain error vs FP64 oracle {err[k]:.2e}, chi2 excess "
              f"{chi[k]:+.2e}")
    for k in ("stream c64", "gram c64"):
        assert err[k] <= max(1e-4, 1.5 * err["oracle fp32"]), k
        assert chi[k] <= max(1e-3, 1.5 * chi["oracle fp32"]), k
    for k in ("stream fp16", "gram fp16"):
        assert chi[k] > 0.1, k
