"""The oracle's smoothness constraint against the reference's own test cases
(ddecal/test/unit/tSmoothnessConstraint.cc), whose inputs and expected
outputs are restated here as fixtures, plus invariances for the paths those
cases leave open. CPU only."""
import numpy as np
import pytest

from oracle import pyoracle as O

FREQ = np.array([1.0e6, 2.0e6, 3.0e6, 4.0e6, 5.0e6])
BANDWIDTH = 2.01e6  # tSmoothnessConstraint.cc:20
SPECTRAL_EXPONENT = -1.0

# [channel block][direction][polarization], one antenna, real inputs
# (tSmoothnessConstraint.cc:88-94)
INPUT = np.array([
    [[0.0, 1.0], [0.0, 1.0], [0.0, 5.0]],
    [[0.0, 0.0], [0.0, 1.0], [0.0, 4.0]],
    [[1.0, 0.0], [0.0, 1.0], [0.0, 3.0]],
    [[0.0, 0.0], [0.0, 1.0], [0.0, 2.0]],
    [[0.0, 0.0], [10.0, 1.0], [0.0, 1.0]],
])
SIMPLE = np.array([  # :95-101
    [[0.000121798, 0.902597], [0.0, 1.0], [0.0, 4.90247]],
    [[0.0886568, 0.0886568], [0.0, 1.0], [0.0, 3.99978]],
    [[0.822484, 0.000110987], [0.00110987, 1.0], [0.0, 3]],
    [[0.0886568, 0.0], [0.886568, 1.0], [0.0, 2.00022]],
    [[0.000121798, 0.0], [9.02597, 1.0], [0.0, 1.09753]],
])
DD_FACTORS = np.array([  # :122-128, dd factors 4, 2, 1
    [[0.209986, 0.366484], [0.0, 1.0], [0.0, 4.90247]],
    [[0.262608, 0.262608], [0.0, 1.0], [0.0, 3.99978]],
    [[0.257334, 0.147445], [0.456402, 1.0], [0.0, 3]],
    [[0.262608, 0.0], [2.542337, 1.0], [0.0, 2.00022]],
    [[0.234536, 0.0], [5.949716, 1.0], [0.0, 1.09753]],
])
NO_TRUNCATION = np.array([  # :146-152, dd factors 4, 2, 1, kernel not truncated
    [[0.202006, 0.352558], [0.000799632, 1.0], [0.0, 4.90247]],
    [[0.241765, 0.241765], [0.0294459, 1.0], [0.0, 3.99978]],
    [[0.257334, 0.147445], [0.456402, 1.0], [0.0, 3]],
    [[0.241765, 0.0793708], [2.534851, 1.0], [0.0, 2.00022]],
    [[0.202006, 0.0379986], [5.925774, 1.0], [0.0, 1.09753]],
])
# flagged_channels (:159-181): [channel block][sub-solution], one polarization,
# the first two channel blocks weighted 0. The reference writes the second
# expected row as "{0.99875}, {0.0}, {0.0, 1.0}, {0.0}, {2.99875}": five
# entries for six sub-solutions; read here as 0.99875, 0, 0, 1, 0, 2.99875,
# the only reading consistent with the other rows (column 4 is constant 1).
FLAGGED_INPUT = np.array([
    [0.0, 1.0, 0.0, 1.0, 0.0, 5.0],
    [0.0, 0.0, 0.0, 1.0, 0.0, 4.0],
    [1.0, 0.0, 0.0, 1.0, 0.0, 3.0],
    [0.0, 0.0, 0.0, 1.0, 0.0, 2.0],
    [0.0, 0.0, 10.0, 1.0, 0.0, 1.0],
])
FLAGGED = np.array([
    [1.0, 0.0, 0.0, 1.0, 0.0, 3.0],
    [0.99875, 0.0, 0.0, 1.0, 0.0, 2.99875],
    [0.902597, 0.0, 0.00121798, 1.0, 0.0, 2.90247],
    [0.0886666, 0.0, 0.886666, 1.0, 0.0, 2.0],
    [0.000121798, 0.0, 9.02597, 1.0, 0.0, 1.09753],
])


def smoothed(values, *, truncate=True, weights=None, dd_factors=None, n_pol=2, **kw):
    sol = np.ascontiguousarray(values.reshape(5, 1, -1, n_pol).astype(np.complex128))
    c = O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, truncate,
                     weights=None if weights is None else np.asarray(weights, float),
                     antenna_factors=np.array([1.0]), dd_factors=dd_factors, **kw)
    O.smooth(sol, c)
    return sol


def check(got, want):
    # the reference checks each entry to a fraction of 1e-4 (expected zeros
    # exactly) and the tensors with xt::allclose
    np.testing.assert_allclose(got.real.reshape(want.shape), want, rtol=1e-4, atol=0)
    np.testing.assert_array_equal(got.imag, 0.0)


def test_simple_case():
    check(smoothed(INPUT), SIMPLE)


def test_dd_factors():
    check(smoothed(INPUT, dd_factors=np.array([4.0, 2.0, 1.0])), DD_FACTORS)


def test_no_truncation():
    check(smoothed(INPUT, truncate=False, dd_factors=np.array([4.0, 2.0, 1.0])), NO_TRUNCATION)


def test_flagged_channels():
    check(smoothed(FLAGGED_INPUT, weights=[0.0, 0.0, 1.0, 1.0, 1.0], n_pol=1), FLAGGED)


def test_non_finite_values_enter_with_zero_weight():
    """A NaN or Inf solution is smoothed as a zero-weight point
    (SmoothnessConstraint.cc:73-80): the same output as weight 0."""
    bad = INPUT.astype(np.complex128).copy()
    bad[1, 0, 0] = complex(np.nan, 0.0)
    bad[3, 2, 1] = complex(0.0, np.inf)
    got = np.ascontiguousarray(bad.reshape(5, 1, 3, 2))
    O.smooth(got, O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, True,
                               antenna_factors=np.array([1.0])))
    # per sub-solution weights with 0 where the values are not finite
    w = np.ones((6, 1, 5))
    w[0, 0, 1] = 0.0      # sub-solution 0 (direction 0, pol 0): block 1
    w[5, 0, 3] = 0.0      # sub-solution 5 (direction 2, pol 1): block 3
    zeroed = INPUT.copy()
    zeroed[1, 0, 0] = 0.0
    zeroed[3, 2, 1] = 0.0
    # one polarization per series: reshape so that sub-solution = (direction, pol)
    want = np.ascontiguousarray(zeroed.reshape(5, 1, 6, 1).astype(np.complex128))
    O.smooth(want, O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, True,
                                sub_solution_weights=w, antenna_factors=np.array([1.0])))
    np.testing.assert_array_equal(got.reshape(-1), want.reshape(-1))


def test_zero_weights_give_nan():
    got = smoothed(INPUT, weights=np.zeros(5))
    assert np.isnan(got.real).all() and np.isnan(got.imag).all()


def test_antenna_factor_acts_as_dd_factor():
    """The kernel size factor is (1 / dd_factor) / antenna_factor
    (SmoothnessConstraint.cc:82-89): antenna factor 4 equals dd factor 4."""
    sol = np.ascontiguousarray(np.tile(INPUT.reshape(5, 1, 3, 2), (1, 2, 1, 1)).astype(np.complex128))
    a = sol.copy()
    O.smooth(a, O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, True,
                             antenna_factors=np.array([4.0, 1.0])))
    b = sol[:, :1].copy()
    O.smooth(b, O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, True,
                             antenna_factors=np.array([1.0]), dd_factors=np.full(3, 4.0)))
    np.testing.assert_array_equal(a[:, :1], b)
    c = sol[:, 1:].copy()
    O.smooth(c, O.Smoothness(FREQ, BANDWIDTH, 0.0, SPECTRAL_EXPONENT, True,
                             antenna_factors=np.array([1.0])))
    np.testing.assert_array_equal(a[:, 1:], c)


@pytest.mark.parametrize("truncate", [True, False])
def test_reference_frequency_without_exponent(truncate):
    """The reference-frequency path, which the reference's cases leave open:
    with spectral exponent 0 and no truncation it has no effect; with a
    bandwidth that spans the grid, truncation on and off agree at any
    reference frequency; a constant series is a fixed point."""
    sol = np.ascontiguousarray(INPUT.reshape(5, 1, 3, 2).astype(np.complex128))
    plain = sol.copy()
    O.smooth(plain, O.Smoothness(FREQ, BANDWIDTH, 0.0, 0.0, truncate, antenna_factors=np.array([1.0])))
    if not truncate:
        ref = sol.copy()
        O.smooth(ref, O.Smoothness(FREQ, BANDWIDTH, 2.5e6, 0.0, False,
                                   antenna_factors=np.array([1.0])))
        np.testing.assert_array_equal(ref, plain)
    # a very large bandwidth covers the grid in every window: truncation off
    # and on agree, at any reference frequency
    wide_on = sol.copy()
    O.smooth(wide_on, O.Smoothness(FREQ, 1e9, 3e6, 0.0, True, antenna_factors=np.array([1.0])))
    wide_off = sol.copy()
    O.smooth(wide_off, O.Smoothness(FREQ, 1e9, 0.0, 0.0, False, antenna_factors=np.array([1.0])))
    np.testing.assert_allclose(wide_on, wide_off, rtol=1e-15, atol=0)
    # and a constant series is a fixed point
    const = np.full_like(sol, 2.0 - 1.0j)
    O.smooth(const, O.Smoothness(FREQ, BANDWIDTH, 3e6, -1.0, truncate, antenna_factors=np.array([1.0])))
    np.testing.assert_allclose(const, 2.0 - 1.0j, rtol=1e-15)
