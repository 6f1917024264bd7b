"""DDECal::CorrectAndSubtractModels on the GPU against the oracle
(steps/DDECal.cc:992-1124): scalar, diagonal and full-Jones solutions,
uneven channel blocks, autocorrelations, subtract and keep-model flags."""
import numpy as np
import pytest

from dp3_amd import hip as H
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def case(n_pol, seed=5, n_t=3, n_ant=7, n_ch=9, n_dir=4, starts=(0, 2, 6)):
    rng = np.random.default_rng(seed)
    bl = [(a, b) for a in range(n_ant) for b in range(a, n_ant)]
    a1 = np.array([x for x, _ in bl], np.int32)
    a2 = np.array([y for _, y in bl], np.int32)
    n_bl = len(bl)

    def cplx(*s):
        return rng.normal(size=s) + 1j * rng.normal(size=s)

    data = cplx(n_t, n_bl, n_ch, 4).astype(np.complex64)
    model = cplx(n_dir, n_t, n_bl, n_ch, 4).astype(np.complex64)
    cbs = np.array(list(starts) + [n_ch], np.int32)
    sol = cplx(len(cbs) - 1, n_ant, n_dir, n_pol)
    return data, model, sol, a1, a2, cbs, n_ant


def within_one_ulp(got, ref, model, sol):
    scale = np.abs(model).max() * np.abs(sol).max() ** 2 * model.shape[0]
    for g, r in ((got.real, ref.real), (got.imag, ref.imag)):
        g, r = g.astype(np.float64), r.astype(np.float64)
        assert np.all(np.abs(g - r) <= 2.0 ** -23 * np.abs(r) + 1e-14 * scale)


@pytest.mark.parametrize("n_pol", [1, 2, 4])
@pytest.mark.parametrize("subtract,keep", [(1, 0), (1, 1), (0, 1)])
def test_correct_and_subtract_matches_oracle(n_pol, subtract, keep):
    print(f"case: correct and subtract n_pol={n_pol} subtract={subtract} keep={keep}", flush=True)
    data, model, sol, a1, a2, cbs, n_ant = case(n_pol)
    d_ref, m_ref = O.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, subtract, keep)
    d, m, ms = H.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, subtract, keep)
    # FP64 products, sums and differences in both, rounded to complex64 once:
    # equal up to one ulp where the FP64 operation order moves a value across a
    # rounding boundary
    within_one_ulp(d, d_ref, model, sol)
    within_one_ulp(m, m_ref, model, sol)
    if not subtract:
        np.testing.assert_array_equal(d, data)
    if not keep:
        np.testing.assert_array_equal(m, model)


@pytest.mark.parametrize("n_pol", [1, 2])
def test_correct_and_subtract_single_channel_blocks(n_pol):
    """Blocks of one channel at both ends and inside: every channel finds its
    block (DDECal.cc:1106-1108, one block step per channel); a block without
    channels is rejected."""
    print(f"case: correct and subtract single-channel blocks n_pol={n_pol}", flush=True)
    data, model, sol, a1, a2, cbs, n_ant = case(n_pol, n_ch=11, starts=(0, 1, 3, 4, 10))
    d_ref, m_ref = O.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)
    d, m, ms = H.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)
    within_one_ulp(d, d_ref, model, sol)
    within_one_ulp(m, m_ref, model, sol)
    data, model, sol, a1, a2, cbs, n_ant = case(n_pol, n_ch=11, starts=(0, 3, 3, 10))
    with pytest.raises(H.HipError, match="channel"):
        H.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)


@pytest.mark.parametrize("n_pol", [1, 2, 4])
def test_correct_and_subtract_many_blocks(n_pol):
    """100 uneven channel blocks of 120 channels and 4 directions: the scalar
    solutions' factors of three rows fit one workgroup's LDS table, the
    diagonal and full-Jones ones (25.6 and 51.2 KB per row) take the kernel
    that reads the solution entries from global memory; both against the
    oracle."""
    print(f"case: correct and subtract 100 blocks n_pol={n_pol}", flush=True)
    rng = np.random.default_rng(8)
    inner = np.sort(rng.choice(np.arange(1, 120), size=99, replace=False))
    starts = tuple([0] + inner.tolist())
    data, model, sol, a1, a2, cbs, n_ant = case(n_pol, seed=9, n_t=2, n_ant=6, n_ch=120,
                                                starts=starts)
    d_ref, m_ref = O.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)
    d, m, ms = H.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)
    within_one_ulp(d, d_ref, model, sol)
    within_one_ulp(m, m_ref, model, sol)


def _hip_runtime():
    import ctypes as C
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            rt = C.CDLL(name)
            break
        except OSError:
            rt = None
    assert rt is not None, "HIP runtime not found"
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    return rt


@pytest.mark.parametrize("n_pol", [2, 4])
def test_device_entry_equals_host_entry(n_pol):
    """ddh_correct_and_subtract_device on caller-owned device buffers (default
    stream) gives the host entry's outputs bit for bit."""
    import ctypes as C
    print(f"case: correct and subtract device entry n_pol={n_pol}", flush=True)
    data, model, sol, a1, a2, cbs, n_ant = case(n_pol, n_ch=30, starts=(0, 7, 19))
    d_host, m_host, _ = H.correct_and_subtract(data, model, sol, a1, a2, cbs, n_ant, 1, 1)
    rt = _hip_runtime()
    arrays = [np.ascontiguousarray(a1), np.ascontiguousarray(a2), np.ascontiguousarray(cbs),
              np.ascontiguousarray(sol, np.complex128), data.copy(), model.copy()]
    ptrs = []
    try:
        for a in arrays:
            p = C.c_void_p()
            assert rt.hipMalloc(C.byref(p), a.nbytes) == 0
            ptrs.append(p)
            assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        n_t, n_bl, n_ch = data.shape[:3]
        rc = H.lib().ddh_correct_and_subtract_device(n_t, n_bl, n_ch, model.shape[0], n_ant,
                                                     len(cbs) - 1, n_pol, *ptrs, 1, 1, None)
        assert rc == 0
        d_dev, m_dev = np.empty_like(data), np.empty_like(model)
        assert rt.hipMemcpy(d_dev.ctypes.data, ptrs[4], d_dev.nbytes, 2) == 0
        assert rt.hipMemcpy(m_dev.ctypes.data, ptrs[5], m_dev.nbytes, 2) == 0
    finally:
        for p in ptrs:
            rt.hipFree(p)
    np.testing.assert_array_equal(d_dev, d_host)
    np.testing.assert_array_equal(m_dev, m_host)
