"""The device GSO step's math (lkl_gsomath.hpp), on the host: correctly
rounded sin and acos by double-double evaluation, pinned against 80-bit long
double, and the slerp built on them against the host-libm slerp the host GSO
driver uses (qt.rs:80-104).  No GPU needed."""
import ctypes as C

import numpy as np
import pytest

import lklight
from lklight import _capi

_d = np.float64


def _lib():
    try:
        return lklight.lib()
    except lklight.LibraryMissing:
        pytest.skip("liblklight_gpu.so not built")


def _nearest_double_or_none(r):
    """The double nearest to long double r, or None when r sits within 2^-60
    relative of a midpoint between doubles (long double cannot decide)."""
    d = np.float64(r)
    lo, hi = np.nextafter(d, -np.inf), np.nextafter(d, np.inf)
    for a, b in ((lo, d), (d, hi)):
        mid = (np.longdouble(a) + np.longdouble(b)) / 2
        if abs(r - mid) <= abs(r) * np.longdouble(2.0) ** -60:
            return None
    return d


def test_sin_acos_correctly_rounded():
    L = _lib()
    rng = np.random.default_rng(11)
    decided = 0
    for x in rng.uniform(0.0, 1.6, 20000):
        want = _nearest_double_or_none(np.sin(np.longdouble(x)))
        if want is None:
            continue
        decided += 1
        assert L.lkl_sin_cr(float(x)) == want, x
    for d in rng.uniform(0.0, 1.0, 20000):
        want = _nearest_double_or_none(np.arccos(np.longdouble(d)))
        if want is None:
            continue
        assert L.lkl_acos_cr(float(d)) == want, d
    assert decided > 19500  # about 1 % sit too close to a midpoint to decide
    # exact cases
    assert L.lkl_sin_cr(0.0) == 0.0 and L.lkl_acos_cr(1.0) == 0.0


def test_slerp_cr_against_host_libm():
    """Same formula, same operation order; only acos/sin may differ in the last
    bit (the host libm is within one ulp, not always correctly rounded)."""
    L = _lib()
    rng = np.random.default_rng(3)
    q = rng.normal(size=(4000, 2, 4))
    q /= np.linalg.norm(q, axis=2, keepdims=True)
    same = 0
    for a, b in q:
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        r1, r2 = np.zeros(4), np.zeros(4)
        L.lkl_quat_slerp(_capi.ptr(a, C.c_double), _capi.ptr(b, C.c_double), 0.5, _capi.ptr(r1, C.c_double))
        L.lkl_quat_slerp_cr(_capi.ptr(a, C.c_double), _capi.ptr(b, C.c_double), 0.5, _capi.ptr(r2, C.c_double))
        ulps = np.abs(r1.view(np.int64) - r2.view(np.int64))
        assert ulps.max() <= 4, (a, b, r1, r2)
        same += int(np.array_equal(r1, r2))
    assert same >= 0.99 * len(q), same
