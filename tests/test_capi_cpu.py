"""Host-side checks of the C ABI that need no GPU: exports, loud failure without a
device, fault injection, host logic and the synthetic generator."""
import glob
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import lklight
from lklight import _capi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(lkl_\w+)\s*\(", text))
    return sorted(names)


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 12
    nm = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lkl_\w+)", nm))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = lklight.lib()
    for n in names:
        assert getattr(L, n) is not None


def _no_gpu():
    return not lklight.device_available()


@pytest.mark.skipif(not _no_gpu(), reason="a gfx950 device is present")
def test_no_device_fails_loudly():
    cx = synth.make_config("tiny")
    with pytest.raises(lklight.GpuError, match="no gfx950 device"):
        lklight.Context(cx, with_field=False)
    # the stateless wrapper reports -1 (gpu_score.rs:317-319 maps it to None)
    L = lklight.lib()
    rx, lx = cx.rec.f32("xyz"), cx.lig.f32("xyz")
    f = lambda a: _capi.ptr(a, _capi.C.c_float)  # noqa: E731
    poses = np.array([[1, 0, 0, 0, 0, 0, 0]], np.float64)
    out = np.zeros(2)
    dummy = np.zeros(1, np.float32)
    rc = L.lkl_batch_score(f(dummy), 0, 0, 0, 0, 0, 0, 0, f(rx), f(cx.rec.f32("ele")),
                           f(cx.rec.f32("seps")), f(cx.rec.f32("rad")),
                           _capi.ptr(cx.rec.heavy, _capi.C.c_ubyte), cx.rec.n, None, None, 0, 0,
                           0, 0, 0, 0, 0, f(lx), _capi.ptr(poses, _capi.C.c_double),
                           f(cx.lig.f32("ele")), f(cx.lig.f32("seps")), f(cx.lig.f32("rad")),
                           _capi.ptr(cx.lig.heavy, _capi.C.c_ubyte), cx.lig.n, 1, 0,
                           _capi.ptr(out, _capi.C.c_double))
    assert rc == -1
    assert "no gfx950 device" in lklight.last_error()


def test_fault_injection_probe():
    """LKL_FAIL_AT=probe forces the 'no device' answer (drives the fallback path)."""
    code = ("import sys; sys.path.insert(0, %r); import lklight; "
            "print(lklight.device_available())" % os.path.join(ROOT, "lklight-gpu_amd"))
    env = dict(os.environ, LKL_FAIL_AT="probe")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True)
    assert out.stdout.strip() == "False", out.stderr


def test_family_flags():
    """gpu_family.rs:187-194."""
    assert lklight.family_flags("dna") == 7
    assert lklight.family_flags("pydock") == 3
    assert lklight.family_flags("cpydock") == 3
    assert lklight.family_flags("vdw") == 0
    assert lklight.family_flags("sipper") == 7


def test_synth_generator():
    a = synth.make_config("tiny")
    b = synth.make_config("tiny")
    for m in ("rec", "lig"):
        ma, mb = getattr(a, m), getattr(b, m)
        for f in ("xyz", "ele", "rad", "seps", "des", "asa"):
            assert np.array_equal(getattr(ma, f), getattr(mb, f))
        # f32-exact values, so the GPU's f32 arrays equal the oracle's f64 arrays
        for f in ("xyz", "ele", "rad", "seps", "des", "asa"):
            v = getattr(ma, f)
            assert np.array_equal(v.astype(np.float32).astype(np.float64), v)
        d2 = ((ma.xyz[:, None] - ma.xyz[None]) ** 2).sum(-1)
        np.fill_diagonal(d2, 1e9)
        assert d2.min() >= 9.0 - 1e-9
    ax = a.meta["rec_axes"]
    x = a.rec.xyz
    assert ((x[:, 0] / ax[0]) ** 2 + (x[:, 1] / ax[1]) ** 2 + (x[:, 2] / ax[2]) ** 2 <= 1 + 1e-9).all()
    assert math.isclose(a.rec.ele.sum(), 10.0, abs_tol=1e-4)
    assert math.isclose(a.lig.ele.sum(), -10.0, abs_tol=1e-4)
    assert np.allclose(a.lig.xyz.mean(0), 0, atol=1e-3)
    p = synth.make_poses(50, ax, 3)
    assert np.allclose(np.linalg.norm(p[:, :4], axis=1), 1, atol=1e-6)


def test_load_balance_intervals():
    """Busy time per CU is the union of its tiles' intervals; times scale by
    the clock rate (kHz)."""
    rec = np.array([[10, 110, 7, 0], [60, 160, 7, 0], [210, 310, 7, 1], [10, 310, 9, 1]], np.uint64)
    lb = lklight.load_balance(rec, 100_000)  # 100 MHz: 100 ticks = 1 us
    assert lb["cus"] == 2 and lb["tiles"] == 4
    assert math.isclose(lb["span_us"], 3.0)
    assert math.isclose(lb["busy_frac"], (150 + 100 + 300) / (2 * 300))
    assert math.isclose(lb["tiles_in_flight_per_cu"], 600 / 600)
    assert math.isclose(lb["cu_finish_us"]["min"], 3.0) and math.isclose(lb["tile_us"]["max"], 3.0)
