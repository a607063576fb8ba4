"""Shared checks of the GPU parity tests: folded scores and their components
against the oracle's f64 CPU grid values.

Bar: per-pose score within 1e-5 relative of the oracle (absolute floor 1
kcal/mol for scores near 0), each component within the same budget; where the
oracle also returns the FP32 rounding bound of the same sum (its "bound"), the
error must lie within it too.
"""
import sys

import numpy as np

import lklight

REL = 1e-5
FLOOR = 1.0


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def assert_scores(method, gpu, orc, rel=REL, floor=FLOOR):
    """gpu/orc: dicts e, v, s -> folded scores + component budgets."""
    g = lklight.fold_scores(method, gpu["e"], gpu["v"], gpu.get("s"))
    o = lklight.fold_scores(method, orc["e"], orc["v"], orc.get("s"))
    tol = rel * np.maximum(np.abs(o), floor)
    err = np.abs(g - o)
    say(method, "max rel err %.3e" % float((err / np.maximum(np.abs(o), floor)).max()))
    assert (err <= tol).all(), (method, err.max(), np.argmax(err / tol))
    assert (np.abs(gpu["e"] - orc["e"]) * 83.0 <= tol).all()
    assert (np.abs(gpu["v"] - orc["v"]) <= tol).all()
    if gpu.get("s") is not None:
        assert (np.abs(gpu["s"] - orc["s"]) <= tol).all()
    if "bound" in orc:  # the FP32 rounding bound of the same sum (oracle.fp32_bound)
        say(method, "max error / FP32 bound %.3f" % float((err / orc["bound"]).max()))
        assert (err <= orc["bound"]).all(), (method, float((err / orc["bound"]).max()))


