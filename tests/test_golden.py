"""Golden fixture (tests/golden/tiny_scores.json, written by
tests/golden/make_golden.py): the oracle reproduces its committed outputs, and
the HIP path matches them at the parity bar."""
import json
import os

import numpy as np
import pytest

import lklight
from oracle import lkl_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "tiny_scores.json")

import importlib.util  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)


@pytest.fixture(scope="module")
def golden():
    with open(GOLD) as f:
        g = json.load(f)
    cx, poses = make_golden.inputs()
    assert make_golden.digest(cx, poses) == g["inputs_sha256"], "synthetic generator changed"
    return g, cx, poses


def test_oracle_reproduces_golden(golden):
    g, cx, poses = golden
    orc = O.Oracle(cx)
    for m, ref in g["families"].items():
        r = orc.batch(m, poses)
        for k in ("e", "v", "s"):
            np.testing.assert_allclose(r[k], ref[k], rtol=1e-12, atol=1e-12, err_msg=f"{m}.{k}")


@pytest.mark.gpu
def test_gpu_matches_golden(gpu, golden):
    g, cx, poses = golden
    ctx = lklight.Context(cx, desolv=True)
    for m, ref in g["families"].items():
        flags = lklight.family_flags(m) | (lklight.F_DESOLV if m == "cpydock" else 0)
        e, v, s = ctx.score(poses, flags)
        sc = lklight.fold_scores(m, e, v, s)
        want = np.asarray(ref["score"])
        err = np.abs(sc - want) / np.maximum(np.abs(want), 1.0)
        assert err.max() <= 1e-5, (m, err.max())
    ctx.close()
