"""lkl_ctx_score_async_noalloc (the entry point that allocates nothing, its
scratch in one static device array) against lkl_ctx_score_async, through the
C ABI: bit-identical scores on two calls in a row, the second smaller, for the
one-tile-per-pose launch with longest-first dispatch, multi-tile poses (tile
accumulators in the scratch), cpydock with its minima in LDS and in the global
scratch; oversized calls are rejected by the host before any launch."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import lklight
from lklight import _capi, synth
from oracle import lkl_oracle as O
from scoring_checks import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(gpu):
    cx = synth.make_config("c1")
    say("c1 field (oracle build)")
    phi, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    cx.set_field(phi, o, d)
    return cx


def _scores(ctx, fn, d_poses, n, flags, desolv):
    out = torch.full((n, 2), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda") if desolv else None
    rc = fn(ctx.h, C.c_void_p(d_poses.data_ptr()), n, flags, C.c_void_p(out.data_ptr()),
            C.c_void_p(s.data_ptr()) if desolv else None, None)
    _capi.check(rc, "score")
    torch.cuda.synchronize()
    return out.cpu().numpy(), (s.cpu().numpy() if desolv else None)


def _compare(ctx, flags, poses, desolv=False):
    L = _capi.lib()
    d_poses = torch.from_numpy(poses).cuda()
    for n in (poses.shape[0], poses.shape[0] // 3 + 1):  # two calls in a row, the second smaller
        say("  %d poses" % n)
        a, sa = _scores(ctx, L.lkl_ctx_score_async_noalloc, d_poses, n, flags, desolv)
        b, sb = _scores(ctx, L.lkl_ctx_score_async, d_poses, n, flags, desolv)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), n
        if desolv:
            assert np.isfinite(sa).all() and np.array_equal(sa, sb), n


def test_noalloc_matches_async_dna(c1):
    poses = synth.make_poses(1500, c1.meta["rec_axes"], 3)
    ctx = lklight.Context(c1)
    say("dna, one tile per pose, longest-first dispatch")
    _compare(ctx, lklight.FLAGS_DNA, poses)
    say("dna, 200 ligand atoms per tile (tile accumulators in the scratch)")
    ctx.set_tiles(5.0, 256, 200)
    _compare(ctx, lklight.FLAGS_DNA, poses)
    ctx.close()


@pytest.mark.parametrize("global_min", [False, True])
def test_noalloc_matches_async_cpydock(c1, global_min, monkeypatch):
    if global_min:
        monkeypatch.setenv("LKL_FORCE_GLOBAL_MIN", "1")
    poses = synth.make_poses(300, c1.meta["rec_axes"], 4)
    ctx = lklight.Context(c1, desolv=True)
    say("cpydock, minima in %s" % ("global scratch" if global_min else "LDS"))
    _compare(ctx, lklight.FLAGS_PYDOCK | lklight.F_DESOLV, poses, desolv=True)
    ctx.close()


def test_noalloc_rejects_oversized_calls(c1):
    """500,000 poses need more than the 16 MiB scratch: the host returns -1
    with a message before launching (the buffers are real and large enough,
    so nothing could read out of bounds either way)."""
    n = 500_000
    d_poses = torch.zeros((n, 7), dtype=torch.float64, device="cuda")
    d_poses[:, 0] = 1.0
    d_out = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    ctx = lklight.Context(c1)
    rc = _capi.lib().lkl_ctx_score_async_noalloc(ctx.h, C.c_void_p(d_poses.data_ptr()), n, lklight.FLAGS_DNA,
                                                 C.c_void_p(d_out.data_ptr()), None, None)
    assert rc == -1
    assert "maximum" in lklight.last_error()
    torch.cuda.synchronize()
    ctx.close()
