"""Multi-problem batch: one launch per step over complexes of uneven
sizes and families must reproduce every problem's own launch bit for bit,
also when replayed from a hipGraph."""
import sys

import numpy as np
import pytest

import lklight
from lklight import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def batch(gpu):
    say("build 12 scaled config-4 complexes (device-built phi)")
    cxs = synth.make_config4(12, seed0=100, scale=0.15)
    ctxs, flags, poses = [], [], []
    for k, cx in enumerate(cxs):
        fam = cx.meta["family"]
        ctx = lklight.Context(cx, with_field=False, device_field=(fam != "vdw"))
        ctxs.append(ctx)
        flags.append(lklight.family_flags(fam))
        poses.append(synth.make_poses(cx.meta["glowworms"], cx.meta["rec_axes"], 7 + k))
    return cxs, ctxs, flags, poses


def test_device_field_equals_host_field(gpu):
    cx = synth.make_config("small")
    cx.build_field()  # host copy of the GPU build (bit-equal to the oracle's)
    poses = synth.make_poses(32, cx.meta["rec_axes"], 3)
    a = lklight.Context(cx)
    b = lklight.Context(cx, with_field=False, device_field=True)
    assert b.field_dims == cx.dims and b.field_origin == cx.origin
    ea, va, _ = a.score(poses, lklight.FLAGS_DNA)
    eb, vb, _ = b.score(poses, lklight.FLAGS_DNA)
    assert np.array_equal(ea, eb) and np.array_equal(va, vb)
    a.close()
    b.close()


def test_multi_matches_single(batch):
    cxs, ctxs, flags, poses = batch
    ref = [np.stack(ctx.score(p, f)[:2], 1) for ctx, p, f in zip(ctxs, poses, flags)]
    ref = np.concatenate(ref)
    allp = np.concatenate(poses)
    m = lklight.Multi(ctxs, flags, [p.shape[0] for p in poses])
    say("multi: graph replays with eager work between them")
    for rep in range(3):
        out = m.score(allp)
        assert np.array_equal(out, ref), rep
        ctxs[rep].score(poses[rep][:5], flags[rep])  # eager work between replays
    m.use_graph(False)
    assert np.array_equal(m.score(allp), ref)
    m.use_graph(True)
    for block, mult in [(64, 1), (512, 0), (128, 4), (256, 2)]:
        say("multi tiles", block, mult)
        m.set_tiles(block, mult)
        assert np.array_equal(m.score(allp), ref), (block, mult)
    # a re-tiled context is picked up by the batch
    ctxs[0].set_tiles(6.0, 128, 0)
    assert np.array_equal(m.score(allp), ref)
    m.close()


def test_multi_rejects_desolvation(batch):
    cxs, ctxs, flags, poses = batch
    with pytest.raises(lklight.GpuError, match="desolvation"):
        lklight.Multi(ctxs[:2], [lklight.F_DESOLV, 0], [4, 4])
