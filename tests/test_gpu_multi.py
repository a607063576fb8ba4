"""Multi-problem batch: one launch per step over complexes of uneven
sizes and families must reproduce every problem's own launch bit for bit,
also when replayed from a hipGraph."""
import sys

import numpy as np
import torch
import pytest

import lklight
from lklight import synth
from oracle import lkl_oracle as O
from scoring_checks import assert_scores, say

pytestmark = pytest.mark.gpu


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
    for block, mult in [(64, 1), (512, 0), (128, 4), (256, 2), (1024, 0), (1024, 1)]:
        say("multi tiles", block, mult)
        m.set_tiles(block, mult)
        assert np.array_equal(m.score(allp), ref), (block, mult)
    # a re-tiled context is picked up by the batch
    ctxs[0].set_tiles(6.0, 128, 0)
    assert np.array_equal(m.score(allp), ref)
    # so is a parameter-type level: pair-type tables, receptor tables, none
    for level in (1, 0, 2):
        say("multi parameter-type level", level)
        for ctx in ctxs:
            ctx.set_param_types(level)
        assert np.array_equal(m.score(allp), ref), level
    m.close()


def test_multi_trace(batch):
    """A traced launch scores the same bits and records every tile once: start
    before end on the wall clock, a CU id, and per problem as many tiles as it
    has poses (whole-ligand tiles)."""
    cxs, ctxs, flags, poses = batch
    ref = np.concatenate([np.stack(ctx.score(p, f)[:2], 1) for ctx, p, f in zip(ctxs, poses, flags)])
    allp = np.concatenate(poses)
    m = lklight.Multi(ctxs, flags, [p.shape[0] for p in poses])
    m.set_tiles(256, 0)
    m.set_trace(True)
    for rep in range(2):  # the graph is captured with the trace buffer, then replayed
        assert np.array_equal(m.score(allp), ref), rep
        rec, khz = m.trace()
        assert khz > 0 and rec.shape == (allp.shape[0], 4)
        assert (rec[:, 0] <= rec[:, 1]).all() and (rec[:, 0] > 0).all()
        counts = np.bincount(rec[:, 3].astype(np.int64), minlength=len(poses))
        assert sorted(counts.tolist()) == sorted(p.shape[0] for p in poses)
        lb = lklight.load_balance(rec, khz)
        say("multi trace: %d tiles on %d CUs, span %.1f us, busy %.3f"
            % (lb["tiles"], lb["cus"], lb["span_us"], lb["busy_frac"]))
        assert 0.0 < lb["busy_frac"] <= 1.0 and lb["cus"] >= 1
    # a launch on another stream: the readback waits for it
    d_p = torch.from_numpy(allp).cuda()
    d_o = torch.zeros((allp.shape[0], 2), dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    m.score_async(d_p.data_ptr(), d_o.data_ptr(), st.cuda_stream)
    rec, khz = m.trace()
    assert (rec[:, 0] > 0).all() and (rec[:, 0] <= rec[:, 1]).all()
    assert np.array_equal(d_o.cpu().numpy(), ref)
    m.set_trace(False)
    assert np.array_equal(m.score(allp), ref)
    m.close()


def test_multi_rejects_desolvation(batch):
    cxs, ctxs, flags, poses = batch
    with pytest.raises(lklight.GpuError, match="desolvation"):
        lklight.Multi(ctxs[:2], [lklight.F_DESOLV, 0], [4, 4])


def test_multi_full_size(gpu):
    """The 64 complexes at full size (receptors 2k-20k atoms, ligands 500-5k,
    100-500 glowworms, families mixed; 20,180 poses per step): the batch
    launch, replayed from its graph, equals the 64 per-complex launches bit
    for bit at the stored tiles and at 1,024-thread tiles, and the first
    complex of each family matches the oracle on 8 poses."""
    say("build the 64 full-size complexes (device-built phi)")
    cxs = synth.make_config4(64)
    ctxs, flags, poses = [], [], []
    for k, cx in enumerate(cxs):
        fam = cx.meta["family"]
        ctxs.append(lklight.Context(cx, with_field=False, device_field=(fam != "vdw"), cell_size=4.0))
        flags.append(lklight.family_flags(fam))
        poses.append(synth.make_poses(cx.meta["glowworms"], cx.meta["rec_axes"], 500 + k))
    allp = np.concatenate(poses)
    assert allp.shape[0] == 20180
    say("64 sequential launches")
    ref = np.concatenate([np.stack(ctx.score(p, f)[:2], 1) for ctx, p, f in zip(ctxs, poses, flags)])
    m = lklight.Multi(ctxs, flags, [p.shape[0] for p in poses])
    t = lklight.tuned_tiles("c4")
    for block, mult in [(t["block"], t["chunk_mult"]), (1024, 0), (1024, 1)]:
        say("batch tiles", block, mult)
        m.set_tiles(block, mult)
        for rep in range(2):
            assert np.array_equal(m.score(allp), ref), (block, mult, rep)
    m.close()
    off = np.concatenate([[0], np.cumsum([p.shape[0] for p in poses])])
    for fam in ("dna", "pydock", "vdw"):
        k = next(i for i, cx in enumerate(cxs) if cx.meta["family"] == fam)
        cx = cxs[k]
        say("oracle:", fam, "complex", k, "(%d x %d atoms)" % (cx.rec.n, cx.lig.n))
        if fam != "vdw":
            cx.build_field()  # host copy of the GPU build (bit-equal to the oracle's)
        got = ref[off[k]:off[k] + 8]
        assert_scores(fam, {"e": got[:, 0], "v": got[:, 1]}, O.Oracle(cx).batch(fam, poses[k][:8], nthreads=8))
        cx.phi = None
    for c in ctxs:
        c.close()
