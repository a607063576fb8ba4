"""A dump_metal_dataset directory scored on the GPU (lkl_ctx_score_coords).

The dataset is written as the reference writes it (dna.rs:473-663): the
seeded pose set, the placed f32 coordinates, and per-pose energies of the CPU
grid path in f64 (the oracle, which reproduces the reference's 1AZP value to
its last digit).  Bars: the device's scores of the placements within 1e-5
relative (floor 1) of ref.bin; the coordinate-input path within the same bar of
the pose-input path on the same poses.
"""
import sys

import numpy as np
import pytest

import lklight
from lklight import dataset as D
from lklight import synth
from oracle import lkl_oracle as O

import azp_fixture as A

pytestmark = pytest.mark.gpu

REL, FLOOR = 1e-5, 1.0


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _written(tmp_path, cx, seed, n):
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    poses = D.seeded_poses(seed, n)
    ref = O.Oracle(cx).batch("dna", poses, nthreads=8)
    D.write(str(tmp_path), cx, poses, lklight.fold_scores("dna", ref["e"], ref["v"]))
    return poses


@pytest.mark.parametrize("name", ["1azp", "small"])
def test_dataset_scores_match_ref(gpu, tmp_path, name):
    cx = A.load("dna") if name == "1azp" else synth.make_config("small")
    say("dataset", name, ": write 96 seeded poses with oracle energies")
    poses = _written(tmp_path, cx, 2024, 96)
    ds = D.read(str(tmp_path))
    g = ds.score()
    err = np.abs(g - ds.ref) / np.maximum(np.abs(ds.ref), FLOOR)
    say(name, "placements vs ref.bin: max rel err %.3e (scores %.1f .. %.1f)"
        % (float(err.max()), float(ds.ref.min()), float(ds.ref.max())))
    assert (err <= REL).all(), (float(err.max()), int(np.argmax(err)))
    # the same poses through the pose-input path of a context on the same data
    ctx = lklight.Context(ds.complex())
    e, v, _ = ctx.score_coords(ds.coords, lklight.family_flags("dna"))
    assert np.array_equal(lklight.fold_scores("dna", e, v), g)  # deterministic
    ctx.close()
    cxr = synth.Complex(rec=ds.complex().rec, lig=cx.lig)
    cxr.set_field(ds.phi, ds.origin, ds.dims)
    ctx = lklight.Context(cxr)
    e2, v2, _ = ctx.score(poses, lklight.family_flags("dna"))
    ctx.close()
    p = lklight.fold_scores("dna", e2, v2)
    d = np.abs(p - g) / np.maximum(np.abs(p), FLOOR)
    say(name, "pose input vs coordinate input: max rel diff %.3e" % float(d.max()))
    assert (d <= REL).all()


def test_score_coords_refuses_bad_coordinates(gpu):
    cx = synth.make_config("tiny")
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    ctx = lklight.Context(cx)
    xyz = np.tile(cx.lig.xyz[None], (2, 1, 1))
    xyz[1, 3, 1] = np.nan
    with pytest.raises(lklight.GpuError):
        ctx.score_coords(xyz, lklight.family_flags("dna"))
    with pytest.raises(ValueError):
        ctx.score_coords(xyz[:, :-1], lklight.family_flags("dna"))
    e, v, _ = ctx.score_coords(xyz[:1], lklight.family_flags("dna"))  # still usable
    assert np.isfinite(e).all() and np.isfinite(v).all()
    ctx.close()
