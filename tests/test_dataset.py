"""The dump_metal_dataset directory format (dna.rs:473-663): writer, reader and
the pieces that must equal the reference's bit for bit (cell list, placed
coordinates, the seeded pose set), checked on the CPU against the oracle.
"""
import math
import os

import numpy as np
import pytest

from lklight import dataset as D
from lklight import synth
from oracle import lkl_oracle as O

import azp_fixture as A


def test_cell_list_matches_oracle():
    """gpu_score.rs:73-145 restated twice (numpy in the package, C in the
    oracle, which is pinned by the all-pairs window test): identical."""
    for xyz in (synth.make_config("small").rec.xyz, np.load(A.PATH)["rec_xyz"]):
        start, atoms, nc, co = D.cell_list(xyz)
        cs, ca, nco, cco = O.crec_build(xyz)
        assert nc == tuple(nco) and np.array_equal(start, cs) and np.array_equal(atoms, ca)
        assert co[:3] == tuple(float(v) for v in cco) and co[3] == 10.0


def test_placed_matches_oracle_transform():
    """dna.rs:546-556: f64 rigid transform (qt.rs order), then f32."""
    cx = A.load("dna")
    rng = np.random.default_rng(3)
    poses = np.concatenate([synth.random_quaternions(12, rng), rng.uniform(-20, 20, (12, 3))], 1)
    got = D.placed(cx.lig.xyz, poses)
    orc = O.Oracle(cx)
    for k in range(poses.shape[0]):
        assert np.array_equal(got[k], orc.transform(poses[k]).astype(np.float32)), k


def test_seeded_poses_follow_the_reference_recipe():
    """lkl_dataset_poses against dna.rs:497-519 restated on the oracle's
    StdRng stream (pinned by the qt.rs:474-487 KAT)."""
    for seed in (0, 7, 324324324):
        got = D.seeded_poses(seed, 40)
        r = O.rng_f64s(seed, 40 * 7 + 64)
        i, want = 0, []
        for _ in range(40):
            t = [r[i] * 40.0 - 20.0, r[i + 1] * 40.0 - 20.0, r[i + 2] * 40.0 - 20.0]
            i += 3
            while True:
                q = [r[i + a] * 2.0 - 1.0 for a in range(4)]
                i += 4
                n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]
                if n2 > 1e-12:
                    inv = 1.0 / math.sqrt(n2)
                    want.append([v * inv for v in q] + t)
                    break
        assert np.array_equal(got, np.asarray(want)), seed


def test_write_read_round_trip(tmp_path):
    cx = synth.make_config("tiny")
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    poses = D.seeded_poses(11, 24)
    orc = O.Oracle(cx)
    ref = np.array([O.fold("dna", p.e_raw, p.v) for p in (orc.parts("dna", q) for q in poses)])
    meta = D.write(str(tmp_path), cx, poses, ref, cpu_grid_ms=1.5)
    nr, nl, n, nphi = cx.rec.n, cx.lig.n, poses.shape[0], int(np.prod(cx.dims))
    size = lambda f: os.path.getsize(os.path.join(str(tmp_path), f))  # noqa: E731
    assert size("field.bin") == 12 + 24 + 8 + 4 * nphi
    assert size("rec.bin") == 4 + 25 * nr + 12 + 16 + 4 * meta["cell_start_len"] + 4 * nr
    assert size("lig.bin") == 4 + 13 * nl
    assert size("poses.bin") == 8 + 12 * n * nl
    assert size("ref.bin") == 4 + 8 * n
    ds = D.read(str(tmp_path))
    assert ds.dims == cx.dims and ds.origin == cx.origin and ds.spacing == 0.5
    assert np.array_equal(ds.phi, cx.phi) and np.array_equal(ds.ref, ref)
    assert np.array_equal(ds.coords, D.placed(cx.lig.xyz, poses))
    assert np.array_equal(ds.rec["xyz"], cx.rec.xyz.astype(np.float32))
    assert np.array_equal(ds.lig["svdw"], cx.lig.seps.astype(np.float32))
    start, atoms, nc, co = D.cell_list(cx.rec.xyz)
    assert np.array_equal(ds.cells[0], start) and np.array_equal(ds.cells[1], atoms)
    assert ds.meta["cpu_grid_ms"] == 1.5 and ds.meta["cpu_exact_ms"] is None
    # a truncated or padded file is refused
    p = os.path.join(str(tmp_path), "lig.bin")
    b = open(p, "rb").read()
    open(p, "wb").write(b[:-1])
    with pytest.raises(ValueError):
        D.read(str(tmp_path))
    open(p, "wb").write(b + b"\0")
    with pytest.raises(ValueError):
        D.read(str(tmp_path))
