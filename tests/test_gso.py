"""GSO driver: RNG and quaternion KATs of qt.rs, and trajectories against the
oracle's restatement of lib.rs / swarm.rs / glowworm.rs."""
import math
import os

import numpy as np
import pytest

import lklight
from lklight import gso as G
from lklight import synth
from oracle import lkl_oracle as O

EPS = 2.220446049250313e-16  # f64::EPSILON of qt.rs float_equals


def _qeq(a, b):
    return all(abs(x - y) < EPS for x, y in zip(a, b))


def test_rng_kat_product_and_oracle():
    """qt.rs:474-487: StdRng::seed_from_u64(324324324) -> Quaternion::random."""
    a, b = G.rng_f64s(324324324, 3), O.rng_f64s(324324324, 3)
    assert np.array_equal(a, b)
    u1, u2, u3 = a
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    kat = (0.31924330894562036, -0.5980633213833059, 0.5444724265858514, 0.49391674399349367)
    assert _qeq(q, kat), q
    assert _qeq(O.quat_random(324324324, 1)[0], kat)
    # the stream continues across ChaCha blocks identically
    assert np.array_equal(G.rng_f64s(7, 100), O.rng_f64s(7, 100))


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_slerp_kats(impl):
    """qt.rs:424-472."""
    sl = G.slerp if impl == "product" else O.slerp
    assert _qeq(sl([1, 0, 0, 2], [3, -1, 4, 3], 0.0), (0.4472135954999579, 0.0, 0.0, 0.8944271909999159))
    assert _qeq(sl([1, 0, 0, 2], [3, -1, 4, 3], 1.0),
                (0.50709255283711, -0.1690308509457033, 0.6761234037828132, 0.50709255283711))
    q = [0.7071067811865476, 0.0, 0.0, 0.7071067811865476]
    assert _qeq(sl(q, q, 0.1), q)
    assert _qeq(sl([1, 0, 0, 0], [0, 0, 1, 0], 0.5), (0.7071067811865475, 0.0, 0.7071067811865475, 0.0))
    assert _qeq(sl([0.7071067811865475, 0, 0, 0.7071067811865475],
                   [0.0, 0.7071067811865475, 0.7071067811865475, 0.0], 0.5), (0.5, 0.5, 0.5, 0.5))


@pytest.fixture(scope="module")
def small_vdw():
    cx = synth.make_config("small")
    pos = synth.make_swarm_positions(1, 40, cx.meta["rec_axes"], 3, lig_radius=cx.meta["lig_axes"][0])
    return cx, O.Oracle(cx), pos


def test_gso_logic_matches_oracle_on_cpu(small_vdw):
    """Product driver with the caller's CPU scorer (no GPU context: the
    supports_batch() == false branch) vs the oracle's GSO: same scores, so the
    trajectories must agree bit for bit (neighbours, draws, moves, slerp)."""
    cx, orc, pos = small_vdw
    steps = 12

    def cpu_scores(poses):
        r = orc.batch("vdw", poses)
        return -r["v"]

    g = G.GSO(None, "vdw", pos, seed=324324, fallback=cpu_scores)
    ref = orc.gso("vdw", pos, 324324, steps)
    for s in range(steps):
        g.update_luciferin()
        g.movement_phase()
        st = g.state()
        assert np.array_equal(st["pose"], ref["pose"][s]), s
        assert np.array_equal(st["luciferin"], ref["luciferin"][s]), s
        assert np.array_equal(st["neighbors"], ref["neighbors"][s]), s
        assert np.array_equal(st["vision"], ref["vision"][s]), s
    assert ref["neighbors"].max() > 0 and (ref["pose"][-1] != ref["pose"][0]).any()
    # one call per step that had glowworms to re-score (moved, or step 0)
    assert 1 <= g.stats()["fallback_calls"] <= steps and g.stats()["gpu_launches"] == 0
    g.close()


def test_gso_save_format(small_vdw, tmp_path):
    """swarm.rs:232-271 line format."""
    cx, orc, pos = small_vdw
    g = G.GSO(None, "vdw", pos[:3], seed=1, fallback=lambda p: np.arange(p.shape[0], dtype=float))
    g.run(1)
    g.save(0, 1, str(tmp_path))
    lines = open(os.path.join(tmp_path, "gso_1.out")).read().splitlines()
    assert lines[0] == "#Coordinates  RecID  LigID  Luciferin  Neighbor's number  Vision Range  Scoring"
    assert len(lines) == 4
    head, tail = lines[1].split(")")
    assert len(head.strip("(").split(", ")) == 7 and all("." in x for x in head.split(", "))
    f = tail.split()
    assert f[0] == "0" and f[1] == "0" and len(f) == 6
    st = g.state()
    assert f[2] == "%.8f" % st["luciferin"][0] and f[5] == "%.8f" % st["scoring"][0]
    g.close()


def test_gso_failure_without_fallback(small_vdw):
    cx, orc, pos = small_vdw
    with pytest.raises(lklight.GpuError):
        G.GSO(None, "vdw", pos, seed=1)


@pytest.mark.gpu
def test_gso_vdw_step_identical_gpu_vs_cpu(gpu):
    """vdw GSO on the 1,500-atom receptor / 800-atom ligand complex
    (1 swarm x 50 glowworms x 20 steps): GPU-scored trajectory == CPU-grid-scored
    trajectory at every step; per-pose scores within 1e-5 relative."""
    cx = synth.make_config("c1")
    pos = synth.make_swarm_positions(1, 50, cx.meta["rec_axes"], 42, lig_radius=cx.meta["lig_axes"][0])
    orc = O.Oracle(cx)
    steps = 20
    ref = orc.gso("vdw", pos, 42, steps, nthreads=8)
    ctx = lklight.Context(cx, with_field=False)
    g = G.GSO(ctx, "vdw", pos, seed=42)
    for s in range(steps):
        g.update_luciferin()
        g.movement_phase()
        st = g.state()
        assert np.array_equal(st["pose"], ref["pose"][s]), s
        assert np.array_equal(st["neighbors"], ref["neighbors"][s]), s
        err = np.abs(st["scoring"] - ref["scoring"][s]) / np.maximum(np.abs(ref["scoring"][s]), 1.0)
        assert err.max() <= 1e-5, (s, err.max())
    stats = g.stats()
    assert 1 <= stats["gpu_launches"] <= steps and stats["fallback_calls"] == 0
    g.close()
    ctx.close()


@pytest.mark.gpu
def test_gso_dna_multi_swarm_one_launch_per_step(gpu):
    """Several swarms step together with one scoring launch per step; each
    swarm's trajectory equals that swarm run alone (swarms are independent)."""
    cx = synth.make_config("c1")
    cx.build_field()
    pos = synth.make_swarm_positions(3, 20, cx.meta["rec_axes"], 5, lig_radius=cx.meta["lig_axes"][0])
    ctx = lklight.Context(cx)
    g = G.GSO(ctx, "dna", pos, seed=7, swarm_sizes=[20, 20, 20])
    g.run(8)
    both = g.state()
    assert 1 <= g.stats()["gpu_launches"] <= 8
    for s in range(3):
        one = G.GSO(ctx, "dna", pos[20 * s:20 * (s + 1)], seed=7)
        one.run(8)
        assert np.array_equal(one.state()["pose"], both["pose"][20 * s:20 * (s + 1)])
        one.close()
    g.close()
    ctx.close()


def test_plan_then_set_poses_equals_movement_phase(small_vdw):
    """lkl_gso_plan consumes the same draws and writes the same moves as
    movement_phase: following the plan reproduces the plain trajectory."""
    cx, orc, pos = small_vdw

    def cpu_scores(poses):
        return -orc.batch("vdw", poses)["v"]

    a = G.GSO(None, "vdw", pos, seed=11, fallback=cpu_scores)
    b = G.GSO(None, "vdw", pos, seed=11, fallback=cpu_scores)
    for _ in range(6):
        a.update_luciferin()
        a.movement_phase()
        b.update_luciferin()
        plan = b.plan()
        moved = a.state()["moved"]
        b.set_poses(plan, mask=moved)
        assert np.array_equal(a.state()["pose"], b.state()["pose"])
        assert np.array_equal(a.state()["vision"], b.state()["vision"])
    a.close()
    b.close()


def test_commit_sets_pose_and_score(small_vdw):
    cx, orc, pos = small_vdw
    g = G.GSO(None, "vdw", pos[:4], seed=1, fallback=lambda p: np.zeros(p.shape[0]))
    g.update_luciferin()
    plan = g.plan()
    acc = np.array([1, 0, 1, 0], np.uint8)
    g.commit(plan, np.array([1.5, 2.5, 3.5, 4.5]), acc)
    st = g.state()
    assert np.array_equal(st["pose"][[0, 2]], plan[[0, 2]])
    assert list(st["scoring"]) == [1.5, 0.0, 3.5, 0.0]
    assert not st["moved"].any()
    g.close()


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["dna", "vdw"])
def test_gso_device_run_matches_host_driver(gpu, method):
    """Device-resident steps (lkl_gso_device_run) against the host driver from
    the same start, 3 swarms x 40 glowworms on the 1,500 x 800-atom complex, 12 steps:
    at every step the moves, neighbour counts and vision ranges are identical,
    poses agree to 1e-9 A (the device slerp's acos/sin are correctly rounded,
    the host libm's within one ulp) and scores / luciferin to 1e-7 relative.
    One 8-step call equals 8 one-step calls bit for bit."""
    cx = synth.make_config("c1")
    if method != "vdw":
        cx.build_field()
    pos = synth.make_swarm_positions(3, 40, cx.meta["rec_axes"], 17, lig_radius=cx.meta["lig_axes"][0])
    ctx = lklight.Context(cx, with_field=method != "vdw")
    sizes = [40, 40, 40]
    host = G.GSO(ctx, method, pos, seed=11, swarm_sizes=sizes)
    dev = G.GSO(ctx, method, pos, seed=11, swarm_sizes=sizes)
    moved_total = 0
    for s in range(12):
        host.update_luciferin()
        host.movement_phase()
        dev.device_run(1)
        a, b = host.state(), dev.state()
        assert np.array_equal(a["moved"], b["moved"]), s
        assert np.array_equal(a["neighbors"], b["neighbors"]), s
        assert np.array_equal(a["vision"], b["vision"]), s
        assert np.abs(a["pose"] - b["pose"]).max() <= 1e-9, (s, np.abs(a["pose"] - b["pose"]).max())
        assert _rel(b["scoring"], a["scoring"]) <= 1e-7, s
        assert _rel(b["luciferin"], a["luciferin"]) <= 1e-7, s
        moved_total += int(a["moved"].sum())
    assert moved_total > 100  # the swarms do move
    one = G.GSO(ctx, method, pos, seed=11, swarm_sizes=sizes)
    one.device_run(8)
    many = G.GSO(ctx, method, pos, seed=11, swarm_sizes=sizes)
    for _ in range(8):
        many.device_run(1)
    x, y = one.state(), many.state()
    for k in x:
        assert np.array_equal(x[k], y[k]), k
    for g in (host, dev, one, many):
        g.close()
    ctx.close()


@pytest.mark.gpu
def test_gso_with_restraints_on_gpu(gpu):
    """Restraints and a membrane keep the reference's GSO off the GPU batch
    (dna.rs:1047-1056); here the GPU scores them (interface counts + the fold
    of dna.rs:989-995).  The GPU-scored trajectory equals the one scored by the
    oracle's restrained CPU path (a fallback-only driver) pose for pose over 8
    steps, and the device-resident steps make the same moves."""
    cx = synth.make_config("c1")
    cx.build_field()
    rng = np.random.default_rng(9)
    rec_res = [list(range(a, a + 6)) for a in rng.choice(np.arange(0, cx.rec.n - 6, 6), 40, replace=False)]
    lig_res = [list(range(a, a + 4)) for a in rng.choice(np.arange(0, cx.lig.n - 4, 4), 10, replace=False)]
    membrane = sorted(rng.choice(cx.rec.n, 60, replace=False).tolist())
    pos = synth.make_swarm_positions(2, 30, cx.meta["rec_axes"], 21, lig_radius=cx.meta["lig_axes"][0])
    orc = O.Oracle(cx)

    def cpu(poses):
        r = orc.batch("dna", poses, nthreads=8)
        return np.array([orc.restraint_score(lklight.fold_scores("dna", r["e"][k], r["v"][k]), poses[k],
                                             rec_res, lig_res, membrane) for k in range(len(poses))])
    ctx = lklight.Context(cx)
    ctx.set_interface(rec_res, lig_res, membrane)
    sizes = [30, 30]
    gpu_g = G.GSO(ctx, "dna", pos, seed=5, swarm_sizes=sizes)
    cpu_g = G.GSO(None, "dna", pos, seed=5, swarm_sizes=sizes, fallback=cpu)
    dev_g = G.GSO(ctx, "dna", pos, seed=5, swarm_sizes=sizes)
    plain = G.GSO(lklight.Context(cx), "dna", pos, seed=5, swarm_sizes=sizes)
    for s in range(8):
        for g in (gpu_g, cpu_g, plain):
            g.update_luciferin()
            g.movement_phase()
        dev_g.device_run(1)
        a, b, d = gpu_g.state(), cpu_g.state(), dev_g.state()
        assert np.array_equal(a["pose"], b["pose"]), s
        err = np.abs(a["scoring"] - b["scoring"]) / np.maximum(np.abs(b["scoring"]), 1.0)
        assert err.max() <= 1e-5, (s, err.max())
        assert np.array_equal(a["moved"], d["moved"]) and np.array_equal(a["neighbors"], d["neighbors"]), s
        assert np.abs(a["pose"] - d["pose"]).max() <= 1e-9, s
    # the restraints change the scores (and the search) against the plain run
    assert not np.allclose(a["scoring"], plain.state()["scoring"])
    for g in (gpu_g, cpu_g, dev_g, plain):
        g.close()
    ctx.close()


@pytest.fixture(scope="module")
def c1_dna(gpu):
    cx = synth.make_config("c1")
    cx.build_field()
    pos = synth.make_swarm_positions(2, 20, cx.meta["rec_axes"], 13, lig_radius=cx.meta["lig_axes"][0])
    orc = O.Oracle(cx)

    def cpu(poses):  # the caller's CPU grid path (dna.rs:1081-1085 re-scores on it)
        r = orc.batch("dna", poses, nthreads=8)
        return lklight.fold_scores("dna", r["e"], r["v"])
    return cx, pos, cpu


@pytest.mark.gpu
@pytest.mark.parametrize("stage", ["upload", "launch", "sync", "readback"])
def test_gpu_failure_every_step_uses_the_fallback(c1_dna, stage, monkeypatch):
    """A GPU context whose every scoring call fails (LKL_FAIL_AT=<stage>, a
    host-side early return after the stage's own work, no GPU fault) scores
    every step through the registered CPU fallback: the trajectory equals a
    fallback-only driver's bit for bit, one fallback call per step that has
    glowworms to score (a step where none moved makes no call)."""
    cx, pos, cpu = c1_dna
    ctx = lklight.Context(cx)
    steps = 4
    ref = G.GSO(None, "dna", pos, seed=3, swarm_sizes=[20, 20], fallback=cpu)
    g = G.GSO(ctx, "dna", pos, seed=3, swarm_sizes=[20, 20], fallback=cpu)
    monkeypatch.setenv("LKL_FAIL_AT", stage)
    calls = 0  # steps with glowworms to score (step 0, or some glowworm moved)
    for s in range(steps):
        calls += int(s == 0 or g.state()["moved"].any())
        for d in (g, ref):
            d.update_luciferin()
            d.movement_phase()
        a, b = g.state(), ref.state()
        for k in ("pose", "scoring", "luciferin", "vision", "neighbors", "moved"):
            assert np.array_equal(a[k], b[k]), (stage, s, k)
    monkeypatch.delenv("LKL_FAIL_AT")
    st = g.stats()
    assert st["fallback_calls"] == calls >= 3 and st["gpu_launches"] == 0, (st, calls)
    # the context is healthy again once the injection stops
    e, v, _ = ctx.score(a["pose"][:4], lklight.FLAGS_DNA)
    assert np.isfinite(e).all() and np.isfinite(v).all()
    for d in (g, ref):
        d.close()
    ctx.close()


@pytest.mark.gpu
def test_gpu_failure_mid_run_rescored_on_cpu(c1_dna, monkeypatch):
    """One failed GPU call in the middle of a run (LKL_FAIL_AT=launch for step
    3 only): that step's scores are the fallback's (the oracle's values
    exactly), every other step is GPU-scored, and the trajectory makes the same
    moves as the run with no failure (poses identical, scores within 1e-5)."""
    cx, pos, cpu = c1_dna
    ctx = lklight.Context(cx)
    g = G.GSO(ctx, "dna", pos, seed=3, swarm_sizes=[20, 20], fallback=cpu)
    h = G.GSO(ctx, "dna", pos, seed=3, swarm_sizes=[20, 20], fallback=cpu)
    launches = 0  # GPU-scored steps of g: steps with glowworms to score, but the failed one
    for s in range(6):
        launches += int(s != 3 and (s == 0 or g.state()["moved"].any()))
        if s == 3:
            monkeypatch.setenv("LKL_FAIL_AT", "launch")
            before = g.state()
            need = before["moved"].astype(bool)
            assert need.any()
            g.update_luciferin()
            monkeypatch.delenv("LKL_FAIL_AT")
            got = g.state()["scoring"]
            assert np.array_equal(got[need], cpu(before["pose"][need]))
            g.movement_phase()
        else:
            g.update_luciferin()
            g.movement_phase()
        h.update_luciferin()
        h.movement_phase()
        a, b = g.state(), h.state()
        assert np.array_equal(a["pose"], b["pose"]) and np.array_equal(a["neighbors"], b["neighbors"]), s
        err = np.abs(a["scoring"] - b["scoring"]) / np.maximum(np.abs(b["scoring"]), 1.0)
        assert err.max() <= 1e-5, (s, err.max())
    assert g.stats()["fallback_calls"] == 1 and g.stats()["gpu_launches"] == launches >= 4, (g.stats(), launches)
    assert h.stats()["fallback_calls"] == 0
    for d in (g, h):
        d.close()
    ctx.close()


@pytest.mark.gpu
def test_gpu_failure_without_fallback_reports_the_error(c1_dna, monkeypatch):
    """No fallback registered: the failed GPU call fails the step, and the
    GPU's message reaches lkl_last_error()."""
    cx, pos, cpu = c1_dna
    ctx = lklight.Context(cx)
    g = G.GSO(ctx, "dna", pos, seed=3, swarm_sizes=[20, 20])
    monkeypatch.setenv("LKL_FAIL_AT", "readback")
    with pytest.raises(lklight.GpuError, match="no fallback is registered.*readback"):
        g.update_luciferin()
    monkeypatch.delenv("LKL_FAIL_AT")
    g.close()
    ctx.close()


def test_gso_rules_closed_form():
    """One movement phase on four glowworms whose outcome follows from
    glowworm.rs / swarm.rs by hand (scores from a fixed function of the pose
    through the fallback scorer; initial luciferin 5, vision 0.2, rho 0.5,
    gamma 0.4, beta 0.08, 5 neighbours, steps 0.5):
      A (0, 0, 0), B (0.1, 0, 0), C (0, 0.15, 0), D (-0.2, 0, 0) with scores
      0, 1, 0.9, 28 -> luciferin 2.5, 2.9, 2.86, 13.7;
      A's neighbours are B and C (brighter, closer than 0.2); D is brighter
      but exactly 0.2 away, outside the strict `<` of is_neighbor;
      A picks B when its draw is below (2.9 - 2.5) / (0.4 + 0.36), else C;
      C's only neighbour is B (0.18 away); B and D have none;
      a move is 0.5 A towards the partner's pre-move position and slerp(0.5)
      of the rotations, which for unit quaternions at angle w is the rotation
      at angle w / 2; vision = 0.2 + 0.08 (5 - |neighbours|).
    The same test pins the oracle's GSO, which equals the product driver bit
    for bit (test_gso_logic_matches_oracle_on_cpu)."""
    half = lambda a: [math.cos(a), math.sin(a), 0.0, 0.0]  # rotation about x by 2a
    qb, qc = half(0.3), [math.cos(0.2), 0.0, math.sin(0.2), 0.0]
    pos = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
                    [0.1, 0.0, 0.0] + qb,
                    [0.0, 0.15, 0.0] + qc,
                    [-0.2, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]])

    def scores(poses):  # rows (w, x, y, z, tx, ty, tz)
        tx, ty = poses[:, 4], poses[:, 5]
        return 10.0 * tx + 6.0 * ty + 30.0 * (tx < -0.1)

    for seed in (1, 5, 9, 13):
        g = G.GSO(None, "vdw", pos, seed=seed, fallback=scores)
        g.update_luciferin()
        st0 = g.state()
        assert np.allclose(st0["luciferin"], [2.5, 2.9, 2.86, 13.7], rtol=0, atol=1e-12)
        g.movement_phase()
        st = g.state()
        r = G.rng_f64s(seed, 4)[0]  # the step's first draw is glowworm A's
        p_b = (2.9 - 2.5) / ((2.9 - 2.5) + (2.86 - 2.5))
        assert list(st["neighbors"]) == [2, 0, 1, 0]
        assert np.allclose(st["vision"], [0.44, 0.6, 0.52, 0.6], rtol=0, atol=1e-15)
        assert list(st["moved"]) == [1, 0, 1, 0]
        # A: 0.5 A along +x towards B, rotation half-way (angle 0.15), or towards C
        if r < p_b:
            assert np.allclose(st["pose"][0], half(0.15) + [0.5, 0.0, 0.0], rtol=0, atol=1e-15)
        else:
            assert np.allclose(st["pose"][0], [math.cos(0.1), 0.0, math.sin(0.1), 0.0, 0.0, 0.5, 0.0],
                               rtol=0, atol=1e-15)
        # C: towards B's pre-move position; B and D stay
        d = np.array([0.1, -0.15, 0.0])
        tc = np.array([0.0, 0.15, 0.0]) + 0.5 * d / np.linalg.norm(d)
        assert np.allclose(st["pose"][2, 4:], tc, rtol=0, atol=1e-15)
        assert np.allclose(st["pose"][1], qb + [0.1, 0.0, 0.0], rtol=0, atol=1e-15)
        assert np.allclose(st["pose"][3], [1.0, 0.0, 0.0, 0.0, -0.2, 0.0, 0.0], rtol=0, atol=1e-15)
        g.close()
