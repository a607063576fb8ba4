"""Parity of the gfx950 HIP path with the CPU oracle (the reference CPU grid path).

Bar (stated per test): per-pose score within 1e-5 relative of the oracle's f64
CPU grid value (absolute floor 1 kcal/mol for scores near 0), each component
within the same budget, integer/bit outputs (transform, phi, invariances)
bit-exact.  All calls go through the C ABI of liblklight_gpu.so.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import lklight
from lklight import _capi, synth
from oracle import lkl_oracle as O
from scoring_checks import FLOOR, REL, assert_scores, say  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gpu):
    say("build complex 'small' + oracle field")
    cx = synth.make_config("small")
    phi, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    cx.set_field(phi, o, d)
    poses = synth.make_poses(64, cx.meta["rec_axes"], 5)
    return cx, O.Oracle(cx), poses


def run(ctx, poses, flags):
    e, v, s = ctx.score(poses, flags)
    return {"e": e, "v": v, "s": s}


def test_transform_bitwise(small):
    cx, orc, poses = small
    say("transform: ctx create")
    ctx = lklight.Context(cx, with_field=False)
    rng = np.random.default_rng(2)
    extra = np.concatenate([synth.random_quaternions(8, rng), rng.normal(0, 30, (8, 3))], 1)
    P = np.concatenate([poses[:24], extra])
    got = ctx.transform(P)
    for k in range(P.shape[0]):
        assert np.array_equal(got[k], orc.transform(P[k])), k
    ctx.close()


@pytest.mark.parametrize("name", ["tiny", "small", "c1"])
def test_field_build_bitwise(gpu, name):
    cx = synth.make_config(name)
    say("field build", name)
    phi_o, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    phi_g = cx.build_field()
    assert cx.origin == o and cx.dims == d
    assert np.array_equal(phi_g.view(np.uint32), phi_o.view(np.uint32))


@pytest.mark.parametrize("method", ["dna", "vdw", "pydock"])
def test_family_parity(small, method):
    cx, orc, poses = small
    say("family parity", method)
    ctx = lklight.Context(cx)
    gpu = run(ctx, poses, lklight.family_flags(method))
    ref = orc.batch(method, poses, nthreads=4)
    if method == "vdw":
        assert (gpu["e"] == 0).all()
    assert_scores(method, gpu, ref)
    ctx.close()


def test_cpydock_parity(small):
    cx, orc, poses = small
    say("cpydock parity (fused E, V, S)")
    ctx = lklight.Context(cx, desolv=True)
    gpu = run(ctx, poses, lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV)
    ref = orc.batch("cpydock", poses, nthreads=4)
    assert_scores("cpydock", gpu, ref)
    # desolvation-only launch gives the same S bit for bit
    s_only = run(ctx, poses, lklight.F_DESOLV)["s"]
    assert np.array_equal(s_only, gpu["s"])
    # the global-scratch variant of the receptor minima gives the same bits
    os.environ["LKL_FORCE_GLOBAL_MIN"] = "1"
    try:
        say("cpydock global-min path")
        glob_ = run(ctx, poses, lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV)
    finally:
        del os.environ["LKL_FORCE_GLOBAL_MIN"]
    for k in ("e", "v", "s"):
        assert np.array_equal(glob_[k], gpu[k]), k
    ctx.close()


def test_tile_and_order_invariance(small):
    """Fixed-point per-pose sums: bitwise identical for every cell edge (cubic
    and narrower along x), block size (64 to 1,024 threads), ligand tile, neighbour-list spacing (or
    the cell-row walk), receptor parameters from the LDS type table or with
    the records, and input atom order."""
    cx, orc, poses = small
    say("tile invariance")
    ctx = lklight.Context(cx, desolv=True)
    flags = lklight.FLAGS_DNA
    ctx.set_list_spacing(0.0)  # reference run: cell-row walk, phi sampled directly
    ctx.set_phi_packed(False)
    assert ctx.list_info()[0] == 0 and not ctx.phi_packed()
    base = run(ctx, poses, flags)
    base_c = run(ctx, poses, lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV)
    for i, (cell, block, chunk, cell_x, lh) in enumerate(
            [(10.0, 256, 0, 0.0, 1.0), (3.4, 64, 64, 0.0, 0.0), (6.0, 512, 0, 0.0, 0.5),
             (4.0, 128, 256, 0.0, 2.0), (5.0, 64, 128, 0.0, 1.5), (5.0, 256, 0, 1.25, 0.0),
             (7.5, 128, 64, 2.5, 0.75), (3.4, 512, 0, 0.75, 3.0),
             # 1,024-thread tiles: the grouped list walk with the LDS type
             # table and LDS desolvation minima, then the cell-row walk
             (7.5, 1024, 0, 1.25, 1.0), (5.0, 1024, 256, 0.0, 0.0)]):
        packed = i % 2 == 0
        types = (2, 1, 0, 2, 1)[i % 5]  # pair types, receptor types, none
        say(f"  cell {cell} block {block} chunk {chunk} cell_x {cell_x} list spacing {lh} packed phi {packed}"
            f" parameter types {types}")
        ctx.set_tiles(cell, block, chunk, cell_x=cell_x)
        ctx.set_list_spacing(lh)
        ctx.set_phi_packed(packed)
        ctx.set_param_types(types)
        assert (ctx.list_info()[0] > 0) == (lh > 0) and ctx.phi_packed() == packed
        assert ctx.param_types() == (20 if types else 0)  # the synthetic receptors have 20 types
        r = run(ctx, poses, flags)
        rc = run(ctx, poses, lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV)
        for k in ("e", "v"):
            assert np.array_equal(r[k], base[k]), (cell, block, chunk, cell_x, k)
        for k in ("e", "v", "s"):
            assert np.array_equal(rc[k], base_c[k]), (cell, block, chunk, cell_x, k)
        ctx.set_work_sort(False)  # the near field in plain atom order
        r = run(ctx, poses, flags)
        rc = run(ctx, poses, lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV)
        ctx.set_work_sort(True)
        for k in ("e", "v"):
            assert np.array_equal(r[k], base[k]), ("unsorted", cell, block, chunk, cell_x, k)
        for k in ("e", "v", "s"):
            assert np.array_equal(rc[k], base_c[k]), ("unsorted", cell, block, chunk, cell_x, k)
    ctx.close()
    say("order invariance")
    rng = np.random.default_rng(0)
    pr, pl = rng.permutation(cx.rec.n), rng.permutation(cx.lig.n)

    def perm(m, p):
        return synth.Mol(xyz=m.xyz[p], ele=m.ele[p], rad=m.rad[p], seps=m.seps[p],
                         heavy=m.heavy[p], hyd=m.hyd, des=m.des[p], asa=m.asa[p])
    cx2 = synth.Complex(rec=perm(cx.rec, pr), lig=perm(cx.lig, pl))
    cx2.set_field(cx.phi, cx.origin, cx.dims)
    ctx2 = lklight.Context(cx2)
    r2 = run(ctx2, poses, flags)
    assert np.array_equal(r2["e"], base["e"]) and np.array_equal(r2["v"], base["v"])
    ctx2.close()


def _list_entries_host(xyz, h):
    """Neighbour-list size by brute force in f32: per node of the list grid, the
    receptor atoms within rad = sqrt(100.5) + h*sqrt(3)/2 + 0.01 of the node,
    padded to a multiple of 32 (the library's list layout: whole blocks of one
    step of the grouped walk)."""
    f = np.float32
    rad = f(np.sqrt(f(100.5))) + f(0.5) * np.sqrt(f(3.0)) * f(h) + f(0.01)
    lo, hi = xyz.min(0), xyz.max(0)
    o = np.floor(lo - float(rad) - h).astype(np.float32)
    n = (np.ceil((hi + float(rad) + h - o.astype(np.float64)) / h) + 1).astype(int)
    r = xyz.astype(np.float32)
    ax = [o[a] + np.arange(n[a], dtype=np.float32) * f(h) for a in range(3)]
    total = 0
    for z in ax[2]:
        dz2 = (z - r[:, 2]) ** 2
        for y in ax[1]:
            dy2 = (y - r[:, 1]) ** 2
            dx = ax[0][:, None] - r[None, :, 0]
            d2 = (dx * dx + dy2[None, :]) + dz2[None, :]
            c = (d2 <= rad * rad).sum(1)
            total += int(((c + 31) // 32 * 32).sum())
    return total, tuple(int(v) for v in n)


def test_list_build_counts(gpu):
    """The device-built neighbour lists hold exactly the f32 brute-force
    membership (list grid dims and padded entry totals)."""
    cx = synth.make_config("tiny")
    ctx = lklight.Context(cx, with_field=False)
    for h in (1.0, 2.0):
        say("list build counts, spacing", h)
        ctx.set_list_spacing(h)
        got, dims = ctx.list_info()
        want, wdims = _list_entries_host(cx.rec.xyz, h)
        assert dims == wdims, (h, dims, wdims)
        assert abs(got - want) <= 64, (h, got, want)  # a pair exactly on the f32 boundary may flip a block
    ctx.set_list_spacing(0.0)
    assert ctx.list_info()[0] == 0
    with pytest.raises(lklight.GpuError, match="list spacing"):
        ctx.set_list_spacing(0.1)
    ctx.close()


def test_large_receptor_wide_lists(gpu):
    """A 70,000-atom receptor: its dummy index (nr) does not fit 16 bits, so
    the neighbour lists (and the interface pass's short lists) hold 32-bit
    indices.  The list walk stays in use, matches the oracle for dna (far
    field built on the device) and vdw on 12 poses, equals the cell-row walk
    bit for bit, and the interface counts equal the oracle's."""
    say("large receptor: 70,000 atoms")
    ax = synth.scaled_axes(70000, (1.0, 0.95, 0.9))
    rec, _ = synth.make_mol(70000, ax, 123, +10.0)
    lax = synth.scaled_axes(300, (1.0, 1.0, 1.0))
    lig, _ = synth.make_mol(300, lax, 124, -2.0, center=True)
    cx = synth.Complex(rec=rec, lig=lig, meta={"rec_axes": ax})
    say("large receptor: far field on the device")
    cx.build_field()
    ctx = lklight.Context(cx)
    assert ctx.list_info()[0] > 0
    poses = synth.make_poses(12, ax, 9)
    orc = O.Oracle(cx)
    got = {}
    for method in ("dna", "vdw"):
        say("large receptor:", method)
        got[method] = run(ctx, poses, lklight.family_flags(method))
        assert_scores(method, got[method], orc.batch(method, poses, nthreads=8))
    assert (got["vdw"]["v"] != 0).sum() >= 6  # most poses touch the receptor
    rng = np.random.default_rng(5)
    rec_res = [list(range(a, a + 8)) for a in rng.choice(np.arange(0, cx.rec.n - 8, 8), 200, replace=False)]
    lig_res = [list(range(a, a + 5)) for a in rng.choice(np.arange(0, cx.lig.n - 5, 5), 8, replace=False)]
    membrane = sorted(rng.choice(cx.rec.n, 300, replace=False).tolist())
    ctx.set_interface(rec_res, lig_res, membrane)
    say("large receptor: interface counts")
    want = _iface_counts_oracle(orc, poses, rec_res, lig_res, membrane)
    assert np.array_equal(ctx.interface(poses), want)
    assert want[:, 1].sum() > 0
    ctx.set_list_spacing(0.0)
    assert ctx.list_info()[0] == 0
    say("large receptor: cell-row walk")
    rows = run(ctx, poses, lklight.FLAGS_DNA)
    assert np.array_equal(rows["e"], got["dna"]["e"]) and np.array_equal(rows["v"], got["dna"]["v"])
    ctx.close()


def _iface_counts_oracle(orc, poses, rec_res, lig_res, membrane):
    out = np.zeros((len(poses), 3), np.int64)
    for k, pose in enumerate(poses):
        ir, il = orc.interface(pose)
        out[k, 0] = sum(int(ir[r].any()) for r in map(np.asarray, rec_res))
        out[k, 1] = sum(int(il[r].any()) for r in map(np.asarray, lig_res))
        out[k, 2] = int(ir[np.asarray(membrane)].sum())
    return out


def test_interface_counts_match_oracle(small):
    """Restraint / membrane interface counts (pairs within 3.9 A, dna.rs:885-995)
    equal the oracle's per pose as integers; they survive a receptor re-sort;
    the folded restraint scores match the oracle's to 1e-5."""
    cx, orc, poses = small
    rng = np.random.default_rng(4)
    starts = rng.choice(np.arange(0, cx.rec.n - 8, 8), 25, replace=False)
    rec_res = [list(range(a, a + 8)) for a in starts]
    lstarts = rng.choice(np.arange(0, cx.lig.n - 5, 5), 8, replace=False)
    lig_res = [list(range(a, a + 5)) for a in lstarts]
    membrane = sorted(rng.choice(cx.rec.n, 40, replace=False).tolist())
    ctx = lklight.Context(cx)
    assert not lib_has_iface(ctx)
    ctx.set_interface(rec_res, lig_res, membrane)
    assert lib_has_iface(ctx)
    say("interface counts, 64 poses")
    got = ctx.interface(poses)
    want = _iface_counts_oracle(orc, poses, rec_res, lig_res, membrane)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert (want[:, 0] > 0).sum() >= 10 and want[:, 2].sum() > 0  # the poses do touch
    ctx.set_tiles(3.4, 128, 0)  # re-sorted receptor: indices are re-translated
    assert np.array_equal(ctx.interface(poses), want)
    e, v, _ = ctx.score(poses, lklight.FLAGS_DNA)
    sc = lklight.fold_restraints(lklight.fold_scores("dna", e, v), got, len(rec_res), len(lig_res),
                                 len(membrane))
    ref = orc.batch("dna", poses, nthreads=4)
    ref_sc = np.array([orc.restraint_score(lklight.fold_scores("dna", ref["e"][k], ref["v"][k]), poses[k],
                                           rec_res, lig_res, membrane) for k in range(len(poses))])
    err = np.abs(sc - ref_sc) / np.maximum(np.abs(ref_sc), 1.0)
    say("restraint-folded dna max rel err %.3e" % err.max())
    assert err.max() <= 1e-5
    with pytest.raises(lklight.GpuError, match="out of range"):
        ctx.set_interface([], [], [cx.rec.n])
    ctx.set_interface()  # cleared
    assert not lib_has_iface(ctx)
    with pytest.raises(lklight.GpuError, match="no restraints"):
        ctx.interface(poses[:2])
    ctx.close()


def lib_has_iface(ctx):
    return bool(lklight.lib().lkl_ctx_has_interface(ctx.h))


def test_param_types_fallback(small):
    """A receptor with more than 256 distinct parameter records (per-atom vdW
    radii) has no type table: the list walk gathers each candidate's
    parameters with its record, bitwise equal to the cell-row walk and within
    the bar of the oracle."""
    cx = synth.make_config("small", radii=(1.5, 2.0))
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    poses = synth.make_poses(48, cx.meta["rec_axes"], 9)
    ctx = lklight.Context(cx)
    assert ctx.list_info()[0] > 0 and ctx.param_types() == 0
    flags = lklight.family_flags("dna")
    got = run(ctx, poses, flags)
    ctx.set_list_spacing(0.0)
    row = run(ctx, poses, flags)
    ctx.close()
    say("no parameter types: list walk vs cell-row walk, bitwise")
    assert np.array_equal(got["e"], row["e"]) and np.array_equal(got["v"], row["v"])
    assert_scores("dna", got, O.Oracle(cx).batch("dna", poses, nthreads=8))


def test_pose_order_is_bitwise_neutral(small):
    """Longest-first dispatch (launches of 257..4,096 poses) changes only the
    order workgroups start in: every pose's scores are the same bits with it
    on and off, with and without desolvation."""
    cx, orc, _ = small
    poses = synth.make_poses(4096, cx.meta["rec_axes"], 77)
    other = synth.make_poses(4096, cx.meta["rec_axes"], 78)
    ctx = lklight.Context(cx, desolv=True)
    # Between the reference and the ordered launch the device buffers hold
    # another pose set's results, so a pose the order left out would show
    # them (300 poses: below the pre-pass threshold, the same launch; 257
    # and 4,096: the smallest and largest ordered launches, whose per-XCD
    # ranges are 32-33 and exactly 512 poses).
    for flags, n in ((lklight.family_flags("dna"), 1500), (lklight.family_flags("dna"), 2600),
                     (lklight.family_flags("dna"), 300), (lklight.family_flags("pydock"), 257),
                     (lklight.family_flags("vdw"), 4096),
                     (lklight.F_FAR | lklight.F_ELEC | lklight.F_DESOLV, 700)):
        ctx.set_pose_order(0)
        ref = run(ctx, poses[:n], flags)
        run(ctx, other[:n], flags)
        ctx.set_pose_order(1)
        got = run(ctx, poses[:n], flags)
        say("pose order on/off, flags", flags, "poses", n)
        for k in ("e", "v", "s"):
            if ref[k] is not None:
                assert np.array_equal(got[k], ref[k]), (n, k)
    ctx.set_pose_order(1)
    ref = orc.batch("dna", poses[:600], nthreads=8)
    assert_scores("dna", run(ctx, poses[:600], lklight.family_flags("dna")), ref)
    ctx.close()


def test_determinism_and_batch_split(small):
    cx, orc, poses = small
    say("determinism (multi-tile poses: accumulators must re-zero between launches)")
    ctx = lklight.Context(cx)
    ctx.set_tiles(4.0, 64, 64)
    a = run(ctx, poses, lklight.FLAGS_DNA)
    b = run(ctx, poses, lklight.FLAGS_DNA)
    c = run(ctx, poses[:17], lklight.FLAGS_DNA)
    d = run(ctx, poses[17:], lklight.FLAGS_DNA)
    for k in ("e", "v"):
        assert np.array_equal(a[k], b[k])
        assert np.array_equal(a[k], np.concatenate([c[k], d[k]]))
    ctx.close()


def _f(a):
    return _capi.ptr(np.ascontiguousarray(a, np.float32), C.c_float)


def test_stateless_wrappers(small):
    """lkl_batch_score / lkl_cpydock_solv / lkl_full_score / lkl_far_field with the
    reference argument lists (full_score.cu:359-371, 571-581, 140-150,
    far_field.cu:59-63)."""
    cx, orc, poses = small
    L = lklight.lib()
    keep = {}

    def f(name, a, dt=np.float32):
        keep[name] = np.ascontiguousarray(a, dt)
        return keep[name]

    phi = f("phi", cx.phi)
    rx, re_, rs, rr = f("rx", cx.rec.xyz), f("re", cx.rec.ele), f("rs", cx.rec.seps), f("rr", cx.rec.rad)
    rh = f("rh", cx.rec.heavy, np.uint8)
    lx, le, ls, lr = f("lx", cx.lig.xyz), f("le", cx.lig.ele), f("ls", cx.lig.seps), f("lr", cx.lig.rad)
    lh = f("lh", cx.lig.heavy, np.uint8)
    cs, ca, nc, co = O.crec_build(cx.rec.xyz)
    P = np.ascontiguousarray(poses, np.float64)
    fp = lambda a: _capi.ptr(a, C.c_float)  # noqa: E731
    up = lambda a: _capi.ptr(a, C.c_ubyte)  # noqa: E731
    ip = lambda a: _capi.ptr(a, C.c_int)  # noqa: E731
    dp = lambda a: _capi.ptr(a, C.c_double)  # noqa: E731
    ctx = lklight.Context(cx, desolv=True)
    for method in ("dna", "pydock", "vdw"):
        say("lkl_batch_score", method)
        flags = lklight.family_flags(method)
        out = np.zeros((P.shape[0], 2))
        rc = L.lkl_batch_score(fp(phi), *cx.dims, *[float(v) for v in cx.origin], 0.5,
                               fp(rx), fp(re_), fp(rs), fp(rr), up(rh), cx.rec.n, ip(cs), ip(ca),
                               *nc, *co, 10.0, fp(lx), dp(P), fp(le), fp(ls), fp(lr), up(lh),
                               cx.lig.n, P.shape[0], flags, dp(out))
        assert rc == 0, lklight.last_error()
        ref = run(ctx, P, flags)
        assert np.array_equal(out[:, 0], ref["e"]) and np.array_equal(out[:, 1], ref["v"])
    say("lkl_cpydock_solv")
    rf, rd, ra = f("rf", cx.rec.flag, np.uint8), f("rd", cx.rec.des), f("ra", cx.rec.asa)
    lf, ld, la = f("lf", cx.lig.flag, np.uint8), f("ld", cx.lig.des), f("la", cx.lig.asa)
    outS = np.zeros(P.shape[0])
    rc = L.lkl_cpydock_solv(fp(rx), up(rf), cx.rec.n, ip(cs), ip(ca), *nc, *co, 10.0, fp(lx), dp(P),
                            up(lf), fp(le), cx.lig.n, P.shape[0], fp(rd), fp(ra), fp(ld), fp(la),
                            dp(outS))
    assert rc == 0, lklight.last_error()
    assert np.array_equal(outS, run(ctx, P, lklight.F_DESOLV)["s"])
    ctx.close()
    say("lkl_full_score / lkl_far_field")
    for k in range(3):
        tl = np.float32(orc.transform(P[k]))  # pre-transformed f32 coordinates
        oe, ov = np.zeros(cx.lig.n, np.float32), np.zeros(cx.lig.n, np.float32)
        rc = L.lkl_full_score(fp(phi), *cx.dims, *[float(v) for v in cx.origin], 0.5, fp(rx),
                              fp(re_), fp(rs), fp(rr), up(rh), cx.rec.n, ip(cs), ip(ca), *nc, *co,
                              10.0, fp(tl), fp(le), fp(ls), fp(lr), up(lh), cx.lig.n, fp(oe), fp(ov))
        assert rc == 0, lklight.last_error()
        lig_t = synth.Mol(xyz=tl.astype(np.float64), ele=cx.lig.ele, rad=cx.lig.rad,
                          seps=cx.lig.seps, heavy=cx.lig.heavy, hyd=cx.lig.hyd, des=cx.lig.des,
                          asa=cx.lig.asa)
        cxt = synth.Complex(rec=cx.rec, lig=lig_t)
        cxt.set_field(cx.phi, cx.origin, cx.dims)
        ot = O.Oracle(cxt)
        ref = ot.parts("dna", [1, 0, 0, 0, 0, 0, 0])
        sc_g = -(oe.astype(np.float64).sum() * 83.0 + ov.astype(np.float64).sum())
        sc_o = O.fold("dna", ref.e_raw, ref.v)
        assert abs(sc_g - sc_o) <= 1e-5 * max(abs(sc_o), FLOOR)
        res = C.c_double(0.0)
        rc = L.lkl_far_field(fp(phi), *cx.dims, *[float(v) for v in cx.origin], 0.5, fp(tl),
                             fp(le), cx.lig.n, C.byref(res))
        assert rc == 0, lklight.last_error()
        far = sum(cx.lig.ele[j] * ot.field.sample(*tl[j].astype(np.float64))
                  for j in range(cx.lig.n) if cx.lig.ele[j] != 0)
        # per-atom fixed-point quantisation: <= 2^-38 per atom
        assert abs(res.value - far) <= 1e-9 * max(1.0, abs(far))
    L.lkl_cache_clear()  # the cache keys on host pointers that die with this test


def test_edge_cases(small):
    cx, orc, poses = small
    say("edge cases")
    ctx = lklight.Context(cx, desolv=True)
    # a single pose; a pose far from everything (no near pairs, far field off-grid)
    far = np.array([[1, 0, 0, 0, 500.0, -400.0, 300.0]])
    r = run(ctx, far, lklight.FLAGS_DNA)
    assert r["e"][0] == 0.0 and r["v"][0] == 0.0
    one = run(ctx, poses[:1], lklight.FLAGS_DNA)
    ref = orc.batch("dna", poses[:1])
    assert_scores("dna", one, ref)
    # ligand straddling the far-field box edge: 40 A outside the receptor box
    edge = poses[:8].copy()
    edge[:, 4] += 40.0
    assert_scores("pydock", run(ctx, edge, lklight.FLAGS_PYDOCK), orc.batch("pydock", edge))
    # errors: N = 0 through the blocking call, unknown flag bits, F_FAR without a field
    with pytest.raises(lklight.GpuError):
        ctx.score(np.zeros((0, 7)), lklight.FLAGS_DNA)
    with pytest.raises(lklight.GpuError, match="unknown flag"):
        ctx.score(poses[:2], 0x100)
    ctx.close()
    nofield = lklight.Context(cx, with_field=False)
    with pytest.raises(lklight.GpuError, match="no far field"):
        nofield.score(poses[:2], lklight.FLAGS_DNA)
    with pytest.raises(lklight.GpuError, match="no desolvation"):
        nofield.score(poses[:2], lklight.F_DESOLV)
    nofield.close()


def test_config1_parity(gpu):
    """The 1,500-atom receptor / 800-atom ligand complex, all four families, 50 poses."""
    cx = synth.make_config("c1")
    say("c1 oracle field")
    phi, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    cx.set_field(phi, o, d)
    poses = synth.make_poses(50, cx.meta["rec_axes"], 42)
    orc = O.Oracle(cx)
    ctx = lklight.Context(cx, desolv=True)
    for method in ("dna", "vdw", "pydock", "cpydock"):
        say("c1", method)
        flags = lklight.family_flags(method) | (lklight.F_DESOLV if method == "cpydock" else 0)
        assert_scores(method, run(ctx, poses, flags), orc.batch(method, poses, nthreads=8))
    ctx.close()


@pytest.fixture(scope="module")
def c2(gpu):
    cx = synth.make_config("c2")
    say("c2 field build (GPU)")
    cx.build_field()
    return cx, O.Oracle(cx)


@pytest.mark.parametrize("method", ["dna", "pydock", "vdw"])
def test_config2_full_size(c2, method):
    """12,625-atom receptor / 1,500-atom ligand at full size, one 1,200-pose
    launch per family: the oracle on 30 of its poses (phi built on the GPU,
    bit-equal to the oracle's by test_field_build_bitwise), size-independent
    properties on the full batch."""
    cx, orc = c2
    flags = lklight.family_flags(method)
    poses = synth.make_poses(1200, cx.meta["rec_axes"], 7)
    ctx = lklight.Context(cx, with_field=method != "vdw")
    t = lklight.tuned_tiles("c2-" + method)
    ctx.set_tiles(t["cell"], t["block"], t["chunk"], cell_x=t["cell_x"])
    say("c2", method, "score 1200 poses")
    full = run(ctx, poses, flags)
    sub = poses[::40]
    say("c2 oracle subset", sub.shape[0])
    assert_scores(method, {"e": full["e"][::40], "v": full["v"][::40]}, orc.batch(method, sub, nthreads=8,
                                                                               bound=True))
    # batch split and repeat invariance at full size
    again = run(ctx, poses[600:], flags)
    assert np.array_equal(again["e"], full["e"][600:]) and np.array_equal(again["v"], full["v"][600:])
    # the default layouts (neighbour lists, corner-packed phi) against the
    # cell-row walk sampling phi directly: the same bits at full size
    assert ctx.list_info()[0] > 0 and (ctx.phi_packed() or method == "vdw")
    ctx.set_list_spacing(0.0)
    ctx.set_phi_packed(False)
    say("c2", method, "score 1200 poses, cell-row walk and unpacked phi")
    rows = run(ctx, poses, flags)
    assert np.array_equal(rows["e"], full["e"]) and np.array_equal(rows["v"], full["v"])
    ctx.close()


def test_blocking_call_equals_async_and_times_kernels(c2):
    """The blocking host-buffer call (a pinned, device-mapped staging buffer
    that the kernels read the poses from and write the scores into; one
    synchronisation) and the asynchronous call on device buffers give the same
    bits, and the context's kernel timing reports both kernels of a
    longest-first launch.  Calls without the dispatch-order pre-pass (17 poses,
    and 5,000: above its range) read the staging buffer from the scoring
    kernel itself."""
    import torch
    cx, orc = c2
    poses = synth.make_poses(1200, cx.meta["rec_axes"], 11)
    ctx = lklight.Context(cx)
    ctx.set_timing(True)
    e, v, _ = ctx.score(poses, lklight.FLAGS_DNA)
    pre, ker = ctx.kernel_ms()
    assert pre > 0.0 and ker > 0.0
    d_p = torch.from_numpy(poses).cuda()
    d_o = torch.zeros((1200, 2), dtype=torch.float64, device="cuda")
    ctx.score_async(d_p.data_ptr(), 1200, lklight.FLAGS_DNA, d_o.data_ptr())
    torch.cuda.synchronize()
    o = d_o.cpu().numpy()
    assert np.array_equal(o[:, 0], e) and np.array_equal(o[:, 1], v)
    # a smaller call after a larger one reuses the staging buffer
    e2, v2, _ = ctx.score(poses[:17], lklight.FLAGS_DNA)
    assert np.array_equal(e2, e[:17]) and np.array_equal(v2, v[:17])
    # a call above the ordered range grows it and keeps the same bits
    big = np.concatenate([poses] * 4 + [poses[:200]])
    e3, v3, _ = ctx.score(big, lklight.FLAGS_DNA)
    assert np.array_equal(e3[:1200], e) and np.array_equal(v3[3600:4800], v) and np.array_equal(e3[4800:], e[:200])
    ctx.close()


def test_config3_cpydock(gpu):
    """8,000-atom receptor / 3,000-atom ligand (radii 1.5-2.0 A) with contact-SASA
    desolvation at full size (per-pose minima in LDS): oracle on a subset of a
    400-pose step."""
    cx = synth.make_config("c3")
    say("c3 field build (GPU)")
    cx.build_field()
    poses = synth.make_poses(400, cx.meta["rec_axes"], 3)
    ctx = lklight.Context(cx, desolv=True)
    flags = lklight.FLAGS_PYDOCK | lklight.F_DESOLV
    say("c3 score 400 poses")
    full = run(ctx, poses, flags)
    sub = poses[::25]
    say("c3 oracle subset", sub.shape[0])
    assert_scores("cpydock", {k: full[k][::25] for k in ("e", "v", "s")},
                  O.Oracle(cx).batch("cpydock", sub, nthreads=8))
    assert np.abs(full["s"]).max() > 0
    # poses split over ligand tiles (receptor minima merged in global memory
    # by the tiles, summed by each pose's last tile) give the same bits; the
    # second call of each shape, with fewer poses, finds the minima restored
    for block, chunk in ((512, 1024), (1024, 1024), (256, 512)):
        ctx.set_tiles(7.5, block, chunk, cell_x=1.25)
        say(f"c3 split tiles: block {block} chunk {chunk}")
        for n in (400, 150):
            r = run(ctx, poses[:n], flags)
            for k in ("e", "v", "s"):
                assert np.array_equal(r[k], full[k][:n]), (block, chunk, n, k)
    ctx.close()
