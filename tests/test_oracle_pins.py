"""Pins of the CPU oracle against what the reference's tests and the mathematics fix.

Each test is chosen so a plausible restatement mistake (dropped term, wrong
sign, transposed matrix, lost cells, wrong shell test) fails it.
"""
import math

import numpy as np
import pytest

from oracle import lkl_oracle as O
from lklight import synth


def _field_of(xyz, q):
    phi, o, d = O.field_build(np.asarray(xyz, float), np.asarray(q, float))
    return O.OField(phi, o, d), phi, o, d


def test_field_unit_charge_closed_form():
    """grid_dna.rs:216-232: one unit charge at the origin."""
    f, *_ = _field_of([[0.0, 0.0, 0.0]], [1.0])
    at12, atm12 = f.sample(12, 0, 0), f.sample(-12, 0, 0)
    assert abs(at12 - atm12) < 1e-6
    assert abs(at12 - 1 / 144) < 1e-4
    assert abs(f.sample(5, 0, 0)) < 1e-9
    assert abs(f.sample(28, 0, 0) - 1 / 784) < 1e-5
    assert abs(f.sample(35, 0, 0)) < 1e-9
    # off-axis, between nodes: 1/d^2 within interpolation error
    for p in [(7.3, 8.1, -4.4), (-15.2, 3.3, 9.9), (0.25, -20.75, 6.1)]:
        d2 = sum(c * c for c in p)
        assert abs(f.sample(*p) - 1 / d2) < 2e-4 * (144 / d2), p


def test_field_nodes_equal_shell_sum():
    """Every node holds the f32 sum (atom order) of q/d^2 over 10 < d <= 30,
    including nodes the x-slab pruning of grid_dna.rs:120-134 must not drop."""
    rng = np.random.default_rng(3)
    xyz = np.round(rng.uniform(-6, 6, (12, 3)) * 1024) / 1024
    q = np.float32(rng.normal(0, 0.5, 12)).astype(float)
    f, phi, o, d = _field_of(xyz, q)
    nx, ny, nz = d
    idx = rng.integers(0, [nx, ny, nz], size=(400, 3))
    for ix, iy, iz in idx:
        node = np.array([o[0] + ix * 0.5, o[1] + iy * 0.5, o[2] + iz * 0.5])
        acc = np.float32(0.0)
        for k in range(len(q)):
            dz = node[2] - xyz[k, 2]
            dy = node[1] - xyz[k, 1]
            dyz2 = dy * dy + dz * dz
            dx = node[0] - xyz[k, 0]
            d2 = dx * dx + dyz2
            if dyz2 < 900 and 100 < d2 <= 900:
                acc = np.float32(acc + np.float32(q[k] / d2))
        assert phi[(iz * ny + iy) * nx + ix] == acc


def test_field_superposition():
    rng = np.random.default_rng(4)
    xyz = np.round(rng.uniform(-5, 5, (6, 3)) * 1024) / 1024
    q = np.float32(rng.normal(0, 0.4, 6)).astype(float)
    _, phi_all, o, d = _field_of(xyz, q)
    _, phi_a, oa, da = _field_of(xyz, np.where(np.arange(6) < 3, q, 0.0))
    _, phi_b, ob, db = _field_of(xyz, np.where(np.arange(6) >= 3, q, 0.0))
    assert o == oa == ob and d == da == db
    assert np.allclose(phi_all, phi_a + phi_b, rtol=0, atol=1e-6)


def _xorshift_points():
    rng = [7]

    def nxt():
        r = rng[0]
        r ^= (r << 13) & 0xFFFFFFFFFFFFFFFF
        r ^= r >> 7
        r ^= (r << 17) & 0xFFFFFFFFFFFFFFFF
        rng[0] = r
        return (r % 200) / 10.0 - 10.0

    return np.array([[nxt(), nxt(), nxt()] for _ in range(300)])


def test_nearcells_matches_allpairs():
    """nearcell.rs:119-159 (xorshift64 seed 7, 300 points, cell 5 A)."""
    pts = _xorshift_points()
    d2 = ((pts[:, None] - pts[None]) ** 2).sum(-1)
    allpairs = int(np.triu(d2 <= 25.0, 1).sum())
    assert O.nearcells_pair_count(pts, 5.0) == allpairs


def test_crec_window_covers_all_pairs():
    """gpu_score.rs:73-145: every pair within 10 A lies in the 27-cell window."""
    cx = synth.make_config("tiny")
    rec = cx.rec.xyz
    cs, ca, nc, co = O.crec_build(rec)
    assert cs[-1] == len(rec) and sorted(ca.tolist()) == list(range(len(rec)))
    rng = np.random.default_rng(9)
    pts = rng.uniform(rec.min(0) - 12, rec.max(0) + 12, (300, 3))
    for p in pts:
        c = np.clip(np.floor((p - np.array(co, float)) / 10.0).astype(int), 0, np.array(nc) - 1)
        win = set()
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    i3 = c + [dx, dy, dz]
                    if (i3 < 0).any() or (i3 >= nc).any():
                        continue
                    cell = (i3[2] * nc[1] + i3[1]) * nc[0] + i3[0]
                    win.update(ca[cs[cell]:cs[cell + 1]].tolist())
        near = set(np.nonzero(((rec - p) ** 2).sum(1) <= 100.0)[0].tolist())
        assert near <= win


@pytest.fixture(scope="module")
def small():
    cx = synth.make_config("small")
    phi, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    cx.set_field(phi, o, d)
    poses = synth.make_poses(6, cx.meta["rec_axes"], 21)
    return cx, O.Oracle(cx), poses


@pytest.mark.parametrize("method", ["dna", "vdw", "pydock", "cpydock"])
def test_cell_path_equals_bruteforce(small, method):
    """Cell enumeration == all-pairs FP64 re-evaluation of the same sum
    (vdw.rs:351-378 asserts 1e-9 for its cell path vs the all-vs-all scan)."""
    cx, orc, poses = small
    for p in poses:
        a, b = orc.parts(method, p), orc.bruteforce(method, p)
        assert a.n_near == b.n_near
        assert abs(a.e_raw - b.e_raw) <= 1e-9 * max(1.0, abs(b.e_raw))
        assert abs(a.v - b.v) <= 1e-9 * max(1.0, abs(b.v))
        assert abs(a.s - b.s) <= 1e-9 * max(1.0, abs(b.s))


def test_near_terms_closed_form():
    """One receptor atom and one ligand atom at a known distance."""
    def mol(x, q, r, s, heavy=1):
        return synth.Mol(xyz=np.array([[x, 0.0, 0.0]]), ele=np.array([q]), rad=np.array([r]),
                         seps=np.array([s]), heavy=np.array([heavy], np.uint8),
                         hyd=np.array([0], np.int32), des=np.array([0.1]), asa=np.array([30.0]))
    for d, qi, qj in [(4.0, 0.5, -0.4), (2.0, 0.9, 0.9), (9.5, -0.3, 0.2), (1.4, 0.1, 0.2)]:
        rec, lig = mol(0.0, qi, 1.9, 0.3), mol(0.0, qj, 1.8, 0.35)
        cx = synth.Complex(rec=rec, lig=lig)
        orc = O.Oracle(cx)
        pose = [1, 0, 0, 0, d, 0, 0]
        rr = 3.7
        p6 = rr ** 6 / d ** 6
        lj = min(0.3 * 0.35 * (p6 * p6 - 2 * p6), 1.0)
        clash = 6.0 * (0.75 * rr - d) if d < 0.75 * rr else 0.0
        el = max(min(qi * qj / d ** 2, 4 / 332), -4 / 332)
        dna = orc.parts("dna", pose)
        assert math.isclose(dna.e_raw, el, rel_tol=1e-12)
        assert math.isclose(dna.v, lj + clash, rel_tol=1e-12)
        vdw = orc.parts("vdw", pose)
        assert math.isclose(vdw.v, lj, rel_tol=1e-12) and vdw.e_raw == 0.0
        py = orc.parts("pydock", pose)
        assert math.isclose(py.v, lj, rel_tol=1e-12) and math.isclose(py.e_raw, el, rel_tol=1e-12)
        # beyond 10 A nothing near
        far = orc.parts("dna", [1, 0, 0, 0, 10.5, 0, 0])
        assert far.v == 0.0 and far.e_raw == 0.0


def test_cpydock_desolvation_closed_form():
    """cpydock.rs:640-665 on a two-atom complex, with the C flag view of
    cpydock.rs:600-601 (index 0 is even: hydrogens[0] != 0 excludes it)."""
    def mol(hyd, asa, des):
        return synth.Mol(xyz=np.array([[0.0, 0.0, 0.0]]), ele=np.array([0.0]), rad=np.array([1.8]),
                         seps=np.array([0.3]), heavy=np.array([1], np.uint8),
                         hyd=np.array([hyd], np.int32), des=np.array([des]), asa=np.array([asa]))
    for d in (3.0, 5.5, 6.39, 6.41, 9.0):
        cx = synth.Complex(rec=mol(0, 20.0, 0.11), lig=mol(0, 4.0, -0.05))
        s = O.Oracle(cx).parts("cpydock", [1, 0, 0, 0, d, 0, 0]).s
        g = lambda asa: min(-10 * d + 65, asa) if d * d <= 6.4 * 6.4 else 0.0  # noqa: E731
        assert math.isclose(s, g(20.0) * 0.11 + g(4.0) * -0.05, rel_tol=1e-12, abs_tol=1e-15)
        # flag set on the receptor atom: neither side sees the pair
        cx2 = synth.Complex(rec=mol(1, 20.0, 0.11), lig=mol(0, 4.0, -0.05))
        assert O.Oracle(cx2).parts("cpydock", [1, 0, 0, 0, d, 0, 0]).s == 0.0
    # asa < 0 atoms contribute asa * des whatever the contact (clamp outside the test)
    cx3 = synth.Complex(rec=mol(0, -1.0, 0.2), lig=mol(0, 4.0, 0.0))
    assert math.isclose(O.Oracle(cx3).parts("cpydock", [1, 0, 0, 0, 50.0, 0, 0]).s, -0.2)


def test_far_field_vs_exact_shell_sum(small):
    """The grid far field approximates the exact 10-30 A pair sum.  Interpolation
    across the 10 A / 30 A shell edges is the error source, so the bound is 1 % of
    the absolute pair sum (a sign or factor mistake moves the result by ~2|exact|)."""
    cx, orc, poses = small
    for p in poses[:3]:
        lc = orc.transform(p)
        d2 = ((lc[:, None, :] - cx.rec.xyz[None]) ** 2).sum(-1)
        shell = (d2 > 100) & (d2 <= 900)
        terms = cx.lig.ele[:, None] * cx.rec.ele[None] / np.where(shell, d2, 1.0) * shell
        exact, mag = float(terms.sum()), float(np.abs(terms).sum())
        grid = sum(cx.lig.ele[j] * orc.field.sample(*lc[j]) for j in range(cx.lig.n))
        assert abs(grid - exact) <= 0.01 * mag, (grid, exact, mag)


def test_dna_grid_vs_exact(small):
    cx, orc, poses = small
    for p in poses[:3]:
        g = orc.parts("dna", p)
        grid_score = O.fold("dna", g.e_raw, g.v)
        exact = orc.dna_exact(p)
        assert abs(grid_score - exact) <= 0.02 * abs(exact)


def test_transform_kats():
    """qt.rs:384-393 rotation KAT, identity exactness, scipy convention check."""
    from scipy.spatial.transform import Rotation
    lig = synth.Mol(xyz=np.array([[1.0, 0.0, 0.0], [0.3, -2.5, 7.125]]), ele=np.zeros(2),
                    rad=np.ones(2), seps=np.ones(2), heavy=np.ones(2, np.uint8),
                    hyd=np.zeros(2, np.int32), des=np.zeros(2), asa=np.zeros(2))
    orc = O.Oracle(synth.Complex(rec=lig, lig=lig))
    out = orc.transform([0.707106781, 0.0, 0.707106781, 0.0, 0, 0, 0])
    assert np.allclose(out[0], [0.0, 0.0, -1.0], atol=1e-8)
    assert np.array_equal(orc.transform([1, 0, 0, 0, 0, 0, 0]), lig.xyz)
    rng = np.random.default_rng(1)
    for _ in range(5):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        t = rng.normal(size=3) * 10
        ref = Rotation.from_quat([q[1], q[2], q[3], q[0]]).apply(lig.xyz) + t
        assert np.allclose(orc.transform(np.r_[q, t]), ref, atol=1e-12)


def test_fold_scores():
    from lklight import fold_scores
    assert O.fold("dna", 2.0, 3.0) == -(2.0 * 83.0 + 3.0)
    assert O.fold("cpydock", 2.0, 3.0, 5.0) == -(2.0 * 83.0 + 0.3 - 5.0)
    assert fold_scores("vdw", 0.0, 3.0) == -3.0
    assert fold_scores("cpydock", 2.0, 3.0, 5.0) == -(166.0 + 0.30000000000000004) + 5.0


def test_screen_candidates_oracle_pins():
    """The oracle's restatement of lkl_screen_candidates: its mix is
    splitmix64's output function (seed-0 sequence of the published generator:
    e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f); candidate 0 of each
    glowworm is its base pose; perturbed quaternions are unit and lie within
    rot_sigma of the base per component before normalisation (so within a
    bounded angle of it); translations lie
    within trans_radius; the draws are uniform (mean 0, variance 1/3 of the
    radius squared over 20,000 candidates)."""
    golden = 0x9E3779B97F4A7C15
    seq = [int(O.splitmix64_mix(np.uint64((k * golden) % 2 ** 64))) for k in range(3)]
    assert seq == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    rng = np.random.default_rng(2)
    n, C, tr, rs = 625, 32, 2.0, 0.1
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    base = np.concatenate([q, rng.uniform(-20, 20, (n, 3))], 1)
    c = O.screen_candidates(base, C, tr, rs, 77).reshape(n, C, 7)
    assert np.array_equal(c[:, 0], base)
    p = c[:, 1:]
    assert np.allclose(np.linalg.norm(p[..., :4], axis=-1), 1.0, atol=1e-15)
    d = p[..., 4:] - base[:, None, 4:]
    assert np.abs(d).max() <= tr and abs(d.mean()) < 0.02 and abs(d.var() - tr * tr / 3) < 0.02
    # a perturbation of at most rs per component of a unit quaternion keeps
    # the normalised candidate within cos > (1 - 2 rs) / (1 + 2 rs) of it
    cosang = np.einsum("ijk,ijk->ij", p[..., :4], np.broadcast_to(base[:, None, :4], p[..., :4].shape))
    assert (cosang > (1 - 2 * rs) / (1 + 2 * rs)).all()
    assert np.array_equal(O.screen_candidates(base, C, tr, rs, 77), c.reshape(-1, 7))
    assert not np.array_equal(O.screen_candidates(base, C, tr, rs, 78), c.reshape(-1, 7))


def test_fp32_bound_magnitudes_and_soundness():
    """The oracle's rounding magnitudes (orc_parts e_abs, v_abs) equal the
    same sums over all pairs within 10 A computed directly in numpy (dna on
    the small complex), and the bound they give covers an f32 evaluation of
    the pair terms in the device's operation order (hi/lo coordinate split,
    d2 from a product and two fmas, f32 reciprocal and products): the f32
    electrostatic sum stays within 11u sum|E_ij| and the f32 LJ sum within
    89u of its magnitude on every pose."""
    cx = synth.make_config("small")
    orc = O.Oracle(cx)
    poses = synth.make_poses(12, cx.meta["rec_axes"], 21)
    r = orc.batch("dna", poses, bound=True)
    rx = np.asarray(cx.rec.xyz, float)
    f32 = np.float32
    u = 2.0 ** -24
    for k, pose in enumerate(poses):
        lx = orc.transform(pose)
        d = lx[:, None, :] - rx[None, :, :]
        d2 = (d * d).sum(-1)
        near = d2 <= 100.0
        qq = np.outer(cx.lig.ele, cx.rec.ele)[near]
        e64 = np.clip(qq / d2[near], -4 / 332, 4 / 332)
        assert np.isclose(np.abs(e64).sum(), r["mag"][k, 0], rtol=1e-12, atol=0)
        eps = np.outer(cx.lig.seps, cx.rec.seps)[near]
        rr = np.add.outer(cx.lig.rad, cx.rec.rad)[near]
        p6 = rr ** 6 / d2[near] ** 3
        lj = eps * (p6 * p6 - 2 * p6)
        keep = lj <= 2.0
        assert np.isclose((np.abs(eps) * (p6 * p6 + 2 * p6))[keep].sum(), r["mag"][k, 2], rtol=1e-12, atol=0)
        # f32 in the device's order (numpy's f32 division is correctly
        # rounded; the hardware reciprocal's extra ulp is inside the bound)
        xh = lx.astype(f32)
        xl = (lx - xh.astype(float)).astype(f32)
        dd = (xh[:, None, :] - rx.astype(f32)[None, :, :]) + xl[:, None, :]
        d2f = (dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1] + dd[..., 2] * dd[..., 2])[near]
        inv = f32(1.0) / d2f
        qf = np.outer(cx.lig.ele.astype(f32), cx.rec.ele.astype(f32))[near]
        ef = np.clip((qf * inv).astype(float), -4 / 332, 4 / 332)
        assert abs(ef.sum() - e64.sum()) <= 11 * u * np.abs(e64).sum()
        rrf = np.add.outer(cx.lig.rad.astype(f32), cx.rec.rad.astype(f32))[near]
        t = (rrf * rrf) * inv
        p6f = t * t * t
        ljf = (np.outer(cx.lig.seps.astype(f32), cx.rec.seps.astype(f32))[near] * (p6f * p6f - f32(2) * p6f))
        cap = np.minimum(ljf.astype(float), 1.0)
        assert abs(cap.sum() - np.minimum(lj, 1.0).sum()) <= 89 * u * r["mag"][k, 2] + 1e-12
