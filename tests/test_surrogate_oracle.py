"""Pins of the surrogate's CPU oracle: the C pair-histogram counts against an
exact pure-Python restatement, closed-form threshold cases and invariants; the
numpy MLP against central finite differences and torch float64 autograd; the
Adam step against its closed form; the bf16 rounding helpers against exhaustive
neighbour checks."""
from fractions import Fraction
import math

import numpy as np
import pytest

from oracle import lkl_oracle as O
from oracle import lkl_oracle_mlp as M


def _f32(x):
    return float(np.float32(x))


def _round_f32(q: Fraction) -> float:
    """Exact rational -> nearest f32 (ties to even)."""
    x = np.float32(float(q))
    best = None
    for c in (np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        key = (d, int(np.float32(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, float(c))
    return best[1]


def _fmaf(a, b, c):
    return _round_f32(Fraction(a) * Fraction(b) + Fraction(c))


def _counts_python(rxyz, rg, lxyz, lg, pose):
    base = np.ascontiguousarray(lxyz, np.float64)
    out = np.zeros(len(base) * 3)
    O.lib().orc_transform(O._p(base), len(base), O._p(np.ascontiguousarray(pose, np.float64)), O._p(out))
    cnt = np.zeros(6144, np.int64)
    for j in range(len(base)):
        x, y, z = (_f32(v) for v in out[3 * j:3 * j + 3])
        for i in range(len(rxyz)):
            dx = _f32(x - _f32(rxyz[i][0]))
            dy = _f32(y - _f32(rxyz[i][1]))
            dz = _f32(z - _f32(rxyz[i][2]))
            d2 = _fmaf(dx, dx, _fmaf(dy, dy, _f32(dz * dz)))
            if d2 < 900.0:
                b = max(k for k in range(24) if 1.5625 * k * k <= d2)
                cnt[((int(rg[i]) % 16) * 16 + int(lg[j]) % 16) * 24 + b] += 1
    return cnt


def _small_system(seed, nr=40, nl=25):
    rng = np.random.default_rng(seed)
    r = np.round(rng.uniform(-14, 14, (nr, 3)) * 1024) / 1024
    l = np.round(rng.uniform(-8, 8, (nl, 3)) * 1024) / 1024
    rg = rng.integers(0, 20, nr).astype(np.uint8)
    lg = rng.integers(0, 20, nl).astype(np.uint8)
    q = rng.normal(size=(3, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.uniform(-12, 12, (3, 3))
    poses = np.concatenate([q, t], 1).astype(np.float32).astype(np.float64)
    return r, rg, l, lg, poses


def test_feature_counts_match_exact_python():
    r, rg, l, lg, poses = _small_system(3)
    got = O.feat_counts(r, rg, l, lg, poses, nthreads=2)
    for k in range(len(poses)):
        ref = _counts_python(r, rg, l, lg, poses[k])
        assert np.array_equal(got[k].astype(np.int64), ref), k
    assert got.sum() > 0


def test_feature_bins_at_exact_thresholds():
    """Receptor atom at the origin, ligand atoms on the x axis at 1.25 k (d2 =
    1.5625 k^2 exactly, bin k; k = 24 is 30 A and excluded) and 2^-12 A short of
    it (bin k - 1); identity pose."""
    ks = np.arange(1, 25)
    xs = np.concatenate([1.25 * ks, 1.25 * ks - 2.0 ** -12])
    lig = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], 1)
    lg = np.zeros(len(xs), np.uint8)
    rec = np.zeros((1, 3))
    rg = np.zeros(1, np.uint8)
    pose = np.array([[1, 0, 0, 0, 0, 0, 0]], np.float64)
    got = O.feat_counts(rec, rg, lig, lg, pose)[0]
    exp = np.zeros(6144, np.int64)
    for k in range(1, 24):
        exp[k] += 1          # on the threshold
        exp[k - 1] += 1      # just below it
    exp[23] += 1             # 30 A - 2^-12 is in the last bin; 30 A itself is out
    assert np.array_equal(got.astype(np.int64), exp)


def test_feature_invariants():
    r, rg, l, lg, poses = _small_system(9, nr=300, nl=120)
    a = O.feat_counts(r, rg, l, lg, poses)
    # atom order does not matter
    pr, pl = np.random.default_rng(1).permutation(len(r)), np.random.default_rng(2).permutation(len(l))
    b = O.feat_counts(r[pr], rg[pr], l[pl], lg[pl], poses)
    assert np.array_equal(a, b)
    # moving every receptor atom into one group keeps the per-(lig group, bin) totals
    c = O.feat_counts(r, np.zeros_like(rg), l, lg, poses)
    assert np.array_equal(a.reshape(-1, 16, 16, 24).sum(1), c.reshape(-1, 16, 16, 24)[:, 0])
    # total = pairs within 30 A (no pair of this system is near the 30 A edge)
    for k in range(len(poses)):
        lc = np.zeros(len(l) * 3)
        O.lib().orc_transform(O._p(np.ascontiguousarray(l)), len(l), O._p(poses[k].copy()), O._p(lc))
        d2 = ((lc.reshape(-1, 1, 3) - r.reshape(1, -1, 3)) ** 2).sum(-1)
        assert np.abs(d2 - 900).min() > 1e-3
        assert a[k].sum() == (d2 < 900).sum()


def test_bf16_helpers():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-3, 3, 20000)).astype(np.float32)
    import torch
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(M.bf16_bits_f32(x), ref)
    # one rounding from f64: the chosen bf16 is nearest, ties to even
    d = np.abs(rng.standard_normal(5000)) * 20 + 1e-3
    d = np.concatenate([d, np.log1p(np.arange(1, 5000, dtype=np.float64))])
    b = M.bf16_bits_f64(d)
    v, lo, hi = M.bf16_to_f64(b), M.bf16_to_f64(b - 1), M.bf16_to_f64(b + 1)
    err = np.abs(v - d)
    assert (err <= np.abs(lo - d)).all() and (err <= np.abs(hi - d)).all()
    tie = (err == np.abs(lo - d)) | (err == np.abs(hi - d))
    assert ((b[tie] & 1) == 0).all()
    assert M.bf16_bits_f64(np.array([0.0]))[0] == 0


def _tiny_net(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    for k, (o, n) in M.offsets().items():
        fan = {"W1": M.NF, "W2": M.D, "W3": M.D, "W4": M.D}.get(k, M.D)
        p[o:o + n] = rng.uniform(-1, 1, n) * math.sqrt(6.0 / fan) * (0.1 if k[0] == "b" else 1.0)
    x = np.log1p(rng.poisson(2.0, (6, M.NF)).astype(np.float64))
    t = rng.normal(size=6)
    return p, x, t


def test_mlp_gradient_central_differences():
    p, x, t = _tiny_net(1)
    loss, g, _ = M.loss_grad(p, x, t, round=False)
    rng = np.random.default_rng(5)
    for name, (o, n) in M.offsets().items():
        for idx in rng.choice(n, size=min(n, 6), replace=False):
            k = o + idx
            h = 1e-5 * max(1.0, abs(p[k]))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (M.loss_grad(pp, x, t, round=False)[0] - M.loss_grad(pm, x, t, round=False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (name, idx, fd, g[k])


def test_mlp_gradient_matches_torch_autograd():
    import torch
    p, x, t = _tiny_net(2)
    loss, g, y = M.loss_grad(p, x, t, round=False)
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in M.split(p).items()}
    X = torch.tensor(x)
    h = torch.relu(X @ P["W1"] + P["b1"])
    h = torch.relu(h @ P["W2"] + P["b2"])
    h = torch.relu(h @ P["W3"] + P["b3"])
    yt = h @ P["W4"] + P["b4"][0]
    lt = 0.5 * ((yt - torch.tensor(t)) ** 2).mean()
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-12 * max(1.0, abs(loss))
    for k, (o, n) in M.offsets().items():
        ref = P[k].grad.numpy().ravel()
        assert np.allclose(g[o:o + n], ref, rtol=1e-10, atol=1e-14), k


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p0, g = rng.normal(size=1000), rng.normal(size=1000)
    p1, m, v = M.adam(p0, g, np.zeros(1000), np.zeros(1000), 1, 1e-3)
    assert np.allclose(p1, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_forward_rounding_modes_agree_roughly():
    """The bf16-rounded forward stays within bf16 noise of the f64 forward."""
    p, x, _ = _tiny_net(4)
    xb = M.features_bits(np.expm1(x))  # the bf16 features of the same counts
    y64, _ = M.forward(p, M.bf16_to_f64(xb), round=False)
    yb, _ = M.forward(p, xb, round=True)
    assert np.abs(yb - y64).max() <= 0.05 * np.abs(y64).max() + 1e-6


def test_fp32_evaluations_are_the_same_network():
    """The permuted-order fp32 evaluations (the tolerance reference of the
    device checks) compute the same function as the f64 oracle: outputs and
    every gradient tensor within 4e-3 relative (the fp32 evaluations, like the
    kernels, round the dZ operands to bf16: 2^-9 each), and different
    summation orders give (slightly) different bits, as intended."""
    p, x, t = _tiny_net(6)
    xb = M.features_bits(np.expm1(x))
    y64, _ = M.forward(p, xb, round=True)
    ya, _ = M.forward_f32(p, xb, 1)
    yb, _ = M.forward_f32(p, xb, 2)
    assert np.abs(ya - y64).max() <= 1e-3 * np.abs(y64).max()
    _, g64, _ = M.loss_grad(p, xb, t, round=True)
    _, ga, _ = M.loss_grad_f32(p, xb, t, 1)
    _, gb, _ = M.loss_grad_f32(p, xb, t, 2)
    for k, (o, n) in M.offsets().items():
        ref = np.linalg.norm(g64[o:o + n])
        assert np.linalg.norm(ga[o:o + n] - g64[o:o + n]) <= 4e-3 * ref + 1e-12, k
    assert not np.array_equal(ga, gb)


@pytest.mark.parametrize("h", [0.5, 1.0])
def test_grid_node_counts_are_pair_counts_at_the_node(h):
    """A node's tabulated counts equal the exact pair histogram of a single
    ligand atom placed at the node (identity pose, ligand group 0)."""
    r, rg, _, _, _ = _small_system(4, nr=200, nl=1)
    o, dims, n = O.feat_grid_dims(r, h)
    assert n == dims[0] * dims[1] * dims[2]
    assert all(o[a] == math.floor(r[:, a].min() - 30 - h) for a in range(3))
    rng = np.random.default_rng(0)
    for _ in range(12):
        ix, iy, iz = (int(rng.integers(0, d)) for d in dims)
        node = O.feat_grid_node(r, rg, h, ix, iy, iz)
        p = np.array([[o[0] + ix * h, o[1] + iy * h, o[2] + iz * h]])
        ref = O.feat_counts(r, rg, p, np.zeros(1, np.uint8), np.array([[1, 0, 0, 0, 0, 0, 0.0]]))[0]
        assert np.array_equal(node, ref.reshape(16, 16, 24)[:, 0, :].ravel())


@pytest.mark.parametrize("h", [0.5, 1.0])
def test_grid_features_are_sums_of_nearest_node_counts(h):
    r, rg, l, lg, poses = _small_system(6, nr=150, nl=40)
    got = O.feat_grid_counts(r, rg, l, lg, poses, h, nthreads=2)
    o, dims, _ = O.feat_grid_dims(r, h)
    of = o.astype(np.float32)
    inv = np.float32(1.0 / h)
    for k in range(len(poses)):
        lc = np.zeros(len(l) * 3)
        O.lib().orc_transform(O._p(np.ascontiguousarray(l)), len(l), O._p(poses[k].copy()), O._p(lc))
        xf = lc.reshape(-1, 3).astype(np.float32)
        # (x - o) * 2^k is exact in f32, so the fmaf equals a rounded add here
        ix = np.floor(((xf - of) * inv) + np.float32(0.5)).astype(np.int64)
        exp = np.zeros((16, 16, 24), np.int64)
        for j in range(len(l)):
            if (ix[j] < 0).any() or (ix[j] >= np.array(dims)).any():
                continue
            c = O.feat_grid_node(r, rg, h, *ix[j]).reshape(16, 24)
            assert np.abs(xf[j] - (o + ix[j] * h)).max() <= h / 2 + 1e-6  # nearest node
            exp[:, int(lg[j]) % 16, :] += c
        assert np.array_equal(got[k].astype(np.int64), exp.ravel()), k
