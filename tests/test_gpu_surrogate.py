"""HIP surrogate kernels against the CPU oracle: bf16 MFMA GEMM (lane layout,
transposed operands), pair-histogram features (exact integer counts and the
bf16 log1p bits), MLP forward, gradients and Adam."""
import numpy as np
import pytest

from lklight import surrogate as S
from lklight import synth
from oracle import lkl_oracle as O
from oracle import lkl_oracle_mlp as M


def _bits(a):
    return M.bf16_bits_f32(np.asarray(a, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["nn", "tn", "nt", "tt"])
def test_gemm_bf16_layouts(gpu, layout):
    """C = A B for every operand orientation; a transposed or mis-mapped MFMA
    lane layout fails the f64 comparison.  Bound: fp32 accumulation error over
    K products, 2^-22 * K * sum|a||b| (loose) — observed errors are far below."""
    rng = np.random.default_rng(7)
    Mm, N, K = 256, 384, 192
    A = rng.normal(size=(Mm, K))
    B = rng.normal(size=(K, N))
    Ab, Bb = _bits(A), _bits(B)
    Av, Bv = M.bf16_to_f64(Ab), M.bf16_to_f64(Bb)
    if layout[0] == "n":
        a_store, sam, sak = Ab, K, 1
    else:
        a_store, sam, sak = np.ascontiguousarray(Ab.T), 1, Mm
    if layout[1] == "n":
        b_store, sbk, sbn = Bb, N, 1
    else:
        b_store, sbk, sbn = np.ascontiguousarray(Bb.T), 1, K
    C = S.gemm_bf16(a_store, sam, sak, b_store, sbk, sbn, Mm, N, K)
    ref = Av @ Bv
    bound = 2.0 ** -22 * K * (np.abs(Av) @ np.abs(Bv))
    assert (np.abs(C - ref) <= bound).all(), np.abs(C - ref).max()
    assert np.abs(C - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(384, 256, 1024), (384, 256, 1056), (1024, 512, 6144)])
def test_gemm_split_k(gpu, shape):
    """K-contiguous operands take the 256 x 128-tile kernel: split over K into
    8, 2 and 1 chunks for these shapes (partials summed in a fixed order), a
    last row tile half past M when M = 384.  Same f64 comparison and bound as
    above."""
    rng = np.random.default_rng(11)
    Mm, N, K = shape
    A = rng.normal(size=(Mm, K))
    B = rng.normal(size=(K, N))
    Ab, Bb = _bits(A), _bits(B)
    Av, Bv = M.bf16_to_f64(Ab), M.bf16_to_f64(Bb)
    C = S.gemm_bf16(Ab, K, 1, np.ascontiguousarray(Bb.T), 1, K, Mm, N, K)
    ref = Av @ Bv
    bound = 2.0 ** -22 * K * (np.abs(Av) @ np.abs(Bv))
    assert (np.abs(C - ref) <= bound).all(), np.abs(C - ref).max()
    # deterministic: the same bits on a second call
    assert np.array_equal(C, S.gemm_bf16(Ab, K, 1, np.ascontiguousarray(Bb.T), 1, K, Mm, N, K))


@pytest.mark.gpu
def test_gemm_rejects_bad_shapes(gpu):
    a = np.zeros(128 * 64, np.uint16)
    with pytest.raises(Exception):
        S.gemm_bf16(a, 32, 1, a, 128, 1, 128, 128, 40)  # K not a multiple of 64
    with pytest.raises(Exception):
        S.gemm_bf16(a, 64, 1, a[:100], 128, 1, 128, 128, 64)  # B too short


@pytest.fixture(scope="module")
def c1():
    return synth.make_config("c1")


@pytest.mark.gpu
def test_features_exact_thresholds(gpu):
    ks = np.arange(1, 25)
    xs = np.concatenate([1.25 * ks, 1.25 * ks - 2.0 ** -12])
    cx = synth.make_config("tiny")
    cx.rec.xyz = np.zeros((1, 3))
    cx.rec.typ = np.zeros(1, np.int32)
    cx.rec.ele = np.zeros(1)
    cx.lig.xyz = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], 1)
    cx.lig.typ = np.zeros(len(xs), np.int32)
    cx.lig.ele = np.zeros(len(xs))
    f = S.Features(cx)
    got = f.counts(np.array([[1, 0, 0, 0, 0, 0, 0.0]]))[0]
    exp = O.feat_counts(cx.rec.xyz, cx.rec.group, cx.lig.xyz, cx.lig.group,
                        np.array([[1, 0, 0, 0, 0, 0, 0.0]]))[0]
    assert np.array_equal(got, exp) and got.sum() == 47
    f.close()


@pytest.mark.gpu
def test_features_match_oracle_c1(gpu, c1):
    poses = synth.make_poses(24, c1.meta["rec_axes"], 3)
    f = S.Features(c1)
    cnt = f.counts(poses)
    bits = f.bits(poses)
    ref = O.feat_counts(c1.rec.xyz, c1.rec.group, c1.lig.xyz, c1.lig.group, poses, nthreads=8)
    assert np.array_equal(cnt, ref)
    assert np.array_equal(bits, M.features_bits(ref))
    assert cnt.sum(1).min() > 0
    f.close()


@pytest.mark.gpu
def test_features_large_counts_and_batching(gpu):
    """Counts above the 4096-entry log1p table go through the f64 path; results
    do not depend on how poses are batched."""
    cx = synth.make_config("c1")
    cx.rec.typ = np.zeros(cx.rec.n, np.int32)  # one group: large counts per bin
    cx.lig.typ = np.zeros(cx.lig.n, np.int32)
    cx.rec.ele = np.zeros(cx.rec.n)
    cx.lig.ele = np.zeros(cx.lig.n)
    assert len(set(cx.rec.group)) == 1 and len(set(cx.lig.group)) == 1
    poses = synth.make_poses(6, cx.meta["rec_axes"], 11, shell=0.0)
    f = S.Features(cx)
    cnt = f.counts(poses)
    assert cnt.max() >= 4096
    ref = O.feat_counts(cx.rec.xyz, cx.rec.group, cx.lig.xyz, cx.lig.group, poses, nthreads=8)
    assert np.array_equal(cnt, ref)
    assert np.array_equal(f.bits(poses), M.features_bits(ref))
    assert np.array_equal(np.concatenate([f.counts(poses[:2]), f.counts(poses[2:])]), cnt)
    f.close()


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM, np.float32)
    for k, (o, n) in M.offsets().items():
        fan = {"W1": M.NF, "W2": M.D, "W3": M.D, "W4": M.D}.get(k, M.D)
        lim = np.sqrt(6.0 / fan) * (0.1 if k[0] == "b" else 1.0) * scale
        p[o:o + n] = rng.uniform(-lim, lim, n)
    return p


@pytest.fixture(scope="module")
def feats(c1):
    poses = synth.make_poses(256, c1.meta["rec_axes"], 5)
    f = S.Features(c1)
    x = f.bits(poses)
    f.close()
    return poses, x


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(gpu, feats):
    """Device forward (bf16 operands, fp32 MFMA accumulation) against the f64
    forward of the same bf16 operands, held to the spread of other valid fp32
    evaluations (oracle forward_f32: the same operands and bf16 activation
    roundings with fp32 products summed in 6 permuted orders): the largest
    error within 4x the largest error of those evaluations, the RMS within 3x
    theirs."""
    _, x = feats
    net = S.MLP(0)
    p = _params(1)
    net.params = p
    y = net.forward(x)
    yr, _ = M.forward(p, x, round=True)
    sims = np.array([M.forward_f32(p, x, s)[0] - yr for s in range(6)])
    err = y - yr
    print("[gpu-test] forward max err %.3e (fp32 evaluations %.3e), rms %.3e (%.3e)"
          % (np.abs(err).max(), np.abs(sims).max(), np.sqrt((err ** 2).mean()), np.sqrt((sims ** 2).mean())))
    assert np.abs(err).max() <= 4.0 * np.abs(sims).max()
    assert np.sqrt((err ** 2).mean()) <= 3.0 * np.sqrt((sims ** 2).mean())
    # the create() initialisation is also a valid network
    net2 = S.MLP(123)
    assert np.isfinite(net2.forward(x)).all()
    net.close()
    net2.close()


@pytest.mark.gpu
def test_mlp_gradients_match_oracle(gpu, feats):
    """Per-tensor gradient error of the device (bf16 dZ GEMM operands, fp32
    accumulation, split-K partials) against the f64 gradient of the same bf16
    operands, held to 4x the largest error of 4 other valid fp32 evaluations
    (oracle loss_grad_f32: permuted fp32 summation orders, the kernels' bf16
    roundings).  The loss within the same spread."""
    _, x = feats
    rng = np.random.default_rng(2)
    t = rng.normal(size=x.shape[0]).astype(np.float32)
    net = S.MLP(0)
    p = _params(3)
    net.params = p
    loss = net.grad(x, t)
    g = net.grads()
    lr, gr, _ = M.loss_grad(p, x, t, round=True)
    sims = [M.loss_grad_f32(p, x, t, s) for s in range(4)]
    assert abs(loss - lr) <= 4.0 * max(abs(sl - lr) for sl, _, _ in sims) + 1e-7 * lr
    for k, (o, n) in M.offsets().items():
        a, b = g[o:o + n].astype(np.float64), gr[o:o + n]
        e = np.linalg.norm(a - b)
        es = max(np.linalg.norm(sg[o:o + n] - b) for _, sg, _ in sims)
        print("[gpu-test] %s: error norm %.3e, fp32 evaluations %.3e (relative %.2e)"
              % (k, e, es, e / np.linalg.norm(b)))
        assert e <= 4.0 * es + 1e-7 * np.linalg.norm(b), (k, e, es)
    net.close()


@pytest.mark.gpu
def test_adam_steps_match_oracle(gpu, feats):
    _, x = feats
    t = np.random.default_rng(4).normal(size=x.shape[0]).astype(np.float32)
    net = S.MLP(0)
    p = _params(5)
    net.params = p
    ref_p = p.astype(np.float64)
    m = np.zeros(M.NPARAM)
    v = np.zeros(M.NPARAM)
    for step in (1, 2, 3):
        net.grad(x, t)
        g = net.grads().astype(np.float64)
        p_before = net.params.astype(np.float64)
        net.adam(1e-3)
        p_after = net.params.astype(np.float64)
        exp, m, v = M.adam(p_before, g, m, v, step, 1e-3)
        assert np.abs(p_after - exp).max() <= 1e-6, step
        if step == 1:  # closed form: m_hat = g, v_hat = g^2
            cf = p_before - 1e-3 * g / (np.abs(g) + 1e-8)
            assert np.abs(p_after - cf).max() <= 1e-6
    net.close()


@pytest.mark.gpu
def test_training_reduces_loss(gpu, feats):
    """A few Adam steps on a fixed batch lower the loss (the kernels train)."""
    _, x = feats
    t = (M.bf16_to_f64(x) @ np.random.default_rng(6).normal(size=M.NF) / 2000.0).astype(np.float32)
    t = (t - t.mean()) / t.std()
    net = S.MLP(9)
    l0 = net.grad(x, t)
    for _ in range(30):
        net.adam(3e-4)
        last = net.grad(x, t)
    assert last < 0.5 * l0, (l0, last)
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("h", [0.5, 1.0])
def test_grid_table_and_features_match_oracle(gpu, c1, h):
    """Grid mode: the device count table equals the oracle's node counts (random
    nodes and nodes near the receptor), and the pose features equal the
    oracle's nearest-node sums bit for bit."""
    f = S.Features(c1, grid=h)
    o, dims, sp, n = f.grid_info()
    oo, od, on = O.feat_grid_dims(c1.rec.xyz, h)
    assert sp == h and dims == od and n == on and np.array_equal(o, oo)
    rng = np.random.default_rng(1)
    picks = list(rng.integers(0, n, 24))
    ctr = ((np.array(dims) // 2)[::-1] * np.array([dims[0] * dims[1], dims[0], 1])).sum()
    picks += [int(ctr), int(ctr) + 1, int(ctr) + dims[0]]
    for node in picks:
        got = f.grid_nodes(node, 1)[0]
        ix, iy, iz = node % dims[0], (node // dims[0]) % dims[1], node // (dims[0] * dims[1])
        assert np.array_equal(got, O.feat_grid_node(c1.rec.xyz, c1.rec.group, h, ix, iy, iz)), node
    poses = synth.make_poses(12, c1.meta["rec_axes"], 8)
    cnt = f.counts(poses)
    ref = O.feat_grid_counts(c1.rec.xyz, c1.rec.group, c1.lig.xyz, c1.lig.group, poses, h, nthreads=8)
    assert np.array_equal(cnt, ref)
    assert np.array_equal(f.bits(poses), M.features_bits(ref))
    # batching and ordering of poses do not matter
    perm = rng.permutation(len(poses))
    assert np.array_equal(f.counts(poses[perm]), cnt[perm])
    f.close()


@pytest.mark.gpu
def test_grid_rejects_bad_spacing(gpu, c1):
    with pytest.raises(Exception):
        S.Features(c1, grid=0.3)
