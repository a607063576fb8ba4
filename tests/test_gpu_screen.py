"""Surrogate-screened GSO: the candidate and selection kernels against their
host restatements, the screen's invariants (committed scores are exact, the
plan follows the reference's GSO draws), and what the trained network gains
over the same screen without it."""
import numpy as np
import pytest
import torch

import lklight
from lklight import _capi, screen, synth
from lklight import surrogate as S
from lklight.gso import GSO
from oracle import lkl_oracle as O


@pytest.mark.gpu
def test_candidates_and_select_match_host(gpu):
    rng = np.random.default_rng(3)
    n, C, k = 37, 32, 2
    base = np.concatenate([synth.random_quaternions(n, rng), rng.uniform(-20, 20, (n, 3))], 1)
    d_base = torch.from_numpy(base).cuda()
    d_cand = torch.empty((n * C, 7), dtype=torch.float64, device="cuda")
    L = _capi.lib()
    _capi.check(L.lkl_screen_candidates(d_base.data_ptr(), n, C, 2.0, 0.1, 12345, d_cand.data_ptr(), None), "cand")
    torch.cuda.synchronize()
    got = d_cand.cpu().numpy()
    ref = O.screen_candidates(base, C, 2.0, 0.1, 12345)
    assert np.abs(got - ref).max() <= 4e-16 * max(1.0, np.abs(ref).max())
    assert np.array_equal(got[::C], base)
    # selection: top-k per glowworm, ties to the lower index
    pred = rng.integers(0, 5, n * C).astype(np.float32)  # many ties
    d_pred = torch.from_numpy(pred).cuda()
    d_pick = torch.empty(n * k, dtype=torch.int32, device="cuda")
    d_sel = torch.empty((n * k, 7), dtype=torch.float64, device="cuda")
    _capi.check(L.lkl_screen_select(d_pred.data_ptr(), d_cand.data_ptr(), n, C, k, d_pick.data_ptr(),
                                    d_sel.data_ptr(), None), "select")
    torch.cuda.synchronize()
    pick = d_pick.cpu().numpy().reshape(n, k)
    exp = np.argsort(-pred.reshape(n, C), axis=1, kind="stable")[:, :k]
    assert np.array_equal(pick, exp)
    assert np.array_equal(d_sel.cpu().numpy().reshape(n, k, 7), got.reshape(n, C, 7)[np.arange(n)[:, None], exp])


@pytest.fixture(scope="module")
def trained():
    cx = synth.make_config("c1")
    cx.build_field()
    ctx = lklight.Context(cx)
    feats = S.Features(cx, grid=0.5)
    sizes = [20] * 10
    pos_tr = synth.make_swarm_positions(10, 20, cx.meta["rec_axes"], 4242, lig_radius=cx.meta["lig_axes"][0])
    P = np.concatenate([screen.collect_poses(ctx, "dna", pos_tr, 4242, sizes, 60, 32, 2.0, 0.1, every=2),
                        synth.make_poses(20000, cx.meta["rec_axes"], 77)])
    sc = np.concatenate([screen.exact_scores(ctx, "dna", P[i:i + 65536]) for i in range(0, len(P), 65536)])
    X = screen.features_of(feats, P)
    net = S.MLP(7)
    tmap = screen.Target(sc)
    hist = screen.train(net, X[:len(P)], torch.from_numpy(tmap(sc)).cuda(), epochs=6, batch=2048)
    yield cx, ctx, feats, net, hist, tmap
    net.close()
    feats.close()
    ctx.close()


@pytest.mark.gpu
def test_trained_surrogate_ranks_candidates(gpu, trained):
    """Held-out candidates (other starts): overall and within-glowworm rank
    correlation, and the top pick of each glowworm's 32 beats the average
    candidate in exact score (the network's gain over picking at random)."""
    cx, ctx, feats, net, hist, _ = trained
    assert hist[-1] < 0.5 * hist[0]
    pos = synth.make_swarm_positions(10, 20, cx.meta["rec_axes"], 42, lig_radius=cx.meta["lig_axes"][0])
    hold = screen.collect_poses(ctx, "dna", pos, 99, [20] * 10, 21, 32, 2.0, 0.1, every=10)
    ex = screen.exact_scores(ctx, "dna", hold).reshape(-1, 32)
    pr = screen.predict(net, screen.features_of(feats, hold), len(hold)).reshape(-1, 32)
    assert screen.spearman(pr.ravel(), ex.ravel()) > 0.6
    within = np.mean([screen.spearman(pr[i], ex[i]) for i in range(len(ex))])
    assert within > 0.3, within
    top = ex[np.arange(len(ex)), pr.argmax(1)]
    assert top.mean() > ex.mean() + 0.25 * (ex.max(1).mean() - ex.mean()), (top.mean(), ex.mean())


@pytest.mark.gpu
def test_screened_gso_invariants(gpu, trained):
    """Committed scores are the exact scores of the committed poses; with
    accept='improve' no glowworm's score decreases; the plan consumed the same
    draws as a plain movement phase (same vision ranges)."""
    cx, ctx, feats, net, _, _ = trained
    pos = synth.make_swarm_positions(4, 16, cx.meta["rec_axes"], 5, lig_radius=cx.meta["lig_axes"][0])
    sg = screen.ScreenedGSO(ctx, feats, net, "dna", pos, 5, [16] * 4, trans_radius=2.0, rot_sigma=0.1)
    plain = GSO(ctx, "dna", pos, 5, [16] * 4)
    prev = None
    for _ in range(5):
        sg.step()
        st = sg.state()
        again = screen.exact_scores(ctx, "dna", st["pose"])
        assert np.array_equal(again, st["scoring"])
        if prev is not None:
            assert (st["scoring"] >= prev).all()
        prev = st["scoring"].copy()
    # first step: same luciferin update and neighbour/vision rule as plain GSO
    sg2 = screen.ScreenedGSO(ctx, feats, net, "dna", pos, 5, [16] * 4)
    sg2.step()
    plain.update_luciferin()
    plain.movement_phase()
    assert np.array_equal(sg2.state()["vision"], plain.state()["vision"])
    assert np.array_equal(sg2.state()["neighbors"], plain.state()["neighbors"])
    sg.close()
    sg2.close()
    plain.close()


@pytest.mark.gpu
def test_online_learning_fits_the_screens_exact_scores(gpu, trained):
    """The screen's online Adam steps train on the exact scores of its own
    picks (a replay buffer on the device): after 20 screened steps the
    updated network predicts the buffered picks' targets with a smaller mean
    squared error than the offline network does, and the run's committed
    scores are still exact."""
    cx, ctx, feats, net, _, tmap = trained
    pos = synth.make_swarm_positions(10, 20, cx.meta["rec_axes"], 11, lig_radius=cx.meta["lig_axes"][0])
    net_s = S.MLP(7)
    net_s.params = net.params
    sg = screen.ScreenedGSO(ctx, feats, net_s, "dna", pos, 11, [20] * 10, trans_radius=2.0, rot_sigma=0.1,
                            online_steps=4, target=tmap)
    for _ in range(20):
        sg.step()
    st = sg.state()
    assert np.array_equal(screen.exact_scores(ctx, "dna", st["pose"]), st["scoring"])
    n = sg.fill
    assert n == 20 * 200 * 2
    npad = (n + 127) // 128 * 128
    X = torch.zeros((npad, S.NF), dtype=torch.int16, device="cuda")
    X[:n] = sg.Xr[:n]
    t = sg.Tr[:n].cpu().numpy()
    mse = {}
    for name, m in (("offline", net), ("online", net_s)):
        mse[name] = float(np.mean((screen.predict(m, X, n) - t) ** 2))
    print("[gpu-test] mse on the screen's exact-scored picks: offline %.4f, online %.4f"
          % (mse["offline"], mse["online"]))
    assert mse["online"] < 0.9 * mse["offline"], mse
    sg.close()
    net_s.close()
