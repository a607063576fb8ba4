"""Surrogate-screened GSO: the learned screening step (no reference
counterpart; the plain exact GSO of lklight.gso stays the fallback and the yardstick).

One step, for every glowworm of every swarm:
  1. update_luciferin (scores are already exact from the previous commit);
  2. plan the GSO move (lkl_gso_plan: the reference's neighbour choice, draws
     and vision rule) and expand it on the GPU into n_cand candidates
     (candidate 0 = the planned move, the rest perturbations of it);
  3. features + MLP for every candidate on the GPU, top n_exact per glowworm by
     predicted score (lkl_screen_select);
  4. the picks are exact-scored by the scoring kernel; a glowworm takes its best
     pick (accept="best") or takes it only when it beats its current exact
     score (accept="improve");
  5. with online_steps > 0 the network keeps learning from those exact scores:
     the picks' features (already computed) and targets go into a replay
     buffer on the device and online_steps Adam steps of 4,096 rows drawn from
     it follow every screening step.  The offline-trained network ranks the
     candidates of poor poses well but not those near a swarm's optima, which
     the swarm reaches only during the run; the replay buffer holds exactly
     those.
The picker "random" replaces step 3's ranking by random predictions (the same
procedure without the network), for comparison.

Device memory comes from torch (plumbing); every kernel is in liblklight_gpu.so.
"""
import time

import numpy as np
import torch

from . import _capi, family_flags, fold_scores
from .gso import GSO
from .surrogate import NF


def spearman(a, b):
    """Spearman rank correlation (average ranks for ties)."""
    def rank(x):
        x = np.asarray(x, np.float64)
        o = np.argsort(x, kind="stable")
        r = np.empty(len(x))
        r[o] = np.arange(len(x), dtype=np.float64)
        xs = x[o]
        i = 0
        while i < len(xs):  # average ties
            j = i
            while j + 1 < len(xs) and xs[j + 1] == xs[i]:
                j += 1
            if j > i:
                r[o[i:j + 1]] = (i + j) / 2.0
            i = j + 1
        return r
    ra, rb = rank(a), rank(b)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra @ rb) / np.sqrt((ra @ ra) * (rb @ rb)))


class Target:
    """Monotone map of scores to training targets: clip(score, lo, hi),
    standardised.  Clipping the clash tail keeps the loss's resolution where
    candidate moves differ (the attractive range); ranking above lo is
    unchanged."""

    def __init__(self, scores, lo=-30.0, hi=30.0):
        y = np.clip(np.asarray(scores, np.float64), lo, hi)
        self.lo, self.hi, self.mean, self.std = lo, hi, float(y.mean()), float(y.std() + 1e-12)

    def __call__(self, scores):
        y = np.clip(np.asarray(scores, np.float64), self.lo, self.hi)
        return ((y - self.mean) / self.std).astype(np.float32)


def exact_scores(ctx, method, poses):
    e, v, s = ctx.score(poses, family_flags(method))
    return fold_scores(method, e, v, s)


def features_of(feats, poses, chunk=65536):
    """Features (device int16 tensor, rows padded to a multiple of 128)."""
    n = poses.shape[0]
    npad = (n + 127) // 128 * 128
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):  # the kernels run on the stream the buffers belong to
        X = torch.zeros((npad, NF), dtype=torch.int16, device="cuda")
        for i in range(0, n, chunk):
            d_p = torch.from_numpy(np.ascontiguousarray(poses[i:i + chunk], np.float64)).cuda()
            feats.compute_device(d_p, X[i:i + d_p.shape[0]], st.cuda_stream)
    st.synchronize()
    return X


def predict(net, X, n):
    y = torch.empty(X.shape[0], dtype=torch.float32, device="cuda")
    net.forward_device(X, y)
    net.sync()
    return y[:n].cpu().numpy()


def train(net, X, T, epochs=4, batch=4096, lr=1e-3, seed=0, log=None):
    """Adam on (X, T) device tensors, shuffled each epoch; the learning rate
    decays linearly to lr / 10.  Returns the per-epoch mean loss."""
    n = int(T.shape[0]) // batch * batch
    g = torch.Generator().manual_seed(seed)
    steps, k = epochs * (n // batch), 0
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(int(T.shape[0]), generator=g)[:n].cuda()
        tot = 0.0
        for b in range(n // batch):
            idx = perm[b * batch:(b + 1) * batch]
            xb = X.index_select(0, idx)
            tb = T.index_select(0, idx)
            torch.cuda.synchronize()
            loss = net.grad_device(xb, tb, want_loss=(b % 50 == 0))
            if loss is not None:
                tot += loss
            net.adam(lr * (1.0 - 0.9 * k / max(1, steps - 1)))
            net.sync()  # xb / tb are torch allocations: keep them until the kernels are done
            k += 1
        net.sync()
        hist.append(tot / max(1, (n // batch + 49) // 50))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return hist


class ScreenedGSO:
    """Surrogate-screened GSO over the swarms of one complex (see module doc)."""

    def __init__(self, ctx, feats, net, method, positions, seed, swarm_sizes=None, n_cand=32,
                 n_exact=2, trans_radius=1.0, rot_sigma=0.05, picker="surrogate", cand_seed=None,
                 accept="improve", online_steps=0, target=None, online_lr=2e-4, replay=65536,
                 online_batch=4096):
        self.gso = GSO(ctx, method, positions, seed, swarm_sizes)
        self.ctx, self.feats, self.net, self.method = ctx, feats, net, method
        self.n, self.C, self.k = self.gso.n, int(n_cand), int(n_exact)
        self.tr, self.rs, self.picker = float(trans_radius), float(rot_sigma), picker
        assert accept in ("improve", "best")
        self.accept = accept
        self.seed = int(seed if cand_seed is None else cand_seed)
        self.step_no = 0
        self.rng = np.random.default_rng(self.seed)
        nc = self.n * self.C
        npad = (nc + 127) // 128 * 128
        dev = "cuda"
        self.d_base = torch.empty((self.n, 7), dtype=torch.float64, device=dev)
        self.d_cand = torch.zeros((npad, 7), dtype=torch.float64, device=dev)
        self.d_cand[:, 0] = 1.0
        self.X = torch.zeros((npad, NF), dtype=torch.int16, device=dev)
        self.y = torch.empty(npad, dtype=torch.float32, device=dev)
        self.d_pick = torch.empty(self.n * self.k, dtype=torch.int32, device=dev)
        self.d_sel = torch.empty((self.n * self.k, 7), dtype=torch.float64, device=dev)
        self.d_out = torch.empty((self.n * self.k, 2), dtype=torch.float64, device=dev)
        self.stream = torch.cuda.Stream()
        self.flags = family_flags(method)
        self.exact_evals = 0
        self.surrogate_evals = 0
        self.online_steps = int(online_steps) if picker == "surrogate" else 0
        if self.online_steps:
            if target is None:
                raise ValueError("online fine-tuning needs the training target map")
            self.target, self.lr, self.batch = target, float(online_lr), int(online_batch)
            self.Xr = torch.zeros((int(replay), NF), dtype=torch.int16, device=dev)
            self.Tr = torch.zeros(int(replay), dtype=torch.float32, device=dev)
            self.fill = self.head = 0
            self.gen = torch.Generator().manual_seed(self.seed)
        self.times = {"screen": 0.0, "exact": 0.0, "online": 0.0, "host": 0.0}

    def _learn(self, scores):
        """Add the step's exact-scored picks to the replay buffer and take
        online_steps Adam steps on batches drawn from it."""
        m = self.n * self.k
        rows = (torch.arange(self.n, device="cuda", dtype=torch.int64)[:, None] * self.C
                + self.d_pick.view(self.n, self.k).long()).reshape(-1)
        at = (torch.arange(m, device="cuda") + self.head) % self.Xr.shape[0]
        self.Xr[at] = self.X.index_select(0, rows)
        self.Tr[at] = torch.from_numpy(self.target(scores.reshape(-1))).cuda()
        self.head = (self.head + m) % self.Xr.shape[0]
        self.fill = min(self.Xr.shape[0], self.fill + m)
        for _ in range(self.online_steps):
            bi = torch.randint(0, self.fill, (self.batch,), generator=self.gen).cuda()
            xb, tb = self.Xr.index_select(0, bi), self.Tr.index_select(0, bi)
            torch.cuda.synchronize()
            self.net.grad_device(xb, tb, want_loss=False)
            self.net.adam(self.lr)
            self.net.sync()  # xb / tb are torch allocations: keep them until the kernels are done

    def step(self):
        L = _capi.lib()
        st = self.stream.cuda_stream
        t0 = time.perf_counter()
        self.gso.update_luciferin()
        base = self.gso.plan()
        cur = self.gso.state()["scoring"]
        t1 = time.perf_counter()
        with torch.cuda.stream(self.stream):
            self.d_base.copy_(torch.from_numpy(base), non_blocking=False)
            seed = (self.seed * 1000003 + self.step_no) & 0xFFFFFFFFFFFFFFFF
            _capi.check(L.lkl_screen_candidates(self.d_base.data_ptr(), self.n, self.C, self.tr, self.rs,
                                                seed, self.d_cand.data_ptr(), st), "lkl_screen_candidates")
            nc = self.n * self.C
            if self.picker == "surrogate":
                self.feats.compute_device(self.d_cand[:nc], self.X[:nc], st)
                self.net.forward_device(self.X, self.y, st)
                self.surrogate_evals += nc
            else:
                self.y[:nc].copy_(torch.from_numpy(self.rng.random(nc, dtype=np.float32)))
            _capi.check(L.lkl_screen_select(self.y.data_ptr(), self.d_cand.data_ptr(), self.n, self.C, self.k,
                                            self.d_pick.data_ptr(), self.d_sel.data_ptr(), st),
                        "lkl_screen_select")
            self.stream.synchronize()
            t2 = time.perf_counter()
            self.ctx.score_async(self.d_sel.data_ptr(), self.n * self.k, self.flags, self.d_out.data_ptr(),
                                 None, st)
            sel = self.d_sel.cpu().numpy().reshape(self.n, self.k, 7)
            out = self.d_out.cpu().numpy()
        self.stream.synchronize()
        t3 = time.perf_counter()
        self.exact_evals += self.n * self.k
        sc = fold_scores(self.method, out[:, 0], out[:, 1]).reshape(self.n, self.k)
        if self.online_steps:
            self._learn(sc)
        t4 = time.perf_counter()
        best = np.argmax(sc, axis=1)
        bsc = sc[np.arange(self.n), best]
        bpose = sel[np.arange(self.n), best]
        if self.accept == "improve":
            accept = (bsc > cur).astype(np.uint8)
        else:  # always move to the best exact-scored pick, as GSO always moves
            accept = np.ones(self.n, np.uint8)
        self.gso.commit(bpose, bsc, accept)
        self.step_no += 1
        t5 = time.perf_counter()
        self.times["host"] += (t1 - t0) + (t5 - t4)
        self.times["screen"] += t2 - t1
        self.times["exact"] += t3 - t2
        self.times["online"] += t4 - t3
        return accept.mean()

    def state(self):
        return self.gso.state()

    def close(self):
        self.gso.close()


def top_k(scores, k=10):
    return np.sort(np.asarray(scores))[::-1][:k]


def recovered(top, ref, tol=0.005):
    """Every rank of `top` reaches the reference's value within tol (relative)."""
    return bool(np.all(top >= ref - tol * np.abs(ref)))


def run_plain(ctx, method, positions, seed, swarm_sizes, steps):
    """Plain exact GSO; returns (final scores, wall seconds, per-step top-10)."""
    g = GSO(ctx, method, positions, seed, swarm_sizes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    for _ in range(steps):
        g.update_luciferin()
        g.movement_phase()
        hist.append((time.perf_counter() - t0, None))
    g.update_luciferin()  # score the last moves
    wall = time.perf_counter() - t0
    sc = g.state()["scoring"]
    g.close()
    return sc, wall, hist


def collect_poses(ctx, method, positions, seed, swarm_sizes, steps, n_cand, trans_radius, rot_sigma,
                  every=1):
    """Training poses from a plain GSO trajectory: at every `every`-th step the
    planned moves are expanded into candidates like the screen does (GPU), and
    the trajectory follows the planned moves.  Returns (k * n * n_cand, 7)."""
    L = _capi.lib()
    g = GSO(ctx, method, positions, seed, swarm_sizes)
    n = g.n
    d_base = torch.empty((n, 7), dtype=torch.float64, device="cuda")
    d_cand = torch.empty((n * n_cand, 7), dtype=torch.float64, device="cuda")
    out = []
    for s in range(steps):
        g.update_luciferin()
        base = g.plan()
        if s % every:
            g.set_poses(base)
            continue
        d_base.copy_(torch.from_numpy(base))
        _capi.check(L.lkl_screen_candidates(d_base.data_ptr(), n, n_cand, trans_radius, rot_sigma,
                                            (seed * 7919 + s) & 0xFFFFFFFFFFFFFFFF, d_cand.data_ptr(),
                                            None), "lkl_screen_candidates")
        torch.cuda.synchronize()
        out.append(d_cand.cpu().numpy())
        g.set_poses(base)
    g.close()
    return np.concatenate(out, 0)
