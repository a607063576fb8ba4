"""Benchmark of the batched GSO pose-scoring path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only c2-dna,...] [--no-cpu]

Headline workload (one step = one blocking scoring call of one GSO step of
6 swarms x 200 glowworms): synthetic 12,625-atom receptor (RSA, 3.0 A
separation, (60, 55, 50) A ellipsoid) x 1,500-atom ligand, 1,200 poses per
step, dna flags (far field + electrostatics + clash), FP32 pair terms.  Each
step is lkl_ctx_score on host buffers: poses H2D, kernels, scores D2H, so
`value` is the steady-state rate of the whole scoring call.  The rate with the
poses already resident in HBM is reported beside it.

The same line carries one object per further workload (`configs`): the same
complex under the pydock and vdw families, the 8,000 x 3,000-atom cpydock
workload and the 64-complex batch, each with its roofline, CPU baseline and
maximum error against the oracle.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lklight-gpu_amd"))
sys.path.insert(0, ROOT)

METRIC = "pose-evals/s (dna, 12.6k-atom rec); surrogate/plain speedup at equal top-10 E"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = "profiles/score_kernel_counters.json"
ALL_CONFIGS = ("c2-dna", "c2-pydock", "c2-vdw", "c3-cpydock", "c4")
TIME_EVERY = 8  # kernel events on every 8th timed call


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pose_bytes(nl, n_far, n_near, desolv=False, nr=0):
    """Algorithmic bytes of one pose evaluation (no-reuse gather model):
    pose 7 f64 + out 2 f64 + ligand 25 B/atom + 8 phi corners per charged atom
    sampled in the grid + one 25 B receptor record per pair within 10 A
    (+ asa and des per atom with desolvation)."""
    b = 56 + 16 + 25 * nl + 32 * n_far + 25 * n_near
    if desolv:
        b += 8 * (nr + nl)
    return b


def measured_traffic(key):
    """HBM bytes per launch of the scoring kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy-kernel calibration), for
    exactly this workload, family, poses per launch and launch shape; None
    otherwise."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    for e in d.get("entries", []):
        if all(e.get(k) == v for k, v in key.items()):
            return e["traffic_bytes_per_launch"]
    return None


def host_info():
    """Host cores this process may use (affinity, cgroup quota) and the CPU model."""
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"nproc": nproc, "affinity_cpus": aff, "cgroup_cpu_quota": quota, "model": model,
            "threads": threads}


class CpuLeg:
    """The oracle (f64 C restatement of the reference CPU grid path, parallel
    over poses like the reference's rayon loop) timed on the host: all cores
    this process may use at -O3 -march=x86-64 (the reference's ISA,
    .cargo/config.toml:1-2), the same at -march=native (built here, on this
    host), and one thread.  Each figure is the median of repeated runs over
    the same pose sample after a warm-up."""

    def __init__(self, info, enabled=True):
        self.info, self.enabled = info, enabled
        self.native = None
        if enabled:
            try:
                from oracle import lkl_oracle as O
                self.native = O.build_native(tempfile.mkdtemp(prefix="lkl_oracle_native_"))
            except (OSError, subprocess.SubprocessError) as e:
                log("native oracle build failed: %s" % e)

    @staticmethod
    def _rate(orc, method, poses, threads, reps):
        rates = []
        r = None
        for _ in range(reps):
            t = time.perf_counter()
            r = orc.batch(method, poses, nthreads=threads, counts=True)
            rates.append(poses.shape[0] / (time.perf_counter() - t))
        return statistics.median(rates), rates, r

    def run(self, cx, method, poses, budget_s, label):
        """-> (cpu_baseline dict, oracle results on the sample poses)."""
        from oracle import lkl_oracle as O
        th = self.info["threads"]
        orc = O.Oracle(cx)
        n0 = min(poses.shape[0], max(2 * th, 32))
        t = time.perf_counter()
        orc.batch(method, poses[:n0], nthreads=th)  # warm-up + size estimate
        est = n0 / max(time.perf_counter() - t, 1e-4)
        reps = 5
        n = int(min(poses.shape[0], max(n0, est * budget_s * 0.5 / reps)))
        sample = poses[:n]
        med, rates, r = self._rate(orc, method, sample, th, reps)
        out = {"value": med, "unit": "pose-evals/s", "cores": th, "kind": "port",
               "sample": "%s: %d poses of the timed workload, median of %d runs (%s), %s host, "
                         "f64 C restatement of the reference CPU grid path, -O3 -march=x86-64"
                         % (label, n, reps, ", ".join("%.0f" % x for x in rates), self.info["model"]),
               "host": self.info}
        if self.native:
            orc_n = O.Oracle(cx, lib_path=self.native)
            orc_n.batch(method, sample[:n0], nthreads=th)
            med_n, rates_n, _ = self._rate(orc_n, method, sample, th, 3)
            out["march_native"] = {"value": med_n, "cores": th, "runs": rates_n}
        n1 = int(min(n, max(8, med / th * budget_s * 0.2 / 3)))
        med_1, rates_1, _ = self._rate(orc, method, sample[:n1], 1, 3)
        out["single_thread"] = {"value": med_1, "cores": 1, "poses": n1, "runs": rates_1}
        return out, r


def rel_err(method, g_e, g_v, g_s, r):
    import lklight
    import numpy as np
    n = r["e"].shape[0]
    a = lklight.fold_scores(method, g_e[:n], g_v[:n], None if g_s is None else g_s[:n])
    b = lklight.fold_scores(method, r["e"], r["v"], r["s"] if method == "cpydock" else None)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def set_tiles(ctx, key, poses, flags, tune):
    """Apply the stored autotuner choice for `key` (or run the autotuner)."""
    import lklight
    if tune:
        cell, cell_x, block, chunk, table = ctx.autotune(poses, flags)
        t = {"cell": cell, "cell_x": cell_x, "block": block, "chunk": chunk, "source": "autotuned now"}
        t["table_best_ms"] = min(row[4] for row in table)
        return t
    t = dict(lklight.tuned_tiles(key))
    ctx.set_tiles(t["cell"], t["block"], t["chunk"], cell_x=t["cell_x"])
    t["source"] = "lklight-gpu_amd/lklight/tiles.json"
    return t


def bench_single(key, cx, K, W, args, cpu, cpu_budget):
    """One workload on one context: K blocking host-buffer scoring calls (the
    metric), the same poses resident in HBM, the CPU baseline and the error."""
    import numpy as np
    import torch
    import lklight
    from lklight import _capi, synth

    workload, method = key.split("-")
    desolv = method == "cpydock"
    ctx = lklight.Context(cx, desolv=desolv, with_field=method != "vdw")
    flags = lklight.family_flags(method) | (lklight.F_DESOLV if desolv else 0)
    N = 400 if workload == "c3" else 1200
    poses = synth.make_poses(N * (K + W), cx.meta["rec_axes"], 2024)
    tiles = set_tiles(ctx, key, poses[:N], flags, args.tune)
    n_list, list_dims = ctx.list_info()
    ctx.set_timing(True)  # creates the events; switched on per timed call
    ctx.set_timing(False)

    # the metric: blocking C-ABI calls on host buffers (H2D poses, kernels,
    # D2H scores), made as a compiled host makes them: preallocated buffers,
    # one lkl_ctx_score per step.  Every TIME_EVERY-th call also records HIP
    # events around its kernels (about 5 us per event on the call's stream).
    L = _capi.lib()
    host = [np.ascontiguousarray(poses[i * N:(i + 1) * N]) for i in range(K + W)]
    outs = np.zeros((K + W, N, 2))
    outS = np.zeros((K + W, N))
    p_in = [_capi.ptr(h, C.c_double) for h in host]
    p_out = [_capi.ptr(outs[i], C.c_double) for i in range(K + W)]
    p_s = [_capi.ptr(outS[i], C.c_double) if desolv else None for i in range(K + W)]

    def call(i):
        _capi.check(L.lkl_ctx_score(ctx.h, p_in[i], N, flags, p_out[i], p_s[i]), "lkl_ctx_score")
    for i in range(W):
        call(i)
    pre_ms, ker_ms = [], []
    t0 = time.perf_counter()
    for k in range(K):
        timed = k % TIME_EVERY == 0
        if timed:
            L.lkl_ctx_set_timing(ctx.h, 1)
        call(W + k)
        if timed:
            a, b = ctx.kernel_ms()
            pre_ms.append(a)
            ker_ms.append(b)
            L.lkl_ctx_set_timing(ctx.h, 0)
    wall = time.perf_counter() - t0
    value = N * K / wall
    g_e, g_v = outs[W:, :, 0].reshape(-1), outs[W:, :, 1].reshape(-1)
    g_s = outS[W:].reshape(-1) if desolv else None
    # steady state over a longer run of the same calls (no events)
    reps = args.steady
    steady = None
    if reps > 0:
        t0 = time.perf_counter()
        for k in range(reps):
            call(W + k % K)
        steady = N * reps / (time.perf_counter() - t0)

    # the same steps with the poses resident in HBM (no copies)
    dev = torch.device("cuda:0")
    d_poses = torch.from_numpy(poses).to(dev)
    d_out = torch.zeros((K + W, N, 2), dtype=torch.float64, device=dev)
    d_s = torch.zeros((K + W, N), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def step(i):
        ctx.score_async(d_poses[i * N:(i + 1) * N].data_ptr(), N, flags, d_out[i].data_ptr(),
                        d_s[i].data_ptr() if desolv else None, sh)
    for i in range(W):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        step(W + k)
    torch.cuda.synchronize()
    resident = N * K / (time.perf_counter() - t0)
    same = bool(np.array_equal(d_out[W:].reshape(-1, 2)[:, 0].cpu().numpy(), g_e)
                and np.array_equal(d_out[W:].reshape(-1, 2)[:, 1].cpu().numpy(), g_v))
    log("%s: host-buffer calls %.0f pose-evals/s (steady %s), resident %.0f, score kernel %.4f ms"
        % (key, value, "%.0f" % steady if steady else "-", resident, sum(ker_ms) / len(ker_ms)))

    cpu_b, r = (None, None)
    if cpu.enabled:
        cpu_b, r = cpu.run(cx, method, poses[W * N:(W + K) * N], cpu_budget, key)
        cpu_b["gpu_over_cpu"] = value / cpu_b["value"]
        cpu_b["max_rel_err_vs_cpu"] = rel_err(method, g_e, g_v, g_s, r)
        log("%s: cpu %.0f pose-evals/s on %d threads (native %s, 1 thread %.0f), max rel err %.2e"
            % (key, cpu_b["value"], cpu_b["cores"],
               "%.0f" % cpu_b["march_native"]["value"] if "march_native" in cpu_b else "-",
               cpu_b["single_thread"]["value"], cpu_b["max_rel_err_vs_cpu"]))
    n_near = float(r["n_near"].mean()) if r is not None else None
    n_far = float(r["n_far"].mean()) if r is not None else None
    kern_s = sum(ker_ms) / len(ker_ms) / 1e3
    b_pose = pose_bytes(cx.lig.n, n_far or 0.0, n_near or 0.0, desolv, cx.rec.n) if r is not None else None
    achieved = b_pose * N / kern_s / 1e9 if b_pose else None
    tkey = {"workload": workload, "family": method, "poses_per_launch": N,
            "tiles": [tiles["cell"], tiles["cell_x"], tiles["block"], tiles["chunk"]]}
    traffic = measured_traffic(tkey)
    ctx.close()
    return {
        "value": value, "unit": "pose-evals/s", "poses_per_step": N, "steps": K, "warmup": W,
        "ms_per_step": wall / K * 1e3, "dtype": "fp32",
        "steady_state": {"pose_evals_per_s": steady, "calls": reps} if steady else None,
        "resident_pose_evals_per_s": resident, "resident_equals_host_calls": same,
        "tiles": tiles, "flags": flags,
        "near_field": ({"neighbour_lists": {"spacing_A": 1.0, "entries": n_list, "nodes": list(list_dims)}}
                       if n_list else "cell-row walk"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": TRAFFIC_FILE if traffic else None,
                     "kernel": "score_kernel", "kernel_ms": kern_s * 1e3,
                     "prepass_ms": sum(pre_ms) / len(pre_ms), "timed_launches": len(ker_ms),
                     "bytes_per_pose": b_pose,
                     "n_near_per_pose": n_near, "n_far_per_pose": n_far},
        "cpu_baseline": cpu_b,
        "max_rel_err_vs_oracle": cpu_b["max_rel_err_vs_cpu"] if cpu_b else None,
    }


def bench_c4(K, W, args, cpu):
    """64 complexes (receptors 2k-20k atoms, ligands 500-5k, 100-500
    glowworms, dna/pydock/vdw mix): one batch launch per GSO step.  The value
    is the host-buffer step (poses H2D, kernel, scores D2H, replayed from a
    captured graph); the resident batch and 64 sequential per-complex launches
    are reported beside it."""
    import numpy as np
    import torch
    import lklight
    from lklight import synth
    from oracle import lkl_oracle as O

    dev = torch.device("cuda:0")
    t = time.perf_counter()
    cxs = synth.make_config4(64)
    ctxs, flags, counts = [], [], []
    for cx in cxs:
        fam = cx.meta["family"]
        ctxs.append(lklight.Context(cx, with_field=False, device_field=(fam != "vdw"), cell_size=4.0))
        flags.append(lklight.family_flags(fam))
        counts.append(cx.meta["glowworms"])
    log("c4: 64 complexes + device phi in %.1f s" % (time.perf_counter() - t))
    P = sum(counts)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    step_poses = np.stack([np.concatenate([synth.make_poses(n, cx.meta["rec_axes"], 1000 * s_ + k)
                                           for k, (n, cx) in enumerate(zip(counts, cxs))])
                           for s_ in range(K + W)])
    m = lklight.Multi(ctxs, flags, counts)
    if args.tune:
        tb, tm, _ = m.autotune(step_poses[0])
        tiles = {"block": tb, "chunk_mult": tm, "source": "autotuned now"}
    else:
        tiles = dict(lklight.tuned_tiles("c4"))
        m.set_tiles(tiles["block"], tiles["chunk_mult"])
        tiles["source"] = "lklight-gpu_amd/lklight/tiles.json"
    # the metric: host-buffer steps replayed from the captured graph
    m.use_graph(True)
    for i in range(W):
        m.score(step_poses[i])
    outs = []
    t0 = time.perf_counter()
    for k in range(K):
        outs.append(m.score(step_poses[W + k]))
    wall = time.perf_counter() - t0
    value = P * K / wall
    # resident batch (kernel events) and 64 sequential launches
    d_poses = torch.from_numpy(step_poses).to(dev)
    d_out = torch.zeros((K + W, P, 2), dtype=torch.float64, device=dev)
    d_seq = torch.zeros((K + W, P, 2), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]

    def run_batch(i, k=None):
        if k is not None:
            ev[k][0].record(stream)
        m.score_async(d_poses[i].data_ptr(), d_out[i].data_ptr(), sh)
        if k is not None:
            ev[k][1].record(stream)

    def run_seq(i, k=None):
        for j, ctx in enumerate(ctxs):
            ctx.score_async(d_poses[i, offs[j]:offs[j + 1]].data_ptr(), counts[j], flags[j],
                            d_seq[i, offs[j]:offs[j + 1]].data_ptr(), None, sh)

    res = {}
    for name, fn in (("batch", run_batch), ("sequential", run_seq)):
        for i in range(W):
            fn(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(K):
            fn(W + k, k)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
    kern_ms = sum(a.elapsed_time(b) for a, b in ev) / K
    # load balance of one traced resident launch (per-tile start/end and CU)
    m.set_trace(True)
    run_batch(W)
    rec, khz = m.trace()
    balance = lklight.load_balance(rec, khz)
    m.set_trace(False)
    log("c4: traced launch %d tiles on %d CUs, span %.1f us, CU busy %.3f, CU finish min/median/max %.0f/%.0f/%.0f us"
        % (balance["tiles"], balance["cus"], balance["span_us"], balance["busy_frac"],
           balance["cu_finish_us"]["min"], balance["cu_finish_us"]["median"], balance["cu_finish_us"]["max"]))
    same_seq = bool(torch.equal(d_out[W:], d_seq[W:]))
    same_host = bool(np.array_equal(d_out[W:].cpu().numpy(), np.stack(outs)))
    log("c4: host-buffer graph steps %.0f pose-evals/s, resident batch %.0f, sequential %.0f, kernel %.3f ms"
        % (value, P * K / res["batch"], P * K / res["sequential"], kern_ms))
    # oracle sample: 16 poses of every complex from the first timed step
    # (host phi built for the sampled complex, then dropped)
    cpu_b, max_rel, b_launch = None, None, None
    if cpu.enabled:
        th = cpu.info["threads"]
        g = outs[0]
        bytes_est, t_cpu, n_cpu, max_rel = 0.0, 0.0, 0, 0.0
        for j, cx in enumerate(cxs):
            fam = cx.meta["family"]
            if fam != "vdw":
                cx.build_field()
            orc = O.Oracle(cx)
            ps = step_poses[W, offs[j]:offs[j] + 16]
            t = time.perf_counter()
            r = orc.batch(fam, ps, nthreads=th, counts=True)
            t_cpu += time.perf_counter() - t
            n_cpu += ps.shape[0]
            gj = g[offs[j]:offs[j] + 16]
            max_rel = max(max_rel, rel_err(fam, gj[:, 0], gj[:, 1], None, r))
            bp = pose_bytes(cx.lig.n, float(r["n_far"].mean()), float(r["n_near"].mean()))
            bytes_est += bp * counts[j]
            cx.phi = None
            del orc
        b_launch = bytes_est
        cpu_b = {"value": n_cpu / t_cpu, "unit": "pose-evals/s", "cores": th, "kind": "port",
                 "sample": "16 poses of each of the 64 complexes (%d), one oracle batch per complex, %s host, "
                           "-O3 -march=x86-64" % (n_cpu, cpu.info["model"]),
                 "gpu_over_cpu": value / (n_cpu / t_cpu), "max_rel_err_vs_cpu": max_rel}
        log("c4: cpu %.0f pose-evals/s, max rel err %.2e" % (cpu_b["value"], max_rel))
    achieved = b_launch / (kern_ms / 1e3) / 1e9 if b_launch else None
    traffic = measured_traffic({"workload": "c4", "poses_per_launch": P,
                                "tiles": [tiles["block"], tiles["chunk_mult"]]})
    fams = {str(f): int(n) for f, n in zip(*np.unique([c.meta["family"] for c in cxs], return_counts=True))}
    m.close()
    for c in ctxs:
        c.close()
    return {
        "value": value, "unit": "pose-evals/s", "poses_per_step": P, "steps": K, "warmup": W,
        "ms_per_step": wall / K * 1e3, "dtype": "fp32",
        "workload": "c4: 64 complexes, receptors 2k-20k atoms, ligands 500-5k, %d poses/step" % P,
        "families": fams, "tiles": tiles,
        "resident_pose_evals_per_s": P * K / res["batch"],
        "sequential_pose_evals_per_s": P * K / res["sequential"],
        "batch_over_sequential": res["sequential"] / res["batch"],
        "bitwise_equal_to_sequential": same_seq, "host_steps_equal_resident": same_host,
        "load_balance": balance,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": TRAFFIC_FILE if traffic else None,
                     "kernel": "multi_kernel", "kernel_ms": kern_ms, "bytes_per_launch": b_launch},
        "cpu_baseline": cpu_b, "max_rel_err_vs_oracle": max_rel,
    }


def retune(args):
    """Run the autotuners on every workload and print a fresh tiles table."""
    import lklight
    from lklight import synth
    out = {}
    cx2 = synth.make_config("c2")
    cx2.build_field()
    cx3 = synth.make_config("c3")
    cx3.build_field()
    for key, cx in (("c2-dna", cx2), ("c2-pydock", cx2), ("c2-vdw", cx2), ("c3-cpydock", cx3)):
        workload, method = key.split("-")
        desolv = method == "cpydock"
        ctx = lklight.Context(cx, desolv=desolv, with_field=method != "vdw")
        flags = lklight.family_flags(method) | (lklight.F_DESOLV if desolv else 0)
        N = 400 if workload == "c3" else 1200
        cell, cell_x, block, chunk, table = ctx.autotune(synth.make_poses(N, cx.meta["rec_axes"], 2024), flags)
        out[key] = {"cell": cell, "cell_x": cell_x, "block": block, "chunk": chunk,
                    "best_ms": min(r[4] for r in table)}
        log("%s: %s" % (key, out[key]))
        ctx.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--only", default="",
                    help="comma list of %s (default: all; the first listed is the headline)"
                         % ", ".join(ALL_CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of oracle work for the headline CPU baseline (others get a third)")
    ap.add_argument("--steady", type=int, default=200,
                    help="host-buffer calls after the timed steps, for a steady-state rate (0: skip)")
    ap.add_argument("--tune", action="store_true", help="run the autotuner instead of the stored tiles")
    ap.add_argument("--retune", action="store_true", help="only run the autotuners and print the tiles")
    ap.add_argument("--workload", default="", choices=["", "c5"],
                    help="c5: only the surrogate-screened GSO against plain exact GSO")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate-screened GSO comparison")
    ap.add_argument("--no-gso", action="store_true",
                    help="skip the host-driven vs device-resident GSO step comparison")
    ap.add_argument("--gso-only", action="store_true",
                    help="only the host-driven vs device-resident GSO step comparison (one JSON line)")
    ap.add_argument("--sur-train", type=int, default=2_000_000, help="c5: training poses")
    ap.add_argument("--sur-epochs", type=int, default=6)
    ap.add_argument("--sur-swarms", type=int, default=40)
    ap.add_argument("--sur-glow", type=int, default=50)
    ap.add_argument("--sur-plain-steps", type=int, default=100)
    ap.add_argument("--sur-max-steps", type=int, default=100)
    ap.add_argument("--sur-grid", type=float, default=1.0)
    ap.add_argument("--sur-online", type=int, default=8, help="c5: online Adam steps per screening step")
    ap.add_argument("--sur-seeds", default="42,7,99", help="c5: start seeds (the first is the headline)")
    ap.add_argument("--sur-trans", type=float, default=2.0)
    ap.add_argument("--sur-rot", type=float, default=0.1)
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a plain 512 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibrate:
        import torch
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                          "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus %d is not supported "
                          "(poses shard across GPUs with no exchange, but no multi-GPU path is "
                          "built)" % args.gpus}))
        sys.exit(2)

    import torch
    import lklight
    from lklight import synth

    if not lklight.device_available():
        raise SystemExit("no gfx950 device: the benchmark measures the HIP path only")
    torch.cuda.init()
    if args.retune:
        return retune(args)
    if args.workload == "c5":
        print(json.dumps(surrogate_experiment(args)))
        return
    if args.gso_only:
        cx = synth.make_config("c2")
        cx.build_field()
        ctx = lklight.Context(cx)
        t = lklight.tuned_tiles("c2-dna")
        ctx.set_tiles(t["cell"], t["block"], t["chunk"], cell_x=t["cell_x"])
        res = gso_step_compare(cx, ctx, "dna")
        print(json.dumps({"metric": "GSO step time, dna, 6 swarms x 200 glowworms (device-resident)",
                          "value": res["device"]["ms_per_step"], "unit": "ms/step", "higher_is_better": False,
                          "n_gpus": 1, "gso_step": res}))
        return
    K, W = args.steps, args.warmup
    keys = [k for k in (args.only.split(",") if args.only else ALL_CONFIGS) if k]
    for k in keys:
        if k not in ALL_CONFIGS:
            raise SystemExit("unknown workload %r (choose from %s)" % (k, ", ".join(ALL_CONFIGS)))
    info = host_info()
    cpu = CpuLeg(info, enabled=not args.no_cpu)
    if cpu.enabled:
        log("cpu legs: %d threads (nproc %d, affinity %d, cgroup quota %s), %s"
            % (info["threads"], info["nproc"], info["affinity_cpus"], info["cgroup_cpu_quota"], info["model"]))
    cache = {}

    def complex_of(workload):
        if workload not in cache:
            log("generating the %s complex" % workload)
            cx = synth.make_config(workload)
            t = time.perf_counter()
            cx.build_field()
            log("phi %s built on GPU in %.2f s" % (cx.dims, time.perf_counter() - t))
            cache[workload] = cx
        return cache[workload]

    results = {}
    for i, key in enumerate(keys):
        budget = args.cpu_budget if i == 0 else args.cpu_budget / 3
        if key == "c4":
            results[key] = bench_c4(K, W, args, cpu)
        else:
            results[key] = bench_single(key, complex_of(key.split("-")[0]), K, W, args, cpu, budget)
    head_key = keys[0]
    head = results[head_key]
    if head_key == "c2-dna":
        metric = METRIC
        workload = ("c2-dna: 12,625-atom receptor x 1,500-atom ligand, 1200 poses/step (6 swarms x 200 "
                    "glowworms), phi %s at 0.5 A, blocking host-buffer scoring call per step"
                    % "x".join(map(str, cache["c2"].dims)))
    else:
        metric = "pose-evals/s (%s)" % head_key
        workload = head.get("workload", head_key)
    out = {
        "metric": metric, "value": head["value"], "unit": "pose-evals/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": workload, "poses_per_step": head["poses_per_step"],
                   "tiles": head.get("tiles"), "near_field": head.get("near_field")},
        "roofline": head["roofline"], "cpu_baseline": head["cpu_baseline"],
        "max_rel_err_vs_oracle": head["max_rel_err_vs_oracle"],
        "resident_pose_evals_per_s": head["resident_pose_evals_per_s"],
        "steady_state": head.get("steady_state"),
        "configs": {k: v for k, v in results.items() if k != head_key},
    }
    if head.get("load_balance"):
        out["load_balance"] = head["load_balance"]
    if "c2" in cache and not args.no_gso:
        cx = cache["c2"]
        ctx = lklight.Context(cx)
        t = lklight.tuned_tiles("c2-dna")
        ctx.set_tiles(t["cell"], t["block"], t["chunk"], cell_x=t["cell_x"])
        out["gso_step"] = gso_step_compare(cx, ctx, "dna")
        stream = torch.cuda.Stream()
        d_poses = torch.from_numpy(synth.make_poses(1200, cx.meta["rec_axes"], 2024)).cuda()
        out["restraints"] = interface_timing(cx, ctx, d_poses, 1200, stream)
        ctx.close()
    if not args.no_surrogate:
        out["surrogate"] = surrogate_experiment(args)
    print(json.dumps(out))


def interface_timing(cx, ctx, d_poses, N, stream, reps=20):
    """Cost of the restraint / membrane interface pass (lkl_ctx_interface_async)
    on one step's poses: 100 receptor residues of 8 atoms, 20 ligand residues
    of 5 atoms and 200 membrane beads, chosen at random."""
    import numpy as np
    import torch
    from lklight import _capi
    rng = np.random.default_rng(1)
    rec_res = [list(range(a, a + 8)) for a in rng.choice(np.arange(0, cx.rec.n - 8, 8), 100, replace=False)]
    lig_res = [list(range(a, a + 5)) for a in rng.choice(np.arange(0, cx.lig.n - 5, 5), 20, replace=False)]
    membrane = sorted(rng.choice(cx.rec.n, 200, replace=False).tolist())
    ctx.set_interface(rec_res, lig_res, membrane)
    d_cnt = torch.zeros((N, 3), dtype=torch.int32, device="cuda")
    L = _capi.lib()

    def launch():
        _capi.check(L.lkl_ctx_interface_async(ctx.h, d_poses.data_ptr(), N, d_cnt.data_ptr(), stream.cuda_stream),
                    "lkl_ctx_interface_async")
    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    cnt = d_cnt.cpu().numpy()
    ctx.set_interface()
    log("interface pass: %.3f ms per %d poses" % (ms, N))
    return {"interface_ms_per_launch": ms, "poses": N, "rec_residues": 100, "lig_residues": 20,
            "membrane_beads": 200, "mean_satisfied_rec_residues": float(cnt[:, 0].mean())}


def gso_step_compare(cx, ctx, family, steps=50):
    """One GSO step at the workload's size (6 swarms x 200 glowworms) driven
    from the host (lkl_gso_update_luciferin + lkl_gso_movement_phase: a
    blocking scoring call and the movement on the CPU) against the
    device-resident step (lkl_gso_device_run: scoring, luciferin and movement
    kernels on one stream, one call for all steps).  Both start from the same
    state after 2 warm-up steps; their trajectories agree to the last bit of
    slerp's acos/sin, so the moved counts are reported side by side."""
    import numpy as np
    from lklight import gso as G
    from lklight import synth
    sizes = [200] * 6
    pos = synth.make_swarm_positions(6, 200, cx.meta["rec_axes"], 42, lig_radius=cx.meta["lig_axes"][0])
    res = {}
    for mode in ("host", "device"):
        g = G.GSO(ctx, family, pos, seed=42, swarm_sizes=sizes)
        if mode == "host":
            g.run(2)
            t0 = time.perf_counter()
            for _ in range(steps):
                g.update_luciferin()
                g.movement_phase()
            dt = time.perf_counter() - t0
        else:
            g.device_run(2)
            t0 = time.perf_counter()
            g.device_run(steps)
            dt = time.perf_counter() - t0
        st = g.state()
        res[mode] = {"ms_per_step": dt / steps * 1e3, "moved_last_step": int(st["moved"].sum()),
                     "best_score": float(np.max(st["scoring"]))}
        g.close()
    res["speedup"] = res["host"]["ms_per_step"] / res["device"]["ms_per_step"]
    res["glowworms"] = 1200
    res["steps"] = steps
    # the configuration's run is 100 steps of the 6 swarms
    res["wall_s_100_steps"] = {m: res[m]["ms_per_step"] * 0.1 for m in ("host", "device")}
    log("gso step: host-driven %.3f ms, device-resident %.3f ms (%.2fx)"
        % (res["host"]["ms_per_step"], res["device"]["ms_per_step"], res["speedup"]))
    return res


def surrogate_cost_by_complex(args, net, workloads=("c1", "c2", "c3")):
    """What a screened candidate costs against an exact evaluation on complexes
    of increasing size: per 64,000-pose batch, the grid features and the MLP
    forward against the exact score of the same poses (dna at c1 and c2,
    cpydock with desolvation at c3, each with its stored tiles), and the share
    of the plain run's steps within which a screen of 32 candidates and 2 exact
    scores per glowworm must reach the plain top-10 for a 3x gain."""
    import numpy as np
    import torch
    import lklight
    from lklight import synth
    from lklight import surrogate as S

    nc, C, k = 64000, 32, 2
    rows = []
    for wl in workloads:
        method = "cpydock" if wl == "c3" else "dna"
        cx = synth.make_config(wl)
        cx.build_field()
        desolv = method == "cpydock"
        ctx = lklight.Context(cx, desolv=desolv)
        key = "%s-%s" % (wl, method)
        t = lklight.tuned_tiles(key)
        if t:  # c1 has no stored tiles: the context's defaults
            ctx.set_tiles(t["cell"], t["block"], t["chunk"], cell_x=t["cell_x"])
        flags = lklight.family_flags(method) | (lklight.F_DESOLV if desolv else 0)
        t0 = time.perf_counter()
        feats = S.Features(cx, grid=args.sur_grid)
        t_grid = time.perf_counter() - t0
        cand = torch.from_numpy(synth.make_poses(nc, cx.meta["rec_axes"], 5)).cuda()
        Xc = torch.zeros(((nc + 127) // 128 * 128, S.NF), dtype=torch.int16, device="cuda")
        yc = torch.empty(Xc.shape[0], dtype=torch.float32, device="cuda")
        d_o = torch.zeros((nc, 2), dtype=torch.float64, device="cuda")
        d_s = torch.zeros(nc, dtype=torch.float64, device="cuda") if desolv else None
        tf, tm, te = [], [], []
        for rep in range(4):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            feats.compute_device(cand, Xc[:nc], 0)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            net.forward_device(Xc, yc)
            net.sync()
            t2 = time.perf_counter()
            ctx.score_async(cand.data_ptr(), nc, flags, d_o.data_ptr(), d_s.data_ptr() if desolv else None)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if rep:
                tf.append(t1 - t0), tm.append(t2 - t1), te.append(t3 - t2)
        f_ns, m_ns, e_ns = (float(np.median(v)) / nc * 1e9 for v in (tf, tm, te))
        ratio = (f_ns + m_ns) / e_ns
        rows.append({"workload": wl, "family": method, "receptor_atoms": int(cx.rec.n), "ligand_atoms": int(cx.lig.n),
                     "grid_build_s": t_grid, "features_ns_per_candidate": f_ns, "mlp_ns_per_candidate": m_ns,
                     "exact_ns_per_pose": e_ns, "candidate_over_exact": ratio,
                     "screened_step_over_plain_step": C * ratio + k,
                     "steps_share_for_3x": 1.0 / (3.0 * (C * ratio + k))})
        log("surrogate cost at %s (%s, %d x %d atoms): features %.1f ns + MLP %.1f ns per candidate, exact %.1f ns "
            "per pose (ratio %.2f)" % (wl, method, cx.rec.n, cx.lig.n, f_ns, m_ns, e_ns, ratio))
        feats.close()
        ctx.close()
        del Xc, yc, d_o, d_s, cand
    return rows


def library_mlp_reference(rows, reps=10):
    """The vendor library's time for the screen's forward shapes (torch.matmul
    in bf16, which calls hipBLASLt): rows x 6144 -> 512 -> 512 -> 512 with bias
    and ReLU, random weights.  A reference point for the hand-written MFMA
    GEMMs, not used by the screen."""
    import torch
    dev = torch.device("cuda:0")
    x = torch.randn(rows, 6144, dtype=torch.bfloat16, device=dev)
    w = [torch.randn(512, k, dtype=torch.bfloat16, device=dev) * 0.01 for k in (6144, 512, 512)]
    bias = torch.zeros(512, dtype=torch.bfloat16, device=dev)

    def layer1():
        return torch.relu(x @ w[0].t() + bias)

    def forward():
        h = layer1()
        for k in (1, 2):
            h = torch.relu(h @ w[k].t() + bias)
        return h
    out = {}
    for name, fn in (("forward_ms", forward), ("layer1_ms", layer1)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    out["layer1_tflops"] = 2.0 * rows * 512 * 6144 / (out["layer1_ms"] * 1e-3) / 1e12
    out["rows"] = rows
    return out


def surrogate_experiment(args):
    """Config-5 comparison on the 1,500-atom / 800-atom complex: train the
    surrogate on the GPU from exact dna scores, then, from the same starts
    (one set per seed), run plain exact GSO (host-driven and device-resident),
    the screened GSO with random picks and the surrogate-screened GSO (the
    network keeps learning from the exact scores of its picks).  Reports rank
    correlation, top-10 recovery, wall times, the per-candidate costs and the
    break-even they imply."""
    import numpy as np
    import torch
    import lklight
    from lklight import gso as G
    from lklight import screen, synth
    from lklight import surrogate as S

    method = "dna"
    cx = synth.make_config("c1")
    cx.build_field()
    ctx = lklight.Context(cx)
    feats = S.Features(cx, grid=args.sur_grid)
    ns, ng = args.sur_swarms, args.sur_glow
    sizes = [ng] * ns
    n = ns * ng

    def starts(seed):
        return synth.make_swarm_positions(ns, ng, cx.meta["rec_axes"], seed, lig_radius=cx.meta["lig_axes"][0])
    # training poses: candidates along a 100-step plain GSO trajectory from
    # other starts (every k-th step, ~80 %) + random poses in the 15 A shell
    t0 = time.perf_counter()
    n_tr = args.sur_train
    every = max(1, (100 * n * 32) // max(1, (n_tr * 4) // 5))
    traj = screen.collect_poses(ctx, method, starts(4242), 4242, sizes, 100, 32, args.sur_trans, args.sur_rot,
                                every=every)[:n_tr]
    rnd = synth.make_poses(max(0, n_tr - traj.shape[0]), cx.meta["rec_axes"], 77)
    P = np.concatenate([traj, rnd], 0)
    scores = np.concatenate([screen.exact_scores(ctx, method, P[i:i + 65536])
                             for i in range(0, P.shape[0], 65536)])
    X = screen.features_of(feats, P)
    tmap = screen.Target(scores)
    T = torch.from_numpy(tmap(scores)).cuda()
    t_data = time.perf_counter() - t0
    log("surrogate data: %d poses (%d trajectory candidates) in %.1f s" % (P.shape[0], traj.shape[0], t_data))
    net = S.MLP(2024)
    t0 = time.perf_counter()
    hist = screen.train(net, X[:P.shape[0]], T, epochs=args.sur_epochs, log=log)
    t_train = time.perf_counter() - t0
    del X
    trained = net.params
    # held-out: candidates at steps 0, 20 and 40 of a plain GSO from the benchmark start
    hold = screen.collect_poses(ctx, method, starts(42), 99, sizes, 41, 32, args.sur_trans, args.sur_rot, every=20)
    hs = screen.exact_scores(ctx, method, hold)
    hp = screen.predict(net, screen.features_of(feats, hold), hold.shape[0])
    rho = screen.spearman(hp, hs)
    rho_pg = float(np.mean([screen.spearman(hp[i:i + 32], hs[i:i + 32]) for i in range(0, len(hs), 32)]))
    log("held-out spearman %.3f (within-glowworm mean %.3f)" % (rho, rho_pg))

    # cost of one screening batch (n x 32 candidates) against exact scoring of the same poses
    nc = n * 32
    cand = torch.from_numpy(synth.make_poses(nc, cx.meta["rec_axes"], 5)).cuda()
    Xc = torch.zeros(((nc + 127) // 128 * 128, S.NF), dtype=torch.int16, device="cuda")
    yc = torch.empty(Xc.shape[0], dtype=torch.float32, device="cuda")
    d_o = torch.zeros((nc, 2), dtype=torch.float64, device="cuda")
    cost = {}
    for rep in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        feats.compute_device(cand, Xc[:nc], 0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        net.forward_device(Xc, yc)
        net.sync()
        t2 = time.perf_counter()
        ctx.score_async(cand.data_ptr(), nc, lklight.FLAGS_DNA, d_o.data_ptr())
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        cost = {"features_ms": (t1 - t0) * 1e3, "mlp_ms": (t2 - t1) * 1e3, "exact_ms": (t3 - t2) * 1e3}
    per_cand = (cost["features_ms"] + cost["mlp_ms"]) / nc * 1e6
    per_exact = cost["exact_ms"] / nc * 1e6
    log("per candidate: features + MLP %.1f ns, exact score %.1f ns" % (per_cand, per_exact))
    del Xc, yc, d_o, cand
    lib_ref = library_mlp_reference((nc + 127) // 128 * 128)
    log("library (torch/hipBLASLt) bf16 forward of the same shapes: %.3f ms, layer 1 %.0f TFLOP/s"
        % (lib_ref["forward_ms"], lib_ref["layer1_tflops"]))

    def plain_device(pos, seed):
        g = G.GSO(ctx, method, pos, seed, sizes)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.device_run(args.sur_plain_steps)
        g.update_luciferin()  # score the last moves
        wall = time.perf_counter() - t0
        top = screen.top_k(g.state()["scoring"])
        g.close()
        return wall, top

    def screened(pos, seed, picker, ref10, online):
        net_s = S.MLP(2024)
        net_s.params = trained
        sg = screen.ScreenedGSO(ctx, feats, net_s, method, pos, seed, sizes, trans_radius=args.sur_trans,
                                rot_sigma=args.sur_rot, picker=picker, online_steps=online, target=tmap)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reached, reached_step = None, None
        for k in range(args.sur_max_steps):
            sg.step()
            top = screen.top_k(sg.state()["scoring"])
            if reached is None and screen.recovered(top, ref10):
                reached, reached_step = time.perf_counter() - t0, k + 1
        wall = time.perf_counter() - t0
        res = {"steps": args.sur_max_steps, "wall_s": wall, "ms_per_step": wall / args.sur_max_steps * 1e3,
               "reached_s": reached, "reached_step": reached_step, "top10": top.tolist(),
               "exact_evals": sg.exact_evals, "surrogate_evals": sg.surrogate_evals, "online_adam_steps": online,
               "phase_ms_per_step": {k: v / args.sur_max_steps * 1e3 for k, v in sg.times.items()}}
        sg.close()
        net_s.close()
        return res

    runs = []
    for seed in [int(v) for v in args.sur_seeds.split(",")]:
        pos = starts(seed)
        plain_sc, plain_wall, _ = screen.run_plain(ctx, method, pos, seed, sizes, args.sur_plain_steps)
        ref10 = screen.top_k(plain_sc)
        dev_wall, dev_top = plain_device(pos, seed)
        rnd_res = screened(pos, seed, "random", ref10, 0)
        sur = screened(pos, seed, "surrogate", ref10, args.sur_online)
        beats = bool(np.all(np.asarray(sur["top10"]) > np.asarray(rnd_res["top10"])))
        run = {"seed": seed,
               "plain": {"steps": args.sur_plain_steps, "wall_s": plain_wall, "top10": ref10.tolist(),
                         "ms_per_step": plain_wall / args.sur_plain_steps * 1e3},
               "plain_device": {"wall_s": dev_wall, "top10": dev_top.tolist(),
                                "ms_per_step": dev_wall / args.sur_plain_steps * 1e3},
               "random_picks": rnd_res, "screened": sur,
               "screened_beats_random_at_every_rank": beats,
               "speedup_vs_plain": plain_wall / sur["reached_s"] if sur["reached_s"] else None,
               "speedup_vs_plain_device": dev_wall / sur["reached_s"] if sur["reached_s"] else None}
        if seed == int(args.sur_seeds.split(",")[0]):
            run["screened_offline_network"] = screened(pos, seed, "surrogate", ref10, 0)
        log("seed %d: plain %.3f s (device %.3f s) top-10 %s; screened %.3f s reached %s top-10 %s; random top-10 %s; "
            "beats random at every rank: %s" % (seed, plain_wall, dev_wall, np.round(ref10, 2), sur["wall_s"],
                                                sur["reached_step"], np.round(sur["top10"], 2),
                                                np.round(rnd_res["top10"], 2), beats))
        runs.append(run)
    head = runs[0]
    plain_ms = head["plain"]["ms_per_step"]
    scr_ms = head["screened"]["ms_per_step"]
    out = {
        "workload": "c5: 1,500-atom receptor x 800-atom ligand, %d swarms x %d glowworms, 32 candidates per "
                    "glowworm, top 2 exact-scored, grid features h=%.2g A, %d online Adam steps per screening "
                    "step" % (ns, ng, args.sur_grid, args.sur_online),
        "speedup_vs_plain": head["speedup_vs_plain"],
        "speedup_vs_plain_device": head["speedup_vs_plain_device"],
        "seeds": runs,
        "beats_random_at_every_rank": [r["screened_beats_random_at_every_rank"] for r in runs],
        "reaches_plain_top10": [r["screened"]["reached_step"] is not None for r in runs],
        "spearman_heldout": rho, "spearman_within_glowworm": rho_pg,
        "cost_per_batch": dict(cost, candidates=nc),
        "library_forward_reference": lib_ref,
        "break_even": {
            "surrogate_ns_per_candidate": per_cand, "exact_ns_per_pose": per_exact,
            "candidate_over_exact": per_cand / per_exact,
            "screened_ms_per_step": scr_ms, "plain_ms_per_step": plain_ms,
            "plain_device_ms_per_step": head["plain_device"]["ms_per_step"],
            # a 3x wall-time gain needs the plain run's top-10 within this many screened steps
            "steps_for_3x_vs_plain": args.sur_plain_steps * plain_ms / (3.0 * scr_ms),
            "steps_for_3x_vs_plain_device": args.sur_plain_steps * head["plain_device"]["ms_per_step"]
            / (3.0 * scr_ms)},
        "train": {"poses": int(P.shape[0]), "epochs": args.sur_epochs, "data_s": t_data,
                  "train_s": t_train, "loss": hist},
        "cost_by_complex": surrogate_cost_by_complex(args, net),
    }
    feats.close()
    net.close()
    ctx.close()
    return out


if __name__ == "__main__":
    main()
