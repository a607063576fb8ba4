"""Benchmark of the batched GSO pose-scoring path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--family dna] [--no-cpu]

Workload (one step = one scoring launch of one GSO step of 6 swarms x 200
glowworms): synthetic 12,625-atom receptor (RSA, 3.0 A separation, (60, 55, 50) A
ellipsoid) x 1,500-atom ligand, 1,200 poses per step, dna flags (far field +
electrostatics + clash), FP32 pair terms.  Poses for all steps are resident in
HBM before the timed region; each step scores a fresh 1,200-pose slice into a
device buffer.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lklight-gpu_amd"))
sys.path.insert(0, ROOT)

METRIC = "pose-evals/s (dna, 12.6k-atom rec); surrogate/plain speedup at equal top-10 E"
HBM_PEAK_GBS = 8000.0
POSES_PER_STEP = 1200


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pose_bytes(nl, n_far, n_near, desolv=False, nr=0):
    """Algorithmic bytes of one pose evaluation (no-reuse gather model):
    pose 7 f64 + out 2 f64 + ligand 25 B/atom + 8 phi corners per charged atom
    sampled in the grid + one 25 B receptor record per pair within 10 A."""
    b = 56 + 16 + 25 * nl + 32 * n_far + 25 * n_near
    if desolv:
        b += 8 * (nr + nl)
    return b


TRAFFIC_FILE = "profiles/score_kernel_c2_dna_counters.json"


def measured_traffic(workload, family, n):
    """HBM bytes per launch of the scoring kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy-kernel calibration), for
    the workload they were taken on; None otherwise."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != "%s-%s" % (workload, family) or d.get("poses_per_launch") != n:
        return None
    return d["traffic_bytes_per_launch"]


def cpu_baseline(cx, poses, method, threads, budget_s=15.0):
    """The oracle (C restatement of the reference CPU grid path, f64) on the
    host cores, parallel over poses like the reference's rayon loop."""
    from oracle import lkl_oracle as O
    orc = O.Oracle(cx)
    n0 = max(threads * 2, 32)
    t = time.perf_counter()
    r = orc.batch(method, poses[:n0], nthreads=threads, counts=True)
    dt = time.perf_counter() - t
    n = int(min(poses.shape[0], max(n0, n0 * budget_s / max(dt, 1e-3))))
    t = time.perf_counter()
    r = orc.batch(method, poses[:n], nthreads=threads, counts=True)
    dt = time.perf_counter() - t
    return n / dt, n, dt, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--family", default="dna", choices=["dna", "pydock", "vdw", "cpydock"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cell", type=float, default=0.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=-1)
    ap.add_argument("--cell-x", type=float, default=0.0, help="receptor cell edge along x (0 = cubic)")
    ap.add_argument("--list-h", type=float, default=-1.0,
                    help="neighbour-list node spacing in A (0 = cell-row walk; default: the library's 1 A)")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: 12.6k-atom metric workload; c3: 8k x 3k cpydock, 400 poses/step; "
                         "c4: 64-complex batch; c5: surrogate-screened GSO vs plain exact GSO")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="c2: skip the surrogate-screened GSO comparison")
    ap.add_argument("--no-gso", action="store_true",
                    help="c2: skip the host-driven vs device-resident GSO step comparison")
    ap.add_argument("--gso-only", action="store_true",
                    help="c2: only the host-driven vs device-resident GSO step comparison (one JSON line)")
    ap.add_argument("--sur-train", type=int, default=2_000_000, help="c5: training poses")
    ap.add_argument("--sur-epochs", type=int, default=6)
    ap.add_argument("--sur-swarms", type=int, default=40)
    ap.add_argument("--sur-glow", type=int, default=50)
    ap.add_argument("--sur-plain-steps", type=int, default=100)
    ap.add_argument("--sur-max-steps", type=int, default=100)
    ap.add_argument("--sur-grid", type=float, default=0.5)
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

    import numpy as np
    import torch
    import lklight
    from lklight import synth

    if not lklight.device_available():
        raise SystemExit("no gfx950 device: the benchmark measures the HIP path only")
    torch.cuda.init()
    dev = torch.device("cuda:0")
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        print(json.dumps(surrogate_experiment(args)))
        return

    if args.workload == "c3":
        args.family = "cpydock"
        log("generating 8,000-atom receptor / 3,000-atom ligand (radii 1.5-2.0 A)")
        cx = synth.make_config("c3")
    else:
        log("generating 12,625-atom receptor / 1,500-atom ligand")
        cx = synth.make_config("c2")
    t = time.perf_counter()
    cx.build_field()
    log("phi %s built on GPU in %.2f s" % (cx.dims, time.perf_counter() - t))
    desolv = args.family == "cpydock"
    ctx = lklight.Context(cx, desolv=desolv)
    flags = lklight.family_flags(args.family) | (lklight.F_DESOLV if desolv else 0)

    K, W, N = args.steps, args.warmup, (400 if args.workload == "c3" else POSES_PER_STEP)
    poses = synth.make_poses(N * (K + W), cx.meta["rec_axes"], 2024)
    if args.list_h >= 0:
        ctx.set_list_spacing(args.list_h)
    tune = None
    if args.cell or args.block or args.chunk >= 0 or args.cell_x:
        tiles = (args.cell or 5.0, args.block or 256, max(args.chunk, 0), args.cell_x or (args.cell or 5.0))
        ctx.set_tiles(tiles[0], tiles[1], tiles[2], cell_x=args.cell_x)
    elif not args.no_tune:
        t = time.perf_counter()
        cell, cell_x, block, chunk, table = ctx.autotune(poses[:N], flags)
        tiles = (cell, block, chunk, cell_x)
        tune = {"table": [[round(a, 2), round(b, 2), int(c), int(d), round(e, 4)] for a, b, c, d, e in table],
                "columns": ["cell_A", "cell_x_A", "block", "chunk", "ms"],
                "seconds": round(time.perf_counter() - t, 2)}
        log("autotuner: cell %.2f A (x %.2f A), block %d, chunk %d (%.1f s)"
            % (cell, cell_x, block, chunk, tune["seconds"]))
    else:
        tiles = (5.0, 256, 0, 5.0)
    n_list, list_dims = ctx.list_info()
    log("near field: %s" % ("neighbour lists, %d entries over %s nodes" % (n_list, list_dims) if n_list
                            else "cell-row walk"))
    if args.gso_only:
        res = gso_step_compare(cx, ctx, args.family)
        print(json.dumps({"metric": "GSO step time, %s, 6 swarms x 200 glowworms (device-resident)" % args.family,
                          "value": res["device"]["ms_per_step"], "unit": "ms/step", "higher_is_better": False,
                          "n_gpus": 1, "gso_step": res}))
        return
    d_poses = torch.from_numpy(poses).to(dev)
    d_out = torch.zeros((K + W, N, 2), dtype=torch.float64, device=dev)
    d_s = torch.zeros((K + W, N), dtype=torch.float64, device=dev)
    # a dedicated (default-priority) stream: its handle is non-null, so the
    # kernels and the timing events share it
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    assert sh, "expected a non-null stream handle"

    def step(i):
        ctx.score_async(d_poses[i * N:(i + 1) * N].data_ptr(), N, flags, d_out[i].data_ptr(),
                        d_s[i].data_ptr() if desolv else None, sh)

    torch.cuda.synchronize()
    for i in range(W):
        step(i)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        ev[k][0].record(stream)
        step(W + k)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    avg_kernel_s = sum(kern_ms) / len(kern_ms) / 1e3
    value = N * K / wall
    log("timed %d steps: %.3f s wall, kernel avg %.3f ms" % (K, wall, avg_kernel_s * 1e3))

    # The same steps through the blocking host-buffer call (poses H2D, kernel,
    # scores D2H, synchronise): the PCIe-inclusive rate, reported beside
    # `value` (which has the poses resident in HBM), never as it.
    host_poses = [np.ascontiguousarray(poses[(W + k) * N:(W + k + 1) * N]) for k in range(min(K, 20))]
    ctx.score(host_poses[0], flags)
    t0 = time.perf_counter()
    for hp in host_poses:
        ctx.score(hp, flags)
    host_io_rate = N * len(host_poses) / (time.perf_counter() - t0)
    log("host buffers (H2D + kernel + D2H per call): %.0f pose-evals/s" % host_io_rate)

    # CPU baseline (oracle) on a bounded sample of the same poses; its pair
    # counts also give the algorithmic bytes of the roofline figure.
    threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
    cpu = None
    n_near = n_far = None
    sample_poses = poses[W * N:]
    if not args.no_cpu:
        log("cpu baseline: oracle on %d threads" % threads)
        rate, n, dt, r = cpu_baseline(cx, sample_poses, args.family, threads)
        n_near, n_far = float(r["n_near"].mean()), float(r["n_far"].mean())
        g = d_out[W:].reshape(-1, 2)[:n].cpu().numpy()
        sc_g = lklight.fold_scores(args.family, g[:, 0], g[:, 1],
                                   d_s[W:].reshape(-1)[:n].cpu().numpy() if desolv else None)
        sc_o = lklight.fold_scores(args.family, r["e"], r["v"], r["s"] if desolv else None)
        max_rel = float(np.max(np.abs(sc_g - sc_o) / np.maximum(np.abs(sc_o), 1.0)))
        model = "unknown"
        try:
            for ln in open("/proc/cpuinfo"):
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": rate, "unit": "pose-evals/s", "cores": threads, "kind": "port",
               "sample": "first %d poses of the timed workload (%.1f s), %s, f64 C restatement "
                         "of the reference CPU grid path, -O3 -march=x86-64" % (n, dt, model),
               "gpu_over_cpu": value / rate, "max_rel_err_vs_cpu": max_rel}
    b_pose = pose_bytes(cx.lig.n, n_far or 0.0, n_near or 0.0, desolv, cx.rec.n)
    achieved = b_pose * N / avg_kernel_s / 1e9 if n_near is not None else None
    out = {
        "metric": METRIC if (args.workload, args.family) == ("c2", "dna")
        else "pose-evals/s (%s, %s)" % (args.family, "8k-atom rec" if args.workload == "c3" else "12.6k-atom rec"),
        "value": value, "unit": "pose-evals/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": ("c2-%s: 12,625-atom receptor x 1,500-atom ligand, %d poses/step "
                                "(6 swarms x 200 glowworms), phi %s at 0.5 A"
                                % (args.family, N, "x".join(map(str, cx.dims))))
                   if args.workload == "c2" else
                   ("c3-cpydock: 8,000-atom receptor x 3,000-atom ligand (radii 1.5-2.0 A), %d poses/step, "
                    "contact-SASA desolvation over Lee & Richards areas (probe 1.4 A), phi %s at 0.5 A"
                    % (N, "x".join(map(str, cx.dims)))),
                   "family": args.family, "flags": flags, "poses_per_step": N,
                   "cell_edge_A": tiles[0], "cell_edge_x_A": tiles[3], "block": tiles[1],
                   "ligand_atoms_per_tile": tiles[2],
                   "near_field": ({"neighbour_lists": {"spacing_A": args.list_h if args.list_h > 0 else 1.0,
                                                       "entries": n_list, "nodes": list(list_dims)}}
                                  if n_list else "cell-row walk"),
                   "autotune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": measured_traffic(args.workload, args.family, N),
                     "traffic_source": TRAFFIC_FILE if measured_traffic(args.workload, args.family, N) else None,
                     "kernel_ms": avg_kernel_s * 1e3, "bytes_per_pose": b_pose,
                     "n_near_per_pose": n_near, "n_far_per_pose": n_far},
        "cpu_baseline": cpu,
        "host_io_pose_evals_per_s": host_io_rate,
    }
    if args.workload == "c2" and not args.no_gso:
        out["gso_step"] = gso_step_compare(cx, ctx, args.family)
        out["restraints"] = interface_timing(cx, ctx, d_poses[W * N:(W + 1) * N], N, stream)
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


def surrogate_experiment(args):
    """Config-5 comparison on the 1,500-atom / 800-atom complex: train the
    surrogate on the GPU from exact dna scores, then run plain exact GSO and the
    surrogate-screened GSO (and the same screen with random picks) from the
    same start; report rank correlation, top-10 recovery and wall times."""
    import numpy as np
    import torch
    import lklight
    from lklight import screen, synth
    from lklight import surrogate as S

    method = "dna"
    cx = synth.make_config("c1")
    cx.build_field()
    ctx = lklight.Context(cx)
    feats = S.Features(cx, grid=args.sur_grid)
    ns, ng = args.sur_swarms, args.sur_glow
    sizes = [ng] * ns
    pos = synth.make_swarm_positions(ns, ng, cx.meta["rec_axes"], 42, lig_radius=cx.meta["lig_axes"][0])
    # training poses: candidates along a 100-step plain GSO trajectory from
    # other starts (every k-th step, ~80 %) + random poses in the 15 A shell
    t0 = time.perf_counter()
    n_tr = args.sur_train
    per_step = ns * ng * 32
    every = max(1, (100 * per_step) // max(1, (n_tr * 4) // 5))
    pos_tr = synth.make_swarm_positions(ns, ng, cx.meta["rec_axes"], 4242, lig_radius=cx.meta["lig_axes"][0])
    traj = screen.collect_poses(ctx, method, pos_tr, 4242, sizes, 100, 32, args.sur_trans, args.sur_rot,
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
    # held-out: candidates from the benchmark start (different swarms)
    # held-out: candidates at steps 0, 20 and 40 of a plain GSO from the benchmark start
    hold = screen.collect_poses(ctx, method, pos, 99, sizes, 41, 32, args.sur_trans, args.sur_rot, every=20)
    hs = screen.exact_scores(ctx, method, hold)
    hp = screen.predict(net, screen.features_of(feats, hold), hold.shape[0])
    rho = screen.spearman(hp, hs)
    rho_pg = float(np.mean([screen.spearman(hp[i:i + 32], hs[i:i + 32]) for i in range(0, len(hs), 32)]))
    log("held-out spearman %.3f (within-glowworm mean %.3f)" % (rho, rho_pg))
    # plain exact GSO (the yardstick the screened run is compared with)
    plain_sc, plain_wall, _ = screen.run_plain(ctx, method, pos, 42, sizes, args.sur_plain_steps)
    ref10 = screen.top_k(plain_sc)
    log("plain GSO %d steps: %.3f s, top-10 %s" % (args.sur_plain_steps, plain_wall, np.round(ref10, 2)))

    def screened(picker):
        sg = screen.ScreenedGSO(ctx, feats, net, method, pos, 42, sizes, trans_radius=args.sur_trans,
                                rot_sigma=args.sur_rot, picker=picker)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reached, steps_done = None, 0
        for k in range(args.sur_max_steps):
            sg.step()
            steps_done = k + 1
            top = screen.top_k(sg.state()["scoring"])
            if screen.recovered(top, ref10):
                reached = time.perf_counter() - t0
                break
        wall = time.perf_counter() - t0
        top = screen.top_k(sg.state()["scoring"])
        res = {"steps": steps_done, "wall_s": wall, "reached_s": reached, "top10": top.tolist(),
               "exact_evals": sg.exact_evals, "surrogate_evals": sg.surrogate_evals}
        sg.close()
        return res

    sur = screened("surrogate")
    rnd_res = screened("random")
    log("screened: %s" % {k: v for k, v in sur.items() if k != "top10"})
    log("random picks: %s" % {k: v for k, v in rnd_res.items() if k != "top10"})
    # None when the screened run never reached the plain run's top-10 (+-0.5 %)
    speedup = plain_wall / sur["reached_s"] if sur["reached_s"] else None
    sur["ms_per_step"] = sur["wall_s"] / sur["steps"] * 1e3
    rnd_res["ms_per_step"] = rnd_res["wall_s"] / rnd_res["steps"] * 1e3
    out = {
        "workload": "c5: 1,500-atom receptor x 800-atom ligand, %d swarms x %d glowworms, 32 candidates "
                    "per glowworm, top 2 exact-scored, grid features h=%.2g A" % (ns, ng, args.sur_grid),
        "speedup_vs_plain": speedup,
        "plain": {"steps": args.sur_plain_steps, "wall_s": plain_wall, "top10": ref10.tolist(),
                  "ms_per_step": plain_wall / args.sur_plain_steps * 1e3},
        "screened": sur, "random_picks": rnd_res,
        "spearman_heldout": rho, "spearman_within_glowworm": rho_pg,
        "train": {"poses": int(P.shape[0]), "epochs": args.sur_epochs, "data_s": t_data,
                  "train_s": t_train, "loss": hist},
    }
    feats.close()
    net.close()
    ctx.close()
    return out


def bench_c4(args):
    """64 complexes (receptors 2k-20k atoms, ligands 500-5k, 100-500 glowworms,
    dna/pydock/vdw mix): one batch launch per GSO step vs 64
    sequential per-complex launches, both on device-resident poses."""
    import numpy as np
    import torch
    import lklight
    from lklight import synth

    dev = torch.device("cuda:0")
    t = time.perf_counter()
    cxs = synth.make_config4(64)
    log("64 complexes generated in %.1f s" % (time.perf_counter() - t))
    t = time.perf_counter()
    ctxs, flags, counts = [], [], []
    for cx in cxs:
        fam = cx.meta["family"]
        ctxs.append(lklight.Context(cx, with_field=False, device_field=(fam != "vdw"), cell_size=4.0))
        flags.append(lklight.family_flags(fam))
        counts.append(cx.meta["glowworms"])
    log("contexts + device phi in %.1f s" % (time.perf_counter() - t))
    K, W = args.steps, args.warmup
    P = sum(counts)
    step_poses = []
    for s_ in range(K + W):
        step_poses.append(np.concatenate([synth.make_poses(n, cx.meta["rec_axes"], 1000 * s_ + k)
                                          for k, (n, cx) in enumerate(zip(counts, cxs))]))
    m = lklight.Multi(ctxs, flags, counts)
    t = time.perf_counter()
    tb, tm, table = m.autotune(step_poses[0])
    log("batch autotune: block %d, chunk_mult %d (%.1f s)" % (tb, tm, time.perf_counter() - t))
    d_poses = torch.from_numpy(np.stack(step_poses)).to(dev)
    d_out = torch.zeros((K + W, P, 2), dtype=torch.float64, device=dev)
    d_seq = torch.zeros((K + W, P, 2), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    offs = np.concatenate([[0], np.cumsum(counts)])

    def run_batch(i):
        m.score_async(d_poses[i].data_ptr(), d_out[i].data_ptr(), sh)

    def run_seq(i):
        for k, ctx in enumerate(ctxs):
            ctx.score_async(d_poses[i, offs[k]:offs[k + 1]].data_ptr(), counts[k], flags[k],
                            d_seq[i, offs[k]:offs[k + 1]].data_ptr(), None, sh)

    res = {}
    for name, fn in (("batch", run_batch), ("sequential", run_seq)):
        torch.cuda.synchronize()
        for i in range(W):
            fn(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(K):
            fn(W + k)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
    same = bool(torch.equal(d_out[W:], d_seq[W:]))
    # host-driven steps replayed from the captured graph (H2D + kernel + D2H)
    m.use_graph(True)
    host_p = np.stack(step_poses)
    m.score(host_p[0])
    t0 = time.perf_counter()
    for k in range(min(K, 20)):
        m.score(host_p[W + k])
    graph_s = (time.perf_counter() - t0) / min(K, 20)
    fams = {str(f): int(n) for f, n in zip(*np.unique([c.meta["family"] for c in cxs],
                                                      return_counts=True))}
    out = {"metric": "aggregate pose-evals/s, 64-complex batch (one launch per step)",
           "value": P * K / res["batch"], "unit": "pose-evals/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": res["batch"] / K * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "c4: 64 complexes, receptors 2k-20k atoms, ligands 500-5k, "
                                  "%d poses/step" % P, "families": fams,
                      "block": tb, "chunk_mult": tm},
           "sequential_pose_evals_per_s": P * K / res["sequential"],
           "batch_over_sequential": res["sequential"] / res["batch"],
           "bitwise_equal_to_sequential": same,
           "graph_step_ms_host_io": graph_s * 1e3}
    print(json.dumps(out, default=str))


if __name__ == "__main__":
    main()
