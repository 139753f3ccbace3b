"""Benchmark: IPCSolver steps/s and narrow-phase contact pairs/s on one MI355X.

Workload (default, the metric's single-scene configuration): 512 rigid icospheres
(subdivision 3, 642 verts / 1280 faces, r ~ U[0.1, 0.3] m, density 1000) falling in an
open-top container of 5 static boxes, dt 5e-3, dhat 1e-3, mu 0.4, FP64
(scenes.config2). The scene is uploaded once; every step runs entirely on the device
(candidate build, stiffness control, CCD gate, narrow phase, Newton/ALM) with state
resident in HBM. `value` is steps/s of that scene.

The same JSON line carries the 256-scene batch (scenes.config3): aggregate scene-steps/s
of one batched engine, and the speedup over running the same scenes one by one through
the plain single-scene path (a bounded sample of scenes, steps/s summed over scenes),
and the adjoint: gradients/s of the differentiation workload (scenes.config4, 100 steps,
checkpoint every 10) and of the 128-scene batch (scenes.config5, 50 steps), the
backward's cost relative to a plain forward of the same rollout, and a central-difference
check (h = 1e-6) of a few components, and the ipc mode (the primitive-distance barrier
and its maxStep) on c2's own contact candidates: pairs/s and k_ipc_pairs' roofline.

Prints one JSON line. --gpus above 1 is not supported: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s & contact pairs/s: plain vs 256-scene batch; adjoint grads/s (FD-checked)"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TIMED = ("newton", "scene_min_flag", "topk", "np_count", "ccd_fv", "bp_step", "lbvh_build", "alm_post")

# Narrow phase, per candidate (reference formulas, sqrt = div = 1 flop): the FV plane
# test is 34 flops (the barycentric test, 53 more, only runs after it passes); the EE
# segment-segment distance is 67 flops.
FLOPS_FV, FLOPS_EE = 34.0, 67.0
# Newton kernel, compulsory unique HBM bytes of one launch (one Newton solve of every
# scene, IPCSolver.cpp:137-262 forms; perfect on-chip reuse across its iterations and
# line-search trials): per body x, x_hat (6 doubles each), mass and body-frame inertia
# (10 doubles), the step-start pose (7 doubles) read, x written (6 doubles) = 280 B; per
# contact its record (5 ints + 7 doubles) and lagged normal force = 84 B; per joint its
# data, multipliers and stiffness (16 + 6 + 1 doubles) read, multipliers written (6) =
# 232 B.
NEWTON_B_BODY, NEWTON_B_CONTACT, NEWTON_B_JOINT = 280.0, 84.0, 232.0
# Latency floor of the Newton kernel: the objective value and the line-search slopes are
# ordered left-to-right FP64 sums (the reference's loop order, kept for bit parity), so
# every Newton iteration holds at least two chains of (bodies + contacts + joints)
# dependent adds; 4.5 ns per dependent FP64 add on this part (measured with a chained-add
# probe kernel).
FP64_ADD_LATENCY_S = 4.5e-9


def cpu_baseline(sc, stride, steps=1):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    o = O.OracleScene(sc, 0)
    o.set_sampling(stride)
    for _ in range(steps):
        o.step(float(sc.dt[0]))
    t_step, t_min_scaled = o.timing()
    est = t_step - t_min_scaled / stride + t_min_scaled
    return steps / est, t_step


def cpu_baseline_c1(steps=200):
    """The oracle's full C1 run (8 cubes + ground, 200 steps) on one host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from nexdynipc_amd import scenes

    sc = scenes.make_config("c1")
    o = O.OracleScene(sc, 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(float(sc.dt[0]))
    T = time.perf_counter() - t0
    return steps / T, T


def load_traffic(kernel, workload="c2"):
    """HBM bytes per launch of `kernel` (a name or a list of names, summed) from the
    committed counter summary of `workload`, or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    total = 0.0
    for name in [kernel] if isinstance(kernel, str) else kernel:
        k = t.get("kernels", {}).get("k_" + name)
        if not k:
            return None
        total += 1024.0 * (k["fetch_kib"] * t["fetch_factor"] + k["write_kib"] * t["write_factor"])
    return total


def run_plain(Engine, torch, sc, steps, warmup, timing=False, graph=True, autotune=False):
    """steps timed steps after warmup. With `graph` each step replays the captured step
    graph, else (or with `timing`, where HIP events bracket each kernel) the kernels are
    enqueued one by one."""
    e = Engine(0)
    e.upload(sc)
    tune = e.autotune(trial_steps=2) if autotune else None
    e.use_graph(graph)
    e.step(warmup)
    torch.cuda.synchronize()
    ev0 = np.array(e.eval_counter())
    c0 = e.solver_counters()
    if timing:
        e.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    e.stats()  # synchronises the context stream
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    ev1 = np.array(e.eval_counter())
    c1 = e.solver_counters()
    stats = e.stats()
    kern = {name: e.timing_get(name) for name in TIMED} if timing else None
    if timing:
        e.timing_enable(False)
    pairs = e.pair_count()
    e.close()
    return dict(T=T, evals=float(np.sum(ev1 - ev0)), stats=stats, kern=kern, pairs=pairs,
                newton_iters=float(np.sum(c1[:, 6] - c0[:, 6])), trials=float(np.sum(c1[:, 7] - c0[:, 7])),
                tune=None if tune is None else {"grids": [int(v) for v in tune[0]], "ms_per_step": tune[1]})


def run_adjoint(Engine, sc, targets, n_steps, every):
    """Forward with checkpoints + reverse sweep of an n_steps rollout, against the plain
    forward of the same rollout (loss: squared distance of the final positions to the
    targets). Second repetition timed (the first allocates)."""
    from nexdynipc_amd.autograd import Simulation

    sim = Simulation(sc, n_steps, every, engine=Engine(0))
    for _ in range(2):
        sim.reset(sc.v, sc.w, sc.mass, sc.mu)
        sim.e.stats()
        t0 = time.perf_counter()
        sim.e.step(n_steps)
        sim.e.stats()
        t_plain = time.perf_counter() - t0
        sim.reset(sc.v, sc.w, sc.mass, sc.mu)
        sim.e.stats()
        t0 = time.perf_counter()
        sim.e.adjoint_forward(n_steps, every)
        sim.e.stats()
        t_fwd = time.perf_counter() - t0
        p = sim.e.get_state()[0]
        t0 = time.perf_counter()
        g = sim.e.adjoint_backward(dl_dp=2.0 * (p - targets))
        t_bwd = time.perf_counter() - t0
    sim.e.close()
    return dict(plain_s=t_plain, forward_s=t_fwd, backward_s=t_bwd, grads=g,
                gradients_per_s=sc.n_scenes / (t_fwd + t_bwd),
                backward_over_forward=t_bwd / t_plain, total_over_forward=(t_fwd + t_bwd) / t_plain)


def adjoint_fd_check(Engine, sc, targets, n_steps, comps, h=1e-6):
    """Central differences of the loss for a few initial-velocity components (device
    forward runs) against the adjoint: largest relative difference."""
    from nexdynipc_amd.autograd import Simulation

    sim = Simulation(sc, n_steps, n_steps, engine=Engine(0))
    sim.reset(sc.v, sc.w, sc.mass, sc.mu)
    sim.e.adjoint_forward(n_steps, n_steps)
    p = sim.e.get_state()[0]
    g = sim.e.adjoint_backward(dl_dp=2.0 * (p - targets))["v0"]
    worst = 0.0
    for b, c in comps:
        vals = []
        for sgn in (1.0, -1.0):
            v = sc.v.copy()
            v[b, c] += sgn * h
            pp = sim.run(v, sc.w, sc.mass, sc.mu)[0]
            vals.append(float(((pp - targets) ** 2).sum()))
        fd = (vals[0] - vals[1]) / (2.0 * h)
        worst = max(worst, abs(g[b, c] - fd) / abs(fd))
    sim.e.close()
    return worst


# ipc mode, compulsory unique HBM bytes. Filter pass (k_ipc_filter + k_ipc_seg_*), per
# pair: the pair record (5 int32) read, its segment key (8 B) and activity flag (4 B)
# written, its list entry (4 B) when within dhat = 36 B; the bodies' Newton variables and
# orientations and the shared meshes are read once per launch (a few hundred KB, not
# counted). Pair pass (k_ipc_pairs), per pair within dhat: its record and list entry read
# (24 B), its energy, body gradient (12) and Hessian (144 doubles) written (1256 B).
IPC_B_FILTER, IPC_B_ACTIVE = 36.0, 1280.0


def ipc_pairs_from_candidates(sc, ee, fv):
    """The ipc pair list (5 int32 each) of one scene's candidate tuples: each FV candidate
    (face body id, vertex body id, face, vertex) a point-triangle pair (vertex body, face
    body, vertex, face); each EE candidate (body ids, the two edges' vertex ids) an
    edge-edge pair with the edges' unique-edge indices (face order, (min, max))."""
    idx = {int(b): i for i, b in enumerate(sc.body_id)}
    F = sc.faces.reshape(-1, 3)
    cache = {}

    def edge_index(b):
        key = (int(sc.face_off[b]), int(sc.n_faces[b]))
        if key not in cache:
            d = {}
            for f in F[key[0]:key[0] + key[1]]:
                for a, c in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
                    k = (min(a, c), max(a, c))
                    if k not in d:
                        d[k] = len(d)
            cache[key] = d
        return cache[key]

    pt = np.stack([np.zeros(len(fv), dtype=np.int64), [idx[int(v)] for v in fv[:, 1]],
                   [idx[int(v)] for v in fv[:, 0]], fv[:, 3], fv[:, 2]], axis=1) if len(fv) else np.zeros((0, 5))
    rows = []
    for b0, b1, a0, a1, c0, c1 in ee:
        i0, i1 = idx[int(b0)], idx[int(b1)]
        rows.append((1, i0, i1, edge_index(i0)[(min(a0, a1), max(a0, a1))], edge_index(i1)[(min(c0, c1), max(c0, c1))]))
    pe = np.array(rows, dtype=np.int64).reshape(-1, 5)
    return np.ascontiguousarray(np.concatenate([pt, pe]).astype(np.int32))


def run_ipc(Engine, torch, steps=200, dhat=1e-2, kappa=1e5, max_pairs=1 << 20, reps=5):
    """ipc mode on c2's own contact candidates: the scene after `steps` device steps, every
    FV and EE candidate of its close body pairs (top-K lists) as an ipc pair (at most
    max_pairs), evaluated at the current poses (theta = 0) with the PSD projection; then
    maxStep along the next step's free-flight displacement v dt. Kernel times from HIP
    events on the context stream, averaged over `reps` calls."""
    from nexdynipc_amd import scenes

    sc = scenes.make_config("c2")
    e = Engine(0)
    e.upload(sc)
    e.step(steps)
    p, q, v, w = e.get_state()[:4]
    vv, ev, ee, fv = e.candidates(0, float(sc.dhat[0]))
    pairs = ipc_pairs_from_candidates(sc, ee, fv)[:max_pairs]
    x = np.concatenate([p, np.zeros_like(p)], axis=1).reshape(-1)
    dt = float(sc.dt[0])
    dx = np.concatenate([v * dt, w * dt], axis=1).reshape(-1)
    n_close = int(e.pair_count())
    e.ipc_barrier_eval(0, pairs, x, dhat, kappa, True, cap=4 * n_close + 16)  # warm-up
    e.ipc_max_step(0, pairs, x, dx)
    torch.cuda.synchronize()
    e.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        val, segb, _, _, active = e.ipc_barrier_eval(0, pairs, x, dhat, kappa, True, cap=4 * n_close + 16,
                                                     with_active=True)
    t_call = (time.perf_counter() - t0) / reps
    for _ in range(reps):
        alpha, toi = e.ipc_max_step(0, pairs, x, dx)
    kf, nk = e.timing_get("ipc_filter")
    kp, _ = e.timing_get("ipc_pairs")
    kr, _ = e.timing_get("ipc_reduce")
    ka, _ = e.timing_get("ipc_advance")
    e.timing_enable(False)
    e.close()
    n = len(pairs)
    # the pair pass (the dominant kernel) writes the terms of the pairs within dhat; the
    # filter pass streams every pair
    b_pairs = IPC_B_ACTIVE * active
    b_filter = IPC_B_FILTER * n
    achieved = b_pairs / (kp * 1e-3) / 1e9
    achieved_f = b_filter / (kf * 1e-3) / 1e9
    return {"workload": f"c2 after {steps} steps: its FV and EE candidates (top-K lists of every close body "
                        f"pair, {n_close} pairs) as {n} ipc pairs ({int((pairs[:, 0] == 0).sum())} PT, "
                        f"{int((pairs[:, 0] == 1).sum())} EE), dhat {dhat} m, kappa {kappa:g}, PSD projection",
            "pairs": n, "active_pairs": active, "segments": int(len(segb)),
            "pairs_per_s": n / ((kf + kp + kr) * 1e-3),
            "kernel_ms": {"ipc_filter": kf, "ipc_pairs": kp, "ipc_reduce": kr},
            "active_pairs_per_s": active / (kp * 1e-3) if kp > 0 else None,
            "host_call_ms": 1e3 * t_call,
            "host_call_note": "ndipc_ipc_barrier_eval from host buffers: allocation, PCIe copies, launches",
            "max_step": {"alpha": alpha, "toi": toi, "kernel_ms": ka, "pairs_per_s": n / (ka * 1e-3)},
            "roofline": {"kernel": "k_ipc_pairs", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": b_pairs,
                         "avg_launch_ms": kp, "launches": nk,
                         "traffic": load_traffic("ipc_pairs", "ipc"),
                         "note": "latency-bound: per pair a 12 x 12 Jacobi eigen-decomposition whose rotations "
                                 "are dependent LDS steps, one wavefront per SIMD"},
            "roofline_filter": {"kernel": "k_ipc_filter + k_ipc_seg_*", "bound": "hbm", "achieved": achieved_f,
                                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_f / HBM_PEAK_GBS,
                                "algorithmic_bytes_per_launch": b_filter, "avg_launch_ms": kf, "launches": nk,
                                "traffic": load_traffic(["ipc_filter", "ipc_seg_count", "ipc_seg_write"], "ipc")}}


def run_ipc_step(Engine, torch, steps=100, warmup=5, cpu_steps=2):
    """The ipc mode's time step on a grid of 32 cubes dropped onto a ground box
    (scenes.ipc_drop_scene: 4 x 4 x 2 cubes of 0.2 m), dhat 1 cm: steps/s on the device
    (inputs resident; each step's Newton loop driven from the host), and the oracle's
    same steps on one host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from nexdynipc_amd import scenes

    sc = scenes.ipc_drop_scene()
    kw = dict(dhat=1e-2, kappa=1e5, inflation=0.05)
    e = Engine(0)
    e.upload(sc)
    for _ in range(warmup):
        e.ipc_step(0, 1e-2, **kw)
    torch.cuda.synchronize()
    newton = trials = pcg = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = e.ipc_step(0, 1e-2, **kw)
        newton += st[0]
        trials += st[1]
        pcg += st[3]
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    pairs = st[2]
    e.close()
    import _oracle as O

    o = O.OracleScene(sc)
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        o.ipc_step(1e-2, **kw)
    Tc = time.perf_counter() - t0
    return {"workload": f"ipc step: {sc.n_bodies - 1} cubes (4 x 4 x 2, 0.2 m) dropped onto a ground box, dt 1e-2, "
                        "dhat 1 cm, kappa 1e5, pairs from the top-K lists at inflation 5 cm",
            "steps": steps, "steps_per_s": steps / T, "ms_per_step": 1e3 * T / steps,
            "newton_iters_per_step": newton / steps, "line_search_trials_per_step": trials / steps,
            "pcg_iters_per_step": pcg / steps, "pair_records": pairs,
            "cpu_steps_per_s": cpu_steps / Tc, "cpu_cores": 1, "cpu_kind": "port",
            "cpu_sample": f"the first {cpu_steps} steps of the same scene by the C++ oracle on 1 core (the device "
                          f"figure is steps {warmup + 1}-{warmup + steps})"}


def roofline_for(name, r, sc):
    """Roofline of kernel `name` over the timed region: algorithmic work / total time."""
    avg_ms, cnt = r["kern"][name]
    t = avg_ms * cnt * 1e-3
    if t <= 0:
        return None
    stats = r["stats"]
    if name == "np_count":
        n_fv, n_ee = float(np.sum(stats[:, 3])), float(np.sum(stats[:, 4]))
        f_fv = n_fv / max(n_fv + n_ee, 1.0)
        flops = r["evals"] * (FLOPS_FV * f_fv + FLOPS_EE * (1.0 - f_fv))
        achieved = flops / t / 1e12
        roof = {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS}
    elif name == "newton":
        contacts = float(np.sum(stats[:, 2]))
        per_launch = NEWTON_B_BODY * sc.n_bodies + NEWTON_B_CONTACT * contacts + NEWTON_B_JOINT * sc.n_joints
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": per_launch}
        # the kernel is latency-bound: its ordered FP64 sums against the dependent-add floor
        chain = sc.n_bodies + contacts + sc.n_joints
        floor_s = r["newton_iters"] * 2.0 * chain * FP64_ADD_LATENCY_S
        roof["latency_floor"] = {"dependent_fp64_adds_per_launch": 2.0 * chain * r["newton_iters"] / cnt,
                                 "add_latency_ns": FP64_ADD_LATENCY_S * 1e9,
                                 "floor_ms_per_launch": 1e3 * floor_s / cnt,
                                 "achieved_over_floor": t / floor_s if floor_s > 0 else None}
    else:
        return None
    roof["traffic"] = load_traffic(name)
    if roof["traffic"] is not None and name == "newton":
        roof["traffic_over_algorithmic"] = roof["traffic"] / roof["algorithmic_bytes_per_launch"]
    roof["kernel"] = "k_" + name
    roof["avg_launch_ms"] = avg_ms
    roof["launches"] = cnt
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", action="store_true", help="measure only the 256-scene batch (config 3)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch / one-by-one comparison")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--one-by-one", type=int, default=24, help="scenes sampled for the one-by-one comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint legs (c4, c5)")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--cpu-stride", type=int, default=2000)
    ap.add_argument("--ipc-only", action="store_true", help="only the ipc-mode leg (for profiling)")
    ap.add_argument("--no-ipc", action="store_true", help="skip the ipc-mode leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 20 device copies of 256 MiB (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs on one GPU "
              "(batched scenes shard by scene across GPUs, which is not built here).", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        a = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_copy": a.numel() * 8, "copies": 20}))
        return

    from nexdynipc_amd import scenes
    from nexdynipc_amd._lib import Engine

    if args.ipc_only:
        print(json.dumps({"ipc": run_ipc(Engine, torch), "ipc_step": run_ipc_step(Engine, torch)}))
        return

    batch = None
    if not args.no_batch:
        specs = scenes.config3()
        scb = scenes.pack(specs)
        rb = run_plain(Engine, torch, scb, args.batch_steps, args.warmup, autotune=True)
        batch_rate = args.batch_steps * scb.n_scenes / rb["T"]
        # the same scenes one by one through the single-scene path: a seeded sample of
        # scenes, each on its own engine, scene-steps/s = steps / (sum of wall times)
        rng = np.random.default_rng(7)
        pick = rng.choice(len(specs), size=min(args.one_by_one, len(specs)), replace=False)
        one = {}
        for graph in (False, True):
            t_sum, st_sum = 0.0, 0
            for i in pick:
                r1 = run_plain(Engine, torch, scenes.pack([specs[int(i)]]), args.batch_steps, args.warmup,
                               graph=graph)
                t_sum += r1["T"]
                st_sum += args.batch_steps
            one[graph] = st_sum / t_sum if t_sum > 0 else None
        one_rate, one_rate_graph = one[False], one[True]
        batch = {"workload": "c3: 256 scenes, log-U[4,128] bodies each (cube / icosphere s2 / 2x1x0.5 box), "
                             "ground box, dt 1e-2", "steps": args.batch_steps,
                 "scene_steps_per_s": batch_rate, "ms_per_step": 1e3 * rb["T"] / args.batch_steps,
                 "contact_pairs_per_s": rb["evals"] / rb["T"],
                 "one_by_one_scene_steps_per_s": one_rate,
                 "one_by_one_sample": f"{len(pick)} of {len(specs)} scenes, the same {args.batch_steps} steps "
                                      "after the same warm-up, each scene on its own context through the "
                                      "single-scene path",
                 "speedup_vs_one_by_one": batch_rate / one_rate if one_rate else None,
                 "autotune": rb["tune"],
                 "one_by_one_graph_scene_steps_per_s": one_rate_graph,
                 "speedup_vs_one_by_one_graph": batch_rate / one_rate_graph if one_rate_graph else None,
                 "note": "batch: one launch per stage for all scenes, step replayed from a HIP graph; "
                         "one_by_one: the plain single-scene call, kernels enqueued per step; "
                         "one_by_one_graph: the single-scene call with the step graph as well"}
        if args.batch:
            out = {"metric": METRIC, "value": batch_rate, "unit": "scene-steps/s", "n_gpus": 1,
                   "steps": args.batch_steps, "warmup": args.warmup, "ms_per_step": batch["ms_per_step"],
                   "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                   "data": "synthetic", "config": {"workload": batch["workload"], "scenes": int(scb.n_scenes),
                                                    "bodies": int(scb.n_bodies)},
                   "batch": batch, "roofline": None, "cpu_baseline": None}
            print(json.dumps(out))
            return

    adjoint = None
    if not args.no_adjoint:
        spec4, t4 = scenes.config4()
        sc4 = scenes.pack([spec4])
        a4 = run_adjoint(Engine, sc4, np.asarray(t4), args.adjoint_steps, 10)
        fd_worst = adjoint_fd_check(Engine, sc4, np.asarray(t4), 10, [(20, 0), (33, 2), (47, 1), (58, 0)])
        sc5 = scenes.pack(scenes.config5(128))
        p5 = sc5.p + 0.0
        a5 = run_adjoint(Engine, sc5, p5, 50, 10)
        adjoint = {
            "c4": {"workload": "c4: 16-link hinge chain (link 0 pinned) + 48 cubes on a 20 deg ramp, "
                               f"{args.adjoint_steps} steps, checkpoint every 10; loss sum |p - target|^2",
                   "gradients_per_s": a4["gradients_per_s"], "plain_forward_s": a4["plain_s"],
                   "recorded_forward_s": a4["forward_s"], "backward_s": a4["backward_s"],
                   "backward_over_forward": a4["backward_over_forward"],
                   "total_over_forward": a4["total_over_forward"],
                   "fd_check": {"h": 1e-6, "steps": 10, "components": 4, "max_rel_err": fd_worst}},
            "c5": {"workload": "c5: 128 scenes x 32 cubes thrown at a wall and a floor, 50 steps, one batched "
                               "forward + adjoint; loss sum |p - p0|^2",
                   "scene_gradients_per_s": a5["gradients_per_s"], "plain_forward_s": a5["plain_s"],
                   "recorded_forward_s": a5["forward_s"], "backward_s": a5["backward_s"],
                   "backward_over_forward": a5["backward_over_forward"],
                   "total_over_forward": a5["total_over_forward"]},
        }

    sc = scenes.make_config("c2")
    workload = "c2: 512 icospheres (subdiv 3) + 5-box open container, single scene"
    r = run_plain(Engine, torch, sc, args.steps, args.warmup, autotune=True)
    T = r["T"]
    stats = r["stats"]
    # the same steps again, kernels enqueued one by one between HIP events
    rt = run_plain(Engine, torch, sc, args.steps, args.warmup, timing=True)
    totals = {k: v[0] * v[1] for k, v in rt["kern"].items()}
    dominant = max(totals, key=totals.get)
    roof = roofline_for(dominant, rt, sc)
    if roof is not None:
        roof["share_of_timed_kernels"] = totals[dominant] / max(sum(totals.values()), 1e-30)
    cpu = None
    c1 = None
    if not args.no_cpu_baseline:
        v, wall = cpu_baseline(sc, args.cpu_stride)
        cpu = {"value": v, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"1 step of c2 by the C++ oracle on 1 host core; computeSceneMinDistance's vertex loop "
                         f"sampled 1/{args.cpu_stride} and its time scaled x{args.cpu_stride} "
                         f"(measured wall {wall:.1f} s)"}
        # C1 (the reference's own CPU-runnable case) fully measured on both sides
        sc1 = scenes.make_config("c1")
        r1 = run_plain(Engine, torch, sc1, 200, 0)
        v1, wall1 = cpu_baseline_c1(200)
        c1 = {"workload": "c1: 8 cubes + ground box, 200 steps, dt 1e-2", "gpu_steps_per_s": 200 / r1["T"],
              "cpu_steps_per_s": v1, "cpu_cores": 1, "cpu_kind": "port",
              "cpu_sample": f"all 200 steps of c1 by the C++ oracle on 1 host core (wall {wall1:.2f} s)",
              "gpu_note": "the same 200 steps from the start state (graph capture inside the timing), "
                          "single-scene path with step graphs"}
    out = {
        "metric": METRIC,
        "value": args.steps / T,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * T / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "bodies": int(sc.n_bodies), "scenes": int(sc.n_scenes),
                   "verts": int(sc.total_verts()), "dt": float(sc.dt[0]), "dhat": float(sc.dhat[0])},
        "contact_pairs_per_s": r["evals"] / T,
        "last_step": {"alm_iters": float(np.mean(stats[:, 0])), "newton_iters": float(np.mean(stats[:, 1])),
                      "contacts": float(np.sum(stats[:, 2])), "fv_candidates": float(np.sum(stats[:, 3])),
                      "ee_candidates": float(np.sum(stats[:, 4])), "close_pairs": int(r["pairs"])},
        "newton_iterations": r["newton_iters"],
        "line_search_trials": r["trials"],
        "kernel_total_ms": totals,
        "ms_per_step_kernels_one_by_one": 1e3 * rt["T"] / args.steps,
        "autotune": r["tune"],
        "roofline": roof,
        "roofline_narrow_phase": roofline_for("np_count", rt, sc),
        "batch": batch,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "c1": c1,
        "ipc": None if args.no_ipc else run_ipc(Engine, torch),
        "ipc_step": None if args.no_ipc else run_ipc_step(Engine, torch),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
