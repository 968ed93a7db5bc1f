"""Benchmark: edge-state collision checks per second on the 4 x 7-DoF arm workload (one MI355X).

Workload (seed 0): 4 arms x 7 revolute DH joints (28-D), 4..8 spheres per link, 50 boxes on a
1.6 m table; 131,072 edges with uniform starts and lognormal lengths (median 0.3 rad, sigma 1.0),
resolution 0.05 rad.  A step is one batched check of all edges (length-sorted persistent queue,
hipGraph replay) with inputs resident in HBM.  value = logical edge-states per second: every
interior state of every edge counts once, whatever the early exit (sum over edges of N_e - 2).

Also measured in the same run: the plain per-call form (one launch per edge), the unsorted single
launch, 1,000,000 configuration checks, an all-free edge variant (no early exit) whose kernel time
gives the roofline fraction, and the CPU oracle on the host cores (bounded sample) with the
boolean mismatch count of the GPU against it.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baselines' OpenMP workers sleep between parallel regions instead of spinning, so they do
# not hold the host cores that launch the next measurements' GPU work.
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
PKG = "multirobot-pathplanning-benchmark_amd"

PEAK_FP32_TFLOPS = 157.3  # MI355X vector FP32 (AMD published)
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def arm_flops_per_state(desc):
    """Algorithmic FP32-equivalent flops of one state, brute-force pair set (FMA = 2)."""
    dof, na = desc.dof, desc.n_arms
    cnt = np.asarray(desc.link_sphere_count).reshape(na, dof)
    S = int(cnt.sum())
    nb = len(desc.boxes)
    self_pairs = 0
    for a in range(na):
        for i in range(dof):
            for j in range(i + 2, dof):
                self_pairs += int(cnt[a, i] * cnt[a, j])
    per_arm = cnt.sum(axis=1)
    inter_pairs = sum(int(per_arm[a] * per_arm[b]) for a in range(na) for b in range(a + 1, na))
    D = na * dof
    flops = 2 * D + 72 * D + 18 * S + 15 * S * nb + 11 * (self_pairs + inter_pairs)
    return flops, {"spheres": S, "boxes": nb, "self_pairs": self_pairs, "inter_pairs": inter_pairs}


def logical_states(q1, q2, res):
    N = np.maximum(2, (np.max(np.abs(q2 - q1), axis=1) / res).astype(np.int64) + 1)
    return int(np.sum(N - 2)), N


def roofline_all_free(args, backend, sc, desc, rng, dev, res, fl_state, pair_info):
    """Roofline workload of the persistent arm kernel: edges whose every state is free, so no
    state exits early and the kernel runs its whole (broad phase + narrow phase) test set.
    The kernel time comes from HIP events recorded around the kernel on the run's stream."""
    import torch

    stream = torch.cuda.current_stream()
    # every state of these edges is free, so no early exit: the kernel runs its whole pair set.
    fq1, fq2 = free_edges(rng, desc, sc, args.edges, res, dev)
    f_states, fN = logical_states(fq1, fq2, res)
    fa, fb = torch.from_numpy(fq1).to(dev), torch.from_numpy(fq2).to(dev)
    fout = torch.empty(len(fq1), dtype=torch.uint8, device=dev)
    fbatch = sc.edge_batch(fa, fb, fout, res)

    def timed(plan):
        plan.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            plan.run()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    # (a) as shipped (link broad phase on): executed tests counted by the kernel
    f_ms = timed(backend.EdgePlan(sc, fbatch, backend.make_tune(use_graph=1)))
    f_split = timed_split(backend.EdgePlan(sc, fbatch, backend.make_tune(use_graph=0, time_kernels=1)),
                          args.steps)
    k_ms = f_split["kernel_ms"]
    assert fout.cpu().numpy().all(), "all-free workload reported a collision"
    cst = backend.EdgePlan(sc, fbatch, backend.make_tune(use_graph=0, collect_stats=1))
    cst.run()
    fst = cst.stats()
    del cst
    S_, L_, D_ = pair_info["spheres"], desc.n_arms * desc.dof, desc.dim
    per_state_fixed = 2 * D_ + 72 * D_ + 18 * S_ + 18 * L_
    exec_flops = (fst["states"] * per_state_fixed + 15 * fst["link_box_tests"] + 15 * fst["sphere_box_tests"]
                  + 11 * fst["link_pair_tests"] + 11 * fst["sphere_pair_tests"])
    achieved = exec_flops / (k_ms * 1e-3) / 1e12
    free_rate = f_states / (f_ms * 1e-3)
    log(f"all-free (broad phase): {len(fq1)} edges, {f_states} states, {f_ms:.3f} ms/run "
        f"(persistent kernel {k_ms:.3f} ms) -> {free_rate:.4g} states/s, "
        f"executed {exec_flops / max(f_states, 1):.0f} flop/state -> {achieved:.2f} TFLOP/s "
        f"({achieved / PEAK_FP32_TFLOPS * 100:.1f}% of FP32 VALU peak); stats {fst}")
    # (b) brute force: the full pair set of every state
    sc.set_broadphase(False)
    bf_split = timed_split(backend.EdgePlan(sc, fbatch, backend.make_tune(use_graph=0, time_kernels=1)),
                           args.steps)
    sc.set_broadphase(True)
    bf_ms = bf_split["kernel_ms"]
    bf_achieved = f_states * fl_state / (bf_ms * 1e-3) / 1e12
    log(f"all-free (brute force): {bf_ms:.3f} ms -> {bf_achieved:.2f} TFLOP/s "
        f"({bf_achieved / PEAK_FP32_TFLOPS * 100:.1f}% of FP32 VALU peak)")

    return {"achieved": achieved, "k_ms": k_ms, "f_ms": f_ms, "f_split": f_split, "exec_flops": exec_flops,
            "f_states": f_states, "free_rate": free_rate, "n_edges": len(fq1), "bf_achieved": bf_achieved,
            "bf_ms": bf_ms, "fst": fst, "fq1": fq1, "fq2": fq2}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--edges", type=int, default=131072)
    ap.add_argument("--configs", type=int, default=1000000)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--per-call-edges", type=int, default=4096)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-secondary", action="store_true")
    ap.add_argument("--mixed-instances", type=int, default=65536)
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline workload of the persistent kernel (for profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one device copy of 256 MiB (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--cost-only", action="store_true",
                    help="only the configuration-cost workload (for profiling runs)")
    ap.add_argument("--clearance-only", action="store_true",
                    help="only the clearance cost + gradient workload (for profiling runs)")
    ap.add_argument("--ur10-only", action="store_true",
                    help="only the four-UR10e mujoco-world workload (for profiling runs)")
    ap.add_argument("--per-call-only", action="store_true",
                    help="only the plain per-call form on the headline edges (for kernel traces)")
    ap.add_argument("--lane-per-call-only", action="store_true",
                    help="only the plain per-call form on the C1 and C2 edges (for kernel traces)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline batch: warmup + timed graph replays (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 (sharding edges over "
                                   "several GPUs is not built)"}))
        sys.exit(2)

    import torch

    pkg = importlib.import_module(PKG)
    backend = importlib.import_module(f"{PKG}.backend")
    scenes = pkg.scenes
    assert torch.cuda.is_available(), "bench needs a HIP device"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.clearance_only:
        ent = clearance_workload(args, backend, scenes, dev, len(os.sched_getaffinity(0)))
        print(json.dumps({"clearance_only": True, **ent}), flush=True)
        return
    if args.ur10_only:
        ent = ur10_workload(args, backend, scenes, dev, len(os.sched_getaffinity(0)))
        print(json.dumps({"ur10_only": True, **(ent or {"error": "fixture missing"})}), flush=True)
        return
    if args.lane_per_call_only:
        for name, desc, q1, q2, res in lane_workloads(scenes):
            n_states, N = logical_states(q1, q2, res)
            rate, _, m, us = per_call_rate(backend.Scene(desc), q1, q2, res, N, dev, m=args.per_call_edges)
            print(json.dumps({"lane_per_call_only": True, "workload": name, "edges": m, "us_per_edge": us,
                              "edge_states_per_s": rate}), flush=True)
        return
    if args.cost_only:
        ent = cost_workload(args, dev, len(os.sched_getaffinity(0)))
        print(json.dumps({"cost_only": True, **ent}), flush=True)
        if args.copy_probe:
            copy_probe(dev)
        return

    rng = np.random.default_rng(0)
    desc = scenes.quad_arms(rng=rng)
    res = 0.05
    sc = backend.Scene(desc)
    fl_state, pair_info = arm_flops_per_state(desc)

    # ---- headline workload -----------------------------------------------------------------
    Qc = scenes.uniform_configs(rng, desc, args.configs).astype(np.float32).astype(np.float64)
    q1, q2 = scenes.lognormal_edges(rng, desc, args.edges, res)
    q1 = q1.astype(np.float32).astype(np.float64)
    q2 = q2.astype(np.float32).astype(np.float64)
    n_states, N = logical_states(q1, q2, res)
    log(f"workload: {args.edges} edges, {n_states} logical edge-states (mean N {N.mean():.2f}), "
        f"{fl_state} flops/state ({pair_info})")
    if args.roofline_only:  # same random stream as the full run up to this point
        rf = roofline_all_free(args, backend, sc, desc, rng, dev, res, fl_state, pair_info)
        print(json.dumps({"roofline_only": True, "kernel": "k_arm_persistent (all-free edges)",
                          "ms_per_launch": rf["k_ms"], "run_split_ms": rf["f_split"],
                          "achieved_tflops": rf["achieved"], "flops_per_launch": rf["exec_flops"],
                          "states_per_launch": rf["f_states"], "brute_force_ms": rf["bf_ms"],
                          "stats": rf["fst"]}), flush=True)
        if args.copy_probe:
            copy_probe(dev)
        return
    a = torch.from_numpy(q1).to(dev)
    b = torch.from_numpy(q2).to(dev)
    if args.per_call_only:  # one launch per edge over the first per_call_edges edges, twice
        m = min(args.per_call_edges, args.edges)
        o1 = torch.empty(m, dtype=torch.uint8, device=dev)
        b1 = sc.edge_batch(a[:m], b[:m], o1, res)
        sc.check_edges_per_call_device(b1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sc.check_edges_per_call_device(b1)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(json.dumps({"per_call_only": True, "edges": m, "us_per_edge": dt / m * 1e6,
                          "edge_states_per_s": int(np.sum(N[:m] - 2)) / dt}), flush=True)
        return
    out = torch.empty(args.edges, dtype=torch.uint8, device=dev)
    batch = sc.edge_batch(a, b, out, res)
    plan = backend.EdgePlan(sc, batch, backend.make_tune(use_graph=1))
    for _ in range(args.warmup):
        plan.run()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.run()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    gpu_free = out.cpu().numpy().astype(bool)
    value = n_states / (wall / args.steps)
    log(f"batched persistent: {ms_step:.3f} ms/step (events), {value:.4g} edge-states/s (wall), "
        f"{gpu_free.mean():.3f} edges free")
    if args.headline_only:
        print(json.dumps({"headline_only": True, "ms_per_step": ms_step, "value": value,
                          "edges_free": float(gpu_free.mean())}), flush=True)
        return

    # stats of one run (separate plan with counters)
    splan = backend.EdgePlan(sc, batch, backend.make_tune(use_graph=0, collect_stats=1))
    splan.run()
    stats = splan.stats()
    del splan
    # queue build / persistent kernel split of a run (HIP events on the run's stream, no graph)
    head_split = timed_split(backend.EdgePlan(sc, batch, backend.make_tune(use_graph=0, time_kernels=1)),
                             args.steps)
    log(f"headline run split: {head_split}")

    # ---- unsorted single launch (ablation) ---------------------------------------------------
    out2 = torch.empty_like(out)
    batch2 = sc.edge_batch(a, b, out2, res)
    plan2 = backend.EdgePlan(sc, batch2, backend.make_tune(persistent=0, sorted=0, use_graph=1))
    plan2.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(3, args.steps // 4)
    for _ in range(reps):
        plan2.run()
    torch.cuda.synchronize()
    single_rate = n_states / ((time.perf_counter() - t0) / reps)
    single_match = bool(np.array_equal(out2.cpu().numpy().astype(bool), gpu_free))

    # ---- plain per-call form -----------------------------------------------------------------
    m = min(args.per_call_edges, args.edges)
    pc_states = int(np.sum(N[:m] - 2))
    out3 = torch.empty(m, dtype=torch.uint8, device=dev)
    batch3 = sc.edge_batch(a[:m], b[:m], out3, res)
    sc.check_edges_per_call_device(batch3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sc.check_edges_per_call_device(batch3)
    torch.cuda.synchronize()
    pc_time = time.perf_counter() - t0
    per_call_value = pc_states / pc_time
    per_call_match = bool(np.array_equal(out3.cpu().numpy().astype(bool), gpu_free[:m]))
    log(f"plain per-call: {m} launches in {pc_time * 1e3:.1f} ms -> {per_call_value:.4g} edge-states/s; "
        f"single launch {single_rate:.4g}")
    # the same per-call form as the planners reach it: MrccEnvironment.is_edge_collision_free from
    # Python on host configurations (rows and verdict in the scene's pinned block, one launch per call)
    env = backend.MrccEnvironment(desc, collision_resolution=res)
    mp = min(1000, m)
    env.is_edge_collision_free(q1[0], q2[0], None)
    t0 = time.perf_counter()
    py_free = [env.is_edge_collision_free(q1[i], q2[i], None) for i in range(mp)]
    py_time = time.perf_counter() - t0
    py_rate = int(np.sum(N[:mp] - 2)) / py_time
    py_match = bool(np.array_equal(np.array(py_free), gpu_free[:mp]))
    log(f"per-call through MrccEnvironment.is_edge_collision_free: {py_time / mp * 1e6:.1f} us/call -> "
        f"{py_rate:.4g} edge-states/s")

    # ---- configuration checks ----------------------------------------------------------------
    qc = torch.from_numpy(Qc).to(dev)
    outc = torch.empty(len(Qc), dtype=torch.uint8, device=dev)
    sc.check_configs_device(qc, outc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        sc.check_configs_device(qc, outc)
    torch.cuda.synchronize()
    cfg_rate = 3 * len(Qc) / (time.perf_counter() - t0)
    cfg_free = outc.cpu().numpy().astype(bool)
    log(f"configs: {cfg_rate:.4g} checks/s, {cfg_free.mean():.4f} free")

    # ---- all-free variant: roofline of the persistent kernel ----------------------------------
    rf = roofline_all_free(args, backend, sc, desc, rng, dev, res, fl_state, pair_info)
    achieved, k_ms, f_ms = rf["achieved"], rf["k_ms"], rf["f_ms"]
    free_check = None
    if not args.skip_cpu:
        # the all-free set is certified by the oracle, not by the GPU that filtered it
        from oracle import oracle as O

        kf = min(8192, len(rf["fq1"]))
        t0 = time.perf_counter()
        fref, fn = O.Scene(desc).check_edges(rf["fq1"][:kf], rf["fq2"][:kf], res, threads=len(os.sched_getaffinity(0)),
                                              count=True)
        free_check = {"edges": kf, "states_evaluated_by_oracle": int(fn), "not_free_by_oracle": int(np.sum(~fref)),
                      "seconds": time.perf_counter() - t0}
        log(f"all-free set vs oracle: {free_check}")
        # the planners' per-call form on long free edges (every state evaluated): the oracle-free
        # edges of this set with more than 20 states
        _, fN8 = logical_states(rf["fq1"][:kf], rf["fq2"][:kf], res)
        sel = np.where(fref & (fN8 > 20))[0][:300]
        if len(sel):
            env.is_edge_collision_free(rf["fq1"][sel[0]], rf["fq2"][sel[0]], None)
            t0 = time.perf_counter()
            got = [env.is_edge_collision_free(rf["fq1"][i], rf["fq2"][i], None) for i in sel]
            dt = time.perf_counter() - t0
            free_check["python_env_long_free"] = {
                "edges": int(len(sel)), "mean_N": float(fN8[sel].mean()), "us_per_call": dt / len(sel) * 1e6,
                "matches": bool(all(got)),
                "note": "MrccEnvironment.is_edge_collision_free on oracle-free edges of more than 20 states"}
            log(f"per-call on long free edges: {free_check['python_env_long_free']}")

    # ---- CPU baseline: the oracle on the host cores -------------------------------------------
    cpu = None
    mism = None
    checked = 0
    if not args.skip_cpu:
        from oracle import oracle as O  # the CPU baseline / checker

        o = O.Scene(desc)
        threads = len(os.sched_getaffinity(0))
        # calibrate on 2048 edges, then a sample of about cpu-seconds of work
        t0 = time.perf_counter()
        cal = min(2048, args.edges)
        o.check_edges(q1[:cal], q2[:cal], res, threads=threads)
        dt = max(time.perf_counter() - t0, 1e-4)
        k = int(min(args.edges, max(cal, cal * args.cpu_seconds / dt)))
        reps = 0
        t0 = time.perf_counter()
        while True:  # repeat the sample until about cpu-seconds of work
            ref, n_phys = o.check_edges(q1[:k], q2[:k], res, threads=threads, count=True)
            reps += 1
            cpu_t = time.perf_counter() - t0
            if cpu_t >= args.cpu_seconds or reps >= 100:
                break
        ks = int(np.sum(N[:k] - 2))
        cpu = {"value": ks * reps / cpu_t, "unit": "edge-states/s", "cores": threads, "kind": "port",
               "sample": f"first {k} of the {args.edges} benchmark edges ({ks} edge-states) x {reps} passes, "
                         f"FP64 oracle, OpenMP over edges, {cpu_t:.1f} s"}
        mism = int(np.sum(ref != gpu_free[:k]))
        checked += ks
        cpu["compared"] = {"edges": k, "logical_edge_states": ks, "states_evaluated_by_oracle": int(n_phys)}
        # the same oracle on one thread, over a 2 s sample of the same edges
        k1 = max(1, min(k, int(k * 2.0 / (cpu_t / reps) / threads)))
        t0 = time.perf_counter()
        o.check_edges(q1[:k1], q2[:k1], res, threads=1)
        t1 = time.perf_counter() - t0
        cpu["one_thread"] = {"value": int(np.sum(N[:k1] - 2)) / t1, "cores": 1, "edges": k1}
        log(f"cpu oracle, one thread: {cpu['one_thread']['value']:.4g} edge-states/s ({k1} edges)")
        # every configuration of the workload
        refc = o.check_configs(Qc, threads=threads)
        mism += int(np.sum(refc != cfg_free))
        checked += len(Qc)
        log(f"cpu oracle: {cpu['value']:.4g} edge-states/s on {threads} threads; mismatches {mism} / {checked}")

    secondary = {} if args.skip_secondary else secondary_workloads(args, backend, scenes, dev)
    if args.copy_probe:
        copy_probe(dev)

    line = {
        "metric": "edge-state collision checks/sec (bit-exact), batched persistent, 4x7-DoF arms",
        "value": value,
        "unit": "edge-states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 narrow phase, f64 kinematics and near-contact recheck",
        "data": "synthetic (seeded generator: DH arms, spheres, boxes, uniform starts, lognormal lengths)",
        "config": {"workload": "quad_arms_7dof: 4 arms x 7 DoF, 50 boxes, 131072 lognormal edges @ 0.05 rad",
                   "edges": args.edges, "edge_states": n_states, "resolution": res,
                   "spheres": pair_info["spheres"], "parallelism": "single GPU"},
        # Roofline of the dominant kernel on the all-free workload with the full pair set (link
        # broad phase off), where algorithmic flops = executed flops; the shipped form (broad phase
        # on) is reported with the flops it executes, never with the full pair set's count.
        "roofline": {"bound": "valu", "achieved": rf["bf_achieved"], "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": rf["bf_achieved"] / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "traffic_source": TRAFFIC_FILE + " (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                                                      "scaled by a 256 MiB copy)",
                     "algorithmic_bytes_per_launch": args.edges * (2 * desc.dim * 8 + 4 + 1),
                     "kernel": "k_arm_persistent, all-free edges (no early exit), full pair set (link broad "
                               "phase off)",
                     "flops_per_state": fl_state, "states_per_launch": rf["f_states"],
                     "flops_per_launch": fl_state * rf["f_states"], "ms_per_launch": rf["bf_ms"],
                     "broad_phase_on": {"achieved": achieved, "frac": achieved / PEAK_FP32_TFLOPS,
                                        "flops_per_launch": rf["exec_flops"], "ms_per_launch": k_ms,
                                        "run_split_ms": rf["f_split"],
                                        "note": "as shipped; flops are the tests it executes"}},
        "cpu_baseline": cpu,
        "plain_per_call": {"value": per_call_value, "unit": "edge-states/s", "edges": m, "matches": per_call_match,
                           "us_per_edge": pc_time / m * 1e6,
                           "note": "one kernel launch per edge on device rows (mrcc_check_edges_per_call)",
                           "python_env": {"value": py_rate, "edges": mp, "us_per_call": py_time / mp * 1e6,
                                          "matches": py_match,
                                          "note": "MrccEnvironment.is_edge_collision_free per edge from Python, "
                                                  "host rows (the unchanged planners' call)"}},
        "single_launch": {"value": single_rate, "unit": "edge-states/s", "matches": single_match},
        "all_free": {"value": rf["free_rate"], "unit": "edge-states/s", "edges": rf["n_edges"],
                     "edge_states": rf["f_states"], "oracle_check": free_check},
        "configs": {"value": cfg_rate, "unit": "configs/s", "n": len(Qc), "free_frac": float(cfg_free.mean())},
        "stats": stats,
        "run_split_ms": head_split,
        "mismatches": None if mism is None else mism + sum(v.get("mismatches", 0) for v in secondary.values()),
        "checks_compared": checked + sum(v.get("checks_compared", 0) for v in secondary.values()),
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_per_call": value / per_call_value,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


TRAFFIC_FILE = "profiles/k_arm_persistent_c3_allfree_traffic.json"
COST_TRAFFIC_FILE = "profiles/k_config_cost_2p20x28_traffic.json"


def measured_traffic(name=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter summary of this same
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a copy of known size);
    None when the summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), name)
    try:
        with open(path) as f:
            return float(json.load(f)["traffic_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def copy_probe(dev):
    """One device copy of 256 MiB: its FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    import torch

    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("copy probe: 268435456 bytes read, 268435456 bytes written")


def timed_split(plan, steps):
    """Mean queue-build / main-kernel / whole-run milliseconds over `steps` runs of a plan made
    with time_kernels=1 (HIP events recorded by the library on the run's stream)."""
    plan.run()
    acc = {"build_ms": 0.0, "kernel_ms": 0.0, "run_ms": 0.0}
    for _ in range(steps):
        plan.run()
        t = plan.timing()
        for k in acc:
            acc[k] += t[k] / steps
    return acc


def lane_flops_per_state(desc):
    """Algorithmic FP32 flops of one state of the lane-per-state scenes (FMA = 2, min/max/abs/cmp =
    1; interpolation 2 per coordinate): disks 6 per robot pair and 10 per disk-rectangle test (C1:
    8 + 6 + 40 * 10 = 414); planar boxes 50 per separating-axis test over robot pairs and robot x
    obstacle pairs (C2: 24 + (6 + 400) * 50 = 20,324; the 4 binary64 sincos are not counted)."""
    D = desc.dim
    n = desc.n_robots
    pairs = n * (n - 1) // 2
    if desc.kind == "disks":
        return 2 * D + 6 * pairs + 10 * n * len(desc.rect_center) + 6 * n * len(desc.sphere_radius)
    return 2 * D + 50 * (pairs + n * len(desc.obstacle_pose))


def lane_free_edges(rng, desc, sc, n, res, dev, short=0.2):
    """n short edges of a lane-per-state scene whose every state is free (filtered on the GPU):
    uniform starts, steps of at most `short` in the max metric."""
    import torch

    q1s, q2s, have = [], [], 0
    while have < n:
        m = 4 * n
        a = scenes_uniform(rng, desc, m)
        d = rng.uniform(-1, 1, size=(m, desc.dim))
        d *= (rng.uniform(0.02, short, m) / np.max(np.abs(d), axis=1))[:, None]
        b = np.clip(a + d, desc.limits[0], desc.limits[1])
        A, B = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
        oa, ob, oe = (torch.empty(m, dtype=torch.uint8, device=dev) for _ in range(3))
        sc.check_configs_device(A, oa)
        sc.check_configs_device(B, ob)
        sc.check_edges_device(sc.edge_batch(A, B, oe, res))
        torch.cuda.synchronize()
        ok = (oa.cpu().numpy() & ob.cpu().numpy() & oe.cpu().numpy()).astype(bool)
        q1s.append(a[ok])
        q2s.append(b[ok])
        have += int(ok.sum())
    return np.concatenate(q1s)[:n], np.concatenate(q2s)[:n]


def scenes_uniform(rng, desc, m):
    return rng.uniform(desc.limits[0], desc.limits[1], size=(m, desc.dim))


def per_call_rate(sc, q1, q2, res, N, dev, m=4096):
    """The plain per-call form: one launch per edge (mrcc_check_edges_per_call) over the first m
    edges; logical edge-states per second and the verdicts."""
    import torch

    m = min(m, len(q1))
    a, b = torch.from_numpy(q1[:m]).to(dev), torch.from_numpy(q2[:m]).to(dev)
    o = torch.empty(m, dtype=torch.uint8, device=dev)
    batch = sc.edge_batch(a, b, o, res)
    sc.check_edges_per_call_device(batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sc.check_edges_per_call_device(batch)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return int(np.sum(N[:m] - 2)) / dt, o.cpu().numpy().astype(bool), m, dt / m * 1e6


def lane_roofline(args, backend, sc, desc, rng, dev, res, threads, oracle_scene):
    """Roofline of the persistent lane-per-state kernel on an all-free edge set (every state
    evaluated, no early exit; the set is checked against the oracle), from HIP events around the
    kernel; and, for the disk scene, of the configuration kernel (HBM: 8 D + 1 bytes per
    configuration) on 2^22 configurations."""
    import torch

    fl = lane_flops_per_state(desc)
    fq1, fq2 = lane_free_edges(rng, desc, sc, 65536, res, dev)
    f_states, fN = logical_states(fq1, fq2, res)
    fa, fb = torch.from_numpy(fq1).to(dev), torch.from_numpy(fq2).to(dev)
    fo = torch.empty(len(fq1), dtype=torch.uint8, device=dev)
    split = timed_split(backend.EdgePlan(sc, sc.edge_batch(fa, fb, fo, res),
                                         backend.make_tune(use_graph=0, time_kernels=1)), args.steps)
    ach = f_states * fl / (split["kernel_ms"] * 1e-3) / 1e12
    ent = {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": ach / PEAK_FP32_TFLOPS, "flops_per_state": fl, "states_per_launch": f_states,
           "edges": len(fq1), "ms_per_launch": split["kernel_ms"],
           "kernel": "k_lane_persistent, all-free edges (every state evaluated)"}
    # the single launch (one workgroup of 64 lanes per edge, states in rounds of 64) on the same set
    dsplit = timed_split(backend.EdgePlan(sc, sc.edge_batch(fa, fb, fo, res),
                                          backend.make_tune(persistent=0, sorted=0, use_graph=0, time_kernels=1)),
                         args.steps)
    dach = f_states * fl / (dsplit["kernel_ms"] * 1e-3) / 1e12
    ent["direct"] = {"achieved": dach, "frac": dach / PEAK_FP32_TFLOPS, "ms_per_launch": dsplit["kernel_ms"],
                     "kernel": "k_lane_direct, the same all-free edges"}
    if oracle_scene is not None:
        k = min(len(fq1), 16384)
        ref, n_ev = oracle_scene.check_edges(fq1[:k], fq2[:k], res, threads=threads, count=True)
        ent["oracle_checked_edges"] = k
        ent["oracle_states_evaluated"] = int(n_ev)
        ent["oracle_mismatches"] = int(np.sum(~ref))  # the set is all free
    if desc.kind == "disks":
        n = 1 << 22
        Q = torch.from_numpy(scenes_uniform(rng, desc, n)).to(dev)
        o = torch.empty(n, dtype=torch.uint8, device=dev)
        sc.check_configs_device(Q, o)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            sc.check_configs_device(Q, o)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        nbytes = n * (8 * desc.dim + 1)
        gbs = nbytes / (ms * 1e-3) / 1e9
        ent["configs"] = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": gbs / PEAK_HBM_GBS, "configs": n, "bytes_per_launch": nbytes, "ms": ms,
                          "valu_tflops": n * (fl - 2 * desc.dim) / (ms * 1e-3) / 1e12,
                          "kernel": "k_lane_configs<DiskView, 2, 2>"}
    return ent


def lane_workloads(scenes):
    """C1 (2 disks, 20 rectangles, 10,000 random pairs) and C2 (4 planar boxes, 100 obstacles,
    262,144 lognormal edges), both at resolution 0.01."""
    for name, mk, gen in (("c1_planar_disks_10k_edges", scenes.planar_disks, "pairs"),
                          ("c2_planar_boxes_262k_edges", scenes.planar_boxes, "lognormal")):
        rng = np.random.default_rng(1)
        desc = mk(rng=rng)
        if gen == "pairs":
            q1, q2 = scenes.random_pair_edges(rng, desc, 10000)
        else:
            q1, q2 = scenes.lognormal_edges(rng, desc, 262144, 0.01)
        yield name, desc, q1, q2, 0.01


def secondary_workloads(args, backend, scenes, dev):
    """C1 (the reference's abstract environment) and C2 (planar boxes): batched GPU rate, plain
    per-call rate, all-free roofline, CPU oracle on the host cores, and for C1 the numpy
    restatement of abstract_env.py called per edge."""
    import torch
    from oracle import oracle as O
    from oracle.abstract_np import AbstractNp

    out = {}
    threads = len(os.sched_getaffinity(0))
    for name, desc, q1, q2, res in lane_workloads(scenes):
        gen = "pairs" if name.startswith("c1") else "lognormal"
        sc = backend.Scene(desc)
        n_states, N = logical_states(q1, q2, res)
        a, b = torch.from_numpy(q1).to(dev), torch.from_numpy(q2).to(dev)
        o_ = torch.empty(len(q1), dtype=torch.uint8, device=dev)
        # execution shape from the autotuner (queue shapes and the single launch, timed on this batch)
        autotune = importlib.import_module(f"{PKG}.autotune")
        lbatch = sc.edge_batch(a, b, o_, res)
        ltune, ltune_ms = autotune.autotune(sc, lbatch, float(np.mean(N)), use_cache=False)
        lbest = min(ltune_ms, key=ltune_ms.get)
        plan = backend.EdgePlan(sc, lbatch, ltune)
        plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 10
        for _ in range(reps):
            plan.run()
        torch.cuda.synchronize()
        gpu_rate = n_states / ((time.perf_counter() - t0) / reps)
        got = o_.cpu().numpy().astype(bool)
        pc_rate, pc_got, pc_m, pc_us = per_call_rate(sc, q1, q2, res, N, dev)
        ent = {"edges": len(q1), "edge_states": n_states, "gpu_batched": gpu_rate, "unit": "edge-states/s",
               "autotuned": {"choice": lbest, "ms": ltune_ms[lbest], "candidates": len(ltune_ms),
                             "default_shape_ms": ltune_ms.get("G8/t4/w16")},
               "gpu_per_call": {"value": pc_rate, "edges": pc_m, "us_per_edge": pc_us,
                                "matches": bool(np.array_equal(pc_got, got[:pc_m])),
                                "note": "one launch per edge (mrcc_check_edges_per_call)"}}
        o = None if args.skip_cpu else O.Scene(desc)
        ent["roofline"] = lane_roofline(args, backend, sc, desc, np.random.default_rng(11), dev, res, threads, o)
        if not args.skip_cpu:
            t0 = time.perf_counter()
            ref, n_ev = o.check_edges(q1, q2, res, threads=threads, count=True)  # every edge of the workload
            dt = time.perf_counter() - t0
            ent["cpu_oracle"] = {"value": n_states / dt, "cores": threads, "edges": len(q1)}
            ent["mismatches"] = int(np.sum(ref != got))
            ent["checks_compared"] = n_states
            ent["states_evaluated_compared"] = int(n_ev)
            ent["edges_compared"] = len(q1)
            if gen == "pairs":
                mirror = AbstractNp(desc)
                t0 = time.perf_counter()
                kk = 0
                st = 0
                while time.perf_counter() - t0 < 3.0 and kk < len(q1):
                    mirror.is_edge_collision_free(q1[kk], q2[kk], res)
                    st += int(N[kk] - 2)
                    kk += 1
                ent["numpy_per_call"] = {"value": st / (time.perf_counter() - t0), "cores": 1, "edges": kk,
                                         "note": "numpy restatement of abstract_env.py, one call per edge"}
        out[name] = ent
        log(f"{name}: {ent}")
    # the reference's C++ per-call pattern (cpp/main.cpp:83-104) through the C ABI without Python:
    # abstract.test, 10,000 configuration calls and 10,000 edge calls (a child process)
    exe = os.path.join(ROOT, PKG, "cpp", "percall_bench")
    if os.path.exists(exe):
        import subprocess
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        if r.returncode == 0:
            out["cpp_per_call_abstract_test"] = {**json.loads(r.stdout.strip().splitlines()[-1]),
                                                 "note": "cpp/percall_bench: mrcc_check_config_host / "
                                                         "mrcc_check_edge_host per call, host buffers"}
        else:
            out["cpp_per_call_abstract_test"] = {"error": r.stderr[-400:]}
        log(f"cpp per-call: {out['cpp_per_call_abstract_test']}")
    # C5: mixed arm counts and held boxes, one launch over all instances (paths of 1..64 edges)
    n_inst = args.mixed_instances
    rng = np.random.default_rng(5)
    desc = scenes.mixed_arms(rng=rng, n_modes=n_inst)
    sc = backend.Scene(desc)
    V, i1, i2, modes, inst = scenes.mixed_instances(rng, desc, n_inst, 0.05)
    n_states, N = logical_states(V[i1], V[i2], 0.05)
    Vd = torch.from_numpy(V).to(dev)
    I1, I2 = torch.from_numpy(i1).to(dev), torch.from_numpy(i2).to(dev)
    Md = torch.from_numpy(modes).to(dev)
    o_ = torch.empty(len(i1), dtype=torch.uint8, device=dev)
    batch = sc.edge_batch(Vd, Vd, o_, 0.05, modes=Md, idx1=I1, idx2=I2)
    # execution shape from the autotuner (its search runs on this batch; cache in a scratch file)
    autotune = importlib.import_module(f"{PKG}.autotune")
    tune, tune_ms = autotune.autotune(sc, batch, float(np.mean(N)), use_cache=False)
    plan = backend.EdgePlan(sc, batch, tune)
    plan.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        plan.run()
    torch.cuda.synchronize()
    gpu_rate = n_states / ((time.perf_counter() - t0) / 5)
    got = o_.cpu().numpy().astype(bool)
    best = min(tune_ms, key=tune_ms.get)
    ent = {"instances": n_inst, "edges": len(i1), "edge_states": n_states, "gpu_batched": gpu_rate,
           "unit": "edge-states/s", "free_frac": float(got.mean()),
           "autotuned": {"choice": best, "ms": tune_ms[best], "candidates": len(tune_ms)}}
    # executed-flop roofline of the persistent kernel on this mixed batch (early exits and culling:
    # the flops are the tests the kernel counts, never the brute-force count): kernel time from
    # HIP events of a plan with the same shape, test counts from a plan with statistics
    tt = backend.make_tune(tune.slots_per_pass, tune.slots_per_ticket, tune.waves_per_cu, use_graph=0,
                           time_kernels=1)
    split = timed_split(backend.EdgePlan(sc, batch, tt), 5)
    cst = backend.EdgePlan(sc, batch, backend.make_tune(tune.slots_per_pass, tune.slots_per_ticket,
                                                        tune.waves_per_cu, use_graph=0, collect_stats=1))
    cst.run()
    fst = cst.stats()
    del cst
    fl_full, pinfo = arm_flops_per_state(desc)
    S_, L_, D_ = pinfo["spheres"], desc.n_arms * desc.dof, desc.dim
    per_state_fixed = 2 * D_ + 72 * D_ + 18 * S_ + 18 * L_  # all 8 arms: an upper bound per state
    ex = (fst["states"] * per_state_fixed + 15 * fst["link_box_tests"] + 15 * fst["sphere_box_tests"]
          + 11 * fst["link_pair_tests"] + 11 * fst["sphere_pair_tests"])
    ach = ex / (split["kernel_ms"] * 1e-3) / 1e12
    ent["roofline_executed"] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                "frac": ach / PEAK_FP32_TFLOPS, "ms_per_launch": split["kernel_ms"],
                                "run_split_ms": split, "stats": fst,
                                "note": "executed tests counted by the kernel; kinematics of all 8 arms per "
                                        "state (an upper bound for modes with fewer arms)"}
    # plain per-instance form: one single-launch check per instance (first 256 instances)
    k_inst = min(256, n_inst)
    k_e = int(np.searchsorted(inst, k_inst))
    t0 = time.perf_counter()
    for i in range(k_inst):
        lo, hi = int(np.searchsorted(inst, i)), int(np.searchsorted(inst, i + 1))
        bi = sc.edge_batch(Vd, Vd, o_[lo:hi], 0.05, modes=Md[lo:hi], idx1=I1[lo:hi], idx2=I2[lo:hi])
        sc.check_edges_device(bi, backend.make_tune(persistent=0, sorted=0, use_graph=0))
    torch.cuda.synchronize()
    ent["gpu_per_instance"] = int(np.sum(N[:k_e] - 2)) / (time.perf_counter() - t0)
    ent["per_instance_matches"] = bool(np.array_equal(o_[:k_e].cpu().numpy().astype(bool), got[:k_e]))
    if not args.skip_cpu:
        o = O.Scene(desc)
        k = min(len(i1), 60000)
        t0 = time.perf_counter()
        ref = o.check_edges(V[i1[:k]], V[i2[:k]], 0.05, modes=modes[:k], threads=threads)
        dt = time.perf_counter() - t0
        ent["cpu_oracle"] = {"value": int(np.sum(N[:k] - 2)) / dt, "cores": threads, "edges": k}
        ent["mismatches"] = int(np.sum(ref != got[:k]))
        ent["checks_compared"] = int(np.sum(N[:k] - 2))
    out["c5_mixed_arms"] = ent
    log(f"c5_mixed_arms: {ent}")
    out["c4_clearance"] = clearance_workload(args, backend, scenes, dev, threads)
    out["config_cost"] = cost_workload(args, dev, threads)
    out["rrt_call_pattern"] = call_pattern_workload(dev)
    ur = ur10_workload(args, backend, scenes, dev, threads)
    if ur is not None:
        out["mujoco_four_ur10"] = ur
    return out


def ur10_workload(args, backend, scenes, dev, threads):
    """The reference's four-UR10e mujoco world (mujoco_env.py:1160-1204) on its spherized model
    (tests/golden/mj_four_ur10.npz, built by mjcf.load_mjcf_arms from the world's MJCF files):
    65,536 edges around the start pose at the environment's resolution 0.01, batched persistent
    launch against the CPU oracle."""
    import torch
    from oracle import oracle as O

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "mj_four_ur10.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    n_arms, dof = int(z["n_arms"]), int(z["dof"])
    desc = scenes.SceneDesc(kind="arms", limits=z["limits"], n_arms=n_arms, dof=dof, base=z["base"],
                            dh=np.zeros((n_arms, dof, 3)), link_sphere_count=z["link_sphere_count"],
                            sphere_local=z["sphere_local"], boxes=z["boxes"], joint_fixed=z["joint_fixed"])
    sc = backend.Scene(desc)
    rng = np.random.default_rng(7)
    n, res = 65536, 0.01
    home = np.tile([0, -2, 1.0, -1.0, -1.57, 1.0], n_arms)
    q1 = np.clip(home + rng.normal(0, 0.3, (n, desc.dim)), desc.limits[0], desc.limits[1])
    step = rng.normal(0, 1, (n, desc.dim))
    step *= (np.exp(rng.normal(np.log(0.1), 0.7, n)) / np.max(np.abs(step), axis=1))[:, None]
    q2 = np.clip(q1 + step, desc.limits[0], desc.limits[1])
    n_states, N = logical_states(q1, q2, res)
    a, b = torch.from_numpy(q1).to(dev), torch.from_numpy(q2).to(dev)
    o_ = torch.empty(n, dtype=torch.uint8, device=dev)
    plan = backend.EdgePlan(sc, sc.edge_batch(a, b, o_, res), backend.make_tune(use_graph=1))
    plan.run()
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.run()
    torch.cuda.synchronize()
    gpu_rate = n_states / ((time.perf_counter() - t0) / reps)
    got = o_.cpu().numpy().astype(bool)
    ent = {"edges": n, "edge_states": n_states, "mean_N": float(np.mean(N)), "gpu_batched": gpu_rate,
           "unit": "edge-states/s", "free_frac": float(got.mean()), "spheres": int(len(desc.sphere_local))}
    if not args.skip_cpu:
        o = O.Scene(desc)
        k = 8192
        t0 = time.perf_counter()
        ref = o.check_edges(q1[:k], q2[:k], res, threads=threads)
        dt = time.perf_counter() - t0
        ent["cpu_oracle"] = {"value": int(np.sum(N[:k] - 2)) / dt, "cores": threads, "edges": k}
        ent["mismatches"] = int(np.sum(ref != got[:k]))
        ent["checks_compared"] = int(np.sum(N[:k] - 2))
    log(f"mujoco_four_ur10: {ent}")
    return ent


def call_pattern_workload(dev, iters=200, k=30):
    """The RRT* call pattern (rrtstar_base.py:1339-1446) on the 4 x 7-DoF arm environment: per
    new node, the parent choice over k near candidates, then the rewire checks of the eligible
    ones; per-call loop (one environment call per edge, as the planner does) against the batched
    helpers (one launch per decision).  Both decide the same; the time per iteration is reported."""
    import torch

    pkg = importlib.import_module(PKG)
    importlib.import_module(f"{PKG}.envs")
    callers = importlib.import_module(f"{PKG}.callers")
    env = pkg.get_env_by_name("mrcc.quad_arms")
    rng = np.random.default_rng(12)
    D = env.desc.dim
    cases = []
    for _ in range(iters):
        centre = np.clip(rng.normal(0.0, 0.4, D), -2.9, 2.9)
        rows = np.clip(centre + rng.normal(0.0, 0.35, (k, D)), -2.9, 2.9)
        q_new = env.start_pos.from_flat(centre)
        near = [env.start_pos.from_flat(r) for r in rows]
        ids = rng.permutation(10 * k)[:k]
        cases.append((q_new, near, ids, ids[rng.integers(k)], env.batch_config_cost(q_new, rows),
                      rng.uniform(0, 3, k), rng.random(k) < 0.7))

    def loop(c):
        q_new, near, ids, nearest, bc, nc, elig = c
        idx = int(np.where(ids == nearest)[0][0])
        cn = nc + bc
        chosen = idx
        mask = cn < cn[idx]
        if np.any(mask):
            for i in np.where(mask)[0][np.argsort(cn[mask])]:
                if env.is_edge_collision_free(near[i], q_new, env.start_mode):
                    chosen = int(i)
                    break
        rew = [bool(e) and env.is_edge_collision_free(q_new, near[i], env.start_mode) for i, e in enumerate(elig)]
        return chosen, rew

    def batched(c):
        q_new, near, ids, nearest, bc, nc, elig = c
        chosen = callers.find_parent_batched(env, env.start_mode, near, q_new, ids, nearest, bc, nc)[0]
        return chosen, list(callers.rewire_batched(env, env.start_mode, q_new, near, elig))

    loop(cases[0]), batched(cases[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a = [loop(c) for c in cases]
    t_loop = (time.perf_counter() - t0) / iters
    t0 = time.perf_counter()
    b = [batched(c) for c in cases]
    t_b = (time.perf_counter() - t0) / iters
    edges = sum(int(np.sum(c[6])) for c in cases)
    ent = {"iterations": iters, "near_candidates": k, "rewire_edges": edges,
           "per_call_ms_per_iteration": t_loop * 1e3, "batched_ms_per_iteration": t_b * 1e3,
           "speedup": t_loop / t_b, "same_decisions": a == b}
    log(f"rrt_call_pattern: {ent}")
    return ent


def cost_workload(args, dev, threads):
    """The planners' configuration cost beside the collision path (abstract_env.py:238-253 with
    the default cost_metric "euclidean" and cost_reduction "max", w = 0.01): one configuration
    against 2^20 stored configurations of the 4 x 7-DoF layout, as a nearest-neighbour query does.
    HBM-bound: each row is read once (algorithmic bytes = n * (8 D + 8) + 8 D per launch)."""
    import torch
    from oracle import oracle as O

    cfg = importlib.import_module(f"{PKG}.configuration")
    n, D = 1 << 20, 28
    sl = np.array([[0, 7], [7, 14], [14, 21], [21, 28]])
    rng = np.random.default_rng(9)
    Q = rng.uniform(-2.9, 2.9, (n, D))
    pt = rng.uniform(-2.9, 2.9, D)
    Qd, Pd = torch.from_numpy(Q).to(dev), torch.from_numpy(pt).to(dev)
    for _ in range(3):
        got = cfg.batch_config_cost(Pd, Qd, "euclidean", "max", tmp_agent_slice=sl)
    torch.cuda.synchronize()
    reps = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        got = cfg.batch_config_cost(Pd, Qd, "euclidean", "max", tmp_agent_slice=sl)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = n * (8 * D + 8) + 8 * D
    gbs = nbytes / (ms * 1e-3) / 1e9
    ent = {"rows": n, "dim": D, "robots": 4, "metric": "euclidean", "reduction": "max", "ms": ms,
           "rows_per_s": n / (ms * 1e-3),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "bytes_per_launch": nbytes,
                        "traffic": measured_traffic(COST_TRAFFIC_FILE),
                        "traffic_source": COST_TRAFFIC_FILE + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by "
                                                              "a 256 MiB copy)"}}
    if not args.skip_cpu:
        t0 = time.perf_counter()
        ref = O.config_cost(pt, Q, sl, "euclidean", "max", threads=threads)
        dt = time.perf_counter() - t0
        ent["cpu_oracle"] = {"value": n / dt, "unit": "rows/s", "cores": threads}
        ent["mismatches"] = int(np.sum(ref != got.cpu().numpy()))
        ent["checks_compared"] = n
    log(f"config_cost: {ent}")
    return ent


def clearance_workload(args, backend, scenes, dev, threads):
    """C4: clearance cost + hand-written backward on 4096 paths x 64 waypoints x 28 DoF (FP32),
    gradient checked against FP64 central differences (h = 1e-6) of the oracle on 64 paths."""
    import torch
    from oracle import oracle as O

    clr = importlib.import_module(f"{PKG}.clearance")
    rng = np.random.default_rng(4)
    desc = scenes.quad_arms(rng=rng)
    sc = backend.Scene(desc)
    P, W, D = 4096, 64, desc.dim
    a = np.clip(rng.normal(0, 0.4, (P, D)), -2.9, 2.9)
    b = np.clip(rng.normal(0, 0.4, (P, D)), -2.9, 2.9)
    t = np.linspace(0, 1, W)[None, :, None]
    paths = np.clip(a[:, None] + t * (b - a)[:, None] + rng.normal(0, 0.1, (P, W, D)), -2.9, 2.9)
    q = torch.from_numpy(paths.astype(np.float32)).to(dev).requires_grad_(True)
    # warm-up as timed: the second pass accumulates into q.grad, so torch's accumulation kernel is
    # loaded here (its first launch in a process costs tens of ms), not inside the timed loop
    for _ in range(3):
        c = clr.clearance_cost(sc, q, 0.02)
        c.sum().backward()
    q.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        c = clr.clearance_cost(sc, q, 0.02)
        c.sum().backward()
    torch.cuda.synchronize()
    fb = (time.perf_counter() - t0) / reps
    g = (q.grad / reps).detach().cpu().numpy().astype(np.float64)
    # the two kernels alone: HIP events around the library calls on the current stream
    lib = importlib.import_module(f"{PKG}._lib")
    import ctypes
    qf = q.detach().reshape(-1, D).contiguous()
    n = qf.shape[0]
    cost = torch.empty(n, dtype=torch.float32, device=dev)
    gq = torch.empty_like(qf)
    gc = torch.ones(n, dtype=torch.float32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fwd = lambda: lib.check(lib.lib().mrcc_clearance_forward(sc.handle, ctypes.c_void_p(qf.data_ptr()), n, D,
                                                             ctypes.c_float(0.02), ctypes.c_void_p(cost.data_ptr()), st))
    bwd = lambda: lib.check(lib.lib().mrcc_clearance_backward(sc.handle, ctypes.c_void_p(qf.data_ptr()), n, D,
                                                              ctypes.c_float(0.02), ctypes.c_void_p(gc.data_ptr()),
                                                              ctypes.c_void_p(gq.data_ptr()), st))
    def kernel_times():
        out = {}
        for name, f in (("forward", fwd), ("backward", bwd)):
            f()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                f()
            e1.record()
            torch.cuda.synchronize()
            out[name] = e0.elapsed_time(e1) / reps
        return out

    kms = kernel_times()  # as shipped: pairs culled with the links' bounding spheres
    cost_culled = cost.clone()
    sc.set_broadphase(False)  # the full pair set, for the roofline
    kms_full = kernel_times()
    sc.set_broadphase(True)
    full_cost_diff = float((cost - cost_culled).abs().max())
    fl_fwd, pinfo = arm_flops_per_state(desc)
    # backward: the forward's pair set again (recomputed), then per sphere c x f and f summed into
    # its link (12 flops), per joint the suffix sums over its arm's links (6 flops a link, dof/2
    # links on average) and z . (T - o x F) (15 flops)
    L_ = desc.n_arms * desc.dof
    fl_bwd = fl_fwd + 12 * pinfo["spheres"] + 6 * L_ * (desc.dof + 1) // 2 + 15 * L_
    ach_f = n * fl_fwd / (kms_full["forward"] * 1e-3) / 1e12
    ach_b = n * fl_bwd / (kms_full["backward"] * 1e-3) / 1e12
    ent = {"paths": P, "waypoints": P * W, "dof": D, "ms_forward_backward": fb * 1e3,
           "waypoints_per_s": P * W / fb, "mean_cost": float(c.detach().mean()),
           "kernel_ms": kms,
           "kernel_ms_full_pair_set": kms_full,
           "full_pair_set_max_cost_difference": full_cost_diff,
           "roofline": {"bound": "valu", "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "forward": {"achieved": ach_f, "frac": ach_f / PEAK_FP32_TFLOPS, "flops_per_waypoint": fl_fwd},
                        "backward": {"achieved": ach_b, "frac": ach_b / PEAK_FP32_TFLOPS,
                                     "flops_per_waypoint": fl_bwd},
                        "kernel": "k_clearance with the broad phase off (the full pair set: lane per sphere, "
                                  "every partner sphere and box, square root only for pairs within reach); "
                                  "as shipped the pairs are culled with the links' bounding spheres "
                                  "(kernel_ms)"},
           "note": "ms_forward_backward is wall time of the autograd pair (forward, sum, backward, "
                   "gradient accumulation) from Python; kernel_ms are HIP events around the two library "
                   "kernels alone"}
    # repair: 80 Adam steps on clearance + smoothness, every result re-validated by the exact checker
    env = backend.MrccEnvironment(desc, collision_resolution=0.05)
    ra = np.clip(rng.normal(0, 0.4, (1024, D)), -2.9, 2.9)
    rb = np.clip(rng.normal(0, 0.4, (1024, D)), -2.9, 2.9)
    keep = env.batch_is_collision_free(ra) & env.batch_is_collision_free(rb)
    ra, rb = ra[keep], rb[keep]
    tr = np.linspace(0, 1, 32)[None, :, None]
    rp = ra[:, None] + tr * (rb - ra)[:, None]
    rp[:, 1:-1] += rng.normal(0, 0.1, rp[:, 1:-1].shape)
    rp = np.clip(rp, -2.9, 2.9)
    before = clr.validate_paths(env, rp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rout, rok = clr.repair_paths(env, rp, steps=80)
    rms = (time.perf_counter() - t0) * 1e3
    bad = ~before
    ent["repair"] = {"paths": int(len(rp)), "waypoints": 32, "steps": 80, "in_collision": int(bad.sum()),
                     "repaired": int((rok & bad).sum()),
                     "acceptance_rate": float(rok[bad].mean()) if bad.any() else None, "ms": rms}
    if not args.skip_cpu:
        o = O.Scene(desc)
        sel = np.nonzero(rok & bad)[0][:16]
        ent["repair"]["oracle_checked_paths"] = int(len(sel))
        ent["repair"]["oracle_invalid"] = int(sum(
            not (all(o.config_free(v) for v in rout[p]) and
                 all(o.edge_free(rout[p][i], rout[p][i + 1], 0.05) for i in range(31))) for p in sel))
    if not args.skip_cpu:
        o = O.Scene(desc)
        pick = rng.choice(P, 64, replace=False)
        q64 = paths.astype(np.float32).astype(np.float64)[pick]          # [64, W, D]
        h = 1e-6
        base = q64.reshape(-1, D)
        E = np.eye(D)
        plus = (base[:, None, :] + h * E[None]).reshape(-1, D)
        minus = (base[:, None, :] - h * E[None]).reshape(-1, D)
        fd = ((o.clearance_cost_batch(plus, 0.02, threads) - o.clearance_cost_batch(minus, 0.02, threads))
              / (2 * h)).reshape(64, W * D)
        gp = g[pick].reshape(64, W * D)
        rel = np.linalg.norm(gp - fd, axis=1) / np.maximum(np.linalg.norm(fd, axis=1), 1e-12)
        ent["grad_rel_err_max"] = float(rel[np.linalg.norm(fd, axis=1) > 1e-9].max()) if np.any(np.linalg.norm(fd, axis=1) > 1e-9) else None
        ent["grad_rel_err_median"] = float(np.median(rel))
        ent["fd_paths"] = 64
    log(f"c4_clearance: {ent}")
    return ent


def free_edges(rng, desc, sc, n, res, dev):
    """n edges whose endpoints and interior states are all collision free (filtered on the GPU):
    starts drawn from a pool of free configurations near the zero pose, lognormal steps."""
    import torch

    def check_cfg(a):
        A = torch.from_numpy(a).to(dev)
        o = torch.empty(len(a), dtype=torch.uint8, device=dev)
        sc.check_configs_device(A, o)
        torch.cuda.synchronize()
        return o.cpu().numpy().astype(bool)

    pool = []
    have = 0
    while have < max(4096, n // 8):
        a = np.clip(rng.normal(0, 0.4, size=(262144, desc.dim)), -2.9, 2.9)
        f = a[check_cfg(a)]
        pool.append(f)
        have += len(f)
    pool = np.concatenate(pool)
    q1s, q2s = [], []
    have = 0
    while have < n:
        m = 262144
        a = pool[rng.integers(0, len(pool), m)]
        L = rng.lognormal(np.log(0.3), 1.0, size=m)
        d = rng.uniform(-1, 1, size=(m, desc.dim))
        d /= np.max(np.abs(d), axis=1, keepdims=True)
        b = np.clip(a + (L[:, None] * d), -2.9, 2.9)
        A, B = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
        ob = torch.empty(m, dtype=torch.uint8, device=dev)
        oe = torch.empty_like(ob)
        sc.check_configs_device(B, ob)
        sc.check_edges_device(sc.edge_batch(A, B, oe, res))
        torch.cuda.synchronize()
        ok = (ob.cpu().numpy() & oe.cpu().numpy()).astype(bool)
        q1s.append(a[ok])
        q2s.append(b[ok])
        have += int(ok.sum())
    return np.concatenate(q1s)[:n], np.concatenate(q2s)[:n]

if __name__ == "__main__":
    main()
