"""Benchmark: stochastic block partitioning on one MI355X.

Default workload: the 1,000,000-vertex, 200-block DC-SBM graph (p_in 0.6, Dirichlet(1) block
sizes, out-degrees on [10, 200], about 28M edges, seed 0), partitioned by hybrid MCMC with the
published run flags (batches 1, approximate merges), graph resident in HBM.

A step is one plain hybrid MCMC sweep (one iteration of finetune.cpp:356-405: the sequential
Metropolis-Hastings pass over the top 15% of vertices by degree, then the asynchronous Gibbs batch
over the rest, then the MDL) from a recorded state: the blockmodel after the first merge phase
from the identity partition (B = N/2). W warmup sweeps, then K timed sweeps continue the chain
from there. value = vertex-move evaluations per second over the K timed sweeps (proposals with
s != r whose delta MDL, Hastings correction and accept were computed); ms_per_step = time per
sweep. The roofline object covers the evaluation kernel class with the most device time in the
timed sweeps (HIP events on the library's stream), with its algorithmic bytes counted on the
device in an untimed replay of the same sweeps from the same recorded state.

In the same run, untimed by `value`: one whole plain partition and one whole learned-proposal
partition of the same graph (sbp::stochastic_block_partition, sbp.cpp:61-98), each timed end to
end, with the time at which the learned run first reaches the plain run's final MDL; and the CPU
oracle's whole partition of a smaller graph of the same family as the CPU baseline. Phases that
would overrun --budget seconds are skipped and marked so.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "vertex-move evals/s + time-to-partition at equal MDL, learned vs plain, 1 GPU"
CPU_SAMPLE_MAX_N = 100000
HBM_PEAK_GBS = 8000.0


def load_pkg():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "sbpx_amd", os.path.join(ROOT, "hybrid-stochastic-block-partitioning_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sbpx_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


KERNELS = ["eval_bin_kernel<64..256> (batched vertex-move evaluation, degree <= 126, LDS tables)",
           "eval_bin_kernel<512..1024> (batched vertex-move evaluation, degree 127-510, LDS tables)",
           "eval_group_kernel (batched vertex-move evaluation, degree > 510, 4-16 wavefronts per vertex)",
           "mh_kernels (sequential Metropolis-Hastings evaluation and commit: mh_wide_kernel<G> with G = 1-8 "
           "wavefronts per entry on up to 512 workgroups, mh_kernel and mh_group_kernel<G> on one)"]


def train_learned(m, cfg, seeds, log_sweeps=8, steps=600):
    """Trains the proposal network (untimed) on accepted moves logged from plain asynchronous Gibbs
    sweeps on other graphs of the same family (graphs of cfg's size, one per seed; one merge phase
    at a time, `log_sweeps` sweeps each, down to about twice the planted block count)."""
    ctx = m.Context(0)
    try:
        ctx.learned_init(1)
        ctx.learned_log(4_000_000)
        t0 = time.perf_counter()
        for seed in seeds:
            op, oi, ip, ii, _ = m.generate_dcsbm(**cfg, seed=seed)
            n = len(op) - 1
            ctx.upload_graph(op, oi, ip, ii)
            ctx.split()
            ctx.build(n, np.arange(n, dtype=np.int32))
            B, call = n, 0
            while B > 2 * cfg["num_blocks"]:
                ctx.merge_blocks(B // 2, seed, call)
                B -= B // 2
                for i in range(log_sweeps):
                    ctx.sweep(m.ALGO_ASYNC_GIBBS, 1, seed, 1000 * call + i)
                call += 1
            print(f"# training graph {seed} logged: {ctx.learned_log_count()} rows, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        ctx.learned_log(0)
        rows = ctx.learned_log_count()
        t1 = time.perf_counter()
        first = ctx.learned_train(50, 4096, 1e-3, 7)
        last = ctx.learned_train(steps - 50, 4096, 1e-3, 8)
        t2 = time.perf_counter()
        params = ctx.learned_params()
    finally:
        ctx.close()
    return params, {"graphs": [int(s) for s in seeds], "vertices_per_graph": cfg["num_vertices"], "rows": rows,
                    "steps": steps, "batch": 4096, "loss_first": first, "loss_last": last, "log_s": t1 - t0,
                    "train_s": t2 - t1}


def measured_traffic(workload):
    """HBM bytes per launch per kernel class from the PMC summary kept in profiles/ (FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy kernel of known size), when it was taken on this workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    out = {}
    if not os.path.exists(path):
        return out
    lines = open(path).read().splitlines()
    if not any(line.startswith("# workload: ") and line[len("# workload: "):] == workload for line in lines):
        return out
    for line in lines:
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#"):
            out[parts[0]] = float(parts[1])
    return out


def load_evaluation():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "sbpx_eval", os.path.join(ROOT, "hybrid-stochastic-block-partitioning_amd", "evaluation.py"))
    ev = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ev)
    return ev


def bench_batch(args, m):
    """The batched workload: `args.graphs` DC-SBM graphs of uneven size partitioned side by side in
    one context (async_gibbs, approximate merges). value = graphs partitioned per second with the
    graphs resident; beside it the same graphs partitioned one at a time on the GPU (a sample),
    the per-graph MDL of the batched run against those runs and against the CPU oracle's runs."""
    import oracle_lib as O

    params = m.c3_graph_params(args.graphs, args.seed)
    graphs, truths = [], []
    for p in params:
        op, oi, ip, ii, truth = m.generate_dcsbm(**p)
        graphs.append((op, oi, ip, ii))
        truths.append(truth)
    sizes = np.array([len(g[0]) - 1 for g in graphs])
    edges = np.array([int(g[0][-1]) for g in graphs])
    seeds = [args.seed * 7919 + g for g in range(len(graphs))]
    print(f"# batch: {len(graphs)} graphs, N={sizes.sum()} E={edges.sum()}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        _, wt = m.partition_batch(graphs[:16], batches=args.batches, seeds=seeds[:16])
        print(f"# warmup {w}: {wt['seconds']:.2f} s for 16 graphs", file=sys.stderr, flush=True)
    times, runs, tot = [], None, None
    for k in range(args.steps):
        runs, tot = m.partition_batch(graphs, batches=args.batches, seeds=seeds)
        times.append(tot["seconds"])
        print(f"# step {k}: {tot['seconds']:.2f} s, {tot['rounds']} rounds, {tot['sweeps']} batched sweeps",
              file=sys.stderr, flush=True)
    t_batch = float(np.mean(times))
    if args.no_count:
        kernel_bytes = np.zeros(4)
    else:
        _, ctot = m.partition_batch(graphs, batches=args.batches, seeds=seeds, count_bytes=True)
        kernel_bytes = np.array(ctot["eval_bytes"])
    ms = np.array(tot["eval_ms"])
    launches = np.array(tot["eval_launches"])
    gbs = np.where(ms > 0, kernel_bytes / np.maximum(ms, 1e-12) / 1e6, 0.0)
    dom = int(np.argmax(ms))
    traffic = measured_traffic(None)
    dom_name = KERNELS[dom].split(" ")[0]
    # the same graphs one at a time on the GPU: a sample, timed, and the per-graph parity
    rng = np.random.default_rng(args.seed)
    sample = sorted(rng.choice(len(graphs), size=min(args.sample, len(graphs)), replace=False).tolist())
    t_alone, same, rel = 0.0, 0, []
    for g in sample:
        rg = m.ResidentGraph(*graphs[g])
        t0 = time.perf_counter()
        a, st = rg.partition(algorithm="async_gibbs", batches=args.batches, approximate=True, seed=seeds[g])
        t_alone += time.perf_counter() - t0
        rg.close()
        same += int((a == runs[g][0]).all() and st.mdl == runs[g][2])
        rel.append(abs(runs[g][2] / st.mdl - 1.0))
    alone_rate = sizes[sample].sum() / t_alone if sample else None  # vertices per second, one graph at a time
    projected_alone = sizes.sum() / alone_rate if sample else None
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "sbpx_eval", os.path.join(ROOT, "hybrid-stochastic-block-partitioning_amd", "evaluation.py"))
    ev = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ev)
    nmis = [ev.nmi(truths[g], runs[g][0]) for g in range(len(graphs))]
    out = {
        "metric": "graphs partitioned/s, batched plain SBP (per-graph MDL parity), 1 GPU",
        "value": len(graphs) / t_batch,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f32 cells, f64 sums",
        "data": "synthetic DC-SBM (seeded generator)",
        "config": {"workload": f"C3 batched plain async_gibbs SBP, {len(graphs)} graphs, batches={args.batches}, "
                               "--approximate",
                   "seed": args.seed, "num_graphs": len(graphs), "total_vertices": int(sizes.sum()),
                   "total_edges": int(edges.sum()), "vertices_per_graph": [int(sizes.min()), int(sizes.max())],
                   "deg": [5, 100]},
        "rounds": tot["rounds"],
        "batched_sweeps": tot["sweeps"],
        "evals_per_s": tot["evals"] / t_batch,
        "alone_on_gpu": {"graphs": len(sample), "seconds": t_alone, "vertices_per_s": alone_rate,
                         "projected_seconds_all": projected_alone,
                         "batched_speedup": projected_alone / t_batch if sample else None,
                         "identical_results": same, "max_mdl_rel_diff": float(max(rel)) if rel else None},
        "nmi_vs_planted": {"mean": float(np.mean(nmis)), "min": float(np.min(nmis))},
        "roofline": {
            "bound": "hbm", "kernel": KERNELS[dom], "achieved": float(gbs[dom]), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": float(gbs[dom]) / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_run": float(kernel_bytes[dom]), "ms_per_run": float(ms[dom]),
            "launches_per_run": float(launches[dom]),
            "avg_launch_us": 1000.0 * float(ms[dom]) / max(1.0, float(launches[dom]))},
        "eval_kernels": [
            {"kernel": KERNELS[i].split(" ")[0], "ms_per_run": float(ms[i]), "launches_per_run": float(launches[i]),
             "bytes_per_run": float(kernel_bytes[i]), "achieved_GBs": float(gbs[i]),
             "frac": float(gbs[i]) / HBM_PEAK_GBS}
            for i in range(4) if launches[i] > 0],
    }
    if dom_name in traffic:
        out["roofline"]["traffic_note"] = "PMC traffic in profiles/ is for the one-graph workload"
    if not args.no_cpu and sample:
        # the CPU oracle's whole runs on graphs of the batch, in the sample's order, for about cpu_seconds
        t_cpu, done, verts, ratios = 0.0, 0, 0, []
        for g in sample:
            og = O.Graph(int(sizes[g]), *graphs[g])
            t0 = time.perf_counter()
            _, orun = O.sbp(og, 1, args.batches, True, seeds[g])
            t_cpu += time.perf_counter() - t0
            done += 1
            verts += int(sizes[g])
            ratios.append(runs[g][2] / orun.mdl)
            if t_cpu > args.cpu_seconds:
                break
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {
            "value": (verts / t_cpu) / float(sizes.mean()), "unit": "graphs/s", "cores": threads, "kind": "port",
            "sample": (f"oracle async_gibbs SBP (approximate merges) run whole on {done} graphs of the batch "
                       f"({verts} vertices) in {t_cpu:.1f} s; graphs/s scaled to the batch's mean graph size "
                       f"({sizes.mean():.0f} vertices); OpenMP threads={threads}"),
            "mdl_batched_over_oracle": [float(r) for r in ratios]}
    print(json.dumps(out), flush=True)


def step_sweeps(m, arrays, args):
    """The timed steps: W warmup + K hybrid sweeps from the state after the first merge phase, then
    an untimed replay of the same sweeps from the same state that counts algorithmic bytes."""
    n = len(arrays[0]) - 1
    algo = m.ALGORITHMS[args.algorithm]
    c = m.Context(0)
    try:
        c.upload_graph(*arrays)
        c.split()
        c.build(n, np.arange(n, dtype=np.int32))
        c.merge_blocks(n // 2, args.seed, 0)
        B0 = c.info()[0]
        state0 = c.state()[0].copy()
        mdl0 = c.mdl()
        for w in range(args.warmup):
            c.sweep(algo, args.batches, args.seed, w)
        c.mdl()  # untimed; its log-posterior kernels also mark where the timed sweeps begin in a trace
        evals = attempts = moves = 0
        ms = np.zeros(4)
        launches = np.zeros(4)
        t0 = time.perf_counter()
        for k in range(args.steps):
            st = c.sweep(algo, args.batches, args.seed, args.warmup + k)  # returns after the stream drained
            evals += st.evals
            attempts += st.attempts
            moves += st.moves
            ms += np.array(st.eval_ms[:4])
            launches += np.array(st.eval_launches[:4])
        elapsed = time.perf_counter() - t0
        mdl1 = c.mdl()
        print(f"# steps: {args.steps} sweeps in {elapsed:.3f} s at B={B0}, evals={evals}", file=sys.stderr, flush=True)
        bytes_ = np.zeros(4)
        replay_evals = None
        if not args.no_count:
            c.build(B0, state0)
            for w in range(args.warmup):
                c.sweep(algo, args.batches, args.seed, w)
            c.set_option(m.OPT_COUNT_BYTES, 1)
            replay_evals = 0
            for k in range(args.steps):
                st = c.sweep(algo, args.batches, args.seed, args.warmup + k)
                bytes_ += np.array(st.eval_bytes[:4])
                replay_evals += st.evals
            c.set_option(m.OPT_COUNT_BYTES, 0)
            print("# byte-counting replay done", file=sys.stderr, flush=True)
    finally:
        c.close()
    return dict(seconds=elapsed, evals=evals, attempts=attempts, moves=moves, ms=ms, launches=launches,
                bytes=bytes_, replay_evals=replay_evals, B0=B0, mdl_start=mdl0, mdl_end=mdl1)


def cpu_whole_partition(m, cfg, n, seed):
    """The oracle's whole partition (sbp::stochastic_block_partition restated on the CPU, hybrid,
    batches 1, approximate merges) of a graph of the same family with n vertices."""
    import oracle_lib as O

    small = dict(cfg, num_vertices=n)
    op, oi, ip, ii, _ = m.generate_dcsbm(**small, seed=seed)
    g = O.Graph(n, op, oi, ip, ii)
    t0 = time.perf_counter()
    _, run = O.sbp(g, 2, 1, True, seed)
    secs = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": run.evals / secs, "unit": "vertex-move evals/s", "cores": threads, "kind": "port",
        "sample": (f"oracle whole hybrid_mcmc partition (batches 1, approximate merges) of a graph of the same "
                   f"family with {n} vertices and {int(op[-1])} edges, seed {seed}: {run.evals} evaluations, "
                   f"{run.sweeps} sweeps, final B={run.B}, {secs:.1f} s to partition; OpenMP threads={threads}"),
        "time_to_partition_s": secs, "num_vertices": n,
    }


def main():
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="C4")
    ap.add_argument("--algorithm", default="hybrid_mcmc", choices=["hybrid_mcmc", "async_gibbs",
                                                                   "metropolis_hastings"])
    ap.add_argument("--batches", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget", type=float, default=510.0,
                    help="seconds for the whole run; later phases that would overrun it are skipped")
    ap.add_argument("--cpu-n", type=int, default=3000, help="vertices of the CPU oracle's whole-partition graph")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="C3 only: CPU oracle sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the untimed byte-counting replay")
    ap.add_argument("--no-partition", action="store_true", help="skip the whole plain and learned partitions")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal partition")
    ap.add_argument("--proposal", type=int, default=2, choices=[1, 2],
                    help="learned mixture in: 1 the batched sweeps, 2 also the sequential pass")
    ap.add_argument("--train-graphs", type=int, default=10, help="training graphs (seeds 0 .. n-1)")
    ap.add_argument("--train-n", type=int, default=200000, help="vertices per training graph")
    ap.add_argument("--train-steps", type=int, default=600, help="Adam steps of the proposal network (>= 100)")
    ap.add_argument("--learn-maxdeg", type=int, default=126, help="degree bound of the learned mixture (<= 510)")
    ap.add_argument("--graphs", type=int, default=512, help="graphs of the batched workload (--config C3)")
    ap.add_argument("--sample", type=int, default=12, help="C3: graphs also run alone (GPU) and by the oracle")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 10 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 4 * n, "launches": 10}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described, not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    m = load_pkg()
    if args.config == "C3":
        return bench_batch(args, m)
    cfg = m.CONFIGS[args.config]
    out_ptr, out_idx, in_ptr, in_idx, truth = m.generate_dcsbm(**cfg, seed=args.seed)
    arrays = (out_ptr, out_idx, in_ptr, in_idx)
    E = int(out_ptr[-1])
    print(f"# graph {args.config}: N={cfg['num_vertices']} E={E}", file=sys.stderr, flush=True)
    steps = step_sweeps(m, arrays, args)
    ms_per = steps["ms"] / args.steps
    launches_per = steps["launches"] / args.steps
    bytes_per = steps["bytes"] / args.steps
    gbs = np.where(ms_per > 0, bytes_per / np.maximum(ms_per, 1e-12) / 1e6, 0.0)
    dom = int(np.argmax(ms_per))
    dom_name = KERNELS[dom].split(" ")[0]
    workload = (f"{args.config} {args.algorithm} sweeps from the state after the first merge phase, batches="
                f"{args.batches}, --approximate")
    traffic = measured_traffic(workload)
    ev = load_evaluation()
    out = {
        "metric": METRIC,
        "value": steps["evals"] / steps["seconds"],
        "unit": "vertex-move evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * steps["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f32 cells, f64 sums",
        "data": "synthetic DC-SBM (seeded generator)",
        "config": {"workload": workload, "seed": args.seed, "num_vertices": cfg["num_vertices"], "num_edges": E,
                   "num_blocks_planted": cfg["num_blocks"], "p_in": cfg["p_in"], "dirichlet": cfg["dirichlet"],
                   "deg": [cfg["deg_min"], cfg["deg_max"]], "blocks_at_step_start": steps["B0"]},
        "attempts_per_s": steps["attempts"] / steps["seconds"],
        "sweeps_per_s": args.steps / steps["seconds"],
        "moves_in_steps": steps["moves"],
        "mdl_step_start": steps["mdl_start"], "mdl_step_end": steps["mdl_end"],
        "roofline": {
            "bound": "hbm", "kernel": KERNELS[dom], "achieved": float(gbs[dom]), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": float(gbs[dom]) / HBM_PEAK_GBS, "traffic": traffic.get(dom_name),
            "traffic_source": ("profiles/hbm_traffic.txt: rocprofv3 PMC FETCH_SIZE + WRITE_SIZE on this workload"
                               if dom_name in traffic else None),
            "bytes_per_launch": float(bytes_per[dom] / max(1.0, launches_per[dom])),
            "avg_launch_us": 1000.0 * float(ms_per[dom]) / max(1.0, float(launches_per[dom])),
            "bytes_per_step": float(bytes_per[dom]), "ms_per_step": float(ms_per[dom]),
            "launches_per_step": float(launches_per[dom]),
            "replay_evals_equal": (steps["replay_evals"] == steps["evals"]) if steps["replay_evals"] is not None
            else None},
        "eval_kernels": [
            {"kernel": KERNELS[i].split(" ")[0], "ms_per_step": float(ms_per[i]),
             "launches_per_step": float(launches_per[i]), "bytes_per_step": float(bytes_per[i]),
             "achieved_GBs": float(gbs[i]), "frac": float(gbs[i]) / HBM_PEAK_GBS}
            for i in range(4) if launches_per[i] > 0],
    }
    budget_left = lambda: args.budget - (time.perf_counter() - t_start)  # noqa: E731
    cpu_reserve = 0.0 if args.no_cpu else 45.0
    os.environ.setdefault("SBP_TRACE", "1")  # the driver reports each outer iteration on stderr
    if not args.no_partition:
        g = m.ResidentGraph(*arrays)
        try:
            kw = dict(algorithm=args.algorithm, batches=args.batches, approximate=True, seed=args.seed)
            # the plain run may take up to 45% of what is left, so the learned run has as long
            cap = max(30.0, 0.45 * (budget_left() - cpu_reserve))
            t0 = time.perf_counter()
            a, st = g.partition(max_seconds=cap, **kw)
            t_plain = time.perf_counter() - t0
            print(f"# plain partition: {t_plain:.1f} s, B={st.num_blocks}, mdl={st.mdl:.1f}", file=sys.stderr, flush=True)
            out["plain_partition"] = {
                "time_to_partition_s": t_plain, "completed": not st.stopped_at_limit, "time_limit_s": cap,
                "final_mdl": st.mdl, "final_num_blocks": int(st.num_blocks), "sweeps": int(st.sweeps),
                "evals_per_s": st.evals / t_plain, "nmi_vs_planted": ev.nmi(truth, a),
                "modularity": ev.modularity(out_ptr, out_idx, in_ptr, a),
                "modularity_planted": ev.modularity(out_ptr, out_idx, in_ptr, truth)}
            train_est = 2.0 * args.train_graphs + 20.0
            if args.no_learned or args.algorithm == "metropolis_hastings":
                pass
            elif budget_left() < train_est + cpu_reserve + 30.0:
                out["learned"] = {"skipped": f"over the {args.budget:.0f} s budget ({budget_left():.0f} s left)"}
            else:
                train_seeds = tuple(range(args.train_graphs))
                params, info = train_learned(m, dict(cfg, num_vertices=args.train_n), train_seeds,
                                             steps=args.train_steps)
                g.load_learned(params)
                g.set_option(m.OPT_LEARN_MAXDEG, args.learn_maxdeg)
                lcap = max(30.0, budget_left() - cpu_reserve - 10.0)
                t0 = time.perf_counter()
                la, lst = g.partition(learned=args.proposal, max_seconds=lcap, **kw)
                tl = time.perf_counter() - t0
                lt, lmdl, _ = g.trajectory()
                hit = np.flatnonzero(lmdl <= st.mdl)
                t_equal = float(lt[hit[0]]) if hit.size else None
                done = not lst.stopped_at_limit
                out["learned"] = {
                    "proposal": ("0.2 * plain + 0.8 * network (Hastings-corrected), vertices of degree <= "
                                 f"{args.learn_maxdeg}: in the batched sweeps"
                                 + (" and in the sequential Metropolis-Hastings pass (each window's entries scored on the "
                                    "current state)" if args.proposal == 2 else "")),
                    "proposal_option": args.proposal, "learn_maxdeg": args.learn_maxdeg,
                    "time_to_partition_s": tl, "completed": done, "time_limit_s": lcap,
                    "speedup_vs_plain": (t_plain / tl) if done else None,
                    "time_to_plain_final_mdl_s": t_equal,
                    "speedup_at_equal_mdl": (t_plain / t_equal) if t_equal else None,
                    "final_mdl": lst.mdl, "mdl_vs_plain": lst.mdl / st.mdl,
                    "final_within_half_percent": abs(lst.mdl / st.mdl - 1.0) <= 5e-3,
                    "final_num_blocks": int(lst.num_blocks), "sweeps": int(lst.sweeps),
                    "evals_per_s": lst.evals / tl, "nmi_vs_planted": ev.nmi(truth, la),
                    "modularity": ev.modularity(out_ptr, out_idx, in_ptr, la), "training": info}
                print(f"# learned partition: {tl:.1f} s ({'complete' if done else 'stopped at its time limit'}), "
                      f"B={lst.num_blocks}, mdl={lst.mdl:.1f}", file=sys.stderr, flush=True)
        finally:
            g.close()
    if not args.no_cpu:
        if budget_left() < 30.0:
            out["cpu_baseline"] = {"value": None, "unit": "vertex-move evals/s", "cores": 0, "kind": "port",
                                   "sample": "skipped: over the time budget"}
        else:
            print("# cpu baseline", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_whole_partition(m, cfg, args.cpu_n, args.seed)
    out["wall_s"] = time.perf_counter() - t_start
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
