"""Benchmark: stochastic block partitioning on one MI355X.

A step is one complete partitioning run (sbp::stochastic_block_partition: block merges, MCMC
sweeps, golden-section search over the number of blocks) of a synthetic DC-SBM graph that is
already resident in HBM. The default workload is the 50,000-vertex, 44-block DC-SBM graph
(p_in 0.6, Dirichlet(1) block sizes, out-degrees on [10, 100]) partitioned by plain hybrid MCMC
with the published run flags (batches 1, approximate merges). value = vertex-move evaluations
per second over the timed steps (proposals with s != r whose delta MDL, Hastings correction
and accept were computed); ms_per_step = time to partition. The roofline object covers the
evaluation kernel with the most device time in the run (the sequential Metropolis-Hastings
kernel or one of the batched kernels by degree bin), with its algorithmic bytes counted on the
device in an untimed run of the same seed and its time taken from HIP events on the library's
stream.
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
           "mh_kernels (sequential Metropolis-Hastings evaluation and commit: mh_wide_kernel on 2-32 workgroups, "
           "mh_kernel and mh_group_kernel<G> on one)"]


def train_learned(m, cfg, seeds, log_sweeps=8, steps=600):
    """Trains the proposal network (untimed) on accepted moves logged from plain asynchronous Gibbs
    sweeps on other graphs of the same family (one merge phase at a time, `log_sweeps` sweeps
    each, down to about twice the planted block count)."""
    ctx = m.Context(0)
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
            print(f"# training graph {seed}: B={B}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx.learned_log(0)
    rows = ctx.learned_log_count()
    t1 = time.perf_counter()
    first = ctx.learned_train(50, 4096, 1e-3, 7)
    last = ctx.learned_train(steps - 50, 4096, 1e-3, 8)
    t2 = time.perf_counter()
    params = ctx.learned_params()
    ctx.close()
    return params, {"graphs": [int(s) for s in seeds], "rows": rows, "steps": steps, "batch": 4096,
                    "loss_first": first, "loss_last": last, "log_s": t1 - t0, "train_s": t2 - t1}


def measured_traffic():
    """HBM bytes per launch per kernel, from the PMC summary kept in profiles/ (FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 2 and not line.startswith("#"):
                out[parts[0]] = float(parts[1])
    return out


def cpu_baseline(arrays, seconds, seed, algorithm):
    """The CPU oracle (C++/OpenMP restatement of the reference) on the host cores: one merge phase
    from the identity partition, then sweeps of the same algorithm (batches=1) for about
    `seconds`."""
    import oracle_lib as O

    n = len(arrays[0]) - 1
    g = O.Graph(n, *arrays)
    g.split()
    bm = O.Blockmodel(g, n)
    t0 = time.perf_counter()
    bm.merge_blocks(True, seed, 0)
    t_merge = time.perf_counter() - t0
    evals = 0
    sweeps = 0
    t1 = time.perf_counter()
    while time.perf_counter() - t1 < seconds and sweeps < 100:
        if algorithm == "hybrid_mcmc":
            st = bm.mh_pass(seed, sweeps)
            evals += st.evals
            st = bm.async_sweep(1, seed, sweeps, hybrid_low=True)
        else:
            st = bm.async_sweep(1, seed, sweeps)
        evals += st.evals
        sweeps += 1
    t_sweeps = time.perf_counter() - t1
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": evals / t_sweeps,
        "unit": "vertex-move evals/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle {algorithm} sweeps on the same graph after one merge phase from the identity "
                   f"partition (B={n}->{bm.B}, merge {t_merge:.1f} s untimed); {sweeps} sweeps, {evals} evals "
                   f"in {t_sweeps:.1f} s, OpenMP threads={threads}"),
    }


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
    traffic = measured_traffic()
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--algorithm", default="hybrid_mcmc", choices=["hybrid_mcmc", "async_gibbs",
                                                                   "metropolis_hastings"])
    ap.add_argument("--batches", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the untimed byte-counting run")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal comparison")
    ap.add_argument("--proposal", type=int, default=2, choices=[1, 2, 3],
                    help="learned mixture in: 1 the batched sweeps, 2 also the sequential pass's one-wavefront "
                         "chunks, 3 also its wavefront-group chunks")
    ap.add_argument("--train-graphs", type=int, default=2, help="graphs the proposal network is trained on")
    ap.add_argument("--train-steps", type=int, default=600, help="Adam steps of the proposal network (>= 100)")
    ap.add_argument("--learn-maxdeg", type=int, default=126,
                    help="degree bound of the learned mixture (<= 510; batched bins 3-4 from 254 and 510)")
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
    g = m.ResidentGraph(*arrays)
    kw = dict(algorithm=args.algorithm, batches=args.batches, approximate=True, seed=args.seed)
    for w in range(args.warmup):
        _, st = g.partition(**kw)
        print(f"# warmup {w}: {st.seconds:.3f} s, B={st.num_blocks}", file=sys.stderr, flush=True)
    times, evals, attempts, sweeps = [], 0, 0, 0
    kernel_ms = np.zeros(4)
    kernel_launches = np.zeros(4)
    last = None
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        a, st = g.partition(**kw)
        times.append(time.perf_counter() - t0)
        evals += st.evals
        attempts += st.attempts
        sweeps += st.sweeps
        kernel_ms += np.array(st.eval_ms[:4])
        kernel_launches += np.array(st.eval_launches[:4])
        last = (a, st)
        print(f"# step {k}: {times[-1]:.3f} s, B={st.num_blocks}, mdl={st.mdl:.1f}, evals={st.evals}",
              file=sys.stderr, flush=True)
    total = time.perf_counter() - t_all
    # untimed counting run of the same seed: algorithmic bytes of the evaluation kernels
    if args.no_count:
        kernel_bytes = np.zeros(4)
    else:
        _, cst = g.partition(count_bytes=True, **kw)
        kernel_bytes = np.array(cst.eval_bytes[:4])
        print("# byte-counting run done", file=sys.stderr, flush=True)
    ms_per_run = kernel_ms / args.steps
    launches_per_run = kernel_launches / args.steps
    gbs = np.where(ms_per_run > 0, kernel_bytes / np.maximum(ms_per_run, 1e-12) / 1e6, 0.0)
    dom = int(np.argmax(ms_per_run))
    achieved = float(gbs[dom])
    a, st = last
    # the PMC summary covers the default workload (C2, hybrid, seed 0) only
    default_workload = args.config == "C2" and args.algorithm == "hybrid_mcmc" and args.seed == 0
    traffic = measured_traffic() if default_workload else {}
    dom_name = KERNELS[dom].split(" ")[0]
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "sbpx_eval", os.path.join(ROOT, "hybrid-stochastic-block-partitioning_amd", "evaluation.py"))
    ev = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ev)
    nmi = ev.nmi(truth, a)
    modularity = ev.modularity(out_ptr, out_idx, in_ptr, a)
    modularity_planted = ev.modularity(out_ptr, out_idx, in_ptr, truth)
    out = {
        "metric": METRIC,
        "value": evals / total,
        "unit": "vertex-move evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * float(np.mean(times)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f32 cells, f64 sums",
        "data": "synthetic DC-SBM (seeded generator)",
        "config": {"workload": f"{args.config} plain {args.algorithm} SBP, batches={args.batches}, --approximate",
                   "seed": args.seed,
                   "num_vertices": cfg["num_vertices"], "num_edges": E, "num_blocks_planted": cfg["num_blocks"],
                   "p_in": cfg["p_in"], "dirichlet": cfg["dirichlet"], "deg": [cfg["deg_min"], cfg["deg_max"]]},
        "time_to_partition_s": float(np.mean(times)),
        "attempts_per_s": attempts / total,
        "sweeps_per_s": sweeps / total,
        "final_mdl": st.mdl,
        "final_num_blocks": st.num_blocks,
        "nmi_vs_planted": nmi,
        "modularity": modularity,
        "modularity_planted": modularity_planted,
        "roofline": {
            "bound": "hbm",
            "kernel": KERNELS[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get(dom_name),
            "traffic_source": "profiles/hbm_traffic.txt (rocprofv3 PMC, same workload)" if dom_name in traffic else None,
            "bytes_per_run": float(kernel_bytes[dom]),
            "ms_per_run": float(ms_per_run[dom]),
            "launches_per_run": float(launches_per_run[dom]),
            "avg_launch_us": 1000.0 * float(ms_per_run[dom]) / max(1.0, float(launches_per_run[dom])),
        },
        "eval_kernels": [
            {"kernel": KERNELS[i].split(" ")[0], "ms_per_run": float(ms_per_run[i]),
             "launches_per_run": float(launches_per_run[i]), "bytes_per_run": float(kernel_bytes[i]),
             "achieved_GBs": float(gbs[i]), "frac": float(gbs[i]) / HBM_PEAK_GBS,
             "traffic": traffic.get(KERNELS[i].split(" ")[0])}
            for i in range(4) if launches_per_run[i] > 0],
    }
    if not args.no_learned and args.algorithm != "metropolis_hastings":
        params, info = train_learned(m, cfg, seeds=tuple(args.seed + 1 + i for i in range(args.train_graphs)),
                                     steps=args.train_steps)
        g.load_learned(params)
        g.set_option(m.OPT_LEARN_MAXDEG, args.learn_maxdeg)
        t0 = time.perf_counter()
        la, lst = g.partition(learned=args.proposal, **kw)
        tl = time.perf_counter() - t0
        plain_t = float(np.mean(times))
        # time at equal MDL: when the learned run's outer loop first reached the plain run's final MDL
        lt, lmdl, _ = g.trajectory()
        hit = np.flatnonzero(lmdl <= st.mdl)
        t_equal = float(lt[hit[0]]) if hit.size else None
        out["learned"] = {
            "proposal": "0.2 * plain + 0.8 * network (Hastings-corrected) for vertices of degree <= 126: in the "
                        "batched sweeps" + {1: "", 2: ", and in the sequential Metropolis-Hastings pass with "
                        "candidate scores fixed per pass (chunks evaluated by wavefront groups keep the plain "
                        "proposal)", 3: ", and in the whole sequential Metropolis-Hastings pass with candidate "
                        "scores fixed per pass"}[args.proposal],
            "proposal_option": args.proposal, "learn_maxdeg": args.learn_maxdeg,
            "time_to_partition_s": tl, "speedup_vs_plain": plain_t / tl,
            "time_to_plain_final_mdl_s": t_equal,
            "speedup_at_equal_mdl": (plain_t / t_equal) if t_equal else None,
            "evals_per_s": lst.evals / tl, "sweeps": int(lst.sweeps),
            "final_mdl": lst.mdl, "mdl_vs_plain": lst.mdl / st.mdl, "final_num_blocks": int(lst.num_blocks),
            "nmi_vs_planted": ev.nmi(truth, la), "modularity": ev.modularity(out_ptr, out_idx, in_ptr, la),
            "training": info}
        print(f"# learned: {tl:.3f} s, B={lst.num_blocks}, mdl={lst.mdl:.1f}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cpu_arrays, cpu_note = arrays, ""
        if cfg["num_vertices"] > CPU_SAMPLE_MAX_N:
            # the oracle's first merge of a graph this large alone takes minutes: time it on a
            # graph of the same family with CPU_SAMPLE_MAX_N vertices
            small = dict(cfg, num_vertices=CPU_SAMPLE_MAX_N)
            cpu_arrays = m.generate_dcsbm(**small, seed=args.seed)[:4]
            cpu_note = f"; graph of the same family with {CPU_SAMPLE_MAX_N} vertices (the oracle's merge of the full graph takes minutes)"
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cpu_arrays, args.cpu_seconds, args.seed, args.algorithm)
        out["cpu_baseline"]["sample"] += cpu_note
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
