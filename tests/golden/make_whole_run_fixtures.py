"""Writes tests/golden/whole_runs_<config>.json: whole partitioning runs of the CPU oracle.

For each seed, the DC-SBM graph of the named shape is generated with that seed (the product's
seeded generator, host code only) and partitioned by the oracle's restatement of
sbp::stochastic_block_partition (sbp.cpp:61-98) with the published run flags
(scripts/synthetic_run.sbatch:17-19: hybrid_mcmc, --batches 1, --approximate). Recorded per seed:
the final number of blocks and float MDL, NMI and modularity against the planted partition, the
sweep/evaluation/move counts, and the per-outer-iteration trace (blocks after the merge phase, MDL
after the sweeps). The GPU whole-run test replays the same seeds on the device and compares.

Usage: python tests/golden/make_whole_run_fixtures.py C2 0 1 2 3 4
Runs in the build container (CPU only; OMP_NUM_THREADS sets the oracle's threads).
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, TESTS)

import numpy as np  # noqa: E402

import oracle_lib as O  # noqa: E402
from pkg import sbpx  # noqa: E402
from test_gpu_whole_runs import evaluation  # noqa: E402

ALGOS = {"metropolis_hastings": 0, "async_gibbs": 1, "hybrid_mcmc": 2}


def run(config, seeds, algorithm="hybrid_mcmc"):
    m = sbpx()
    ev = evaluation()
    path = os.path.join(HERE, f"whole_runs_{config}.json")
    doc = {"config": config, "params": m.CONFIGS[config], "algorithm": algorithm, "batches": 1,
           "approximate": True, "runs": {}}
    if os.path.exists(path):
        with open(path) as f:
            doc = json.load(f)
    for seed in seeds:
        op, oi, ip, ii, truth = m.generate_dcsbm(**m.CONFIGS[config], seed=seed)
        n = len(op) - 1
        g = O.Graph(n, op, oi, ip, ii)
        t0 = time.time()
        a, st, tb, tm = O.sbp_traced(g, ALGOS[algorithm], 1, True, seed)
        secs = time.time() - t0
        row = {
            "num_vertices": n, "num_edges": int(op[-1]), "num_blocks": int(st.B), "mdl": float(st.mdl),
            "nmi": float(ev.nmi(truth, a)), "modularity": float(ev.modularity(op, oi, ip, a)),
            "modularity_planted": float(ev.modularity(op, oi, ip, truth)),
            "outer_iterations": int(st.outer_iterations), "sweeps": int(st.sweeps), "evals": int(st.evals),
            "moves": int(st.moves), "trace_B": [int(x) for x in tb], "trace_mdl": [float(x) for x in tm],
            "assignment_sha256": hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest(),
            "oracle_seconds": round(secs, 1), "oracle_threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0),
        }
        doc["runs"][str(seed)] = row
        print(config, "seed", seed, {k: row[k] for k in ("num_blocks", "mdl", "nmi", "sweeps", "oracle_seconds")},
              flush=True)
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)


if __name__ == "__main__":
    run(sys.argv[1], [int(s) for s in sys.argv[2:]])
