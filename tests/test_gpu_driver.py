"""GPU tests of the C++ driver around the path: checkpoint and resume of the outer loop, and a
graph read from a file as the reference reads it."""
import numpy as np
import pytest

from pkg import sbpx

pytestmark = pytest.mark.gpu


def _arrays(n=1200, K=8, seed=0):
    m = sbpx()
    op, oi, ip, ii, truth = m.generate_dcsbm(n, K, 0.8, 20.0, 10, 100, 2.1, seed=seed)
    return (op, oi, ip, ii), truth


@pytest.mark.parametrize("algorithm", ["hybrid_mcmc", "async_gibbs"])
def test_checkpoint_resume_gives_the_uninterrupted_run(tmp_path, algorithm):
    m = sbpx()
    arrays, _ = _arrays()
    g = m.ResidentGraph(*arrays)
    kw = dict(algorithm=algorithm, batches=1, approximate=True, seed=3)
    full, st_full = g.partition(**kw)
    ck = tmp_path / "run.ckpt"
    g.partition(checkpoint=ck, max_iterations=3, **kw)
    assert ck.exists()
    resumed, st = g.partition(checkpoint=ck, resume=True, **kw)
    assert (resumed == full).all()
    assert st.num_blocks == st_full.num_blocks and st.mdl == st_full.mdl
    g.close()


def test_file_graph_partitions_like_the_arrays_it_holds(tmp_path):
    m = sbpx()
    arrays, truth = _arrays(900, 6, seed=4)
    op, oi = arrays[0], arrays[1]
    lines = [f"{u + 1}\t{v + 1}\n" for u in range(len(op) - 1) for v in oi[op[u]:op[u + 1]]]
    (tmp_path / "g.tsv").write_text("".join(lines))
    (tmp_path / "g_truePartition.tsv").write_text("".join(f"{v + 1}\t{b + 1}\n" for v, b in enumerate(truth)))
    f = m.load_graph(tmp_path / "g")
    for key, ref in zip(("out_ptr", "out_idx", "in_ptr", "in_idx"), arrays):
        assert (f[key] == ref).all(), key
    assert (f["truth"] == truth).all() and f["num_edges"] == len(oi)
    kw = dict(algorithm="async_gibbs", batches=1, approximate=True, seed=5)
    a1, s1 = m.ResidentGraph(*arrays).partition(**kw)
    a2, s2 = m.ResidentGraph(f["out_ptr"], f["out_idx"], f["in_ptr"], f["in_idx"],
                             num_edges=f["num_edges"]).partition(**kw)
    assert (a1 == a2).all() and s1.mdl == s2.mdl


def test_intermediates_carry_the_reference_fields(tmp_path):
    """sbp::Intermediate rows (sbp.cpp:33-59): normalised MDLs, the share of edges between blocks
    and the block-size variation, checked against restatements from the final assignment; then the
    rows written in the reference's CSV format."""
    import importlib.util
    import os

    m = sbpx()
    (op, oi, ip, ii), truth = _arrays(1200, 8, seed=6)
    g = m.ResidentGraph(op, oi, ip, ii)
    a, st = g.partition(algorithm="async_gibbs", batches=1, approximate=True, seed=2)
    rows = g.intermediates()
    E, N = len(oi), len(op) - 1
    assert rows[0]["iteration"] == 0 and rows[0]["num_blocks"] == N and rows[-1]["iteration"] == -1
    src = np.repeat(np.arange(N), np.diff(op))
    self_loops = int(np.sum(src == oi))
    assert rows[0]["interblock_edges"] == pytest.approx(1.0 - self_loops / E, rel=1e-12)
    last = rows[-1]
    assert last["num_blocks"] == st.num_blocks and last["mdl"] == pytest.approx(st.mdl, rel=1e-7)
    assert last["interblock_edges"] == pytest.approx(1.0 - np.sum(a[src] == a[oi]) / E, rel=1e-12)
    sizes = np.bincount(a, minlength=st.num_blocks).astype(np.float64)
    mean = sizes.sum() / st.num_blocks
    var = np.sum((sizes - mean) ** 2) / st.num_blocks
    bsv = np.float32(var / ((sizes.max() - sizes.min()) * (sizes.max() - mean) / 4.0))
    assert last["block_size_variation"] == pytest.approx(float(bsv), rel=1e-6)
    spec = importlib.util.spec_from_file_location("sbpx_results", os.path.join(
        os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hybrid-stochastic-block-partitioning_amd",
        "results.py"))
    R = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(R)
    for r in rows:
        assert r["normalized_mdl_v1"] == pytest.approx(r["mdl"] / R.null_mdl_v1(E), rel=1e-12)
        assert r["normalized_mdl_v2"] == pytest.approx(r["mdl"] / R.null_mdl_v2(N, E), rel=1e-12)
    path = R.write_results(str(tmp_path) + "/csv", "synthetic", N, E, rows, f1_score=0.9, nmi=0.8, true_mdl=1.0,
                           runtime=st.seconds)
    assert len(open(path).read().splitlines()) == len(rows) + 1
    g.close()
