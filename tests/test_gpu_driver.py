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
