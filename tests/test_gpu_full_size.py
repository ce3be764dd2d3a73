"""Property checks at the full sizes of the 1M-vertex (K = 200), 5M-vertex (K = 400) and batched
512-graph workloads, where the oracle's own sweeps are too slow to run beside the device.

After a merge phase from the identity partition and sweeps on the device, the state must equal a
rebuild from the downloaded assignment (initialize_edge_counts, blockmodel.cpp:187-255) computed
on the host with numpy: every block-matrix cell (row-major and column-major) and dout/din (the
total degrees carry the reference's drift of Blockmodel::move_vertex, so they are not compared);
and the MDL the device reports must equal entropy::mdl (entropy.cpp:584-589) recomputed by the
oracle from those host cells to 1e-9. For the batch, each graph's log posterior must equal the
oracle's on that graph's own cells."""
import math
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def host_cells(op, oi, ip, a, B):
    """Cells (row-major sorted), dout, din of the block matrix of assignment a, on the host."""
    op = np.asarray(op, np.int64)
    src = np.repeat(np.arange(len(op) - 1, dtype=np.int64), np.diff(op))
    key = a[src].astype(np.int64) * B + a[np.asarray(oi, np.int64)]
    uk, cnt = np.unique(key, return_counts=True)
    dout = np.bincount(a, weights=np.diff(op), minlength=B).astype(np.int64)
    din = np.bincount(a, weights=np.diff(np.asarray(ip, np.int64)), minlength=B).astype(np.int64)
    return (uk // B).astype(np.int32), (uk % B).astype(np.int32), cnt.astype(np.int32), dout, din


def check_against_host(c, arrays, label):
    op, oi, ip, ii = arrays
    n = len(op) - 1
    B, nnz = c.info()
    a, dout, din, _ = c.state()
    t0 = time.time()
    rows, cols, vals, hdout, hdin = host_cells(op, oi, ip, a, B)
    say(f"{label}: host rebuild of {len(rows)} cells in {time.time() - t0:.1f} s (B={B})")
    assert nnz == len(rows)
    assert (dout == hdout).all() and (din == hdin).all()
    r, k, v = c.coo()
    assert (r == rows).all() and (k == cols).all() and (v == vals).all()
    r2, k2, v2 = c.coo(by_column=True)
    order = np.lexsort((rows, cols))
    assert (r2 == rows[order]).all() and (k2 == cols[order]).all() and (v2 == vals[order]).all()
    want = O.mdl_cells(rows, cols, vals, hdout, hdin, B, n, int(op[-1]))
    assert c.mdl() == pytest.approx(want, rel=1e-9)


@pytest.mark.parametrize("config,hybrid", [("C4", True), ("C5", False)])
def test_full_size_state_equals_host_rebuild(config, hybrid):
    m = sbpx()
    t0 = time.time()
    op, oi, ip, ii, truth = m.generate_dcsbm(**m.CONFIGS[config], seed=0)
    arrays = (op, oi, ip, ii)
    n = len(op) - 1
    say(f"{config}: N={n} E={int(op[-1])} generated in {time.time() - t0:.1f} s")
    c = m.Context(0)
    c.upload_graph(*arrays)
    c.split()
    c.build(n, np.arange(n, dtype=np.int32))
    check_against_host(c, arrays, f"{config} identity")
    c.merge_blocks(n // 2, 1, 0)
    say(f"{config}: merged to B={c.info()[0]}")
    st = c.sweep(m.ALGO_ASYNC_GIBBS, 1, 1, 0)
    assert st.moves > 0
    if hybrid:
        st = c.sweep(m.ALGO_HYBRID, 1, 1, 1)
        assert st.moves > 0
    say(f"{config}: sweeps done")
    check_against_host(c, arrays, f"{config} after sweeps")
    c.close()


def test_batched_512_graphs_state_equals_host_rebuild():
    m = sbpx()
    params = m.c3_graph_params(512, 0)
    graphs = []
    t0 = time.time()
    for p in params:
        op, oi, ip, ii, _ = m.generate_dcsbm(**p)
        graphs.append((op, oi, ip, ii))
    offs, op, oi, ip, ii = m.union_graphs(graphs)
    say(f"batch: {len(graphs)} graphs, N={offs[-1]} E={int(op[-1])}, generated in {time.time() - t0:.1f} s")
    c = m.Context(0)
    c.upload_graph(op, oi, ip, ii)
    c.batch_define(offs)
    G = len(graphs)
    ids = np.arange(G, dtype=np.int32)
    sizes = np.diff(offs)
    c.batch_merge(ids, sizes // 2, np.arange(G) + 11, np.zeros(G, np.uint32))
    st = c.batch_sweep(ids, np.arange(G) + 11, np.ones(G, np.uint32))
    assert st.moves > 0
    nb = c.batch_info()
    a_all = c.batch_download()
    lp = c.batch_log_posterior()
    say("batch: merged and swept; checking every graph")
    for g, (gop, goi, gip, gii) in enumerate(graphs):
        a = a_all[offs[g]:offs[g + 1]]
        B = int(nb[g])
        assert a.min() >= 0 and a.max() < B
        rows, cols, vals, dout, din = host_cells(gop, goi, gip, a, B)
        N, E = len(gop) - 1, int(gop[-1])
        want = O.mdl_cells(rows, cols, vals, dout, din, B, N, E)
        x = B * B / E
        model = E * ((1 + x) * math.log(1 + x) - x * math.log(x)) + N * math.log(B)
        assert model - lp[g] == pytest.approx(want, rel=1e-9), g
    c.close()
