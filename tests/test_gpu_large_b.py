"""Large block counts on the device against the oracle: a 132,000-vertex graph of the
1M-vertex workload's family (K = 200, p_in 0.6, Dirichlet(1) block sizes, out-degrees on
[10, 200]) from the identity partition through two merge phases.

The block-matrix build sorts (block(u), block(v)) keys of 2 * bits bits, bits = ceil(log2 B):
  * B = 132,000 and B = 66,000 take 64-bit keys (bits 18 and 17),
  * B = 33,000 takes 32-bit keys that fill the whole word (bits 16, the padding key is all ones),
so the full rebuild, the incremental update after a batch of moves, and both key widths at
their boundary are compared with the oracle (initialize_edge_counts, blockmodel.cpp:187-255;
the batch application of finetune.cpp:58-71). At each block count: cells (CSR and CSC), degrees
with drift, MDL, merge proposals, one asynchronous Gibbs sweep and one hybrid sweep."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx
from test_gpu_parity import DS_RTOL, MDL_RTOL, assert_same_state, ctx_for

pytestmark = pytest.mark.gpu

N = 132_000


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def test_large_block_counts_match_oracle():
    m = sbpx()
    cfg = dict(m.CONFIGS["C4"])
    cfg["num_vertices"] = N
    op, oi, ip, ii, truth = m.generate_dcsbm(**cfg, seed=3)
    arrays = (op, oi, ip, ii)
    say("large-B: graph N", N, "E", int(op[-1]))
    c = ctx_for(arrays)
    g = O.Graph(N, *arrays)
    a = np.arange(N, dtype=np.int32)
    c.build(N, a)
    bm = O.Blockmodel(g, N, a)
    assert_same_state(c, bm)
    assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)
    B = N
    call = 0
    for target in (66_000, 33_000):
        say("large-B: merge proposals at B", B)
        dS, best = c.merge_proposals(91, call)
        odS, obest = bm.merge_proposals(91, call)
        assert (best == obest).all()
        assert np.allclose(dS, odS, rtol=DS_RTOL, atol=1e-9)
        k = B - target
        say("large-B: merge", B, "->", target)
        c.carry_out_merges(k, odS, obest)  # the same scores on both sides: the sort order is fixed
        bm.num_to_merge = k
        bm.carry_out(odS, obest)
        B = target
        assert c.info()[0] == bm.B == B
        assert_same_state(c, bm)
        assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)
        say("large-B: async sweep at B", B)
        st = c.sweep(m.ALGO_ASYNC_GIBBS, 1, 91, call)
        ost = bm.async_sweep(1, 91, call)
        assert (st.attempts, st.evals, st.moves) == (ost.attempts, ost.evals, ost.moves)
        assert st.moves > 0
        assert_same_state(c, bm)
        assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)
        say("large-B: hybrid sweep at B", B)
        st = c.sweep(m.ALGO_HYBRID, 1, 91, call + 1)
        o1 = bm.mh_pass(91, call + 1)
        o2 = bm.async_sweep(1, 91, call + 1, hybrid_low=True)
        assert (st.attempts, st.evals, st.moves) == (o1.attempts + o2.attempts, o1.evals + o2.evals,
                                                     o1.moves + o2.moves)
        assert_same_state(c, bm)
        assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)
        call += 2
    c.close()
