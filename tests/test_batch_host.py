"""Host side of the batched form: the union of several graphs handed to one context is their
disjoint union (vertex ids shifted by each graph's offset, no edge between graphs)."""
import numpy as np

import oracle_lib as O
from pkg import sbpx


def test_union_graphs_is_the_disjoint_union():
    m = sbpx()
    rng = np.random.default_rng(0)
    graphs = []
    for n in (7, 1, 40, 13):
        e = sorted({(int(a), int(b)) for a, b in rng.integers(0, n, size=(3 * n, 2))})
        graphs.append(O.csr_from_edges(n, e))
    offs, op, oi, ip, ii = m.union_graphs(graphs)
    assert list(offs) == [0, 7, 8, 48, 61]
    edges = []
    for k, (gop, goi, _, _) in enumerate(graphs):
        for u in range(len(gop) - 1):
            edges += [(u + offs[k], int(w) + offs[k]) for w in goi[gop[u]:gop[u + 1]]]
    ref = O.csr_from_edges(int(offs[-1]), sorted(edges))
    for a, b in zip(ref, (op, oi, ip, ii)):
        assert a.dtype == b.dtype and (a == b).all()
