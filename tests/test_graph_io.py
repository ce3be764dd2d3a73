"""Graph files read as the reference reads them (graph.cpp:24-210, utils.cpp:19-60): checked
against hand-written expectations and against a line-by-line Python restatement of the
reference's parsers on random files."""
import numpy as np
import pytest

from pkg import sbpx


def lists_from(res):
    op, oi, ip, ii = res["out_ptr"], res["out_idx"], res["in_ptr"], res["in_idx"]
    n = len(op) - 1
    return ([list(oi[op[v]:op[v + 1]]) for v in range(n)], [list(ii[ip[v]:ip[v + 1]]) for v in range(n)])


def restate(text, undirected, mtx):
    """graph.cpp's parsers restated: first-appearance order, no duplicates, 1-based ids."""
    rows = [line.split() for line in text.splitlines()]
    out, inn = {}, {}

    def ins(l, a, b):
        l.setdefault(a, [])
        if b not in l[a]:
            l[a].append(b)

    n, E = 0, 0
    start = 0
    if mtx:
        undirected = rows[0][4] == "symmetric"
        for i, r in enumerate(rows):
            if not r or r[0][0] == "%":
                continue
            n, E, start = int(r[0]), int(r[2]), i + 1
            break
    for r in rows[start:]:
        if len(r) < 2:
            continue
        a, b = int(r[0]) - 1, int(r[1]) - 1
        n = max(n, a + 1, b + 1)
        ins(out, a, b)
        ins(inn, b, a)
        if undirected and a != b:
            ins(out, b, a)
            ins(inn, a, b)
            if mtx:
                E += 1
    if not mtx:
        E = sum(len(v) for v in out.values())
        if undirected:
            E //= 2
            inn = {k: list(v) for k, v in out.items()}  # parse_undirected: in-lists copy the out-lists
    return [out.get(v, []) for v in range(n)], [inn.get(v, []) for v in range(n)], E


def write(tmp_path, name, text, truth=None):
    base = tmp_path / name
    ext = ".mtx" if text.startswith("%%MatrixMarket") else ".tsv"
    (tmp_path / (name + ext)).write_text(text)
    if truth is not None:
        (tmp_path / (name + "_truePartition.tsv")).write_text(truth)
    return base


def test_tsv_known_answer(tmp_path):
    m = sbpx()
    text = "1\t2\n2\t3\n1\t2\n3\t1\n3\t3\n5\t1\n"  # duplicate 1->2, self-loop on 3, vertex 4 isolated
    base = write(tmp_path, "g", text, truth="1 2\n2 2\n3 1\n5 3\n")
    res = m.load_graph(base)
    out, inn = lists_from(res)
    assert out == [[1], [2], [0, 2], [], [0]]
    assert inn == [[2, 4], [0], [1, 2], [], []]
    assert res["num_edges"] == 5
    assert list(res["truth"]) == [1, 1, 0, -1, 2]


def test_matrix_market_symmetric_known_answer(tmp_path):
    m = sbpx()
    text = "%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n4 4 4\n2 1\n3 1\n3 2\n4 4\n"
    res = m.load_graph(write(tmp_path, "s", text))
    out, inn = lists_from(res)
    assert res["undirected"]
    assert out == [[1, 2], [0, 2], [0, 1], [3]]
    assert inn == [[1, 2], [0, 2], [0, 1], [3]]
    assert res["num_edges"] == 4 + 3  # size line plus one per reverse edge
    assert list(res["truth"]) == [0, 0, 0, 0]


def test_matrix_market_rejects_dense(tmp_path):
    m = sbpx()
    with pytest.raises(RuntimeError):
        m.load_graph(write(tmp_path, "d", "%%MatrixMarket matrix array real general\n2 2\n1\n0\n0\n1\n"))


@pytest.mark.parametrize("kind", ["tsv", "tsv-undirected", "mtx", "mtx-symmetric"])
def test_random_files_match_the_restated_parsers(tmp_path, kind):
    m = sbpx()
    rng = np.random.default_rng(hash(kind) % 2**32)
    for trial in range(5):
        n = int(rng.integers(5, 60))
        e = rng.integers(1, n + 1, size=(int(rng.integers(1, 300)), 2))
        body = "".join(f"{a}\t{b}\t{rng.random():.3f}\n" for a, b in e)
        mtx = kind.startswith("mtx")
        if mtx:
            sym = "symmetric" if kind.endswith("symmetric") else "general"
            text = f"%%MatrixMarket matrix coordinate real {sym}\n% x\n{n} {n} {len(e)}\n" + body
        else:
            text = body
        undirected = kind == "tsv-undirected"
        res = m.load_graph(write(tmp_path, f"r{trial}", text), undirected=undirected)
        out, inn = lists_from(res)
        rout, rinn, rE = restate(text, undirected, mtx)
        assert out == rout and inn == rinn
        assert res["num_edges"] == rE
