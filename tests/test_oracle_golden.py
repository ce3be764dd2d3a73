"""Pins the CPU oracle against the reference's own known-answer tests (test/*.cpp, transcribed in
tests/golden/toy_example.json), against closed forms and invariants, and against the
independent cross-check values in tests/golden/crosscheck_moves.json."""
import itertools
import math

import numpy as np
import pytest

import oracle_lib as O

ULP4 = 4  # EXPECT_FLOAT_EQ in googletest: within 4 float ULPs


def float_eq(a, b, ulps=ULP4):
    fa, fb = np.float32(a), np.float32(b)
    return abs(int(fa.view(np.int32)) - int(fb.view(np.int32))) <= ulps


def toy_graph(toy):
    return O.Graph.from_edges(toy["num_vertices"], toy["edges"])


def hlogf(x):
    return O.logf(np.float32(x))


def cell32(value, din, dout):
    """FP32 cell entropy value * logf(value / (din * dout)) with host logf."""
    value, din, dout = np.float32(value), np.float32(din), np.float32(dout)
    if value == 0:
        return np.float32(0.0)
    return np.float32(value * np.float32(hlogf(np.float32(value / np.float32(din * dout)))))


def mdl_from_assignment(toy_or_n, edges, assignment, B, exact_log=False):
    """MDL from scratch in numpy: the block matrix is counted directly from the edge list."""
    n = len(assignment)
    M = np.zeros((B, B), np.int64)
    seen = set()
    for u, v in edges:
        if (u, v) in seen:
            continue
        seen.add((u, v))
        M[assignment[u], assignment[v]] += 1
    E = len(seen)
    dout = M.sum(1)
    din = M.sum(0)
    lp = 0.0
    for i in range(B):
        for j in range(B):
            if M[i, j]:
                ratio = M[i, j] / (float(dout[i]) * float(din[j]))
                lp += M[i, j] * (math.log(ratio) if exact_log else float(hlogf(np.float32(ratio))))
    x = B ** 2 / E
    h = (1 + x) * math.log(1 + x) - x * math.log(x)
    return E * h + n * math.log(B) - lp


# ----------------------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10 (kat_vectors).
    assert list(O.philox(0, 0, 0, 0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox(0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF)) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    key = 0xA4093822 | (0x299F31D0 << 32)
    assert list(O.philox(key, 0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344)) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ----------------------------------------------------------------------------- graph + build
def test_toy_graph_setup(toy):
    g = toy_graph(toy)
    assert g.n == 11 and g.E == toy["num_edges"]  # entropy_test.cpp:43-48


def test_block_degrees_known_answer(toy):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    _, dout, din, deg = bm.state()
    assert list(dout) == toy["block_degrees_out"]
    assert list(din) == toy["block_degrees_in"]
    assert list(deg) == toy["block_degrees"]


def test_build_matches_direct_count_and_invariants():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(2, 40))
        m = int(rng.integers(0, 4 * n))
        edges = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(m, 2))]
        g = O.Graph.from_edges(n, edges)
        B = int(rng.integers(1, n + 1))
        a = rng.integers(0, B, n).astype(np.int32)
        bm = O.Blockmodel(g, B, a)
        M = bm.dense()
        ref = np.zeros((B, B), np.int64)
        for u, v in set(edges):
            ref[a[u], a[v]] += 1
        assert (M == ref).all()
        _, dout, din, deg = bm.state()
        assert (dout == ref.sum(1)).all() and (din == ref.sum(0)).all()
        assert (deg == dout + din - np.diag(ref)).all()
        assert M.sum() == g.E


# ----------------------------------------------------------------------------- MDL
def test_mdl_known_answer(toy, crosschecks):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    mdl = bm.mdl()
    assert float_eq(mdl, toy["entropy"])  # entropy_test.cpp:50-53
    # The reference computes the log in FP32 (utils.hpp:145-151); an exact log gives ENTROPY.
    exact = mdl_from_assignment(None, toy["edges"], toy["assignment"], 3, exact_log=True)
    assert abs(exact - toy["entropy"]) < 1e-7
    assert abs(mdl - crosschecks["toy_mdl_reference_rule"]) < 1e-7
    assert abs(mdl - exact) > 5e-7  # the FP32 log rule is visible at this precision
    assert mdl == pytest.approx(mdl_from_assignment(None, toy["edges"], toy["assignment"], 3), rel=1e-14)


def test_null_models(toy):
    nm = toy["null_models"]
    assert float_eq(O.null_mdl_v1(50), nm["v1_E50"])
    assert float_eq(O.null_mdl_v2(10, 50), nm["v2_N10_E50"])
    assert float_eq(O.null_mdl_v1(int(5e12)), nm["v1_E5e12"])
    assert float_eq(O.null_mdl_v2(int(1e12), int(5e12)), nm["v2_N1e12_E5e12"])


# ----------------------------------------------------------------------------- vertex Delta
def test_vertex_delta_known_answer(toy):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    mv = toy["move"]
    d, sw = bm.vertex_delta(mv["vertex"], mv["proposal"])
    assert len(d) == 6
    for r, c, v in mv["delta"]:
        assert d[(r, c)] == v
    for r, c in mv["delta_absent"]:
        assert (r, c) not in d
    cx = toy["complex"]
    bm6 = O.Blockmodel(g, cx["B"], cx["assignment"])
    d6, _ = bm6.vertex_delta(cx["vertex"], cx["proposal"])
    assert len(d6) == 6
    for r, c, v in cx["delta"]:
        assert d6[(r, c)] == v


def test_vertex_delta_sums_to_zero(toy):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    for v in range(11):
        for s in range(3):
            if s == toy["assignment"][v]:
                continue
            d, _ = bm.vertex_delta(v, s)
            assert sum(d.values()) == 0  # finetune_test.cpp:105-126


# ----------------------------------------------------------------------------- delta MDL
def _all_moves(toy):
    yield 3, toy["assignment"]
    yield 6, toy["complex"]["assignment"]


def test_delta_mdl_equals_mdl_difference_for_known_move(toy, crosschecks):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    mv = toy["move"]
    dS, H, _ = bm.eval_moves([mv["vertex"]], [mv["proposal"]], [0.5])
    a2 = list(toy["assignment"])
    a2[mv["vertex"]] = mv["proposal"]
    after = O.Blockmodel(g, 3, a2).mdl()
    assert float_eq(dS[0], after - bm.mdl())  # entropy_test.cpp:86-98
    ref = crosschecks["moves"][0]
    assert dS[0] == pytest.approx(ref["dS"], abs=2e-6)
    assert H[0] == pytest.approx(ref["H"], abs=2e-7)


def test_delta_mdl_closed_form_with_net_zero_double_count(toy, crosschecks):
    """Reference delta MDL = exact MDL difference + the entries that a net-zero Delta key
    contributes a second time in the row/column loops (entropy.cpp:317-401)."""
    g = toy_graph(toy)
    fired = 0
    for B, assign in _all_moves(toy):
        bm = O.Blockmodel(g, B, assign)
        before = bm.mdl()
        M = bm.dense()
        _, dout, din, _ = bm.state()
        for v in range(11):
            r = assign[v]
            for s in range(B):
                if s == r:
                    continue
                dS, _, _ = bm.eval_moves([v], [s], [0.5])
                a2 = list(assign)
                a2[v] = s
                exact = O.Blockmodel(g, B, a2).mdl() - before
                d, self_w = bm.vertex_delta(v, s)
                out_v = [w for (x, w) in set(map(tuple, toy["edges"])) if x == v]
                k_out = len(out_v)
                k_in = len([x for (x, w) in set(map(tuple, toy["edges"])) if w == v and x != v])
                din2, dout2 = din.astype(np.int64).copy(), dout.astype(np.int64).copy()
                din2[r] -= k_in + self_w
                din2[s] += k_in + self_w
                dout2[r] -= k_out
                dout2[s] += k_out
                extra = 0.0
                for (i, j), dv in d.items():
                    if dv != 0 or M[i, j] == 0:
                        continue
                    in_rows = i in (r, s)
                    in_cols = j in (r, s) and i not in (r, s)
                    if in_rows or in_cols:
                        extra += float(cell32(M[i, j], din[j], dout[i])) - float(cell32(M[i, j], din2[j], dout2[i]))
                        fired += 1
                assert dS[0] == pytest.approx(exact + extra, abs=2e-5), (B, v, s)
    assert fired > 0
    for c in crosschecks["double_count_moves"]:
        assign = toy["assignment"] if c["B"] == 3 else toy["complex"]["assignment"]
        bm = O.Blockmodel(g, c["B"], assign)
        dS, _, _ = bm.eval_moves([c["v"]], [c["s"]], [0.5])
        assert dS[0] == pytest.approx(c["dS"], abs=2e-6)


# ----------------------------------------------------------------------------- Hastings
def test_hastings_cross_checks(toy, crosschecks):
    g = toy_graph(toy)
    for c in crosschecks["moves"]:
        assign = toy["assignment"] if c["B"] == 3 else toy["complex"]["assignment"]
        bm = O.Blockmodel(g, c["B"], assign)
        dS, H, _ = bm.eval_moves([c["v"]], [c["s"]], [0.5])
        assert H[0] == pytest.approx(c["H"], abs=2e-7)
        assert dS[0] == pytest.approx(c["dS"], abs=2e-6)


def test_hastings_reverse_move_is_reciprocal(toy):
    """With exact block degrees and no self-loop, H(r->s) * H(s->r after the move) = 1."""
    g = toy_graph(toy)
    selfs = toy["self_edges"]
    checked = 0
    for B, assign in _all_moves(toy):
        bm = O.Blockmodel(g, B, assign)
        for v in range(11):
            if selfs[v]:
                continue
            r = assign[v]
            for s in range(B):
                if s == r:
                    continue
                _, H, _ = bm.eval_moves([v], [s], [0.5])
                a2 = list(assign)
                a2[v] = s
                _, H2, _ = O.Blockmodel(g, B, a2).eval_moves([v], [r], [0.5])
                assert H[0] * H2[0] == pytest.approx(1.0, rel=1e-12)
                checked += 1
    assert checked > 50


def test_accept_rule(toy):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    dS, H, _ = bm.eval_moves([7], [0], [0.0])
    p = min(1.0, math.exp(-3.0 * dS[0]) * H[0])
    _, _, acc_lo = bm.eval_moves([7, 7], [0, 0], [p * 0.999999, p])
    _, _, acc_hi = bm.eval_moves([7], [0], [p * 1.000001])
    assert list(acc_lo) == [1, 1] and list(acc_hi) == [0]


# ----------------------------------------------------------------------------- move_vertex
def test_move_vertex_matches_rebuild_and_degree_drift(toy, crosschecks):
    g = toy_graph(toy)
    cx = toy["complex"]
    bm = O.Blockmodel(g, 6, cx["assignment"])
    bm.apply_moves([cx["vertex"]], [cx["proposal"]])
    a2 = list(cx["assignment"])
    a2[cx["vertex"]] = cx["proposal"]
    fresh = O.Blockmodel(g, 6, a2)
    assert (bm.dense() == fresh.dense()).all()  # blockmodel_test.cpp:110-123
    _, dout, din, deg = bm.state()
    _, fdout, fdin, fdeg = fresh.state()
    assert list(fdout) == cx["after_out"] and list(fdin) == cx["after_in"] and list(fdeg) == cx["after_total"]
    assert (dout == fdout).all() and (din == fdin).all()
    # blockmodel.cpp:348-359 recomputes deg from the updated diagonal plus the Delta diagonal.
    q9 = crosschecks["q9_drift"]
    assert deg[3] == q9["reference_deg3"]
    d, _ = O.Blockmodel(g, 6, cx["assignment"]).vertex_delta(cx["vertex"], cx["proposal"])
    for x in range(6):
        drift = d.get((x, x), 0) if x in (cx["current"], cx["proposal"]) else 0
        assert deg[x] == fdeg[x] - drift


def test_apply_moves_matrix_equals_rebuild_random():
    rng = np.random.default_rng(11)
    for trial in range(10):
        n = 60
        edges = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(300, 2))]
        g = O.Graph.from_edges(n, edges)
        B = 7
        a = rng.integers(0, B, n).astype(np.int32)
        bm = O.Blockmodel(g, B, a)
        vs = rng.permutation(n)[:25].astype(np.int32)
        ss = rng.integers(0, B, 25).astype(np.int32)
        bm.apply_moves(vs, ss)
        a2 = a.copy()
        a2[vs] = ss
        fresh = O.Blockmodel(g, B, a2)
        assert (bm.dense() == fresh.dense()).all()
        _, dout, din, _ = bm.state()
        _, fdout, fdin, _ = fresh.state()
        assert (dout == fdout).all() and (din == fdin).all()


# ----------------------------------------------------------------------------- merges
def test_merge_delta_known_answer(toy, crosschecks):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 3, toy["assignment"])
    mg = toy["merge"]
    d = bm.merge_delta(mg["from"], mg["to"])
    assert len(d) == 6
    for r, c, v in mg["delta"]:
        assert d[(r, c)] == v
    for r, c in mg["delta_absent"]:
        assert (r, c) not in d
    merged = O.Blockmodel(g, 3, mg["merged_assignment"])
    _, dout, din, deg = merged.state()
    assert list(dout) == mg["new_out"] and list(din) == mg["new_in"] and list(deg) == mg["new_total"]
    dS = bm.merge_delta_mdl(mg["from"], mg["to"])
    assert float_eq(dS, merged.mdl() - bm.mdl())  # entropy_test.cpp:238-244
    assert dS == pytest.approx(crosschecks["merge_0_to_1_dS"], abs=2e-6)


def test_merge_delta_mdl_is_exact_for_all_pairs(toy):
    g = toy_graph(toy)
    for B, assign in _all_moves(toy):
        bm = O.Blockmodel(g, B, assign)
        before = bm.mdl()
        for r, s in itertools.permutations(range(B), 2):
            merged = [s if b == r else b for b in assign]
            after = O.Blockmodel(g, B, merged).mdl()
            assert bm.merge_delta_mdl(r, s) == pytest.approx(after - before, abs=2e-5)


def test_carry_out_best_merges_hand_case():
    # 8 vertices in 6 blocks; sorted order by dS: 4, 1, 0, 5, 2, 3 (ties broken as std::sort).
    n = 8
    edges = [(i, (i + 1) % n) for i in range(n)]
    g = O.Graph.from_edges(n, edges)
    a = np.array([0, 0, 1, 2, 3, 4, 5, 5], np.int32)
    bm = O.Blockmodel(g, 6, a)
    bm.num_to_merge = 3
    dS = np.array([0.3, 0.2, 0.5, 0.9, 0.1, 0.4])
    best = np.array([1, 0, 3, 2, 1, 4], np.int32)
    bm.carry_out(dS, best)
    # 4 -> 1 (map[4]=1); 1 -> map[0]=0 (blocks 1,4 -> 0); 0 -> map[best[0]=1]=0: skip;
    # 5 -> map[4]=0: merge 5 -> 0. Three merges; survivors {0, 2, 3} -> {0, 1, 2}.
    aa, _, _, _ = bm.state()
    assert bm.B == 3
    assert list(aa) == [0, 0, 0, 1, 2, 0, 0, 0]


def _replay_heap_merges(g, B, assign, dS, best, k):
    """block_merge.cpp:123-217 replayed in plain Python: the minimum-score entry is found by a
    scan (scores are made distinct, so any priority queue pops the same order), each merge is
    re-scored on a blockmodel rebuilt from the current assignment, block_map is rewritten as the
    reference does, and labels are compacted in ascending order at the end."""
    assign = np.array(assign, np.int32)
    block_map = list(range(B))
    queue = [(float(dS[i]), i, int(best[i])) for i in range(B)]
    merged = 0
    rescored = 0
    while merged < k and queue:
        i = min(range(len(queue)), key=lambda j: queue[j][0])
        _, r, b = queue.pop(i)
        to = block_map[b]
        if r == to:
            continue
        actual = O.Blockmodel(g, B, assign).merge_delta_mdl(r, to)
        if queue and actual > min(q[0] for q in queue):
            queue.append((actual, r, b))
            rescored += 1
            continue
        block_map = [to if x == r else x for x in block_map]
        assign[assign == r] = to
        merged += 1
    labels = np.unique(assign)
    return np.searchsorted(labels, assign).astype(np.int32), merged, rescored


def test_carry_out_best_merges_advanced_replays_the_heap_loop():
    rng = np.random.default_rng(12)
    total_rescored = 0
    for trial in range(6):
        n = 240
        edges = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(1500, 2))]
        g = O.Graph.from_edges(n, edges)
        B = 48
        a = rng.integers(0, B, n).astype(np.int32)
        bm = O.Blockmodel(g, B, a)
        dS, best = bm.merge_proposals(7, trial)
        dS = dS + rng.random(B) * 1e-3  # distinct scores
        k = B // 3
        want, merged, rescored = _replay_heap_merges(g, B, a, dS, best, k)
        total_rescored += rescored
        assert merged == k
        bm.num_to_merge = k
        bm.carry_out(dS, best, advanced=True)
        got, dout, din, _ = bm.state()
        assert bm.B == B - k
        assert (got == want).all()
        fresh = O.Blockmodel(g, B - k, want)
        assert (bm.dense() == fresh.dense()).all()
    assert total_rescored > 0  # the re-push branch was exercised


# ----------------------------------------------------------------------------- proposals
def test_vertex_proposal_distribution(toy):
    """s ~ sum over neighbour slots (1/k) (M[t,s] + M[s,t]) / (dout[t] + din[t])."""
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 6, toy["complex"]["assignment"])
    M = bm.dense().astype(float)
    a, dout, din, _ = bm.state()
    edges = sorted(set(map(tuple, toy["edges"])))
    for v in (3, 7, 10):
        nbrs = [w for (x, w) in edges if x == v] + [x for (x, w) in edges if w == v and x != v]
        p = np.zeros(6)
        for w in nbrs:
            t = a[w]
            p += (M[t, :] + M[:, t]) / (dout[t] + din[t]) / len(nbrs)
        n = 20000
        s, _ = bm.propose(np.full(n, v, np.int32), seed=123, call=0)
        # identical vertex => identical draw; vary the call id instead
        s = np.array([bm.propose([v], seed=123, call=c)[0][0] for c in range(n)])
        counts = np.bincount(s, minlength=6)
        expected = p * n
        mask = expected > 0
        assert counts[~mask].sum() == 0
        chi2 = ((counts[mask] - expected[mask]) ** 2 / expected[mask]).sum()
        assert chi2 < 30.0, (v, counts, expected)


def _plain_law(g_edges, a, M, dout, din, v):
    """The sampler's law of test_vertex_proposal_distribution, for every target block."""
    nbrs = [w for (x, w) in g_edges if x == v] + [x for (x, w) in g_edges if w == v and x != v]
    p = np.zeros(M.shape[0])
    for w in nbrs:
        t = a[w]
        p += (M[t, :] + M[:, t]) / (dout[t] + din[t]) / len(nbrs)
    return p


def test_plain_probabilities_are_the_samplers_law_both_ways(toy):
    """The learned mixture's plain part: P(s | now) is the sampler's law pinned above (closed form
    and chi-squared), and P(r | after the move) equals P(r) of a blockmodel rebuilt from the moved
    assignment, so the reverse term, the degrees after the move and the self-loop's slot are pinned
    by an independent build. The reference's closed form (smoothed by 1 and B) is not that law."""
    g = toy_graph(toy)
    edges = sorted(set(map(tuple, toy["edges"])))
    B = 6
    a0 = np.array(toy["complex"]["assignment"], np.int32)
    bm = O.Blockmodel(g, B, a0)
    M = bm.dense().astype(float)
    a, dout, din, _ = bm.state()
    checked = differs = 0
    for v in range(len(a)):
        law = _plain_law(edges, a, M, dout, din, v)
        r = int(a[v])
        for s in range(B):
            if s == r:
                continue
            fwd, bwd = bm.plain_probabilities(v, s)
            assert fwd == pytest.approx(law[s], rel=1e-12, abs=1e-15), (v, s)
            a2 = a.copy()
            a2[v] = s
            bm2 = O.Blockmodel(g, B, a2)
            M2 = bm2.dense().astype(float)
            _, dout2, din2, _ = bm2.state()
            back = _plain_law(edges, a2, M2, dout2, din2, v)[r]
            assert bwd == pytest.approx(back, rel=1e-12, abs=1e-15), (v, s)
            assert bm2.plain_probabilities(v, r)[0] == pytest.approx(bwd, rel=1e-12, abs=1e-15)
            pf, _ = bm.hastings_terms(v, s)
            k = len([w for (x, w) in edges if x == v]) + len([x for (x, w) in edges if w == v and x != v])
            differs += abs(pf / k - fwd) > 1e-3 * max(fwd, 1e-9)
            checked += 1
    assert checked > 40 and differs > checked // 2


def _merge_law(M, r, t_from_row_only=False, s_from_row_only=False, keep_r=False):
    """P(s | r) of propose_merge_sparse (block_merge.cpp:335-368 with common.cpp:131-182,
    block_merge=true): t in proportion to row r + column r (a diagonal cell counts twice), then s
    in proportion to row t + column t with index r zeroed; a uniform other block when row r and
    column r are empty or when t has no weight left after zeroing r. The keyword arguments give
    the laws of plausible mistakes, which the test must be able to tell apart."""
    B = M.shape[0]
    uniform = np.full(B, 1.0 / (B - 1))
    uniform[r] = 0.0
    w = M[r, :] + (0 if t_from_row_only else M[:, r])
    if w.sum() == 0:
        return uniform
    p = np.zeros(B)
    for t in range(B):
        if w[t] == 0:
            continue
        e = M[t, :] + (0 if s_from_row_only else M[:, t])
        e = e.astype(float)
        if not keep_r:
            e[r] = 0.0
        p += w[t] / w.sum() * (e / e.sum() if e.sum() > 0 else uniform)
    return p


def _chi2(counts, p, n):
    expected = p * n
    mask = expected > 0
    if counts[~mask].sum() > 0:
        return float("inf")
    return float(((counts[mask] - expected[mask]) ** 2 / expected[mask]).sum())


def test_merge_proposal_distribution():
    """The oracle's merge sampler against the closed-form law, by chi-squared over many draws,
    on a blockmodel built to reach every branch: diagonal cells (t = r), a block whose only
    neighbour is r (t has nothing left once r is zeroed: uniform fallback), an empty block (k = 0:
    uniform), and asymmetric rows/columns, so that drawing t from row r alone, s from row t
    alone, or s without zeroing r all fail."""
    rng = np.random.default_rng(21)
    B = 7
    sizes = [6, 5, 5, 4, 4, 3, 0]  # block 6 is empty
    a = np.concatenate([np.full(k, b) for b, k in enumerate(sizes)]).astype(np.int32)
    n = len(a)
    members = [np.flatnonzero(a == b) for b in range(B)]
    edges = set()
    # blocks 0-4: random, directed-asymmetric block pattern with in-block edges (diagonal cells)
    pattern = rng.integers(0, 4, (5, 5)) * (rng.random((5, 5)) < 0.7)
    for x in range(5):
        for y in range(5):
            for _ in range(int(pattern[x, y])):
                edges.add((int(rng.choice(members[x])), int(rng.choice(members[y]))))
    # block 5 talks only to block 0, in one direction
    for v in members[5]:
        edges.add((int(v), int(rng.choice(members[0]))))
    g = O.Graph.from_edges(n, sorted(edges))
    bm = O.Blockmodel(g, B, a)
    M = bm.dense().astype(float)
    assert M[6].sum() == 0 and M[:, 6].sum() == 0
    assert M[5, 0] > 0 and M[5].sum() == M[5, 0] and M[:, 5].sum() == 0
    assert np.diag(M)[:5].sum() > 0
    draws = 6000
    mistakes = ({"t_from_row_only": True}, {"s_from_row_only": True}, {"keep_r": True})
    alt = {i: [] for i in range(len(mistakes))}
    for r in range(B):
        s = np.array([bm.propose_merge_draw(r, c % 10, 99, c // 10) for c in range(draws)])
        counts = np.bincount(s, minlength=B).astype(float)
        assert counts[r] == 0
        assert _chi2(counts, _merge_law(M, r), draws) < 35.0, (r, counts, _merge_law(M, r) * draws)
        for i, kw in enumerate(mistakes):
            alt[i].append(_chi2(counts, _merge_law(M, r, **kw), draws))
    # each plausible mistake's law is rejected on at least one block
    for i, kw in enumerate(mistakes):
        assert max(alt[i]) > 200.0, (kw, alt[i])


def test_async_batches_skip_the_truncated_tail():
    """finetune.cpp:44-45: batch_size = ceil(N / batches) and the loop runs N / batch_size
    batches in integer division. Worked by hand: N = 1001, batches = 10 gives batch_size 101 and
    9 batches, so vertices 909-1000 are never proposed (909 attempts); N = 1000, batches = 3 gives
    334 and 2 batches, 668 attempts; N = 1000, batches = 10 divides evenly, 1000 attempts."""
    rng = np.random.default_rng(4)
    for n, batches, want in ((1001, 10, 909), (1000, 3, 668), (1000, 10, 1000)):
        edges = set()
        for u in range(n):
            for w in rng.integers(0, n, 6):
                if int(w) != u:
                    edges.add((u, int(w)))
        g = O.Graph.from_edges(n, sorted(edges))
        bm = O.Blockmodel(g, n // 4, (np.arange(n) % (n // 4)).astype(np.int32))
        before = bm.state()[0].copy()
        st = bm.async_sweep(batches, 5, 0)
        assert st.attempts == want, (n, batches, st.attempts)
        after = bm.state()[0]
        moved = np.flatnonzero(after != before)
        assert len(moved) > 0 and moved.max() < want, (n, batches, moved.max())


def test_merge_proposals_are_in_support(toy):
    g = toy_graph(toy)
    bm = O.Blockmodel(g, 6, toy["complex"]["assignment"])
    dS, best = bm.merge_proposals(seed=7, call=0)
    M = bm.dense()
    for r in range(6):
        assert best[r] != r and 0 <= best[r] < 6
        assert dS[r] == pytest.approx(bm.merge_delta_mdl(r, int(best[r])), abs=1e-12)


# ----------------------------------------------------------------------------- driver
def test_sbp_end_to_end_small(toy):
    rng = np.random.default_rng(3)
    n, K = 120, 3
    truth = rng.integers(0, K, n)
    edges = []
    for u in range(n):
        for _ in range(8):
            if rng.random() < 0.85:
                cand = np.flatnonzero(truth == truth[u])
            else:
                cand = np.flatnonzero(truth != truth[u])
            v = int(rng.choice(cand))
            if v != u:
                edges.append((u, v))
    g = O.Graph.from_edges(n, edges)
    for algo in (0, 1, 2):
        a, run = O.sbp(g, algo, 1, True, 42)
        a2, run2 = O.sbp(g, algo, 1, True, 42)
        assert (a == a2).all() and run.B == run2.B  # deterministic given the seed
        assert 1 <= run.B <= n
        initial = O.Blockmodel(g, n).mdl()
        assert run.mdl < initial


# ----------------------------------------------------------------------------- early stop, golden section
def _early_stop(iteration, bracket, initial, deltas):
    L = O.lib()
    return L.oracle_early_stop(iteration, int(bracket), initial, np.asarray(deltas, np.float64), len(deltas))


def test_early_stop_hand_worked_cases():
    """finetune::early_stop (finetune.cpp:144-162), cases worked by hand from the reference:
    stop at once when the last change is exactly 0; never before the third iteration; otherwise
    stop when minus the mean of the LAST THREE changes is strictly below 5e-4 * S0 before the
    bracket forms and 1e-4 * S0 after."""
    assert _early_stop(1, False, 1000.0, [-5.0, 0.0]) == 1          # last change 0
    assert _early_stop(2, False, 1000.0, [-1.0, -1.0, -1e-9]) == 0  # iteration < 3
    # -mean(-1, -0.4, -0.05) = 0.48333 < 0.5 (no bracket) but not < 0.1 (bracket)
    assert _early_stop(3, False, 1000.0, [-1.0, -0.4, -0.05]) == 1
    assert _early_stop(3, True, 1000.0, [-1.0, -0.4, -0.05]) == 0
    # exactly at the threshold: 0.5 < 0.5 is false
    assert _early_stop(3, False, 1000.0, [-0.75, -0.5, -0.25]) == 0
    # only the last three count: mean of all four would be 25.075
    assert _early_stop(4, False, 1000.0, [-100.0, -0.1, -0.1, -0.1]) == 1
    # a rising MDL (positive changes) has a negative mean decrease, below any threshold
    assert _early_stop(3, True, 1000.0, [0.3, 0.3, 0.3]) == 1


def _golden_trace(n, score):
    L = O.lib()
    out = np.zeros(256, np.int32)
    k = L.oracle_golden_trace(n, np.asarray(score, np.float64), out, len(out))
    assert k > 0
    return list(out[:k - 1]), int(out[k - 1])


def test_golden_section_hand_worked_traces():
    """The driver's search over the number of blocks (sbp.cpp:61-112 with
    BlockmodelTriplet::get_next_blockmodel / update / is_done, blockmodel_triplet.cpp:3-100),
    scored by a table instead of merges and sweeps. Traces worked by hand from the reference."""
    # (B - 10)^2 + 100: halving 32, 16, 8, 4 brackets at 8; then 0.618 steps with int truncation:
    # 16 - 12 = 4 -> 12 (ties with 8 at 104 and takes the middle), 14, 10, 11, 9; done when
    # the outer pair is 2 apart (11 and 9), returning the middle, 10
    score = [(b - 10) ** 2 + 100.0 for b in range(65)]
    assert _golden_trace(64, score) == ([32, 16, 8, 4, 12, 14, 10, 11, 9], 10)
    # MDL increasing in B: halving until num_to_merge = int(1 * 0.5) = 0 ends the search at 1
    assert _golden_trace(8, [float(b) for b in range(9)]) == ([4, 2, 1], 1)
    # best at the first iteration (upper slot empty): steps from the lower side, 20 - 10 = 10 ->
    # 10 + int(6.18) = 16, then 18, 19; done when middle and lower are 1 apart
    assert _golden_trace(40, [abs(b - 20.0) for b in range(41)]) == ([20, 10, 16, 18, 19], 20)


# ----------------------------------------------------------------------------- learned mixture
def _scores(x, seed=3):
    """A fixed deterministic function of the features standing in for the network."""
    A = np.random.default_rng(seed).normal(size=(32, 64)).astype(np.float32)
    return np.tanh(A @ np.asarray(x, np.float32)).astype(np.float32) * 3.0


def _mixture_graph():
    rng = np.random.default_rng(17)
    n, B = 60, 5
    edges = set()
    for u in range(n):
        for w in rng.integers(0, n, 5):
            if int(w) != u:
                edges.add((u, int(w)))
        if u % 7 == 0:  # a few self-loops
            edges.add((u, u))
    a = rng.integers(0, B, n).astype(np.int32)
    return n, B, sorted(edges), a


def test_learned_features_do_not_change_when_the_vertex_moves():
    """The learned proposal's candidates and features of vertex v are the same in the state after v
    moves r -> s: its neighbours' blocks do not change, and block degrees are taken without v."""
    n, B, edges, a = _mixture_graph()
    g = O.Graph.from_edges(n, edges)
    bm = O.Blockmodel(g, B, a)
    checked = 0
    for v in range(n):
        cand, cnt, x = bm.learned_features(v)
        for s in range(B):
            if s == a[v]:
                continue
            a2 = a.copy()
            a2[v] = s
            bm2 = O.Blockmodel(g, B, a2)
            c2, k2, x2 = bm2.learned_features(v)
            assert (c2 == cand).all() and (k2 == cnt).all() and (x2 == x).all(), (v, s)
            checked += 1
    assert checked > 100


def test_mixture_hastings_reduces_and_reverses(toy):
    """The oracle's mixture Hastings ratio: with all weight on the plain proposal it is the ratio of
    the plain sampler's own probabilities (not the reference's smoothed correction, which differs
    on most moves); with uniform scores and all weight on the network it is (candidates other than
    r) / (candidates other than s); and for every move whose target is a candidate, self-loops
    included, H(x -> x') H(x' -> x) = 1 with the reverse evaluated from scratch in the state after
    the move (detailed balance of the mixture proposal)."""
    n, B, edges, a = _mixture_graph()
    g = O.Graph.from_edges(n, edges)
    bm = O.Blockmodel(g, B, a)
    self_loop = {u for (u, w) in edges if u == w}
    rng = np.random.default_rng(2)
    products = looped = differs = moves = 0
    for v in range(n):
        r = int(a[v])
        cand, cnt, x = bm.learned_features(v)
        z = _scores(x)
        for s in range(B):
            if s == r:
                continue
            _, H_ref, _ = bm.eval_moves([v], [s], [0.5])
            H1, _, _ = bm.mixture_hastings(v, s, z, cand, w=1.0)
            pf0, pb0 = bm.plain_probabilities(v, s)
            if pf0 > 0:
                assert H1 == pytest.approx(pb0 / pf0, rel=1e-12)
                moves += 1
                differs += abs(H1 / float(H_ref[0]) - 1.0) > 1e-6
            valid = cand[cand >= 0]
            if s in valid and r in valid:
                Hu, qf, qb = bm.mixture_hastings(v, s, np.zeros(32, np.float32), cand, w=0.0)
                assert qf == pytest.approx(1.0 / (len(valid) - 1)) and qb == pytest.approx(1.0 / (len(valid) - 1))
                assert Hu == pytest.approx(1.0, rel=1e-12)
            if s not in valid:
                continue
            w = float(rng.choice([0.2, 0.5, 0.9]))
            H, _, _ = bm.mixture_hastings(v, s, z, cand, w=w)
            a2 = a.copy()
            a2[v] = s
            bm2 = O.Blockmodel(g, B, a2)
            c2, _, x2 = bm2.learned_features(v)
            Hr, _, _ = bm2.mixture_hastings(v, r, _scores(x2), c2, w=w)
            assert H * Hr == pytest.approx(1.0, rel=1e-12), (v, r, s)
            products += 1
            looped += v in self_loop
    assert products > 50 and looped > 0
    assert differs > moves // 2


def test_mdl_from_cells_equals_blockmodel_mdl(toy):
    """The cell-list MDL used by the full-size checks equals the blockmodel's entropy::mdl (KAT on
    the toy graph; random assignments of a random graph)."""
    g = O.Graph.from_edges(toy["num_vertices"], toy["edges"])
    bm = O.Blockmodel(g, 3, toy["assignment"])
    r, c, v = bm.coo()
    _, dout, din, _ = bm.state()
    got = O.mdl_cells(r, c, v, dout, din, 3, toy["num_vertices"], g.E)
    assert got == pytest.approx(toy["entropy"], rel=1e-7)
    assert got == pytest.approx(bm.mdl(), rel=1e-12)
    rng = np.random.default_rng(8)
    n = 300
    edges = sorted({(int(u), int(w)) for u, w in rng.integers(0, n, (2000, 2))})
    g = O.Graph.from_edges(n, edges)
    for B in (n, 40, 3):
        a = np.arange(n, dtype=np.int32) if B == n else rng.integers(0, B, n).astype(np.int32)
        bm = O.Blockmodel(g, B, a)
        r, c, v = bm.coo()
        _, dout, din, _ = bm.state()
        assert O.mdl_cells(r, c, v, dout, din, B, n, g.E) == pytest.approx(bm.mdl(), rel=1e-12)
