"""GPU parity: the MI355X library, called through its C ABI, against the CPU oracle on the same
seeded inputs. Integer state (block matrix, degrees, drift, assignments, proposals, accept
decisions) must be bit-exact; a vertex-move delta MDL must agree to 1e-6 relative (FP32 cells,
FP64 sums in a different order); the global MDL to 1e-9 relative."""
import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx

pytestmark = pytest.mark.gpu

DS_RTOL = 1e-6
MDL_RTOL = 1e-9


def ctx_for(graph_arrays):
    m = sbpx()
    c = m.Context(0)
    c.upload_graph(*graph_arrays)
    return c


def toy_arrays(toy):
    return O.csr_from_edges(toy["num_vertices"], toy["edges"])


def gen_arrays(n, K, p_in=0.8, alpha=20.0, dmin=10, dmax=100, seed=0):
    m = sbpx()
    out_ptr, out_idx, in_ptr, in_idx, truth = m.generate_dcsbm(n, K, p_in, alpha, dmin, dmax, 2.1, seed=seed)
    return (out_ptr, out_idx, in_ptr, in_idx), truth


def assert_same_state(c, bm, check_deg=True):
    a, dout, din, deg = c.state()
    oa, odout, odin, odeg = bm.state()
    assert (a == oa).all(), "assignment"
    assert (dout == odout).all(), "dout"
    assert (din == odin).all(), "din"
    if check_deg:
        assert (deg == odeg).all(), ("deg", np.flatnonzero(deg != odeg)[:10])
    r, col, v = c.coo()
    orow, ocol, ov = bm.coo()
    assert len(r) == len(orow), "nnz"
    assert (r == orow).all() and (col == ocol).all() and (v == ov).all(), "CSR"
    r2, c2, v2 = c.coo(by_column=True)
    order = np.lexsort((orow, ocol))
    assert (r2 == orow[order]).all() and (c2 == ocol[order]).all() and (v2 == ov[order]).all(), "CSC"


def assignments(n, rng):
    yield n, np.arange(n, dtype=np.int32)
    for B in (n // 2, 37, 8, 1):
        B = max(1, min(B, n))
        a = rng.integers(0, B, n).astype(np.int32)
        yield B, a


# ----------------------------------------------------------------------------- logf
def test_device_logf_matches_host_logf_on_every_positive_float():
    c = sbpx().Context(0)
    L = O.lib()
    chunk = 1 << 26
    lo, hi = 1, 0x7F800001  # all positive subnormals, normals and +inf
    bad = 0
    start = lo
    while start < hi:
        n = min(chunk, hi - start)
        dev = c.selftest_logf(start, n)
        bad += L.oracle_logf_check(start, start + n, dev)
        start += n
    assert bad == 0


# ----------------------------------------------------------------------------- build + MDL
def test_build_and_mdl_toy(toy):
    c = ctx_for(toy_arrays(toy))
    g = O.Graph.from_edges(toy["num_vertices"], toy["edges"])
    for B, a in ((3, toy["assignment"]), (6, toy["complex"]["assignment"]), (11, list(range(11)))):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        assert_same_state(c, bm)
        assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)
    c.build(3, toy["assignment"])
    _, dout, din, deg = c.state()
    assert list(dout) == toy["block_degrees_out"] and list(din) == toy["block_degrees_in"]
    assert list(deg) == toy["block_degrees"]
    assert abs(c.mdl() - toy["entropy"]) < 1e-5


def test_build_and_mdl_generated_graphs():
    rng = np.random.default_rng(0)
    for seed, (n, K) in enumerate(((1000, 8), (3000, 20))):
        arrays, truth = gen_arrays(n, K, seed=seed)
        c = ctx_for(arrays)
        g = O.Graph(n, *arrays)
        for B, a in list(assignments(n, rng)) + [(K, truth)]:
            c.build(B, a)
            bm = O.Blockmodel(g, B, a)
            assert_same_state(c, bm)
            assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)


def test_build_rejects_bad_block_ids(toy):
    c = ctx_for(toy_arrays(toy))
    m = sbpx()
    with pytest.raises(m.SbpxError):
        c.build(3, [0, 1, 2, 3, 0, 0, 0, 0, 0, 0, 0])
    with pytest.raises(m.SbpxError):
        c.build(0, [0] * 11)


def test_options_reject_values_outside_their_range(toy):
    c = ctx_for(toy_arrays(toy))
    m = sbpx()
    for option, bad in ((m.OPT_MH_GROUP, 3), (m.OPT_EVAL_VARIANT, 4), (m.OPT_PROPOSAL, 3), (m.OPT_INCREMENTAL, 2),
                        (m.OPT_MH_WIDE, 3), (m.OPT_MH_WIDE, 1024), (m.OPT_MH_HEADROOM, -1),
                        (m.OPT_LEARN_MAXDEG, -1), (m.OPT_LEARN_MAXDEG, 511), (99, 0)):
        with pytest.raises(m.SbpxError):
            c.set_option(option, bad)
    for option, good in ((m.OPT_MH_WIDE, 32), (m.OPT_MH_WIDE, 0), (m.OPT_MH_HEADROOM, 0), (m.OPT_PROPOSAL, 2),
                         (m.OPT_PROPOSAL, 1), (m.OPT_PROPOSAL, 0), (m.OPT_LEARN_MAXDEG, 510),
                         (m.OPT_LEARN_MAXDEG, 126)):
        c.set_option(option, good)


# ----------------------------------------------------------------------------- moves
def _random_proposals(n, B, a, rng, count):
    v = rng.permutation(n)[:count].astype(np.int32)
    s = rng.integers(0, B, count).astype(np.int32)
    s = np.where(s == a[v], (s + 1) % B, s).astype(np.int32)
    u = rng.random(count)
    return v, s, u


def test_eval_moves_toy_every_move(toy):
    c = ctx_for(toy_arrays(toy))
    g = O.Graph.from_edges(toy["num_vertices"], toy["edges"])
    for B, a in ((3, toy["assignment"]), (6, toy["complex"]["assignment"])):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        vs, ss = zip(*[(v, s) for v in range(11) for s in range(B) if s != a[v]])
        u = np.linspace(0.01, 0.99, len(vs))
        dS, H, acc = c.eval_moves(vs, ss, u)
        odS, oH, oacc = bm.eval_moves(vs, ss, u)
        assert np.allclose(dS, odS, rtol=DS_RTOL, atol=1e-9)
        assert np.allclose(H, oH, rtol=1e-12)
        assert (acc == oacc).all()


def test_eval_moves_generated():
    rng = np.random.default_rng(1)
    n = 2000
    arrays, truth = gen_arrays(n, 12, seed=3)
    c = ctx_for(arrays)
    g = O.Graph(n, *arrays)
    for B, a in list(assignments(n, rng))[:-1] + [(12, truth)]:
        if B < 2:
            continue
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        v, s, u = _random_proposals(n, B, a, rng, 400)
        dS, H, acc = c.eval_moves(v, s, u)
        odS, oH, oacc = bm.eval_moves(v, s, u)
        assert np.allclose(dS, odS, rtol=DS_RTOL, atol=1e-9), np.max(np.abs(dS - odS))
        assert np.allclose(H, oH, rtol=1e-12)
        assert (acc == oacc).sum() >= len(acc) - 1  # a decision may flip only at u == p to 1e-15


def test_proposals_match_oracle_sampler():
    rng = np.random.default_rng(2)
    n = 1500
    arrays, truth = gen_arrays(n, 10, seed=4)
    c = ctx_for(arrays)
    g = O.Graph(n, *arrays)
    for B, a in list(assignments(n, rng))[:-1] + [(10, truth)]:
        if B < 2:
            continue
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        v = np.arange(n, dtype=np.int32)
        for tag in (0, 1):
            s, u = c.propose(v, seed=99, call=7, tag=tag)
            os_, ou = bm.propose(v, seed=99, call=7, tag=tag)
            assert (s == os_).all()
            assert (u == ou).all()


def test_apply_moves_matches_sequential_application_with_degree_drift():
    rng = np.random.default_rng(3)
    n = 1200
    arrays, truth = gen_arrays(n, 9, seed=5)
    c = ctx_for(arrays)
    g = O.Graph(n, *arrays)
    for B, a in list(assignments(n, rng))[:-1] + [(9, truth)]:
        if B < 2:
            continue
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        v, s, _ = _random_proposals(n, B, a, rng, 300)
        c.apply_moves(v, s)
        order = np.argsort(v)
        bm.apply_moves(v[order], s[order])
        assert_same_state(c, bm)
        # a second batch on top keeps the earlier drift where untouched
        v2, s2, _ = _random_proposals(n, B, bm.state()[0], rng, 50)
        c.apply_moves(v2, s2)
        o2 = np.argsort(v2)
        bm.apply_moves(v2[o2], s2[o2])
        assert_same_state(c, bm)


# ----------------------------------------------------------------------------- sweeps
@pytest.mark.parametrize("batches,variant", [(1, 0), (3, 0), (10, 0), (1, 3), (3, 3)])
def test_async_gibbs_sweeps_match_oracle(batches, variant):
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=0)
    c = ctx_for(arrays)
    c.set_option(3, variant)
    g = O.Graph(n, *arrays)
    rng = np.random.default_rng(4)
    for B, a in ((n, np.arange(n, dtype=np.int32)), (64, rng.integers(0, 64, n).astype(np.int32)), (8, truth)):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        for call in range(3):
            st = c.sweep(sbpx().ALGO_ASYNC_GIBBS, batches, 1234, call)
            ost = bm.async_sweep(batches, 1234, call)
            assert (st.attempts, st.evals, st.moves) == (ost.attempts, ost.evals, ost.moves)
            assert_same_state(c, bm)
            assert c.mdl() == pytest.approx(bm.mdl(), rel=MDL_RTOL)


@pytest.mark.parametrize("batches", [1, 4])
def test_incremental_update_equals_full_rebuild(batches):
    """The block matrix updated from the moved vertices' edges after each batch equals the one a
    full sort rebuilds from the assignment: every CSR/CSC array, the degrees and the MDL, through
    sweeps from many moves (full rebuilds) down to few (updates), across merges, at block counts
    that are powers of two (the largest key then fills the sorted bit range: the padding of the
    change list shares it) and at one that is not."""
    m = sbpx()
    n = 3000
    arrays, truth = gen_arrays(n, 12, p_in=0.7, alpha=1.0, seed=6)
    cs = []
    for inc in (0, 1):
        c = ctx_for(arrays)
        c.set_option(m.OPT_INCREMENTAL, inc)
        c.build(n, np.arange(n, dtype=np.int32))
        c.merge_blocks(n - 1024, 4, 0)
        cs.append(c)
    for call in range(18):
        if call in (6, 12):
            for c in cs:
                B = c.info()[0]
                c.merge_blocks(B - (16 if call == 6 else 13), 4, call)
        st = [c.sweep(m.ALGO_ASYNC_GIBBS, batches, 17, call) for c in cs]
        assert (st[0].moves, st[0].evals) == (st[1].moves, st[1].evals)
        s0, s1 = cs[0].state(), cs[1].state()
        for a, b in zip(s0, s1):
            assert (np.asarray(a) == np.asarray(b)).all()
        for by_column in (0, 1):
            for a, b in zip(cs[0].coo(by_column), cs[1].coo(by_column)):
                assert (a == b).all()
        assert cs[0].info() == cs[1].info()
        assert cs[0].mdl() == cs[1].mdl()


# ----------------------------------------------------------------------------- merges
def test_merge_proposals_match_oracle():
    rng = np.random.default_rng(5)
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=1)
    c = ctx_for(arrays)
    g = O.Graph(n, *arrays)
    for B, a in ((n, np.arange(n, dtype=np.int32)), (200, rng.integers(0, 200, n).astype(np.int32)),
                 (16, rng.integers(0, 16, n).astype(np.int32))):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        dS, best = c.merge_proposals(77, 3)
        odS, obest = bm.merge_proposals(77, 3)
        assert (best == obest).all()
        assert np.allclose(dS, odS, rtol=DS_RTOL, atol=1e-9)


@pytest.mark.parametrize("approximate,headroom", [(True, 0), (False, 0), (False, 48)])
def test_merge_blocks_matches_oracle(approximate, headroom):
    """Merge phases from the identity partition: proposals, then the sorted greedy selection
    (--approximate) or the heap-ordered selection with re-scoring on the current matrix. With a
    table pool of `headroom` spare slots the heap selection runs out of room repeatedly and
    resumes with the merge it had accepted."""
    m = sbpx()
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=2)
    c = ctx_for(arrays)
    c.set_option(m.OPT_MH_HEADROOM, headroom)
    g = O.Graph(n, *arrays)
    a = np.arange(n, dtype=np.int32)
    c.build(n, a)
    bm = O.Blockmodel(g, n, a)
    B = n
    mode = m.MERGE_APPROXIMATE if approximate else m.MERGE_ADVANCED
    for call in range(4):
        k = B // 2
        c.merge_blocks(k, 5, call, mode=mode)
        bm.num_to_merge = k
        bm.merge_blocks(approximate, 5, call)
        B -= k
        assert c.info()[0] == bm.B == B
        assert_same_state(c, bm)


def test_heap_merge_selection_with_tied_scores():
    """Heap-ordered selection from fixed scores with many exact ties (scores rounded to whole
    numbers): the heap must pop tied entries in the order std::priority_queue does."""
    m = sbpx()
    n = 1500
    arrays, truth = gen_arrays(n, 10, seed=8)
    g = O.Graph(n, *arrays)
    rng = np.random.default_rng(9)
    for B in (n, 300):
        a = np.arange(n, dtype=np.int32) if B == n else rng.integers(0, B, n).astype(np.int32)
        c = ctx_for(arrays)
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        dS, best = bm.merge_proposals(3, 1)
        dS = np.round(dS)
        k = B // 3
        c.carry_out_merges(k, dS, best, mode=m.MERGE_ADVANCED)
        bm.num_to_merge = k
        bm.carry_out(dS, best, advanced=True)
        assert c.info()[0] == bm.B
        assert_same_state(c, bm)


# ----------------------------------------------------------------------------- driver
def test_sbp_async_gibbs_end_to_end_matches_oracle():
    m = sbpx()
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=0)
    g = O.Graph(n, *arrays)
    a, st = m.run_sbp(*arrays, algorithm="async_gibbs", batches=1, approximate=True, seed=11)
    oa, ost = O.sbp(g, 1, 1, True, 11)
    assert st.num_blocks == ost.B
    assert st.mdl == pytest.approx(ost.mdl, rel=5e-3)
    print("device", st.num_blocks, st.mdl, "oracle", ost.B, ost.mdl, "identical:", bool((a == oa).all()))


# ----------------------------------------------------------------------------- sequential MH
@pytest.mark.parametrize("algo,group,wide", [("hybrid", 0, 0), ("mh", 0, 0), ("mh", 1, 1), ("mh", 2, 1), ("hybrid", 4, 1),
                                             ("mh", 8, 1), ("mh", 16, 0), ("mh", 2, 0), ("hybrid", 4, 0), ("mh", 8, 0),
                                             ("mh", 1, 2), ("mh", 1, 4), ("mh", 1, 16), ("mh", 2, 64), ("mh", 4, 32),
                                             ("hybrid", 8, 128), ("mh", 8, 512), ("mh", 1, 32), ("hybrid", 1, 8),
                                             ("mh", 1, 64), ("mh", 1, 256), ("hybrid", 1, 128), ("mh", 1, 512)])
def test_sequential_sweeps_match_oracle(algo, group, wide):
    """Sequential Metropolis-Hastings (alone, or on the high-degree vertices of the hybrid) equals
    the oracle's one-vertex-at-a-time pass for every number of wavefronts per vertex and every
    number of workgroups sharing a window: wide = 1 runs the one-workgroup kernels (16 / group
    entries per window); otherwise group <= 8 wavefronts per entry run the window-check kernel
    on wide workgroups of 8 / group entries each (wide = 0: as many as the window yield calls
    for; up to 4096 entries, which covers the whole queue here), and group = 16 one workgroup."""
    sequential_case(algo, group, wide)


@pytest.mark.parametrize("algo,group,wide", [("mh", 1, 1), ("mh", 1, 8), ("mh", 16, 1), ("hybrid", 0, 0)])
def test_sequential_sweeps_resume_after_table_pool_runs_out(algo, group, wide):
    """With a table pool of 48 spare slots the pass runs out of room over and over; each time the
    host rebuilds the tables from the moves so far with twice the headroom and resumes at the
    same queue entry. The pass still equals the oracle's."""
    launches = sequential_case(algo, group, wide, headroom=48)
    assert max(launches) > 2, "the pool never ran out"


@pytest.mark.parametrize("algo,wide,team", [("mh", 0, 2), ("mh", 4, 4), ("hybrid", 64, 8), ("mh", 0, 32),
                                             ("hybrid", 0, 16), ("mh", 512, 2)])
def test_sequential_sweeps_with_workgroup_teams_match_oracle(algo, wide, team):
    """Teams of 2 to 32 workgroups sharing each window entry (each scans a slice of the proposal's
    row and of the delta-MDL rows, member 0 reads the neighbour blocks' cells by lookups and adds
    the slices' sums) give the oracle's pass: the same moves and state as one vertex at a time."""
    sequential_case(algo, 8, wide, team=team)


def test_sequential_sweeps_with_teams_resume_after_table_pool_runs_out():
    """A team pass that runs out of table room resumes at the same queue entry, as the others do."""
    launches = sequential_case("mh", 8, 16, headroom=48, team=4)
    assert max(launches) > 2, "the pool never ran out"


def sequential_case(algo, group, wide, headroom=0, team=0):
    m = sbpx()
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=6)
    c = ctx_for(arrays)
    c.set_option(m.OPT_MH_GROUP, group)
    c.set_option(m.OPT_MH_WIDE, wide)
    c.set_option(m.OPT_MH_TEAM, team)
    launches = []
    g = O.Graph(n, *arrays)
    rng = np.random.default_rng(7)
    code = m.ALGO_HYBRID if algo == "hybrid" else m.ALGO_MH
    # a giant block 0 next to ~600 small ones: its row and column hold more than 256 slots, so
    # proposals drawn from it take the multi-pass radix selection
    giant = rng.integers(1, 600, n).astype(np.int32)
    giant[rng.random(n) < 0.4] = 0
    for B, a in ((n, np.arange(n, dtype=np.int32)), (128, rng.integers(0, 128, n).astype(np.int32)), (8, truth),
                 (600, giant)):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        for call in range(3):
            c.set_option(m.OPT_MH_HEADROOM, headroom)  # also restarts the headroom doubling
            st = c.sweep(code, 2, 4321, call)
            launches.append(st.eval_launches[3])
            if algo == "hybrid":
                o1 = bm.mh_pass(4321, call)
                o2 = bm.async_sweep(2, 4321, call, hybrid_low=True)
                ost = (o1.attempts + o2.attempts, o1.evals + o2.evals, o1.moves + o2.moves)
            else:
                o1 = bm.mh_pass(4321, call, order=np.arange(n, dtype=np.int32))
                ost = (o1.attempts, o1.evals, o1.moves)
                assert st.dS_accepted == pytest.approx(o1.dS_sum, rel=1e-9, abs=1e-9)
            assert (st.attempts, st.evals, st.moves) == ost
            assert_same_state(c, bm)
    return launches


@pytest.mark.parametrize("approximate", [True, False])
def test_sbp_hybrid_end_to_end_close_to_oracle(approximate):
    m = sbpx()
    n = 1000
    arrays, truth = gen_arrays(n, 8, seed=0)
    g = O.Graph(n, *arrays)
    a, st = m.run_sbp(*arrays, algorithm="hybrid_mcmc", batches=1, approximate=approximate, seed=5)
    oa, ost = O.sbp(g, 2, 1, approximate, 5)
    assert st.num_blocks == ost.B
    assert st.mdl == pytest.approx(ost.mdl, rel=5e-3)
    print("device", st.num_blocks, st.mdl, "oracle", ost.B, ost.mdl, "identical:", bool((a == oa).all()))


# ----------------------------------------------------------------------------- all degree bins
def heavy_arrays(seed=0):
    """Graph whose vertices fall in every degree bin: two hubs with in-degree > 2047, ten with
    degree in (510, 2047], a few hundred in (126, 510], and a thousand low-degree vertices
    (mostly <= 30) that the hubs and mids do not reach."""
    rng = np.random.default_rng(seed)
    n = 3000
    edges = set()
    for v in range(n):
        for w in rng.integers(0, n, 6 if v < 2000 else 2):
            if w != v:
                edges.add((v, int(w)))
    for hub in (0, 1):
        for v in rng.permutation(n)[:2300]:
            if v != hub:
                edges.add((int(v), hub))
    for mid in range(2, 300):
        for v in rng.permutation(2000)[: int(rng.integers(130, 400))]:
            if v != mid:
                edges.add((mid, int(v)) if v % 2 else (int(v), mid))
    for big in range(300, 310):
        for v in rng.permutation(2000)[: int(rng.integers(600, 1500))]:
            if v != big:
                edges.add((big, int(v)) if v % 3 else (int(v), big))
    for v in (5, 17):
        edges.add((v, v))  # self-loops on binned vertices
    return O.csr_from_edges(n, sorted(edges))


@pytest.mark.parametrize("variant", [1, 2, 3, 0])
def test_every_degree_bin_matches_oracle(variant):
    """Every degree bin and kernel variant (one wavefront per vertex with 64..1024-slot LDS
    tables, more wavefronts per SIMD, 16 lanes per vertex for degrees <= 30, 4 or 16 wavefronts
    sharing a vertex, tables in global memory) against the oracle, through proposals,
    fixed-proposal evaluation and sweeps."""
    m = sbpx()
    arrays = heavy_arrays()
    n = len(arrays[0]) - 1
    deg = np.diff(arrays[0]) + np.diff(arrays[2])
    assert (deg > 2047).sum() >= 2 and ((deg > 510) & (deg <= 2047)).sum() >= 5
    assert ((deg > 126) & (deg <= 510)).sum() >= 50 and (deg <= 30).sum() >= 100
    c = ctx_for(arrays)
    c.set_option(3, variant)
    g = O.Graph(n, *arrays)
    rng = np.random.default_rng(8)
    for B, a in ((n, np.arange(n, dtype=np.int32)), (400, rng.integers(0, 400, n).astype(np.int32)),
                 (20, rng.integers(0, 20, n).astype(np.int32))):
        c.build(B, a)
        bm = O.Blockmodel(g, B, a)
        v = np.arange(n, dtype=np.int32)
        s, u = c.propose(v, seed=3, call=1)
        os_, ou = bm.propose(v, seed=3, call=1)
        assert (s == os_).all() and (u == ou).all()
        vv, ss, uu = _random_proposals(n, B, a, rng, 600)
        hubs = np.array([0, 1, 2, 3, 5, 17], np.int32)
        vv = np.concatenate([hubs, vv[~np.isin(vv, hubs)]])[:600].astype(np.int32)
        ss = rng.integers(0, B, len(vv)).astype(np.int32)
        ss = np.where(ss == a[vv], (ss + 1) % B, ss).astype(np.int32)
        uu = rng.random(len(vv))
        dS, H, acc = c.eval_moves(vv, ss, uu)
        odS, oH, oacc = bm.eval_moves(vv, ss, uu)
        assert np.allclose(dS, odS, rtol=DS_RTOL, atol=1e-9), np.max(np.abs(dS - odS))
        assert np.allclose(H, oH, rtol=1e-12)
        for call in range(2):
            st = c.sweep(m.ALGO_ASYNC_GIBBS, 1, 9, call)
            ost = bm.async_sweep(1, 9, call)
            assert (st.attempts, st.evals, st.moves) == (ost.attempts, ost.evals, ost.moves)
            assert_same_state(c, bm)
        st = c.sweep(m.ALGO_HYBRID, 1, 9, 5)
        o1 = bm.mh_pass(9, 5)
        o2 = bm.async_sweep(1, 9, 5, hybrid_low=True)
        assert (st.attempts, st.evals, st.moves) == (o1.attempts + o2.attempts, o1.evals + o2.evals,
                                                    o1.moves + o2.moves)
        assert_same_state(c, bm)
