"""GPU tests of the learned proposal: the network's forward pass and gradients against an FP32
reference (PyTorch on the CPU, with the device's BF16 roundings applied to inputs, weights and
hidden activations), the state after learned sweeps against a rebuild by the oracle, and what a
trained network gains over the plain proposal."""
import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bf16(a):
    """Round float32 to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    shapes = [(256, 64), (256,), (256, 256), (256,), (32, 256), (32,)]
    out, off = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(params[off:off + n].reshape(sh))
        off += n
    return out


def reference_forward(x, params, requires_grad=False):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(bf16(p) if p.ndim == 2 else p, requires_grad=requires_grad)
                              for p in unpack(params)]
    X = torch.tensor(bf16(x))
    h1 = torch.relu(X @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    return (W1, b1, W2, b2, W3, b3), h1, h2, h2 @ W3.T + b3


def test_forward_matches_fp32_reference():
    c = sbpx().Context(0)
    c.learned_init(3)
    params = c.learned_params()
    x = np.random.default_rng(0).random((300, 64)).astype(np.float32)
    z = c.learned_forward(x)
    # reference with the device's BF16 rounding of the hidden activations
    W1, b1, W2, b2, W3, b3 = [bf16(p) if p.ndim == 2 else p for p in unpack(params)]
    X = bf16(x).astype(np.float64)
    h1 = bf16(np.maximum(X @ W1.T + b1, 0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ W2.T + b2, 0).astype(np.float32)).astype(np.float64)
    ref = h2 @ W3.T + b3
    assert np.abs(z - ref).max() < 2e-2 * max(1.0, np.abs(ref).max())
    assert np.abs(z - ref).mean() < 1e-3 * max(1.0, np.abs(ref).mean())


def test_gradients_match_fp32_reference():
    c = sbpx().Context(0)
    c.learned_init(4)
    params = c.learned_params()
    rng = np.random.default_rng(1)
    n = 512
    x = rng.random((n, 64)).astype(np.float32)
    mask = rng.integers(1, 1 << 16, n).astype(np.uint32)
    y = np.array([rng.choice(np.flatnonzero([(m >> j) & 1 for j in range(16)])) for m in mask], np.int32)
    loss, grads = c.learned_gradients(x, mask, y)
    P, _, _, z = reference_forward(x, params, requires_grad=True)
    bits = torch.tensor([[(int(m) >> j) & 1 for j in range(32)] for m in mask], dtype=torch.bool)
    zm = z.masked_fill(~bits, float("-inf"))
    ref_loss = torch.nn.functional.cross_entropy(zm, torch.tensor(y, dtype=torch.long), reduction="none")
    ref_loss.mean().backward()
    assert np.allclose(loss, ref_loss.detach().numpy(), rtol=2e-2, atol=2e-2)
    got = unpack(grads)
    for g, p in zip(got, P):
        r = p.grad.numpy().reshape(g.shape)
        cos = float((g * r).sum() / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-30))
        assert cos > 0.995, cos
        assert abs(np.linalg.norm(g) / np.linalg.norm(r) - 1.0) < 0.03


def _graph(n, seed, K=20):
    m = sbpx()
    cfg = dict(m.CONFIGS["C2"])
    cfg.update(num_vertices=n, num_blocks=K)
    op, oi, ip, ii, truth = m.generate_dcsbm(**cfg, seed=seed)
    return (op, oi, ip, ii)


def _merge_to(c, n, B0, seed):
    c.build(n, np.arange(n, dtype=np.int32))
    B, call = n, 0
    while B > B0:
        k = min(B // 2, B - B0)
        c.merge_blocks(k, seed, call)
        B -= k
        call += 1


def test_learned_sweeps_keep_the_blockmodel_consistent():
    m = sbpx()
    n = 3000
    arrays = _graph(n, 11)
    c = m.Context(0)
    c.upload_graph(*arrays)
    c.split()
    c.learned_init(5)
    c.set_option(m.OPT_PROPOSAL, 1)
    g = O.Graph(n, *arrays)
    _merge_to(c, n, 400, 3)
    moves = 0
    for call in range(4):
        st = c.sweep(m.ALGO_ASYNC_GIBBS, 1, 21, call)
        moves += st.moves
        a = c.state()[0]
        bm = O.Blockmodel(g, c.info()[0], a)
        from test_gpu_parity import assert_same_state
        assert_same_state(c, bm, check_deg=False)
        assert np.isfinite(c.mdl())
    assert moves > 0
    # the hybrid keeps its sequential part plain and runs the learned mixture on the rest
    st = c.sweep(m.ALGO_HYBRID, 1, 21, 9)
    assert st.attempts == n


def test_trained_network_gains_over_plain_proposal():
    """Trained on accepted moves logged from plain sweeps on two graphs, the mixture reaches a lower
    MDL than the plain proposal after the same number of sweeps from the same state of a held-out
    graph of the same family."""
    m = sbpx()
    n = 8000
    c = m.Context(0)
    c.learned_init(1)
    c.learned_log(1_000_000)
    for seed in (1, 2):
        c.upload_graph(*_graph(n, seed))
        c.split()
        B, call = n, 0
        c.build(n, np.arange(n, dtype=np.int32))
        while B > 40:
            c.merge_blocks(B // 2, seed, call)
            B -= B // 2
            for i in range(8):
                c.sweep(m.ALGO_ASYNC_GIBBS, 1, seed, 100 * call + i)
            call += 1
    c.learned_log(0)
    assert c.learned_log_count() > 10000
    first = c.learned_train(50, 4096, 1e-3, 7)
    last = c.learned_train(400, 4096, 1e-3, 8)
    assert last < first < np.log(32.0)
    c.upload_graph(*_graph(n, 100))
    c.split()
    for B0 in (n // 8, 80):
        final = {}
        for proposal in (0, 1):
            c.set_option(m.OPT_PROPOSAL, proposal)
            _merge_to(c, n, B0, 5)
            for i in range(10):
                c.sweep(m.ALGO_ASYNC_GIBBS, 1, 5, 500 + i)
            final[proposal] = c.mdl()
        print("B", B0, "plain", final[0], "learned", final[1])
        assert final[1] < final[0]


def test_training_is_reproducible():
    """Logged rows keep the order of the vertices they come from, and minibatches are drawn from a
    seeded stream, so two contexts that log and train the same way end with the same network."""
    m = sbpx()
    n = 3000
    params = []
    for _ in range(2):
        c = m.Context(0)
        c.learned_init(2)
        c.learned_log(200_000)
        c.upload_graph(*_graph(n, 3))
        c.split()
        c.build(n, np.arange(n, dtype=np.int32))
        B = n
        for call in range(3):
            c.merge_blocks(B // 2, 3, call)
            B -= B // 2
            for i in range(4):
                c.sweep(m.ALGO_ASYNC_GIBBS, 1, 3, 10 * call + i)
        c.learned_log(0)
        count = c.learned_log_count()
        c.learned_train(30, 1024, 1e-3, 9)
        params.append((count, c.learned_params()))
        c.close()
    assert params[0][0] == params[1][0] > 200
    assert np.array_equal(params[0][1], params[1][1])


@pytest.mark.parametrize("algo,group,maxdeg", [("hybrid", 0, 126), ("mh", 0, 126), ("mh", 4, 126), ("mh", 8, 126),
                                              ("hybrid", 16, 126), ("hybrid", 0, 510), ("mh", 0, 510)])
def test_learned_sequential_pass_keeps_the_blockmodel_consistent(algo, group, maxdeg):
    """Proposal 2: the sequential pass proposes from the mixture too, each evaluation scoring its
    vertex's candidates on the current state (one wavefront per vertex, or the first wavefront of
    a group), for vertices up to degree 126 or 510. Every move is still judged by delta MDL and
    applied as in the plain pass; the state matches a rebuild from the assignment."""
    m = sbpx()
    n = 3000
    arrays = _graph(n, 12)
    c = m.Context(0)
    c.upload_graph(*arrays)
    c.split()
    c.learned_init(6)
    g = O.Graph(n, *arrays)
    _merge_to(c, n, 300, 4)
    c.set_option(m.OPT_PROPOSAL, 2)
    c.set_option(m.OPT_MH_GROUP, group)
    c.set_option(m.OPT_LEARN_MAXDEG, maxdeg)
    code = m.ALGO_HYBRID if algo == "hybrid" else m.ALGO_MH
    moves = 0
    from test_gpu_parity import assert_same_state
    for call in range(3):
        st = c.sweep(code, 1, 31, call)
        assert st.attempts == n
        moves += st.moves
        bm = O.Blockmodel(g, c.info()[0], c.state()[0])
        assert_same_state(c, bm, check_deg=False)
        assert np.isfinite(c.mdl())
    assert moves > 0
    with pytest.raises(m.SbpxError):  # the learned sequential pass needs the window-check kernel
        c.set_option(m.OPT_MH_WIDE, 1)
        c.sweep(code, 1, 31, 9)


@pytest.mark.parametrize("option,values", [("group", (1, 2, 4, 8, 16)), ("wide", (0, 2, 8, 512)),
                                           ("team", (1, 2, 8, 32))])
def test_learned_sequential_pass_same_moves_however_evaluated(option, values):
    """The learned sequential pass gives the same chain whether one wavefront or a group of 2 to 16
    evaluates each vertex, whatever number of workgroups shares a window of speculative
    evaluations, and whether teams of 2 to 32 workgroups share each entry: the proposals are keyed by vertex and pass, the scores come from the state the
    vertex's step starts from (the window checks cover what the scorer reads), and delta MDL is equal
    up to the order of its FP64 sums, so every accept decision and the final state agree."""
    m = sbpx()
    n = 3000
    arrays = _graph(n, 12)
    states = []
    for value in values:
        c = m.Context(0)
        c.upload_graph(*arrays)
        c.split()
        c.learned_init(6)
        _merge_to(c, n, 300, 4)
        c.set_option(m.OPT_PROPOSAL, 2)
        if option == "group":
            c.set_option(m.OPT_MH_GROUP, value)
        elif option == "team":
            c.set_option(m.OPT_MH_GROUP, 8)
            c.set_option(m.OPT_MH_TEAM, value)
        else:
            c.set_option(m.OPT_MH_GROUP, 1)
            c.set_option(m.OPT_MH_WIDE, value)
        moves = [c.sweep(m.ALGO_MH, 1, 31, call).moves for call in range(2)]
        a, dout, din, _ = c.state()
        states.append((moves, a.copy(), dout.copy(), din.copy(), c.mdl()))
        c.close()
    assert states[0][0][0] > 0
    for other in states[1:]:
        assert other[0] == states[0][0]
        assert all(np.array_equal(x, y) for x, y in zip(other[1:4], states[0][1:4]))
        assert other[4] == pytest.approx(states[0][4], rel=1e-9)


def _high_degree_graph(n, seed):
    """A DC-SBM graph whose vertices mostly fall in the degree bins 127-254 and 255-510."""
    m = sbpx()
    op, oi, ip, ii, _ = m.generate_dcsbm(n, 12, 0.6, 1.0, 60, 300, 1.5, seed=seed)
    return (op, oi, ip, ii)


@pytest.mark.parametrize("maxdeg", [254, 510])
def test_learned_batched_bins_above_126_keep_the_same_chain_in_both_kernels(maxdeg):
    """With the degree bound at 254 or 510 the batched sweeps give the vertices of degree 127-254
    (and 255-510) the learned mixture too. Both kernel variants of those bins (one wavefront per
    vertex; four wavefronts sharing a vertex) take the same draws and the same net probabilities,
    so they accept the same moves; the state matches a rebuild from the assignment; and the
    learned run differs from the bound-126 run, which keeps those bins plain."""
    m = sbpx()
    n = 3000
    arrays = _high_degree_graph(n, 13)
    deg = np.diff(arrays[0]) + np.diff(arrays[2])
    assert ((deg > 126) & (deg <= 254)).sum() > 500 and ((deg > 254) & (deg <= 510)).sum() > 100
    g = O.Graph(n, *arrays)
    from test_gpu_parity import assert_same_state
    runs = []
    for bound, variant in ((maxdeg, 1), (maxdeg, 2), (126, 1)):
        c = m.Context(0)
        c.upload_graph(*arrays)
        c.split()
        c.learned_init(5)
        c.set_option(m.OPT_PROPOSAL, 1)
        c.set_option(m.OPT_LEARN_MAXDEG, bound)
        c.set_option(m.OPT_EVAL_VARIANT, variant)
        _merge_to(c, n, 200, 3)
        moves = [c.sweep(m.ALGO_ASYNC_GIBBS, 1, 21, call).moves for call in range(3)]
        a = c.state()[0].copy()
        bm = O.Blockmodel(g, c.info()[0], a)
        assert_same_state(c, bm, check_deg=False)
        runs.append((moves, a, c.mdl()))
        c.close()
    assert sum(runs[0][0]) > 0
    assert runs[1][0] == runs[0][0] and np.array_equal(runs[1][1], runs[0][1])
    assert runs[1][2] == pytest.approx(runs[0][2], rel=1e-9)
    assert not np.array_equal(runs[2][1], runs[0][1])
