"""The learned mixture's Hastings ratio on the device against an independent calculation.

* The device's candidates and scores (batched path) against the oracle's candidates and the
  network run on the oracle's features; the sequential pass's workgroup scorer against the batched
  path, bit for bit.
* For fixed moves, some to a candidate (as the learned branch draws them) and some drawn by the
  plain proposal: the device's net probabilities q_net(s), q_net(r), its Hastings ratio and delta
  MDL against the oracle's mixture ratio with the same scores (plain terms: the plain sampler's
  own probabilities, pinned on the CPU) and the oracle's delta MDL; accept decisions against u <= min(1, exp(-3 dS) H).
* Detailed balance: H(x -> x') H(x' -> x) = 1 on the device, the reverse evaluated from scratch in
  the state after the move, for self-loop-free vertices.
"""
import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx
from test_gpu_learned import _graph, _merge_to

pytestmark = pytest.mark.gpu

W_PLAIN = 0.2


def _setup(n=3000, seed=11, B0=400):
    m = sbpx()
    arrays = _graph(n, seed)
    c = m.Context(0)
    c.upload_graph(*arrays)
    c.split()
    c.learned_init(5)
    _merge_to(c, n, B0, 3)
    g = O.Graph(n, *arrays)
    a = c.state()[0].copy()
    B = c.info()[0]
    bm = O.Blockmodel(g, B, a)
    deg = np.diff(arrays[0]) + np.diff(arrays[2])
    self_loop = np.zeros(n, bool)
    for v in range(n):
        self_loop[v] = v in arrays[1][arrays[0][v]:arrays[0][v + 1]]
    return m, c, g, bm, a, B, deg, self_loop


def test_scores_match_oracle_features_and_both_device_scorers():
    m, c, g, bm, a, B, deg, _ = _setup()
    rng = np.random.default_rng(0)
    v = rng.choice(np.flatnonzero(deg <= 126), 400, replace=False).astype(np.int32)
    cand0, z0 = c.learned_scores(v, scorer=0)
    cand1, z1 = c.learned_scores(v, scorer=1)
    ocand = np.zeros((len(v), 32), np.int32)
    ox = np.zeros((len(v), 64), np.float32)
    for i, x in enumerate(v):
        ocand[i], _, ox[i] = bm.learned_features(int(x))
    assert (cand0 == ocand).all()
    assert (cand1 == ocand).all()
    zo = c.learned_forward(ox)  # the network on the oracle's features
    scale = max(1.0, float(np.abs(zo).max()))
    # features computed with FP32 log1p on both sides; BF16 rounding of the inputs can flip in a
    # rare last bit, so the scores agree to BF16 precision
    assert np.abs(z0 - zo).max() < 2e-2 * scale and np.abs(z0 - zo).mean() < 1e-3 * scale
    # the sequential pass scores its entries with the batched path's matrix-core products
    assert (z1 == z0).all()


def test_learned_hastings_matches_oracle_mixture():
    m, c, g, bm, a, B, deg, self_loop = _setup()
    rng = np.random.default_rng(1)
    v = rng.choice(np.flatnonzero(deg <= 126), 600, replace=False).astype(np.int32)
    cand, z = c.learned_scores(v, scorer=0)
    s = np.full(len(v), -1, np.int32)
    plain_s = [c.propose(v, 77, call)[0] for call in range(4)]
    for i, x in enumerate(v):
        others = [b for b in cand[i] if b >= 0 and b != a[x]]
        drawn = [int(p[i]) for p in plain_s if p[i] != a[x]]
        if i % 2 == 0 and others:  # a move the learned branch can draw
            s[i] = rng.choice(others)
        elif drawn:  # a move the plain proposal drew
            s[i] = drawn[0]
        elif others:
            s[i] = rng.choice(others)
    keep = s >= 0
    v, s, cand, z = v[keep], s[keep], cand[keep], z[keep]
    u = rng.random(len(v))
    dS, H, acc, qf, qb = c.learned_eval_moves(v, s, u)
    odS, _, _ = bm.eval_moves(v, s, u)
    assert np.allclose(dS, odS, rtol=1e-6, atol=1e-9)
    learned_moves = 0
    caught = caught_smoothed = 0
    for i, x in enumerate(v):
        Ho, qfo, qbo = bm.mixture_hastings(int(x), int(s[i]), z[i], cand[i], w=W_PLAIN)
        assert qf[i] == pytest.approx(qfo, rel=1e-5, abs=1e-7), i
        assert qb[i] == pytest.approx(qbo, rel=1e-5, abs=1e-7), i
        assert H[i] == pytest.approx(Ho, rel=1e-5), i
        if qfo > 0:
            learned_moves += 1
            # plausible mistakes, the reverse net term taken from the forward softmax or the plain
            # part taken from the reference's smoothed closed form, would move H by more than the
            # tolerance above on many of these moves
            pf0, pb0 = bm.plain_probabilities(int(x), int(s[i]))
            wrong = (W_PLAIN * pb0 + (1 - W_PLAIN) * qfo) / (W_PLAIN * pf0 + (1 - W_PLAIN) * qfo)
            k = int(deg[x] - self_loop[x])
            pf, pb = bm.hastings_terms(int(x), int(s[i]))
            smoothed = (W_PLAIN * pb / k + (1 - W_PLAIN) * qbo) / (W_PLAIN * pf / k + (1 - W_PLAIN) * qfo)
            if Ho > 0:  # H = 0: the reverse proposal cannot draw r, the move is never accepted
                caught += abs(wrong / Ho - 1.0) > 1e-3
                caught_smoothed += abs(smoothed / Ho - 1.0) > 1e-3
        a_i = min(1.0, np.exp(-3.0 * dS[i]) * H[i])
        assert bool(acc[i]) == (u[i] <= a_i)
    assert learned_moves > 100
    assert caught > 10 and caught_smoothed > 10


def test_learned_hastings_reverses_on_the_device():
    m, c, g, bm, a, B, deg, self_loop = _setup(n=2000, seed=13, B0=150)
    rng = np.random.default_rng(2)
    pool = np.flatnonzero((deg <= 126) & ~self_loop)
    v = rng.choice(pool, 60, replace=False).astype(np.int32)
    cand, _ = c.learned_scores(v, scorer=0)
    checked = impossible = 0
    for i, x in enumerate(v):
        x = int(x)
        r = int(a[x])
        others = [b for b in cand[i] if b >= 0 and b != r]
        if not others:
            continue
        s = int(rng.choice(others))
        _, H1, _, _, _ = c.learned_eval_moves([x], [s], [0.5])
        if H1[0] == 0:
            # neither part of the mixture can propose r from the moved state: the oracle agrees
            # that r is no candidate and the plain sampler's reverse probability is zero
            assert r not in cand[i] and bm.plain_probabilities(x, s)[1] == 0.0
            impossible += 1
            continue
        a2 = a.copy()
        a2[x] = s
        c.build(B, a2)
        _, H2, _, _, _ = c.learned_eval_moves([x], [r], [0.5])
        c.build(B, a)
        assert H1[0] * H2[0] == pytest.approx(1.0, rel=1e-4), (x, r, s, H1, H2)
        checked += 1
    assert checked > 30
