"""The learned scorer on the GPU: BF16 MFMA forward/backward against a CPU
FP32 reference (numpy), training, labelled samples, and the learned beam."""
import numpy as np
import pytest

import synchrowords_amd as sw
from conftest import golden
from synchrowords_amd import gen, replay

pytestmark = pytest.mark.gpu


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def unpack(p, D, H):
    o = 0
    w1 = p[o:o + H * D].reshape(H, D); o += H * D
    b1 = p[o:o + H]; o += H
    w2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o:o + 1]
    return w1, b1, w2, b2, w3, b3


def rank_loss(y, t, G):
    """Pairwise ranking loss within consecutive groups of G rows and its dy."""
    B = len(y)
    groups = B // G
    inv = 1.0 / (groups * (G * (G - 1) // 2))
    dy = np.zeros(B)
    loss = 0.0
    for g in range(groups):
        for a in range(g * G, g * G + G):
            for b in range(g * G, g * G + G):
                if t[a] < t[b]:
                    z = float(y[a]) - float(y[b])
                    loss += np.logaddexp(0.0, z) * inv
                    s = 1.0 / (1.0 + np.exp(-z))
                    dy[a] += s * inv
                    dy[b] -= s * inv
    return loss, dy


def reference(p, x_bits, t, D, H, rank_group=0):
    """FP32 forward/backward with the GPU's BF16 rounding points."""
    w1, b1, w2, b2, w3, b3 = unpack(p, D, H)
    x = (x_bits.astype(np.uint32) << 16).view(np.float32)
    a1 = x @ bf16(w1).T + b1
    h1 = np.maximum(a1, 0)
    h1b = bf16(h1)
    a2 = h1b @ bf16(w2).T + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ w3 + b3[0]
    B = x.shape[0]
    if rank_group:
        loss, dy = rank_loss(y, t, rank_group)
        dy = dy.astype(np.float32)
    else:
        dy = 2 * (y - t) / B
        loss = float(np.sum((y - t) ** 2) / B)
    gw3 = h2.T @ dy
    gb3 = np.array([dy.sum()], dtype=np.float32)
    dh2 = bf16(np.outer(dy, w3) * (h2 > 0))
    gw2 = dh2.T @ h1b
    gb2 = dh2.sum(axis=0)
    dh1 = bf16(dh2 @ bf16(w2)) * (h1 > 0)
    gw1 = dh1.T @ x
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3, gb3]).astype(np.float32)
    return y.astype(np.float32), g, loss


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.mark.parametrize("n,k,hidden,batch,rank", [(300, 2, 512, 300, 0), (48, 3, 64, 77, 0), (100, 2, 128, 1024, 0),
                                                   (300, 2, 512, 256, 8), (64, 2, 64, 96, 4)])
def test_forward_and_gradients_match_fp32_reference(ctx, n, k, hidden, batch, rank):
    m = sw.Mlp(ctx, n, k, hidden=hidden, seed=3)
    m.set_loss(rank)
    D, H = m.in_dim, m.hidden
    rng = np.random.default_rng(n + k)
    p = m.params()
    p[H * D:H * D + H] = rng.normal(0, 0.1, H)           # non-zero biases
    p[H * D + H + H * H:H * D + 2 * H + H * H] = rng.normal(0, 0.1, H)
    p[-1] = 0.5
    m.set_params(p)
    x = np.zeros((batch, D), dtype=np.float32)
    x[:, :n] = rng.random((batch, n)) < 0.3
    x[:, n:n + k] = rng.random((batch, k))
    xb = bf16_bits(x)
    t = rng.random(batch).astype(np.float32) * 3
    if rank:
        t = np.floor(t * 2).astype(np.float32)  # ties too: tied pairs carry no loss
    y_ref, g_ref, loss_ref = reference(m.params(), xb, t, D, H, rank)
    y = m.forward(xb)
    np.testing.assert_allclose(y, y_ref, rtol=2e-2, atol=2e-2)
    g, loss = m.gradients(xb, t)
    assert abs(loss - loss_ref) <= 2e-2 * max(1.0, loss_ref)
    # Per tensor: relative L2 error of the gradient within 3e-2.
    sizes = [H * D, H, H * H, H, H, 1]
    o = 0
    for s in sizes:
        assert rel_err(g[o:o + s], g_ref[o:o + s]) < 3e-2, (o, s)
        o += s


def test_reference_gradient_by_finite_differences():
    """The CPU reference's gradient formulas against central differences (float64)."""
    rng = np.random.default_rng(0)
    D, H, B = 32, 16, 8
    n = D * H + H + H * H + H + H + 1
    p = rng.normal(0, 0.3, n).astype(np.float64)
    x = rng.random((B, D))
    t = rng.random(B)

    def loss_grad(p):
        w1, b1, w2, b2, w3, b3 = unpack(p, D, H)
        a1 = x @ w1.T + b1
        h1 = np.maximum(a1, 0)
        a2 = h1 @ w2.T + b2
        h2 = np.maximum(a2, 0)
        y = h2 @ w3 + b3[0]
        dy = 2 * (y - t) / B
        dh2 = np.outer(dy, w3) * (h2 > 0)
        dh1 = (dh2 @ w2) * (h1 > 0)
        g = np.concatenate([(dh1.T @ x).ravel(), dh1.sum(0), (dh2.T @ h1).ravel(), dh2.sum(0), h2.T @ dy, [dy.sum()]])
        return float(np.sum((y - t) ** 2) / B), g

    _, g = loss_grad(p)
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = 1e-6
        fd = (loss_grad(p + e)[0] - loss_grad(p - e)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(g[i])


def test_training_reduces_loss(ctx):
    m = sw.Mlp(ctx, 64, 2, hidden=128, seed=5)
    rng = np.random.default_rng(1)
    bits = rng.random((4096, 64)) < 0.4
    x = np.zeros((4096, m.in_dim), dtype=np.float32)
    x[:, :64] = bits
    t = (bits[:, :8].sum(axis=1) * 0.5).astype(np.float32)   # a learnable target
    curve = m.train(bf16_bits(x), t, epochs=8, batch=256, lr=2e-3, seed=1)
    assert curve[-1] < 0.25 * curve[0], curve


def _forward_distance(n, k, d, S, cap):
    """Length of the shortest u with delta(Q,u) inside S, by BFS over images (host)."""
    d = np.asarray(d).reshape(n, k)
    start = frozenset(range(n))
    seen = {start}
    level = [start]
    for dist in range(cap + 1):
        if any(A <= S for A in level):
            return dist
        nxt = []
        for A in level:
            for a in range(k):
                B = frozenset(int(d[q, a]) for q in A)
                if B not in seen:
                    seen.add(B)
                    nxt.append(B)
        level = nxt
    return cap + 1


def test_samples_are_labelled_with_forward_distances(ctx):
    n, k = 24, 2
    d = gen.random_dfa(n, k, 11)
    sets, dist = sw.mlp_samples(ctx, n, k, d, fdepth=6, width=32, bdepth=12, per_depth=6, cap=200)
    assert len(sets) > 20
    for row, dd in zip(sets, dist):
        S = frozenset(q for q in range(n) if (int(row[0]) >> q) & 1)
        assert int(dd) == _forward_distance(n, k, d, S, 6)


def test_device_beams_give_valid_bounds(ctx):
    """Cardinality-ordered and network-scored device beams: every bound they
    report below the cap is the length of a real reset word, never below the
    exact length."""
    cases = golden("c1_random48.json")[:10]
    m = sw.Mlp(ctx, 48, 2, hidden=64, seed=2)
    for c in cases:
        cap = c["eppstein"] + 1
        b, w = ctx.beam(48, 2, c["delta"], 64, cap)
        assert c["upper"] <= b <= cap
        if b < cap:
            assert len(w) == b and replay(48, 2, c["delta"], w) == 1
        b, w = m.beam(c["delta"], width=64, upper=cap)
        assert c["upper"] <= b <= cap
        if b < cap:
            assert len(w) == b and replay(48, 2, c["delta"], w) == 1
    m.close()


def guided_features(n, k, delta, sets):
    """Restatement of the scorer's input features (see sw_mlp_features)."""
    d = np.asarray(delta).reshape(n, k)
    full = frozenset(range(n))
    seen, level, rows, ends, last = {full}, [full], [full], [], [full]
    for depth in range(1, 9):
        nxt = []
        for A in level:
            for a in range(k):
                B = frozenset(int(d[q, a]) for q in A)
                if B not in seen:
                    seen.add(B)
                    nxt.append(B)
        rows += nxt
        if nxt:
            last = nxt
        level = nxt
        if depth % 2 == 0:
            ends.append(len(rows))
    freq = np.zeros(n)
    for A in last:
        for q in A:
            freq[q] += 1
    order = np.argsort(-freq, kind="stable")
    perm = np.empty(n, dtype=np.int64)
    perm[order] = np.arange(n)
    D = (n + k + 5 + 31) // 32 * 32
    x = np.zeros((len(sets), D), dtype=np.uint16)
    for i, row in enumerate(sets):
        S = {q for q in range(n) if (int(row[q // 64]) >> (q % 64)) & 1}
        for q in S:
            x[i, perm[q]] = 0x3F80
        for a in range(k):
            x[i, n + a] = bf16_bits(np.float32(8) * np.float32(len({int(d[q, a]) for q in S})) / np.float32(n))
        for j in range(4):
            best = min(len(A - S) for A in rows[:ends[j]])
            x[i, n + k + j] = bf16_bits(np.float32(8) * np.float32(best) / np.float32(n))
        x[i, n + k + 4] = bf16_bits(np.float32(8) * np.float32(len(S)) / np.float32(n))
    return x


@pytest.mark.parametrize("n,k,seed", [(24, 2, 1), (70, 3, 2), (130, 2, 3)])
def test_features_match_restatement(ctx, n, k, seed):
    d = gen.random_dfa(n, k, seed)
    rng = np.random.default_rng(seed)
    W = (n + 63) // 64
    sets = np.zeros((60, W), dtype=np.uint64)
    for i in range(60):
        bits = rng.random(n) < rng.random()
        for q in np.nonzero(bits)[0]:
            sets[i, q // 64] |= np.uint64(1) << np.uint64(q % 64)
    m = sw.Mlp(ctx, n, k, hidden=32, seed=0)
    got = m.features(d, sets)
    m.close()
    np.testing.assert_array_equal(got, guided_features(n, k, d, sets))


def test_samples_come_in_whole_groups(ctx):
    n, k = 40, 2
    d = gen.random_dfa(n, k, 5)
    sets, dist = sw.mlp_samples(ctx, n, k, d, fdepth=8, width=32, bdepth=30, per_depth=8, cap=1000)
    assert len(sets) > 0 and len(sets) % 8 == 0


def test_trained_scorer_gains_over_cardinality_order(ctx):
    """What the trained network gains: on the held-out automata of the
    learned-bound benchmark (n=300, k=2, seeds 1000..1019), a device beam of
    width 64 ordered by the network ends closer to the shortest length than
    the same beam ordered by cardinality (the reference's order).  Every
    bound is the length of a word that replays to a singleton."""
    from pathlib import Path
    from synchrowords_amd.train_scorer import load_scorer
    path = Path(sw.__file__).parent / "models" / "scorer_n300_k2.npz"
    m = load_scorer(ctx, path)
    n, k, width, cap = 300, 2, 64, 300
    learned, cardinality = [], []
    for s in range(1000, 1020):
        d = sw.random_dfa(n, k, s)
        bl, wl = m.beam(d, width, cap)
        bc, wc = ctx.beam(n, k, d, width, cap)
        for b, w in ((bl, wl), (bc, wc)):
            assert b < cap and len(w) == b and replay(n, k, d, w) == 1
        learned.append(bl)
        cardinality.append(bc)
    m.close()
    known = {c["seed"]: c["upper"] for c in golden("c5_random300.json")}
    for s, b in zip(range(1000, 1020), learned):
        if s in known:
            assert b >= known[s]
    assert sum(learned) < sum(cardinality), (learned, cardinality)
    assert sum(l < c for l, c in zip(learned, cardinality)) > sum(l > c for l, c in zip(learned, cardinality))


def _sink_states(n, k, d):
    """States of the sink component: those reachable from every state (the
    automata here are synchronizing, so it is unique)."""
    dd = np.asarray(d, dtype=np.int64).reshape(n, k)
    reach = []
    for s in range(n):
        seen = np.zeros(n, dtype=bool)
        seen[s] = True
        stack = [s]
        while stack:
            q = stack.pop()
            for t in dd[q]:
                if not seen[t]:
                    seen[t] = True
                    stack.append(int(t))
        reach.append(seen)
    return np.logical_and.reduce(reach)


def _learned_beam_restated(m, n, k, d, width, cap):
    """The learned beam (beam.hpp:98-188 with the network's order): from the
    sink singletons with two preimages under some letter, each depth takes
    the preimages (letter-major), stops at the full set, drops repeats (the
    first kept), and keeps the `width` lowest scores (ties by position)."""
    import oracle_lib as orc
    dd = np.asarray(d, dtype=np.int64).reshape(n, k)
    sink = _sink_states(n, k, d)
    indeg = np.zeros((n, k), dtype=np.int64)
    for a in range(k):
        np.add.at(indeg[:, a], dd[:, a], 1)
    W = sw.words(n)
    seeds = [q for q in range(n) if sink[q] and (indeg[q] > 1).any()]
    lst = np.zeros((len(seeds), W), dtype=np.uint64)
    for i, q in enumerate(seeds):
        lst[i, q // 64] = np.uint64(1) << np.uint64(q % 64)
    full = np.zeros(W, dtype=np.uint64)
    for q in range(n):
        full[q // 64] |= np.uint64(1) << np.uint64(q % 64)
    for length in range(1, cap):
        cand = orc.apply(n, k, d, lst, backward=True)
        if (cand == full).all(axis=1).any():
            return length
        _, first = np.unique(cand, axis=0, return_index=True)
        uniq = cand[np.sort(first)]
        if len(uniq) > width:
            y = m.forward(m.features(d, uniq))
            uniq = uniq[np.sort(np.argsort(y, kind="stable")[:width])]
        lst = uniq
        if not len(lst):
            return cap
    return cap


@pytest.mark.parametrize("n,width,trained", [(300, 12000, True), (100, 9000, False), (300, 80000, True)])
def test_learned_beam_past_the_counted_order(ctx, n, width, trained):
    """Depths of more than 16384 distinct candidates select the kept sets on
    the device by radix selection over (score, position): the bound equals
    a host restatement that orders the same network's scores with a stable
    sort, for the committed network and for an untrained one."""
    from pathlib import Path
    from synchrowords_amd.train_scorer import load_scorer
    k = 2
    if trained:
        m = load_scorer(ctx, Path(sw.__file__).parent / "models" / "scorer_n300_k2.npz")
    else:
        m = sw.Mlp(ctx, n, k, hidden=64, seed=3)
    try:
        for s in (11, 12):
            d = gen.random_dfa(n, k, 9100 + s)
            cap = n * 2
            want = _learned_beam_restated(m, n, k, d, width, cap)
            b, w = m.beam(d, width, cap)
            assert b == want
            if b < cap:
                assert len(w) == b and replay(n, k, d, w) == 1
    finally:
        m.close()
