"""GPU checks of the learned prefilter's kernels (no counterpart in the reference).

The screen is checked exactly against numpy on BF16-representable data; the encoder forward
and the training gradients against a float64 numpy model of the same network (with the
GPU's BF16 rounding points on the forward activations); the end-to-end search against the
exhaustive exact pass (every reported pair must be an exact hit with its exact score)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C, E = 128, 32


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def to_bf16_bits(x):
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def unpack(p):
    """Parameter vector -> (emb [32][32], [(W [5 cin][128], b [128])] * 3)."""
    emb = p[:32 * E].reshape(32, E)
    o = 32 * E
    layers = []
    for cin in (E, C, C):
        W = p[o:o + 5 * cin * C].reshape(5 * cin, C)
        o += 5 * cin * C
        b = p[o:o + C]
        o += C
        layers.append((W, b))
    assert o == len(p)
    return emb, layers


def conv(X, W, b):
    """X [L][cin] -> [L][128], zero padding 2 each side, W rows k * cin + ci."""
    L, cin = X.shape
    Xp = np.vstack([np.zeros((2, cin)), X, np.zeros((2, cin))])
    cols = np.stack([Xp[p:p + 5].reshape(-1) for p in range(L)]) if L else np.zeros((0, 5 * cin))
    return cols @ W + b, cols


def forward(params, seqs, round_acts=True):
    emb, layers = unpack(params.astype(np.float64))
    r = bf16 if round_acts else (lambda v: v)
    embr = r(emb)
    out, caches = [], []
    for s in seqs:
        X = embr[s]
        cache = [X]
        for W, b in layers:
            Y, _ = conv(X, r(W), b)
            X = r(np.maximum(Y, 0.0))
            cache.append(X)
        out.append(X.mean(axis=0) if len(s) else np.zeros(C))
        caches.append(cache)
    return np.array(out), caches


def backward(params, seqs, pa, pb, y, w):
    """float64 gradient of sum w (|e_a - e_b|^2 - y)^2 through the same network."""
    emb, layers = unpack(params.astype(np.float64))
    e, caches = forward(params, seqs)
    g_emb = np.zeros_like(emb)
    g_layers = [(np.zeros_like(W), np.zeros_like(b)) for W, b in layers]
    dE = np.zeros_like(e)
    for a, bb, yt, wt in zip(pa, pb, y, w):
        diff = e[a] - e[bb]
        d = (diff ** 2).sum()
        g = 4 * wt * (d - yt) * diff
        dE[a] += g
        dE[bb] -= g
    for s, cache, de in zip(seqs, caches, dE):
        L = len(s)
        dH = np.tile(de / L, (L, 1))
        for li in (2, 1, 0):
            W, b = layers[li]
            H = cache[li + 1]
            dY = dH * (H > 0)
            _, cols = conv(cache[li], bf16(W), b)
            g_layers[li][0][...] += cols.T @ dY
            g_layers[li][1][...] += dY.sum(axis=0)
            cin = W.shape[0] // 5
            dcols = dY @ bf16(W).T  # [L][5 cin]
            dX = np.zeros((L + 4, cin))
            for p in range(L):
                dX[p:p + 5] += dcols[p].reshape(5, cin)
            dH = dX[2:L + 2]
        np.add.at(g_emb, s, dH)
    flat = [g_emb.reshape(-1)]
    for gW, gb in g_layers:
        flat += [gW.reshape(-1), gb]
    return np.concatenate(flat)


def rand_params(rng, n):
    p = np.zeros(n, np.float32)
    emb_n = 32 * E
    p[:emb_n] = rng.normal(0, 0.5, emb_n)
    o = emb_n
    for cin in (E, C, C):
        k = 5 * cin * C
        p[o:o + k] = rng.normal(0, np.sqrt(2.0 / (5 * cin)), k)
        o += k
        p[o:o + C] = rng.normal(0, 0.05, C)
        o += C
    return p


def rand_set(rng, n, lo, hi, A=20):
    lens = rng.integers(lo, hi + 1, size=n)
    seqs = [rng.integers(0, A, size=int(L)).astype(np.uint8) for L in lens]
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    codes = np.concatenate(seqs) if n else np.zeros(0, np.uint8)
    return seqs, codes, off


def test_screen_matches_numpy(gpu):
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.prefilter import screen

    rng = np.random.default_rng(3)
    nq, nr = 300, 2077
    Q = rng.integers(-8, 9, size=(nq, C)) / 8.0
    R = Q[rng.integers(0, nq, size=nr)] + rng.integers(-2, 3, size=(nr, C)) / 8.0
    Q, R = Q.astype(np.float32), R.astype(np.float32)
    qn = (Q.astype(np.float64) ** 2).sum(1)
    rn = (R.astype(np.float64) ** 2).sum(1)
    d = qn[:, None] + rn[None, :] - 2 * (Q.astype(np.float64) @ R.astype(np.float64).T)
    tau = float(np.quantile(d, 0.002) // (1 / 64) * (1 / 64))
    want = np.argwhere(d <= tau)
    dq = DeviceArray.from_numpy(to_bf16_bits(Q))
    dr = DeviceArray.from_numpy(to_bf16_bits(R))
    hq, hr, n = screen(dq, DeviceArray.from_numpy(qn.astype(np.float32)), nq, dr,
                       DeviceArray.from_numpy(rn.astype(np.float32)), nr, tau, 64)
    got = np.stack([hq.numpy()[:n], hr.numpy()[:n]], 1).astype(np.int64)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    assert n == len(want) and np.array_equal(got, want)


@pytest.mark.parametrize("nq,nr", [(1, 1), (1, 15), (3, 16), (17, 63), (65, 64), (64, 65),
                                   (2, 257)])
def test_screen_small_shapes(gpu, nq, nr):
    """Query and reference counts around the screen's blocks (16 queries, 64 per wave, 16-row
    quarters, 64-row tiles, four-tile ring rounds): every pair within tau, exactly."""
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.prefilter import screen

    rng = np.random.default_rng(nq * 1000 + nr)
    Q = (rng.integers(-8, 9, size=(nq, C)) / 8.0).astype(np.float32)
    R = (rng.integers(-8, 9, size=(nr, C)) / 8.0).astype(np.float32)
    qn = (Q.astype(np.float64) ** 2).sum(1)
    rn = (R.astype(np.float64) ** 2).sum(1)
    d = qn[:, None] + rn[None, :] - 2 * (Q.astype(np.float64) @ R.astype(np.float64).T)
    tau = float(np.median(d) // (1 / 64) * (1 / 64))
    want = np.argwhere(d <= tau)
    hq, hr, n = screen(DeviceArray.from_numpy(to_bf16_bits(Q)),
                       DeviceArray.from_numpy(qn.astype(np.float32)), nq,
                       DeviceArray.from_numpy(to_bf16_bits(R)),
                       DeviceArray.from_numpy(rn.astype(np.float32)), nr, tau, 8)
    got = np.stack([hq.numpy()[:n], hr.numpy()[:n]], 1).astype(np.int64)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    assert n == len(want) and np.array_equal(got, want)


def test_encode_matches_numpy_model(gpu):
    from seqtree_amd.prefilter import DeviceSet, Prefilter, _bf16_to_f32

    rng = np.random.default_rng(5)
    model = Prefilter(seed=1)
    params = rand_params(rng, model.n_params)
    model.load(params)
    assert np.array_equal(model.params(), params)
    seqs, codes, off = rand_set(rng, 301, 1, 20)
    emb, norm = model.encode(DeviceSet.upload(codes, off), 20)
    got = _bf16_to_f32(emb.numpy())
    want, _ = forward(params, seqs)
    assert np.allclose(got, want, rtol=3e-2, atol=3e-2 * np.abs(want).max())
    assert np.allclose(norm.numpy(), (got.astype(np.float64) ** 2).sum(1), rtol=1e-4)
    model.close()


def test_encode_empty_and_longest_sequences(gpu):
    """Lengths 0 .. 64: sequences cut by the encoder's 128-row tiles (pooled in two parts) and
    empty ones (zero embedding) next to whole ones."""
    from seqtree_amd.prefilter import DeviceSet, Prefilter, _bf16_to_f32

    rng = np.random.default_rng(13)
    model = Prefilter(seed=4)
    params = rand_params(rng, model.n_params)
    model.load(params)
    seqs, codes, off = rand_set(rng, 700, 0, 64)
    emb, norm = model.encode(DeviceSet.upload(codes, off), 64)
    got = _bf16_to_f32(emb.numpy())
    want, _ = forward(params, seqs)
    assert np.allclose(got, want, rtol=3e-2, atol=3e-2 * np.abs(want).max())
    empty = np.array([len(q) == 0 for q in seqs])
    assert empty.any() and not got[empty].any()
    assert np.allclose(norm.numpy(), (got.astype(np.float64) ** 2).sum(1), rtol=1e-4)
    model.close()


@pytest.mark.parametrize("lens", [[7], [0], [0, 0, 3], [64, 0, 64], [20] * 7])
def test_encode_tiny_sets(gpu, lens):
    """Sets of one to seven sequences: a single tile holding every sequence, empty ones
    included, and sequences that run past the first tile's end."""
    from seqtree_amd.prefilter import DeviceSet, Prefilter, _bf16_to_f32

    rng = np.random.default_rng(sum(lens) + len(lens))
    model = Prefilter(seed=3)
    params = rand_params(rng, model.n_params)
    model.load(params)
    seqs = [rng.integers(0, 20, size=L).astype(np.uint8) for L in lens]
    off = np.zeros(len(lens) + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    codes = np.concatenate(seqs) if sum(lens) else np.zeros(0, np.uint8)
    emb, norm = model.encode(DeviceSet.upload(codes, off), max(max(lens), 1))
    got = _bf16_to_f32(emb.numpy())
    want, _ = forward(params, seqs)
    assert np.allclose(got, want, rtol=3e-2, atol=3e-2 * max(np.abs(want).max(), 1e-6))
    assert np.allclose(norm.numpy(), (got.astype(np.float64) ** 2).sum(1), rtol=1e-4, atol=1e-6)
    model.close()


def test_encode_across_blocks(gpu):
    """450,000 sequences of length 0 .. 20 take three encoder blocks of up to 2^22 rows (the
    persistent kernel's tiles, with their inputs fetched ahead, run many per workgroup): the
    sequences at and beside the block boundaries and a random sample equal the numpy model;
    every norm is its embedding's."""
    from seqtree_amd.prefilter import DeviceSet, Prefilter, _bf16_to_f32

    rng = np.random.default_rng(31)
    model = Prefilter(seed=6)
    params = rand_params(rng, model.n_params)
    model.load(params)
    n = 450_000
    seqs, codes, off = rand_set(rng, n, 0, 20)
    emb, norm = model.encode(DeviceSet.upload(codes, off), 20)
    got_all = emb.numpy()
    norm_all = norm.numpy()
    blk = (1 << 22) // 22  # sequences per block at max_len 20
    edges = [0, blk, 2 * blk, n]
    idx = sorted({i for e in edges for i in range(e - 4, e + 4) if 0 <= i < n} |
                 set(rng.integers(0, n, 400).tolist()))
    got = _bf16_to_f32(got_all[idx])
    want, _ = forward(params, [seqs[i] for i in idx])
    assert np.allclose(got, want, rtol=3e-2, atol=3e-2 * np.abs(want).max())
    sample = np.unique(np.concatenate([idx, rng.integers(0, n, 20_000)]))
    g = _bf16_to_f32(got_all[sample]).astype(np.float64)
    assert np.allclose(norm_all[sample], (g ** 2).sum(1), rtol=1e-4, atol=1e-6)
    empty = np.array([len(seqs[i]) == 0 for i in sample])
    assert empty.any() and not g[empty].any()
    model.close()


def test_training_gradients_match_numpy_model(gpu):
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.prefilter import DeviceSet, Prefilter

    rng = np.random.default_rng(7)
    model = Prefilter(seed=2)
    params = rand_params(rng, model.n_params)
    model.load(params)
    seqs, codes, off = rand_set(rng, 41, 3, 12)
    npairs = 70
    pa = rng.integers(0, 41, size=npairs)
    pb = rng.integers(0, 41, size=npairs)
    y = rng.uniform(0, 2, size=npairs).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=npairs).astype(np.float32)
    loss = model.train_step(DeviceSet.upload(codes, off),
                            DeviceArray.from_numpy(pa.astype(np.uint32)),
                            DeviceArray.from_numpy(pb.astype(np.uint32)),
                            DeviceArray.from_numpy(y), DeviceArray.from_numpy(w), npairs, 0.0, 12)
    e, _ = forward(params, seqs)
    d = ((e[pa] - e[pb]) ** 2).sum(1)
    assert np.isclose(loss, float((w * (d - y) ** 2).mean()), rtol=3e-2)
    got = model.grads().astype(np.float64)
    want = backward(params, seqs, pa, pb, y, w)
    # per parameter block: same direction and size up to BF16 gradient rounding
    o = 0
    for name, size in (("emb", 32 * E), ("W1", 5 * E * C), ("b1", C), ("W2", 5 * C * C),
                       ("b2", C), ("W3", 5 * C * C), ("b3", C)):
        g, h = got[o:o + size], want[o:o + size]
        o += size
        rel = np.linalg.norm(g - h) / max(np.linalg.norm(h), 1e-12)
        assert rel < 0.05, (name, rel)
    assert np.array_equal(model.params(), params)  # lr = 0 leaves the parameters alone
    model.close()


def test_select_le(gpu):
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.prefilter import select_le

    rng = np.random.default_rng(9)
    v = rng.integers(-100, 100, size=200_003).astype(np.int32)
    got = select_le(DeviceArray.from_numpy(v), len(v), -90)
    assert np.array_equal(got, np.flatnonzero(v <= -90))


def test_small_search_reports_only_exact_hits(gpu):
    from seqtree_amd import synthetic
    from seqtree_amd.prefilter import (DeviceSet, GapblockParams, Prefilter, TrainingData,
                                       exact_threshold, label_pairs, recall, sample_pairs,
                                       search, train)

    rng = np.random.default_rng(11)
    codes, off, fam = synthetic.families(rng, 400, 4, 10, 20)
    gp = GapblockParams.blosum62_central(20)
    pa, pb = sample_pairs(rng, fam, 20_000, 0.5)
    score = label_pairs(codes, off, pa, pb, gp)
    model = Prefilter(seed=3)
    losses = train(model, TrainingData(codes, off, pa, pb, score), epochs=3, batch_pairs=2048,
                   max_len=20)
    assert losses[-1] < losses[0]
    qc, qo, _ = synthetic.families(rng, 60, 1, 10, 20)
    q, r = DeviceSet.upload(qc, qo), DeviceSet.upload(codes, off)
    cutoff = 60
    res = search(model, q, r, cutoff, 150.0, gp, 20)
    eq, er, es = exact_threshold(q, r, cutoff, gp)
    exact = {(int(a), int(b)): int(s) for a, b, s in zip(eq, er, es)}
    for a, b, s in zip(res.query, res.ref, res.score):
        assert exact[(int(a), int(b))] == int(s)
    assert len(set(zip(res.query.tolist(), res.ref.tolist()))) == len(res.query)  # no repeats
    model.close()


# ------------------------------------------------------------------- C5 at a scaled size
# The C5 recipe (200,000 synthetic CDR3s -- independent draws and families of 4 -- of which
# 180,000 train on 2M pairs mined and labelled by the exact kernel for 16 epochs and 20,000
# calibrate the margin) and its cutoff, searched with
# 5,000 queries x 100,000 references of two kinds: independent CDR3s (the generator alone) and
# planted families (each query a fresh mutant of one reference family's root).
C5_LEN, C5_CUTOFF = (10, 20), 48


@pytest.fixture(scope="module")
def c5_model(gpu):
    from seqtree_amd import synthetic as syn
    from seqtree_amd.prefilter import (DEFAULT_QUANTILE, GapblockParams, Prefilter,
                                       TrainingData, calibrate_on_set, label_pairs, mine_pairs,
                                       train)

    rng = np.random.default_rng(2025)
    # 180,000 training sequences and 20,000 held out for the calibration
    tc, to = syn.training_set(rng, 90_000, 22_500, 4, *C5_LEN)
    hc, ho = syn.training_set(rng, 10_000, 2_500, 4, *C5_LEN)
    gp = GapblockParams.blosum62_central(C5_LEN[1])
    pa, pb, n_close = mine_pairs(rng, tc, to, gp, 2_000_000, 0.4, C5_CUTOFF, 2 * C5_CUTOFF)
    assert n_close > 10_000
    score = label_pairs(tc, to, pa, pb, gp)
    model = Prefilter(seed=2025)
    losses = train(model, TrainingData(tc, to, pa, pb, score), epochs=16, max_len=C5_LEN[1],
                   seed=2025)
    assert losses[-1] < 0.5 * losses[0]
    tau, n_cal = calibrate_on_set(model, hc, ho, gp, C5_CUTOFF, DEFAULT_QUANTILE, C5_LEN[1])
    assert n_cal > 1000
    yield model, gp, tau
    model.close()


def _c5_sets(kind, rng):
    from seqtree_amd import synthetic as syn

    if kind == "iid":
        rc, ro, _ = syn.cdr3_csr(rng, 100_000, *C5_LEN)
        qc, qo, _ = syn.cdr3_csr(rng, 5_000, *C5_LEN)
        return qc, qo, rc, ro
    roots = []
    rc, ro, _ = syn.families(rng, 25_000, 4, *C5_LEN, roots_out=roots)
    pick = rng.choice(25_000, size=5_000, replace=False)
    qc, qo = syn.join(syn._mutate_many(rng, [roots[f] for f in pick], 1, *C5_LEN))
    return qc, qo, rc, ro


@pytest.mark.parametrize("kind", ["iid", "planted"])
def test_scaled_c5_recall(c5_model, oracle_lib, kind):
    """Every reported hit is an exact hit with its exact score, the exhaustive pass equals the
    oracle on sampled rows, and the prefilter finds at least 99.5 % of the exhaustive hits."""
    from seqtree_amd import synthetic as syn
    from seqtree_amd.prefilter import DeviceSet, exact_threshold, recall, search

    model, gp, tau = c5_model
    rng = np.random.default_rng(7 if kind == "iid" else 8)
    qc, qo, rc, ro = _c5_sets(kind, rng)
    Q, R = DeviceSet.upload(qc, qo), DeviceSet.upload(rc, ro)
    print(f"scaled C5 {kind}: {Q.n} x {R.n}, held-out tau {tau:.1f}", flush=True)
    eq, er, es = exact_threshold(Q, R, C5_CUTOFF, gp)
    assert len(eq) > 1000, len(eq)
    # the exhaustive pass against the oracle on sampled query rows
    qs, rs = syn.decode(qc, qo), syn.decode(rc, ro)
    rows = rng.choice(Q.n, size=4, replace=False)
    dense = np.asarray(oracle_lib.gapblock_matrix([qs[i] for i in rows], rs, gp.pen,
                                                  gp.gap_open, gp.gap_extend, gp.cube,
                                                  gp.width)).reshape(len(rows), R.n)
    for k, i in enumerate(rows):
        want = {(int(j), int(dense[k, j])) for j in np.flatnonzero(dense[k] <= C5_CUTOFF)}
        got = {(int(j), int(s)) for j, s in zip(er[eq == i], es[eq == i])}
        assert got == want, (kind, i)
    res = search(model, Q, R, C5_CUTOFF, tau, gp, C5_LEN[1], audit=0.02)  # default quantile
    exact = dict(zip(((eq.astype(np.int64) << 32) | er).tolist(), es.tolist()))
    keys = ((res.query.astype(np.int64) << 32) | res.ref).tolist()
    assert len(set(keys)) == len(keys)
    assert all(exact.get(k_) == int(s) for k_, s in zip(keys, res.score.tolist()))
    rec = recall(res.query, res.ref, eq, er)
    print(f"  exact hits {len(eq)}, found {len(res.query)}, recall {rec:.4f}, candidates "
          f"{res.candidates}, tau {res.tau:.1f}, audit {res.audit_queries} queries / "
          f"{res.audit_hits} hits", flush=True)
    assert res.candidates < 0.01 * Q.n * R.n  # the screen does screen
    assert rec >= 0.995, (kind, rec)
