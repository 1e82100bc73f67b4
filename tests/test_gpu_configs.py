"""Oracle parity at the benchmark configurations' own workloads, bit for bit.

Every configuration the bench times is run here at its full shape through the
same entry points the bench uses (batch plans, hipGraph replays, the serving
loop), and its results are compared with the oracle on the same inputs:

- n = 2^16, L = 24, dnum = 4: hipGraph-replayed EvalMult+Rescale batches of 16
  and 64 at level 24, and one plan over uneven levels;
- n = 2^15, L = 12, dnum = 3: 16 ciphertexts through a 10-deep
  EvalMult+Rescale chain, one plan per depth, replayed as graphs;
- n = 2^16, L = 20, dnum = 4: 64 rotations at levels 8..20 under 16 resident
  rotation keys (slot offsets 2^j), one plan;
- a serving trace over rings 2^14 / 2^15 / 2^16 at levels 3..20 with the
  60 / 30 / 10 EvalMult+Rescale / Rotate / EvalAdd mix and 16 rotation keys per
  ring: every result equals the same operation run alone, and a sample
  equals the oracle.

Inputs and keys are uniform residues generated on the device (a switching
key is computationally uniform; bit parity needs no valid key) and
downloaded for the oracle, which runs threaded (tests/test_evalmult_rescale.cpp
and bench/bench_evalmult_resident.cpp:80-128 are the reference's
counterparts at n = 2^15, L = 4).
"""
import numpy as np
import pytest

from common import make_params, parallel_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


def filled(G, ctx, level, seed, s):
    d = G.Dev(level * ctx.n)
    G.check(G.lib.gf_fill_uniform(ctx.h, d.ptr, level, seed, s.s))
    return d


class Cfg3:
    """n = 2^16, L = 24 (60 + 23 x 50-bit), dnum = 4, K = 6, autotuned like the bench."""

    def __init__(self, G):
        self.P = make_params(1 << 16, 24, 4)
        P = self.P
        self.s = G.Stream()
        self.ctx = G.Context(P.n, P.q, P.p, P.roots, P.dnum)
        self.ctx.autotune(self.s)
        self.key = G.Key.uniform(self.ctx, P.dnum, 4242, self.s)
        self.av, self.bv = self.key.download(self.s)


_cfg3 = None


def cfg3(G):
    global _cfg3
    if _cfg3 is None:
        _cfg3 = Cfg3(G)
    return _cfg3


def emr_batch(G, c, levels, seed0):
    n = c.P.n
    ins = [[filled(G, c.ctx, lv, seed0 + 10 * i + j, c.s) for j in range(4)] for i, lv in enumerate(levels)]
    outs = [[G.Dev((lv - 1) * n) for _ in range(2)] for lv in levels]
    ws = [c.ctx.workspace() for _ in levels]
    col = lambda j: [x[j].ptr for x in ins]  # noqa: E731
    plan = G.plan_evalmult_rescale(c.ctx, c.key, list(levels), col(0), col(1), col(2), col(3),
                                   [o[0].ptr for o in outs], [o[1].ptr for o in outs], [w.ptr for w in ws])
    return plan, ins, outs, ws


def oracle_emr(c, levels, ins, idx):
    def one(i):
        cts = [x.download() for x in ins[i]]
        return c.P.evalmult_rescale(levels[i], *cts, c.av, c.bv)
    return parallel_map(one, idx)


@pytest.mark.parametrize("B,sample", [(16, 16), (64, 8)])
def test_cfg3_graph_batch_matches_oracle(G, B, sample):
    """The bench's measured form: one plan over B ops at level 24, replayed
    as a hipGraph; graph replays with eager work between them give the same
    results; `sample` ops (all at B = 16) equal the oracle."""
    c = cfg3(G)
    levels = [24] * B
    plan, ins, outs, ws = emr_batch(G, c, levels, 1000 * B)
    plan.run_graph(c.s)
    c.s.sync()
    first = [(o[0].download(), o[1].download()) for o in outs]
    idx = list(range(B)) if sample >= B else sorted({0, B - 1} | set(range(1, B - 1, max(1, B // sample))))[:sample]
    want = oracle_emr(c, levels, ins, idx)
    for i, (w0, w1) in zip(idx, want):
        assert np.array_equal(first[i][0], w0), ("c0", i)
        assert np.array_equal(first[i][1], w1), ("c1", i)
    for rep in range(2):
        for o in outs:
            G.check(G.lib.gf_fill_uniform(c.ctx.h, o[0].ptr, 23, 99 + rep, c.s.s))
            G.check(G.lib.gf_fill_uniform(c.ctx.h, o[1].ptr, 23, 98 + rep, c.s.s))
        plan.run_graph(c.s)
        c.s.sync()
        # unrelated eager work between replays
        G.check(G.lib.gf_add(c.ctx.h, 2, ins[0][0].ptr, ins[0][1].ptr, ins[1][0].ptr, ins[1][1].ptr,
                             ws[0].ptr, ws[0].at(2 * c.P.n), c.s.s))
        c.s.sync()
        for i in range(B):
            assert np.array_equal(outs[i][0].download(), first[i][0]), (rep, i)
            assert np.array_equal(outs[i][1].download(), first[i][1]), (rep, i)


def test_cfg3_uneven_levels_run_and_graph(G):
    """One plan at n = 2^16, L = 24 over uneven levels (the grouping of the
    fused row pass + inner product by part count, the conversion re-chunking
    and the lazy accumulator rows all meet here): run() and run_graph() both
    equal the oracle."""
    c = cfg3(G)
    levels = [24, 23, 12, 2, 7, 18]
    plan, ins, outs, ws = emr_batch(G, c, levels, 77)
    want = oracle_emr(c, levels, ins, range(len(levels)))
    for mode in ("run", "graph"):
        for lv, o in zip(levels, outs):
            G.check(G.lib.gf_fill_uniform(c.ctx.h, o[0].ptr, lv - 1, 5, c.s.s))
            G.check(G.lib.gf_fill_uniform(c.ctx.h, o[1].ptr, lv - 1, 6, c.s.s))
        if mode == "run":
            plan.run(c.s)
        else:
            plan.run_graph(c.s)
        c.s.sync()
        for i, (w0, w1) in enumerate(want):
            assert np.array_equal(outs[i][0].download(), w0), (mode, levels[i])
            assert np.array_equal(outs[i][1].download(), w1), (mode, levels[i])


def test_cfg3_integer_and_fp64_passes(G):
    """The plan with every transform and conversion in integer instructions
    (modup_f64 = 0, the bench's integer_only line), with fp64 passes around
    the fused ModUp and ModDown only, and with the fused inner product's row
    passes in fp64 as well, gives the oracle's results at n = 2^16, L = 24
    over uneven levels (Q rows with one own part, P rows with none, lazy
    accumulator rows).  The fused ModUp and ModDown, whose column passes the
    fp64 setting covers, are forced on for this small batch (modup_min_wgs =
    0), and each plan runs directly and as a graph."""
    c = cfg3(G)
    levels = [24, 13, 2]
    saved = {k: c.ctx.knob(k) for k in ("modup_f64", "ksi_f64", "modup_min_wgs")}
    want = None
    try:
        c.ctx.set_knob("modup_min_wgs", 0)
        for modup_f64, ksi_f64 in ((0, 0), (1, 0), (1, 1)):
            c.ctx.set_knob("modup_f64", modup_f64)
            c.ctx.set_knob("ksi_f64", ksi_f64)
            plan, ins, outs, ws = emr_batch(G, c, levels, 31)
            if want is None:
                want = oracle_emr(c, levels, ins, range(len(levels)))
            for mode in ("run", "graph"):
                for lv, o in zip(levels, outs):
                    G.check(G.lib.gf_fill_uniform(c.ctx.h, o[0].ptr, lv - 1, 5, c.s.s))
                    G.check(G.lib.gf_fill_uniform(c.ctx.h, o[1].ptr, lv - 1, 6, c.s.s))
                if mode == "run":
                    plan.run(c.s)
                else:
                    plan.run_graph(c.s)
                c.s.sync()
                for i, (w0, w1) in enumerate(want):
                    assert np.array_equal(outs[i][0].download(), w0), (modup_f64, ksi_f64, mode, levels[i])
                    assert np.array_equal(outs[i][1].download(), w1), (modup_f64, ksi_f64, mode, levels[i])
    finally:
        for k, v in saved.items():
            c.ctx.set_knob(k, v)


def test_cfg2_chain_16x10_matches_oracle(G):
    """16 ciphertexts, each multiplied (EvalMult+Rescale) 10 times in a row by
    a resident operand of the matching level, n = 2^15, L = 12 -> 2, dnum = 3:
    one plan per depth over the batch, replayed as graphs in order (the
    bench's chain).  Every intermediate equals the oracle's sequential chain."""
    n, L, dnum, batch, depth = 1 << 15, 12, 3, 16, 10
    P = make_params(n, L, dnum)
    s = G.Stream()
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    ctx.autotune(s)
    key = G.Key.uniform(ctx, dnum, 77, s)
    av, bv = key.download(s)
    cts = [[[filled(G, ctx, L - d, 31 * (d * batch + b) + j, s) if d == 0 else G.Dev((L - d) * n) for j in range(2)]
            for b in range(batch)] for d in range(depth + 1)]
    ops = [[[filled(G, ctx, L - d, 7 + 31 * (d * batch + b) + j, s) for j in range(2)] for b in range(batch)]
           for d in range(depth)]
    ws = [ctx.workspace() for _ in range(batch)]
    plans = []
    for d in range(depth):
        src, dst, opd = cts[d], cts[d + 1], ops[d]
        plans.append(G.plan_evalmult_rescale(ctx, key, [L - d] * batch, [x[0].ptr for x in src],
                                             [x[1].ptr for x in src], [x[0].ptr for x in opd],
                                             [x[1].ptr for x in opd], [x[0].ptr for x in dst],
                                             [x[1].ptr for x in dst], [w.ptr for w in ws]))
    for pl in plans:
        pl.run_graph(s)
    s.sync()
    got = [[(cts[d][b][0].download(), cts[d][b][1].download()) for d in range(depth + 1)] for b in range(batch)]
    host_ops = [[(ops[d][b][0].download(), ops[d][b][1].download()) for d in range(depth)] for b in range(batch)]
    def chain(b):
        c0, c1 = got[b][0]
        out = []
        for d in range(depth):
            c0, c1 = P.evalmult_rescale(L - d, c0, c1, host_ops[b][d][0], host_ops[b][d][1], av, bv)
            out.append((c0, c1))
        return out

    want = parallel_map(chain, range(batch))
    for b in range(batch):
        for d in range(depth):
            assert np.array_equal(got[b][d + 1][0], want[b][d][0]), (b, d)
            assert np.array_equal(got[b][d + 1][1], want[b][d][1]), (b, d)


def test_cfg4_rotations_64_mixed_levels_16_keys(G):
    """64 ciphertexts at levels uniform in 8..20 (n = 2^16, L = 20, dnum = 4),
    each rotated by one of 16 power-of-two slot offsets (Galois index
    5^(2^j) mod 2n) under its own resident key, one plan replayed as a graph
    (the bench's rotation batch).  Every result equals the same rotation run
    alone, and one rotation per key equals the oracle."""
    n, L, dnum, count, nkeys = 1 << 16, 20, 4, 64, 16
    P = make_params(n, L, dnum)
    s = G.Stream()
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    ctx.autotune(s)
    rng = np.random.default_rng(9)
    keys = [G.Key.uniform(ctx, dnum, 500 + j, s) for j in range(nkeys)]
    ks = [pow(5, 1 << j, 2 * n) for j in range(nkeys)]
    levels = [int(x) for x in rng.integers(8, L + 1, count)]
    which = [i % nkeys for i in range(count)]
    cin = [[filled(G, ctx, lv, 91 * i + j, s) for j in range(2)] for i, lv in enumerate(levels)]
    cout = [[G.Dev(lv * n) for _ in range(2)] for lv in levels]
    ws = [ctx.workspace() for _ in range(count)]
    plan = G.plan_rotate(ctx, [keys[w] for w in which], levels, [ks[w] for w in which],
                         [x[0].ptr for x in cin], [x[1].ptr for x in cin], [x[0].ptr for x in cout],
                         [x[1].ptr for x in cout], [w.ptr for w in ws])
    plan.run_graph(s)
    s.sync()
    got = [(o[0].download(), o[1].download()) for o in cout]
    # alone: the single-op entry point on the same inputs
    w1 = ctx.workspace()
    a0, a1 = G.Dev(L * n), G.Dev(L * n)
    for i, lv in enumerate(levels):
        G.check(G.lib.gf_rotate(ctx.h, keys[which[i]].h, lv, ks[which[i]], cin[i][0].ptr, cin[i][1].ptr,
                                a0.ptr, a1.ptr, w1.ptr, s.s))
        s.sync()
        assert np.array_equal(a0.download(lv * n), got[i][0]), ("alone c0", i, lv)
        assert np.array_equal(a1.download(lv * n), got[i][1]), ("alone c1", i, lv)
    # oracle: the first rotation under each key
    idx = [which.index(j) for j in range(nkeys)]
    # device downloads first, on this thread; then the oracle on the pool
    jobs = [(i, keys[which[i]].download(s), cin[i][0].download(), cin[i][1].download()) for i in idx]

    def one(job):
        i, (av, bv), c0, c1 = job
        return P.rotate(levels[i], c0, c1, ks[which[i]], av, bv)

    want = parallel_map(one, jobs)
    for i, (w0, w1_) in zip(idx, want):
        assert np.array_equal(got[i][0], w0), ("oracle c0", i, levels[i])
        assert np.array_equal(got[i][1], w1_), ("oracle c1", i, levels[i])


SERVE_RINGS = ((1 << 14, 0.3), (1 << 15, 0.4), (1 << 16, 0.3))
SERVE_L, SERVE_DNUM, SERVE_KEYS = 20, 4, 16


def test_cfg5_serving_trace_mix_equals_alone_and_oracle(G):
    """A serving trace at the configuration's rings (2^14 / 2^15 / 2^16 with
    weights 0.3 / 0.4 / 0.3), levels uniform in 3..20, 60% EvalMult+Rescale,
    30% Rotate (one of 16 resident keys per ring), 10% EvalAdd, arriving
    faster than one batch completes so requests batch and overlap.  Every
    result equals the same operation run alone, a sample of each ring and
    operation equals the oracle, EvalAdd equals the modular sum."""
    rng = np.random.default_rng(55)
    s = G.Stream()
    srv = G.Server(ws_budget_bytes=24 << 30, lanes=4, max_batch=32)
    rings = []
    for ri, (n, _) in enumerate(SERVE_RINGS):
        P = make_params(n, SERVE_L, SERVE_DNUM)
        ctx = G.Context(n, P.q, P.p, P.roots, SERVE_DNUM)
        ctx.autotune(s)
        rlk = G.Key.uniform(ctx, SERVE_DNUM, 1000 * ri + 1, s)
        rot_k = [pow(5, 1 << j, 2 * n) for j in range(SERVE_KEYS)]
        rks = [G.Key.uniform(ctx, SERVE_DNUM, 1000 * ri + 10 + j, s) for j in range(SERVE_KEYS)]
        rid = srv.add_ring(ctx, rlk, rks, rot_k)
        rings.append(dict(P=P, ctx=ctx, rlk=rlk, rks=rks, rot_k=rot_k, id=rid))
    count = 200
    ring = rng.choice(len(SERVE_RINGS), count, p=[w for _, w in SERVE_RINGS])
    u = rng.random(count)
    op = np.where(u < 0.6, G.OP_EVALMULT_RESCALE, np.where(u < 0.9, G.OP_ROTATE, G.OP_ADD))
    lv = rng.integers(3, SERVE_L + 1, count)
    rot = rng.integers(0, SERVE_KEYS, count)
    t = np.cumsum(rng.exponential(0.03, count))
    reqs, keep = [], []
    for i in range(count):
        R = rings[int(ring[i])]
        n, level = R["P"].n, int(lv[i])
        a = [filled(G, R["ctx"], level, 17 * i + j, s) for j in range(2)]
        b = [filled(G, R["ctx"], level, 17 * i + 2 + j, s) for j in range(2)]
        out_l = level - 1 if op[i] == G.OP_EVALMULT_RESCALE else level
        o = [G.Dev(out_l * n) for _ in range(2)]
        reqs.append(G.Request(R["id"], int(op[i]), level, int(rot[i]), a[0].ptr, a[1].ptr, b[0].ptr, b[1].ptr,
                              o[0].ptr, o[1].ptr, float(t[i])))
        keep.append((a, b, o, out_l))
    s.sync()
    comps = srv.run_trace(reqs)
    assert len(comps) == count
    assert max(c.batch_size for c in comps) > 1, "no batching happened"
    st = srv.stats()
    assert st["admitted"] >= count and st["batches"] < count
    got = [(o[0].download(), o[1].download()) for _, _, o, _ in keep]
    # alone: each request's operation through the single-op entry points
    w1 = {}
    for i in range(count):
        R = rings[int(ring[i])]
        ctx, n, level = R["ctx"], R["P"].n, int(lv[i])
        a, b, _, out_l = keep[i]
        if ctx.n not in w1:
            w1[ctx.n] = (ctx.workspace(), G.Dev(SERVE_L * n), G.Dev(SERVE_L * n))
        ws, x0, x1 = w1[ctx.n]
        if op[i] == G.OP_EVALMULT_RESCALE:
            G.check(G.lib.gf_evalmult_rescale(ctx.h, R["rlk"].h, level, a[0].ptr, a[1].ptr, b[0].ptr, b[1].ptr,
                                              x0.ptr, x1.ptr, ws.ptr, s.s))
        elif op[i] == G.OP_ROTATE:
            G.check(G.lib.gf_rotate(ctx.h, R["rks"][int(rot[i])].h, level, R["rot_k"][int(rot[i])], a[0].ptr,
                                    a[1].ptr, x0.ptr, x1.ptr, ws.ptr, s.s))
        else:
            G.check(G.lib.gf_add(ctx.h, level, a[0].ptr, a[1].ptr, b[0].ptr, b[1].ptr, x0.ptr, x1.ptr, s.s))
        s.sync()
        assert np.array_equal(x0.download(out_l * n), got[i][0]), ("alone", i, int(ring[i]), int(op[i]), level)
        assert np.array_equal(x1.download(out_l * n), got[i][1]), ("alone", i, int(ring[i]), int(op[i]), level)
    # oracle: per ring, the first two EvalMult+Rescale and the first two Rotate
    # requests, and every EvalAdd checked as the modular sum
    sample = []
    for r in range(len(rings)):
        for o_ in (G.OP_EVALMULT_RESCALE, G.OP_ROTATE):
            sample += [i for i in range(count) if ring[i] == r and op[i] == o_][:2]
    jobs = []
    for i in sample:
        R = rings[int(ring[i])]
        a, b, _, _ = keep[i]
        key = R["rlk"] if op[i] == G.OP_EVALMULT_RESCALE else R["rks"][int(rot[i])]
        av, bv = key.download(s)
        jobs.append((i, [x.download() for x in a], [x.download() for x in b], av, bv))

    def one(job):
        i, a, b, av, bv = job
        R = rings[int(ring[i])]
        if op[i] == G.OP_EVALMULT_RESCALE:
            return R["P"].evalmult_rescale(int(lv[i]), a[0], a[1], b[0], b[1], av, bv)
        return R["P"].rotate(int(lv[i]), a[0], a[1], R["rot_k"][int(rot[i])], av, bv)

    want = parallel_map(one, jobs)
    for (i, *_), (w0, w1_) in zip(jobs, want):
        assert np.array_equal(got[i][0], w0), ("oracle", i, int(ring[i]), int(op[i]), int(lv[i]))
        assert np.array_equal(got[i][1], w1_), ("oracle", i, int(ring[i]), int(op[i]), int(lv[i]))
    for i in range(count):
        if op[i] != G.OP_ADD:
            continue
        R = rings[int(ring[i])]
        n, level = R["P"].n, int(lv[i])
        a, b, _, _ = keep[i]
        q = np.repeat(R["P"].q[:level], n)
        for c_ in range(2):
            x = a[c_].download() + b[c_].download()   # residues below 2^61: no wrap
            want_add = np.where(x >= q, x - q, x)
            assert np.array_equal(got[i][c_], want_add), ("add", i)
