"""Composed operations through the handle C ABI vs the oracle, bit-exact:
hybrid keyswitch, EvalMult(+Rescale), Rescale, Rotate, EvalAdd; batch plans
over uneven levels; hipGraph replays; decrypt-level correctness."""
import numpy as np
import pytest

import oracle as O
from common import KeySet, decode, encode, make_params, mulmod_vec, poly_mul_negacyclic_small

pytestmark = pytest.mark.gpu

# (n, L, dnum): small ring, two-phase NTT ring, the benchmark configurations' shapes
SHAPES = [(1 << 10, 4, 2), (1 << 13, 6, 3), (1 << 15, 4, 2), (1 << 15, 12, 3), (1 << 16, 24, 4)]
# edge shapes: one digit (alpha = L), one tower per digit (alpha = 1), the
# largest ring the transforms support, the smallest ring; at a two-phase ring
# the most parts the fused row pass + inner product takes (7, more parts than
# waves) and one part more (the separate kernels); the 2^14 ring (the 128 x 128
# transform split); the benchmark ring with two digits (11 special primes), with 6 parts
# (fused, more parts than waves) and with 8 parts (the separate kernels)
EDGE_SHAPES = [(1 << 12, 4, 1), (1 << 12, 6, 6), (1 << 17, 3, 1), (1 << 4, 3, 3), (1 << 13, 7, 7),
               (1 << 13, 8, 8), (1 << 14, 5, 2), (1 << 16, 24, 2), (1 << 16, 24, 6), (1 << 16, 24, 8)]


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


_ctx_cache = {}


def setup(G, n, L, dnum, seed=1):
    """Context, random switching key (bit parity needs no valid key) and a
    random ciphertext generator."""
    key = (n, L, dnum, seed)
    if key in _ctx_cache:
        return _ctx_cache[key]
    P = make_params(n, L, dnum)
    rng = np.random.default_rng(seed)
    rows = P.L + P.K
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    k = G.Key(ctx, av, bv, P.dnum)
    assert av.size == P.dnum * rows * n
    out = (P, ctx, k, av, bv)
    _ctx_cache[key] = out
    return out


def rand_ct(P, level, rng):
    return np.concatenate([rng.integers(0, int(q), P.n, dtype=np.uint64) for q in P.q[:level]])


def levels_for(L, dnum):
    a = -(-L // dnum)
    return sorted({L, max(1, L - 1), min(L, a + 1), 1})


@pytest.mark.parametrize("n,L,dnum", SHAPES)
def test_keyswitch(G, n, L, dnum):
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(88)
    s = G.Stream()
    ws = ctx.workspace()
    for level in levels_for(L, dnum):
        d = rand_ct(P, level, rng)
        dd, b0, b1 = G.Dev.of(d), G.Dev(level * n), G.Dev(level * n)
        G.check(G.lib.gf_keyswitch(ctx.h, k.h, level, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.keyswitch(level, d, av, bv)
        assert np.array_equal(b0.download(), w0), ("ba0", level)
        assert np.array_equal(b1.download(), w1), ("ba1", level)


@pytest.mark.parametrize("n,L,dnum", EDGE_SHAPES)
def test_edge_shapes_keyswitch_evalmult_rotate(G, n, L, dnum):
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(404)
    s = G.Stream()
    ws = ctx.workspace()
    for level in sorted({L, 2, 1}):
        d = rand_ct(P, level, rng)
        dd, b0, b1 = G.Dev.of(d), G.Dev(level * n), G.Dev(level * n)
        G.check(G.lib.gf_keyswitch(ctx.h, k.h, level, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.keyswitch(level, d, av, bv)
        assert np.array_equal(b0.download(), w0) and np.array_equal(b1.download(), w1), ("ks", level)
        c1 = rand_ct(P, level, rng)
        d1 = G.Dev.of(c1)
        kk = 5 % (2 * n) if n > 2 else 1
        G.check(G.lib.gf_rotate(ctx.h, k.h, level, kk, dd.ptr, d1.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.rotate(level, d, c1, kk, av, bv)
        assert np.array_equal(b0.download(), w0) and np.array_equal(b1.download(), w1), ("rot", level)
        if level >= 2:
            cts = [rand_ct(P, level, rng) for _ in range(4)]
            dc = [G.Dev.of(c) for c in cts]
            o0, o1 = G.Dev((level - 1) * n), G.Dev((level - 1) * n)
            G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, level, *[x.ptr for x in dc], o0.ptr, o1.ptr, ws.ptr, s.s))
            s.sync()
            w0, w1 = P.evalmult_rescale(level, *cts, av, bv)
            assert np.array_equal(o0.download(), w0) and np.array_equal(o1.download(), w1), ("emr", level)


@pytest.mark.parametrize("n,L,dnum", SHAPES)
def test_evalmult_rescale(G, n, L, dnum):
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(101)
    s = G.Stream()
    ws = ctx.workspace()
    for level in sorted({L, 2}):
        cts = [rand_ct(P, level, rng) for _ in range(4)]
        d = [G.Dev.of(c) for c in cts]
        o0, o1 = G.Dev((level - 1) * n), G.Dev((level - 1) * n)
        G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, level, *[x.ptr for x in d], o0.ptr, o1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.evalmult_rescale(level, *cts, av, bv)
        assert np.array_equal(o0.download(), w0), level
        assert np.array_equal(o1.download(), w1), level


@pytest.mark.parametrize("n,L,dnum", [(1 << 13, 6, 3), (1 << 16, 5, 2)])
def test_evalmult_rescale_ascending_primes(G, n, L, dnum):
    """Q primes after q0 in ascending order, so the dropped modulus is larger
    than the towers that survive the rescale: the centred switch subtracts
    (q_last - q_i) mod q_i instead of adding q_i - q_last."""
    D = make_params(n, L, dnum)
    q = [int(D.q[0])] + sorted(int(v) for v in D.q[1:])
    P = O.Params(n, q, [int(v) for v in D.p], [int(O.lib.or_native_root(n, m)) for m in q + [int(v) for v in D.p]],
                 dnum)
    rng = np.random.default_rng(23)
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    k = G.Key(ctx, av, bv, P.dnum)
    s = G.Stream()
    ws = ctx.workspace()
    for fused in (True, False):
        ctx.set_fusion(fused)
        for level in sorted({L, 2}):
            cts = [rand_ct(P, level, rng) for _ in range(4)]
            d = [G.Dev.of(c) for c in cts]
            o0, o1 = G.Dev((level - 1) * n), G.Dev((level - 1) * n)
            G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, level, *[x.ptr for x in d], o0.ptr, o1.ptr, ws.ptr, s.s))
            s.sync()
            w0, w1 = P.evalmult_rescale(level, *cts, av, bv)
            assert np.array_equal(o0.download(), w0), (fused, level)
            assert np.array_equal(o1.download(), w1), (fused, level)


@pytest.mark.parametrize("n,L,dnum", [(1 << 13, 5, 2), (1 << 16, 4, 2)])
def test_wide_moduli_keyswitch_evalmult_rescale(G, n, L, dnum):
    """A 61-bit q0 and 61-bit special primes beside 50-bit towers: conversions
    take the full 128-bit products (moduli of 2^60 or more), the 61-bit rows
    keep the reduced transform butterflies and digits, the 50-bit rows take
    the lazy ones; every result equals the oracle."""
    P = make_params(n, L, dnum, qbits=50, q0bits=61, pbits=61)
    assert max(int(m) for m in P.mods) >> 60
    rng = np.random.default_rng(61)
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    k = G.Key(ctx, av, bv, P.dnum)
    s = G.Stream()
    ws = ctx.workspace()
    for level in sorted({L, 2}):
        d = rand_ct(P, level, rng)
        dd, b0, b1 = G.Dev.of(d), G.Dev(level * n), G.Dev(level * n)
        G.check(G.lib.gf_keyswitch(ctx.h, k.h, level, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.keyswitch(level, d, av, bv)
        assert np.array_equal(b0.download(), w0) and np.array_equal(b1.download(), w1), ("ks", level)
        cts = [rand_ct(P, level, rng) for _ in range(4)]
        dc = [G.Dev.of(c) for c in cts]
        o0, o1 = G.Dev((level - 1) * n), G.Dev((level - 1) * n)
        G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, level, *[x.ptr for x in dc], o0.ptr, o1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.evalmult_rescale(level, *cts, av, bv)
        assert np.array_equal(o0.download(), w0) and np.array_equal(o1.download(), w1), ("emr", level)


@pytest.mark.parametrize("n,L,dnum", [SHAPES[1], SHAPES[3]])
def test_evalmult_rescale_word_aligned_buffers(G, n, L, dnum):
    """Operands, outputs and workspace only 8-byte aligned (one word past a
    16-byte boundary): the streaming kernels take their one-residue-per-access
    path and give the same residues as the oracle."""
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(17)
    s = G.Stream()
    ws = G.Dev(ctx.ws_bytes // 8 + 1)
    for level in sorted({L, 2}):
        cts = [rand_ct(P, level, rng) for _ in range(4)]
        d = []
        for c in cts:
            b = G.Dev(c.size + 1)
            b.upload(c, word_off=1)
            d.append(b)
        o0, o1 = G.Dev((level - 1) * n + 1), G.Dev((level - 1) * n + 1)
        G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, level, *[x.at(1) for x in d], o0.at(1), o1.at(1),
                                          ws.at(1), s.s))
        s.sync()
        w0, w1 = P.evalmult_rescale(level, *cts, av, bv)
        assert np.array_equal(o0.download(word_off=1), w0), level
        assert np.array_equal(o1.download(word_off=1), w1), level
        r0, r1 = G.Dev(level * n + 1), G.Dev(level * n + 1)
        G.check(G.lib.gf_evalmult(ctx.h, k.h, level, *[x.at(1) for x in d], r0.at(1), r1.at(1), ws.at(1), s.s))
        s.sync()
        w0, w1 = P.evalmult(level, *cts, av, bv)
        assert np.array_equal(r0.download(word_off=1), w0), level
        assert np.array_equal(r1.download(word_off=1), w1), level


@pytest.mark.parametrize("n,L,dnum", [SHAPES[1], SHAPES[4]])
def test_evalmult_rescale_unfused_sequence(G, n, L, dnum):
    """The reference's stage sequence (ModDown NTTs, combine, INTT/switch/NTT/
    fuse) and the merged default give the same residues as the oracle."""
    P = make_params(n, L, dnum)
    rng = np.random.default_rng(3)
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    ctx.set_fusion(False)
    k = G.Key(ctx, av, bv, P.dnum)
    s = G.Stream()
    ws = ctx.workspace()
    cts = [rand_ct(P, L, rng) for _ in range(4)]
    d = [G.Dev.of(c) for c in cts]
    o0, o1 = G.Dev((L - 1) * n), G.Dev((L - 1) * n)
    G.check(G.lib.gf_evalmult_rescale(ctx.h, k.h, L, *[x.ptr for x in d], o0.ptr, o1.ptr, ws.ptr, s.s))
    s.sync()
    w0, w1 = P.evalmult_rescale(L, *cts, av, bv)
    assert np.array_equal(o0.download(), w0) and np.array_equal(o1.download(), w1)


@pytest.mark.parametrize("n,L,dnum", SHAPES[:3])
def test_evalmult_norescale_and_rescale(G, n, L, dnum):
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(7)
    s = G.Stream()
    ws = ctx.workspace()
    cts = [rand_ct(P, L, rng) for _ in range(4)]
    d = [G.Dev.of(c) for c in cts]
    r0, r1 = G.Dev(L * n), G.Dev(L * n)
    G.check(G.lib.gf_evalmult(ctx.h, k.h, L, *[x.ptr for x in d], r0.ptr, r1.ptr, ws.ptr, s.s))
    s.sync()
    w0, w1 = P.evalmult(L, *cts, av, bv)
    assert np.array_equal(r0.download(), w0) and np.array_equal(r1.download(), w1)
    o0, o1 = G.Dev((L - 1) * n), G.Dev((L - 1) * n)
    G.check(G.lib.gf_rescale(ctx.h, L, r0.ptr, r1.ptr, o0.ptr, o1.ptr, ws.ptr, s.s))
    s.sync()
    assert np.array_equal(o0.download(), P.rescale(w0, L))
    assert np.array_equal(o1.download(), P.rescale(w1, L))


@pytest.mark.parametrize("n,L,dnum", [SHAPES[0], SHAPES[1], (1 << 16, 20, 4)])
def test_rotate(G, n, L, dnum):
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(404)
    s = G.Stream()
    ws = ctx.workspace()
    for level, kk in ((L, 5), (max(1, L // 2), pow(5, 3, 2 * n)), (L, 2 * n - 1)):
        c0, c1 = rand_ct(P, level, rng), rand_ct(P, level, rng)
        d0, d1 = G.Dev.of(c0), G.Dev.of(c1)
        o0, o1 = G.Dev(level * n), G.Dev(level * n)
        G.check(G.lib.gf_rotate(ctx.h, k.h, level, kk, d0.ptr, d1.ptr, o0.ptr, o1.ptr, ws.ptr, s.s))
        s.sync()
        w0, w1 = P.rotate(level, c0, c1, kk, av, bv)
        assert np.array_equal(o0.download(), w0), (level, kk)
        assert np.array_equal(o1.download(), w1), (level, kk)


def test_add(G):
    P, ctx, k, av, bv = setup(G, *SHAPES[1])
    n, L = P.n, P.L
    rng = np.random.default_rng(5)
    cts = [rand_ct(P, L, rng) for _ in range(4)]
    d = [G.Dev.of(c) for c in cts]
    o0, o1 = G.Dev(L * n), G.Dev(L * n)
    s = G.Stream()
    G.check(G.lib.gf_add(ctx.h, L, *[x.ptr for x in d], o0.ptr, o1.ptr, s.s))
    s.sync()
    for o, a, b in ((o0, cts[0], cts[2]), (o1, cts[1], cts[3])):
        w = a.copy()
        for t in range(L):
            O.lib.or_add_into(O.P(w[t * n:(t + 1) * n]), O.P(b[t * n:(t + 1) * n].copy()), int(P.q[t]), n)
        assert np.array_equal(o.download(), w)


def test_batch_plan_uneven_levels_and_graph(G):
    """One plan over ops of uneven level == per-op results; graph replays ==
    eager, with eager work between replays."""
    n, L, dnum = 1 << 13, 6, 3
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(2)
    levels = [6, 2, 5, 3, 6]
    ins, outs, wss = [], [], []
    for lv in levels:
        ins.append([G.Dev.of(rand_ct(P, lv, rng)) for _ in range(4)])
        outs.append([G.Dev((lv - 1) * n), G.Dev((lv - 1) * n)])
        wss.append(ctx.workspace())
    col = lambda j: [x[j].ptr for x in ins]  # noqa: E731
    plan = G.plan_evalmult_rescale(ctx, k, levels, col(0), col(1), col(2), col(3),
                                   [o[0].ptr for o in outs], [o[1].ptr for o in outs], [w.ptr for w in wss])
    s = G.Stream()
    plan.run(s)
    s.sync()
    want = []
    for i, lv in enumerate(levels):
        cts = [x.download() for x in ins[i]]
        w0, w1 = P.evalmult_rescale(lv, *cts, av, bv)
        want.append((w0, w1))
        assert np.array_equal(outs[i][0].download(), w0), i
        assert np.array_equal(outs[i][1].download(), w1), i
    for rep in range(3):
        for o in outs:
            o[0].upload(np.zeros(o[0].words, np.uint64))
            o[1].upload(np.zeros(o[1].words, np.uint64))
        plan.run_graph(s)
        s.sync()
        # unrelated eager work between replays
        G.check(G.lib.gf_add(ctx.h, 2, ins[1][0].ptr, ins[1][1].ptr, ins[1][2].ptr, ins[1][3].ptr,
                             wss[0].ptr, wss[0].at(2 * n), s.s))
        s.sync()
        for i in range(len(levels)):
            assert np.array_equal(outs[i][0].download(), want[i][0]), (rep, i)
            assert np.array_equal(outs[i][1].download(), want[i][1]), (rep, i)
    assert plan.launches < 30
    # the same jobs, one launch each (the per-tower launch structure)
    for o in outs:
        o[0].upload(np.zeros(o[0].words, np.uint64))
        o[1].upload(np.zeros(o[1].words, np.uint64))
    plan.run_per_job(s)
    s.sync()
    for i in range(len(levels)):
        assert np.array_equal(outs[i][0].download(), want[i][0]), ("per job", i)
        assert np.array_equal(outs[i][1].download(), want[i][1]), ("per job", i)


@pytest.mark.parametrize("ksi_extra,conv_chunk,rs_chunk,modup_fused,modup_f64,ksi_f64,moddown_fused",
                         [(0, 8, 8, 1, 1, 1, 1), (1, 16, 32, 1, 0, 0, 1), (1, 32, 16, 0, 1, 1, 1),
                          (0, 5, 3, 0, 0, 0, 0), (0, 8, 16, 1, 1, 0, 0)])
def test_plan_knobs_keep_results(G, ksi_extra, conv_chunk, rs_chunk, modup_fused, modup_f64, ksi_f64,
                                 moddown_fused):
    """Every setting of the autotuner's plan knobs (extra inner-product wave,
    conversion and prologue chunk widths, re-chunking off so the widths take
    effect, the fused or the unfused ModUp, the fp64 passes of the 50-bit
    towers around the fused ModUp and in the fused inner product, the fused
    or the unfused ModDown + rescale prologue, the tensor product fused with
    the row pass after it or apart, the P part's inverse row pass inside the
    P rows' inner product or apart) gives the
    oracle's results, on a batch of uneven levels (parts of 3 and 2 towers;
    ragged chunks with widths 5 and 3); the autotuner's own search leaves a
    report naming its choices."""
    n, L, dnum = 1 << 13, 6, 3
    P, ctx, k, av, bv = setup(G, n, L, dnum, seed=77)   # a context of its own: the knobs stay here
    assert any(int(q) < 1 << 50 for q in P.q)   # the fp64 passes have towers to take
    for name, v in (("ksi_extra", ksi_extra), ("conv_chunk", conv_chunk), ("rs_chunk", rs_chunk),
                    ("modup_fused", modup_fused), ("modup_f64", modup_f64), ("ksi_f64", ksi_f64),
                    ("moddown_fused", moddown_fused), ("tensor_ib", 0 if conv_chunk == 16 else 1),
                    ("ksi_rowpass", 0 if conv_chunk == 32 else 1),
                    ("modup_min_wgs", 0), ("rechunk_wgs", 0)):
        ctx.set_knob(name, v)
        assert ctx.knob(name) == v
    rng = np.random.default_rng(11 + conv_chunk)
    levels = [6, 4, 5, 2]
    ins = [[G.Dev.of(rand_ct(P, lv, rng)) for _ in range(4)] for lv in levels]
    outs = [[G.Dev((lv - 1) * n), G.Dev((lv - 1) * n)] for lv in levels]
    wss = [ctx.workspace() for _ in levels]
    col = lambda j: [x[j].ptr for x in ins]  # noqa: E731
    plan = G.plan_evalmult_rescale(ctx, k, levels, col(0), col(1), col(2), col(3),
                                   [o[0].ptr for o in outs], [o[1].ptr for o in outs], [w.ptr for w in wss])
    s = G.Stream()
    plan.run_graph(s)
    s.sync()
    for i, lv in enumerate(levels):
        w0, w1 = P.evalmult_rescale(lv, *[x.download() for x in ins[i]], av, bv)
        assert np.array_equal(outs[i][0].download(), w0), (i, lv)
        assert np.array_equal(outs[i][1].download(), w1), (i, lv)
    with pytest.raises(G.GpuFheError, match="unknown tuning knob"):
        ctx.set_knob("no_such_knob", 1)
    if ksi_extra == 0 and conv_chunk == 8:
        ctx.set_knob("rechunk_wgs", 8192)
        ctx.autotune(s)
        rep = ctx.tuning_report()
        for name in ("ntt_radix", "ksi_extra", "conv_chunk", "rs_chunk", "modup_fused", "modup_f64"):
            assert rep[name]["chosen"] == ctx.knob(name), name
        if ctx.knob("modup_fused"):   # the fused ModDown is weighed where the fused ModUp runs
            assert rep["moddown_fused"]["chosen"] == ctx.knob("moddown_fused")
        if ctx.knob("modup_f64"):
            assert rep["ksi_f64"]["chosen"] == ctx.knob("ksi_f64")


def test_fused_kernels_largest_ring(G):
    """The fused ModUp, the fused ModDown and the fused row pass + inner
    product (fp64 passes on) at the largest two-phase ring, n = 2^17 (a
    512-row column pass: each thread stages two twiddle pairs per phase),
    forced on for a small batch of uneven levels: the oracle's results."""
    n, L, dnum = 1 << 17, 4, 2
    P, ctx, k, av, bv = setup(G, n, L, dnum, seed=91)
    for name, v in (("modup_fused", 1), ("modup_f64", 1), ("ksi_f64", 1), ("moddown_fused", 1),
                    ("modup_min_wgs", 0), ("rechunk_wgs", 0)):
        ctx.set_knob(name, v)
    rng = np.random.default_rng(17)
    levels = [4, 3]
    ins = [[G.Dev.of(rand_ct(P, lv, rng)) for _ in range(4)] for lv in levels]
    outs = [[G.Dev((lv - 1) * n), G.Dev((lv - 1) * n)] for lv in levels]
    wss = [ctx.workspace() for _ in levels]
    col = lambda j: [x[j].ptr for x in ins]  # noqa: E731
    plan = G.plan_evalmult_rescale(ctx, k, levels, col(0), col(1), col(2), col(3),
                                   [o[0].ptr for o in outs], [o[1].ptr for o in outs], [w.ptr for w in wss])
    kinds = [kd for kd, _ in plan.steps()]
    assert "modup" in kinds and "moddown" in kinds and "ksi" in kinds, kinds
    s = G.Stream()
    plan.run_graph(s)
    s.sync()
    for i, lv in enumerate(levels):
        w0, w1 = P.evalmult_rescale(lv, *[x.download() for x in ins[i]], av, bv)
        assert np.array_equal(outs[i][0].download(), w0), (i, lv)
        assert np.array_equal(outs[i][1].download(), w1), (i, lv)


def test_rotate_plan_multi_keys(G):
    n, L, dnum = 1 << 12, 5, 5
    P = make_params(n, L, dnum)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rng = np.random.default_rng(9)
    keys, hk = [], []
    for _ in range(3):
        av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(dnum) for m in P.mods])
        bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(dnum) for m in P.mods])
        keys.append((av, bv))
        hk.append(G.Key(ctx, av, bv, dnum))
    ks = [5, 25, 125]
    levels = [5, 3, 4]
    c = [(rand_ct(P, lv, rng), rand_ct(P, lv, rng)) for lv in levels]
    dc = [(G.Dev.of(a), G.Dev.of(b)) for a, b in c]
    do = [(G.Dev(lv * n), G.Dev(lv * n)) for lv in levels]
    ws = [ctx.workspace() for _ in levels]
    plan = G.plan_rotate(ctx, hk, levels, ks, [x[0].ptr for x in dc], [x[1].ptr for x in dc],
                         [x[0].ptr for x in do], [x[1].ptr for x in do], [w.ptr for w in ws])
    s = G.Stream()
    plan.run(s)
    s.sync()
    for i, lv in enumerate(levels):
        w0, w1 = P.rotate(lv, c[i][0], c[i][1], ks[i], *keys[i])
        assert np.array_equal(do[i][0].download(), w0) and np.array_equal(do[i][1].download(), w1)


@pytest.mark.parametrize("n,L,dnum", [(1 << 12, 6, 3), (1 << 16, 20, 4)])
def test_rotate_hoisted_vs_oracle(G, n, L, dnum):
    """Hoisted rotations (one decomposition per ciphertext, several indices)
    bit-exact against the oracle's hoisted restatement, uneven levels."""
    P = make_params(n, L, dnum)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rng = np.random.default_rng(31)
    m = 3
    keys, hk = [], []
    for _ in range(m):
        av = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(dnum) for q in P.mods])
        bv = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(dnum) for q in P.mods])
        keys.append((av, bv))
        hk.append(G.Key(ctx, av, bv, dnum))
    ks = [pow(5, 1 << j, 2 * n) for j in range(m)]
    levels = [L, L - 2] if n < (1 << 16) else [L - 7]
    cts = [(rand_ct(P, lv, rng), rand_ct(P, lv, rng)) for lv in levels]
    dc = [(G.Dev.of(a), G.Dev.of(b)) for a, b in cts]
    do = [(G.Dev(lv * n), G.Dev(lv * n)) for lv in levels for _ in range(m)]
    wb = [ctx.workspace() for _ in levels]
    wr = [ctx.workspace() for _ in range(len(levels) * m)]
    plan = G.plan_rotate_hoisted(ctx, levels, [x[0].ptr for x in dc], [x[1].ptr for x in dc], hk, ks,
                                 [x[0].ptr for x in do], [x[1].ptr for x in do], [w.ptr for w in wb],
                                 [w.ptr for w in wr])
    s = G.Stream()
    plan.run(s)
    s.sync()
    for i, lv in enumerate(levels):
        want = P.rotate_hoisted(lv, cts[i][0], cts[i][1], ks, keys)
        for j in range(m):
            assert np.array_equal(do[i * m + j][0].download(), want[j][0]), (i, j)
            assert np.array_equal(do[i * m + j][1].download(), want[j][1]), (i, j)


@pytest.mark.parametrize("n,L,dnum", [(1 << 12, 6, 3), (1 << 16, 20, 4)])
def test_linear_transform_vs_oracle(G, n, L, dnum):
    """Plaintext-weighted sum of hoisted rotations with an unrotated term
    (no key, k = 1), bit-exact against the oracle, two ciphertexts at uneven
    levels in one plan, replayed twice."""
    P = make_params(n, L, dnum)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rng = np.random.default_rng(33)
    ks = [1, pow(5, 1, 2 * n), pow(5, 6, 2 * n)]
    m = len(ks)
    keys, hk = [None], [None]
    for _ in range(m - 1):
        av = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(dnum) for q in P.mods])
        bv = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(dnum) for q in P.mods])
        keys.append((av, bv))
        hk.append(G.Key(ctx, av, bv, dnum))
    pts = [rand_ct(P, L, rng) for _ in range(m)]
    dp = [G.Dev.of(p) for p in pts]
    levels = [L, L - 3] if n < (1 << 16) else [L - 5]
    cts = [(rand_ct(P, lv, rng), rand_ct(P, lv, rng)) for lv in levels]
    dc = [(G.Dev.of(a), G.Dev.of(b)) for a, b in cts]
    do = [(G.Dev(lv * n), G.Dev(lv * n)) for lv in levels]
    wb = [ctx.workspace() for _ in levels]
    wr = [ctx.workspace() for _ in range(len(levels) * m)]
    plan = G.plan_linear_transform(ctx, levels, [x[0].ptr for x in dc], [x[1].ptr for x in dc], hk, ks,
                                   [p.ptr for p in dp], [x[0].ptr for x in do], [x[1].ptr for x in do],
                                   [w.ptr for w in wb], [w.ptr for w in wr])
    s = G.Stream()
    want = [P.linear_transform(lv, cts[i][0], cts[i][1], ks, keys, [p[: lv * n] for p in pts])
            for i, lv in enumerate(levels)]
    for rep in range(2):
        plan.run(s)
        s.sync()
        for i in range(len(levels)):
            assert np.array_equal(do[i][0].download(), want[i][0]), (rep, i)
            assert np.array_equal(do[i][1].download(), want[i][1]), (rep, i)


def test_plaintext_ops_vs_oracle(G):
    """Plaintext multiply and add (src/ptops.cpp:1-26 on the device), batch
    form at uneven levels and single form in place, against per-tower
    oracle products and sums."""
    P = make_params(1 << 12, 6, 3)
    n = P.n
    ctx = G.Context(n, P.q, P.p, P.roots, 3)
    rng = np.random.default_rng(35)
    levels = [6, 4, 1]
    cts = [(rand_ct(P, lv, rng), rand_ct(P, lv, rng)) for lv in levels]
    pts = [rand_ct(P, lv, rng) for lv in levels]
    dc = [(G.Dev.of(a), G.Dev.of(b)) for a, b in cts]
    dp = [G.Dev.of(p) for p in pts]
    om = [(G.Dev(lv * n), G.Dev(lv * n)) for lv in levels]
    oa = [(G.Dev(lv * n), G.Dev(lv * n)) for lv in levels]
    pm = G.plan_mul_plain(ctx, levels, [x[0].ptr for x in dc], [x[1].ptr for x in dc], [p.ptr for p in dp],
                          [x[0].ptr for x in om], [x[1].ptr for x in om])
    pa = G.plan_add_plain(ctx, levels, [x[0].ptr for x in dc], [x[1].ptr for x in dc], [p.ptr for p in dp],
                          [x[0].ptr for x in oa], [x[1].ptr for x in oa])
    s = G.Stream()
    pm.run(s)
    pa.run(s)
    s.sync()

    def per_tower(f, a, b, lv):
        return np.concatenate([f(a[t * n:(t + 1) * n], b[t * n:(t + 1) * n], P.q[t]) for t in range(lv)])

    addv = lambda a, b, q: ((a.astype(object) + b.astype(object)) % int(q)).astype(np.uint64)  # noqa: E731
    for i, lv in enumerate(levels):
        assert np.array_equal(om[i][0].download(), per_tower(mulmod_vec, cts[i][0], pts[i], lv)), i
        assert np.array_equal(om[i][1].download(), per_tower(mulmod_vec, cts[i][1], pts[i], lv)), i
        assert np.array_equal(oa[i][0].download(), per_tower(addv, cts[i][0], pts[i], lv)), i
        assert np.array_equal(oa[i][1].download(), cts[i][1]), i
    # single-op forms, in place
    G.mul_plain(ctx, 6, dc[0][0].ptr, dc[0][1].ptr, dp[0].ptr, dc[0][0].ptr, dc[0][1].ptr, s)
    s.sync()
    assert np.array_equal(dc[0][0].download(), om[0][0].download())
    assert np.array_equal(dc[0][1].download(), om[0][1].download())
    G.add_plain(ctx, 4, dc[1][0].ptr, dc[1][1].ptr, dp[1].ptr, dc[1][0].ptr, dc[1][1].ptr, s)
    s.sync()
    assert np.array_equal(dc[1][0].download(), oa[1][0].download())
    assert np.array_equal(dc[1][1].download(), cts[1][1])


def test_single_op_hoisted_and_linear_transform(G):
    """gf_rotate_hoisted and gf_linear_transform (one ciphertext, plan cached
    by arguments) against the oracle, called twice."""
    P = make_params(1 << 12, 6, 3)
    n, lv = P.n, 5
    ctx = G.Context(n, P.q, P.p, P.roots, 3)
    rng = np.random.default_rng(37)
    ks = [pow(5, 2, 2 * n), pow(5, 9, 2 * n)]
    keys, hk = [], []
    for _ in ks:
        av = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(3) for q in P.mods])
        bv = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for _ in range(3) for q in P.mods])
        keys.append((av, bv))
        hk.append(G.Key(ctx, av, bv, 3))
    c0, c1 = rand_ct(P, lv, rng), rand_ct(P, lv, rng)
    d0, d1 = G.Dev.of(c0), G.Dev.of(c1)
    outs = [(G.Dev(lv * n), G.Dev(lv * n)) for _ in ks]
    wb = ctx.workspace()
    wr = [ctx.workspace() for _ in ks]
    s = G.Stream()
    want = P.rotate_hoisted(lv, c0, c1, ks, keys)
    for _ in range(2):
        G.rotate_hoisted(ctx, lv, d0.ptr, d1.ptr, hk, ks, [o[0].ptr for o in outs], [o[1].ptr for o in outs],
                         wb.ptr, [w.ptr for w in wr], s)
        s.sync()
        for j in range(len(ks)):
            assert np.array_equal(outs[j][0].download(), want[j][0]), j
            assert np.array_equal(outs[j][1].download(), want[j][1]), j
    pts = [rand_ct(P, lv, rng) for _ in range(3)]
    dp = [G.Dev.of(p) for p in pts]
    o0, o1 = G.Dev(lv * n), G.Dev(lv * n)
    lk = [1] + ks
    w0, w1 = P.linear_transform(lv, c0, c1, lk, [None] + keys, pts)
    for _ in range(2):
        G.linear_transform(ctx, lv, d0.ptr, d1.ptr, [None] + hk, lk, [p.ptr for p in dp], o0.ptr, o1.ptr, wb.ptr,
                           [0] + [w.ptr for w in wr], s)
        s.sync()
        assert np.array_equal(o0.download(), w0) and np.array_equal(o1.download(), w1)


def test_linear_transform_rejects_keyless_rotation(G):
    P = make_params(1 << 10, 3, 3)
    ctx = G.Context(P.n, P.q, P.p, P.roots, 3)
    d = G.Dev(3 * P.n)
    ws = ctx.workspace()
    with pytest.raises(RuntimeError):
        G.plan_linear_transform(ctx, [3], [d.ptr], [d.ptr], [None], [5], [d.ptr], [d.ptr], [d.ptr],
                                [ws.ptr], [ws.ptr])


@pytest.mark.parametrize("n,L,dnum", [(1 << 12, 4, 2), (1 << 13, 6, 3)])
def test_device_keygen_matches_oracle(G, n, L, dnum):
    """gf_key_generate (relinearisation and rotation keys from device-resident
    secret and public key) equals the oracle's evalkeygen, and a key uploaded
    from the host reads back unchanged."""
    P = make_params(n, L, dnum)
    ks = KeySet(P)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    dpa, dpb = G.Dev.of(ks.pkAQP), G.Dev.of(ks.pkBQP)
    for sold, seed in ((ks.s2(), 202), (P.automorph(ks.sQP, P.L + P.K, 25, P.K), 77)):
        ds = G.Dev.of(sold)
        key = G.Key.generate(ctx, ds, dpa, dpb, seed)
        av, bv = key.download()
        wa, wb = O.evalkeygen_sold(P.n, P.L, P.K, P.alpha, P.dnum, P.mods, P.roots, sold, ks.pkAQP,
                                   ks.pkBQP, seed=seed)
        assert np.array_equal(av, wa) and np.array_equal(bv, wb), seed
    up = G.Key(ctx, wa, wb, dnum)
    ra, rb = up.download()
    assert np.array_equal(ra, wa) and np.array_equal(rb, wb)


@pytest.mark.parametrize("n,L,dnum", [(1 << 12, 4, 2), (1 << 13, 6, 3), (1 << 16, 24, 4)])
def test_device_keygen_counter_sampler_matches_oracle(G, n, L, dnum):
    """Keys generated entirely on the device (gf_keygen: s, e, a drawn by the
    counter sampler, embedded, transformed, b = ns e - a s; then
    gf_key_generate_ctr: the relinearisation key of s^2 under that public
    key) equal the oracle's counter-sampler key generation bit for bit, up to
    n = 2^16, L = 24, dnum = 4."""
    P = make_params(n, L, dnum)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    s = G.Stream()
    ds, da, db = G.keygen(ctx, 4242, stream=s)
    ws_, wa, wb = O.keygen_ctr(n, P.mods, P.roots, seed=4242)
    assert np.array_equal(ds.download(), ws_), "s"
    assert np.array_equal(da.download(), wa), "pkA"
    assert np.array_equal(db.download(), wb), "pkB"
    W = P.L + P.K
    s2 = np.concatenate([mulmod_vec(ws_[t * n:(t + 1) * n], ws_[t * n:(t + 1) * n], int(P.mods[t])) for t in range(W)])
    key = G.Key.generate_ctr(ctx, G.Dev.of(s2), da, db, 5151, stream=s)
    av, bv = key.download(s)
    ea, eb = O.evalkeygen_ctr(n, P.L, P.K, P.alpha, P.dnum, P.mods, P.roots, s2, wa, wb, seed=5151)
    assert np.array_equal(av, ea) and np.array_equal(bv, eb), "switching key"


@pytest.mark.parametrize("n,L,dnum", [(1 << 12, 3, 3), (1 << 13, 6, 3), (1 << 16, 24, 4)])
def test_device_encrypt_decrypt_match_oracle(G, n, L, dnum):
    """gf_encrypt equals the oracle's encryption (same sampler and seed) bit
    for bit at the top level and at level 1, gf_decrypt equals the oracle's
    decryption, and a fresh ciphertext decrypts to its message up to the
    encryption noise."""
    P = make_params(n, L, dnum)
    ks = KeySet(P)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rng = np.random.default_rng(n + L)
    m = np.ascontiguousarray(rng.integers(-(1 << 20), 1 << 20, n).astype(np.int64))
    dpa, dpb, dsk = G.Dev.of(ks.pkA), G.Dev.of(ks.pkB), G.Dev.of(ks.sQ)
    s = G.Stream()
    for level in sorted({L, 1}):
        seed = 303 + level
        c0, c1 = G.Dev(level * n), G.Dev(level * n)
        G.check(G.lib.gf_encrypt(ctx.h, level, dpa.ptr, dpb.ptr, m.ctypes.data, 1, 3.19, seed, c0.ptr, c1.ptr,
                                 s.s))
        w0, w1 = O.encrypt(n, P.q[:level], P.roots[:level], m, ks.pkA[:level * n], ks.pkB[:level * n], seed=seed)
        assert np.array_equal(c0.download(), w0) and np.array_equal(c1.download(), w1), level
        out = np.zeros(n, np.int64)
        G.check(G.lib.gf_decrypt(ctx.h, c0.ptr, c1.ptr, dsk.ptr, out.ctypes.data, s.s))
        want = O.decrypt(n, int(P.q[0]), int(P.roots[0]), w0[:n], w1[:n], ks.sQ[:n])
        assert np.array_equal(out, want), level
        # encryption noise v e + e1 s + e0: standard deviation about sigma sqrt(4n/3)
        assert int(np.abs(out - m).max()) < 8 * 3.19 * (4 * n / 3) ** 0.5 + 50, level


def test_chain_plans_graphs(G):
    """Multiply-rescale chain (tests/test_chain.cpp shape): a batch of
    ciphertexts multiplied depth times, one plan per step over the batch,
    replayed as graphs; every step equals the oracle's sequential result."""
    n, L, dnum, batch, depth = 1 << 12, 6, 3, 3, 4
    P, ctx, k, av, bv = setup(G, n, L, dnum)
    rng = np.random.default_rng(21)
    cur = [rand_ct(P, L, rng) for _ in range(2 * batch)]
    cts = [[G.Dev.of(cur[2 * b]), G.Dev.of(cur[2 * b + 1])] for b in range(batch)]
    chain = [cts]
    opnds, host_ops = [], []
    for d in range(depth):
        lv = L - d
        ho = [(rand_ct(P, lv, rng), rand_ct(P, lv, rng)) for _ in range(batch)]
        host_ops.append(ho)
        opnds.append([[G.Dev.of(a), G.Dev.of(b)] for a, b in ho])
        chain.append([[G.Dev((lv - 1) * n), G.Dev((lv - 1) * n)] for _ in range(batch)])
    ws = [ctx.workspace() for _ in range(batch)]
    plans = []
    for d in range(depth):
        src, dst, op = chain[d], chain[d + 1], opnds[d]
        plans.append(G.plan_evalmult_rescale(ctx, k, [L - d] * batch, [x[0].ptr for x in src], [x[1].ptr for x in src],
                                             [x[0].ptr for x in op], [x[1].ptr for x in op],
                                             [x[0].ptr for x in dst], [x[1].ptr for x in dst], [w.ptr for w in ws]))
    s = G.Stream()
    for rep in range(2):
        for pl in plans:
            pl.run_graph(s)
        s.sync()
        for b in range(batch):
            c0, c1 = cur[2 * b], cur[2 * b + 1]
            for d in range(depth):
                c0, c1 = P.evalmult_rescale(L - d, c0, c1, host_ops[d][b][0], host_ops[d][b][1], av, bv)
                assert np.array_equal(chain[d + 1][b][0].download(), c0), (rep, b, d)
                assert np.array_equal(chain[d + 1][b][1].download(), c1), (rep, b, d)


def test_decrypt_exact_product_cfg1_shape(G):
    """tests/test_evalmult_rescale.cpp bar on the GPU path: exact small-integer
    product after round(dec * q_l / D^2), n = 2^15, 60 + 3 x 50-bit Q, dnum 2."""
    P = make_params(1 << 15, 4, 2)
    ks = KeySet(P)
    av, bv = ks.rlk()
    n, L = P.n, P.L
    ctx = G.Context(n, P.q, P.p, P.roots, P.dnum)
    key = G.Key(ctx, av, bv, P.dnum)
    D = 1 << 40
    a1 = np.zeros(n, np.int64)
    a2 = np.zeros(n, np.int64)
    a1[0], a1[1], a1[n - 1] = 3, -2, 7
    a2[0], a2[2] = 4, 5
    c0a, c1a = O.encrypt(n, P.q, P.roots[:L], a1 * D, ks.pkA, ks.pkB, seed=303)
    c0b, c1b = O.encrypt(n, P.q, P.roots[:L], a2 * D, ks.pkA, ks.pkB, seed=304)
    d = [G.Dev.of(x) for x in (c0a, c1a, c0b, c1b)]
    o0, o1 = G.Dev((L - 1) * n), G.Dev((L - 1) * n)
    ws = ctx.workspace()
    s = G.Stream()
    G.check(G.lib.gf_evalmult_rescale(ctx.h, key.h, L, *[x.ptr for x in d], o0.ptr, o1.ptr, ws.ptr, s.s))
    s.sync()
    dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0.download(), o1.download(), ks.sQ)
    ql = int(P.q[L - 1])
    got = np.array([round(int(x) * ql / (D * D)) for x in dec], dtype=object)
    assert np.max(np.abs((got - poly_mul_negacyclic_small(a1, a2, n)).astype(np.int64))) == 0


def test_decrypt_slots_rotation(G):
    P = make_params(1 << 12, 4, 2)
    ks = KeySet(P)
    n, L = P.n, P.L
    ctx = G.Context(n, P.q, P.p, P.roots, P.dnum)
    z = np.random.default_rng(4).uniform(-1, 1, n // 2)
    delta = float(1 << 45)
    c0, c1 = O.encrypt(n, P.q, P.roots[:L], encode(z, n, delta), ks.pkA, ks.pkB, seed=505)
    kk = pow(5, 2, 2 * n)
    av, bv = ks.rotkey(kk)
    key = G.Key(ctx, av, bv, P.dnum)
    d0, d1, o0, o1 = G.Dev.of(c0), G.Dev.of(c1), G.Dev(L * n), G.Dev(L * n)
    ws = ctx.workspace()
    s = G.Stream()
    G.check(G.lib.gf_rotate(ctx.h, key.h, L, kk, d0.ptr, d1.ptr, o0.ptr, o1.ptr, ws.ptr, s.s))
    s.sync()
    dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0.download(), o1.download(), ks.sQ)
    assert np.max(np.abs(decode(dec, n, delta).real - np.roll(z, -2))) < 1e-6


def test_key_destroy_evicts_cached_plans(G):
    """A single-op plan holds its key's device pointers: destroying the key
    evicts it, and a new key on the same buffers (whose handle may reuse the
    freed address) gets a plan of its own.  Key B has a different row layout
    (fullQ = L + 2), so a stale plan would read the wrong rows."""
    n, L, dnum = 1 << 12, 6, 3
    P = make_params(n, L, dnum)
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rng = np.random.default_rng(77)
    s = G.Stream()
    ws = ctx.workspace()
    d = rand_ct(P, L, rng)
    dd, b0, b1 = G.Dev.of(d), G.Dev(L * n), G.Dev(L * n)

    def rand_rows(rows_mods):
        return np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(dnum) for m in rows_mods])

    av, bv = rand_rows(P.mods), rand_rows(P.mods)
    ka = G.Key(ctx, av, bv, dnum)
    G.check(G.lib.gf_keyswitch(ctx.h, ka.h, L, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
    s.sync()
    w0, w1 = P.keyswitch(L, d, av, bv)
    assert np.array_equal(b0.download(), w0) and np.array_equal(b1.download(), w1)
    before = ctx.cached_plans()
    assert before >= 1
    del ka   # gf_key_destroy
    assert ctx.cached_plans() == before - 1
    # key B: rows = fullQ + K with fullQ = L + 2 (two unused Q rows before the P rows)
    fullQ = L + 2
    row_mods = list(P.q) + [P.q[0], P.q[1]] + list(P.p)
    av2, bv2 = rand_rows(row_mods), rand_rows(row_mods)
    h = G.lib.gf_key_create_rows(ctx.h, G.P64(av2), G.P64(bv2), dnum, fullQ + P.K, fullQ, None)
    assert h, G.lib.gf_last_error()
    G.check(G.lib.gf_keyswitch(ctx.h, h, L, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
    s.sync()
    # the same key in the context's own row layout for the oracle
    rows = fullQ + P.K
    pick = list(range(L)) + [fullQ + k for k in range(P.K)]
    a3 = av2.reshape(dnum, rows, n)[:, pick, :].reshape(-1)
    b3 = bv2.reshape(dnum, rows, n)[:, pick, :].reshape(-1)
    w0, w1 = P.keyswitch(L, d, a3, b3)
    try:
        assert np.array_equal(b0.download(), w0) and np.array_equal(b1.download(), w1)
    finally:
        G.lib.gf_key_destroy(h)
    assert ctx.cached_plans() == before - 1
