"""The allocation-free EvalMult+Rescale (gf_evalmult_rescale_noalloc): its
output equals gf_evalmult_rescale's (caller workspace, cached plan) bit for
bit, on two calls in a row with the second smaller (fewer towers, another
context), and also the oracle's; contexts whose workspace exceeds the static
array's maximum are rejected; no plan is cached by the call."""
import numpy as np
import pytest

from common import make_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


def run_both(G, ctx, key, level, seed, s):
    n = ctx.n
    ins = []
    for j in range(4):
        d = G.Dev(level * n)
        G.check(G.lib.gf_fill_uniform(ctx.h, d.ptr, level, seed + j, s.s))
        ins.append(d)
    a = [G.Dev((level - 1) * n) for _ in range(2)]
    b = [G.Dev((level - 1) * n) for _ in range(2)]
    ws = ctx.workspace()
    cached = ctx.cached_plans()
    G.check(G.lib.gf_evalmult_rescale_noalloc(ctx.h, key.h, level, *[x.ptr for x in ins], a[0].ptr, a[1].ptr, s.s))
    s.sync()
    assert ctx.cached_plans() == cached, "the allocation-free call cached a plan"
    G.check(G.lib.gf_evalmult_rescale(ctx.h, key.h, level, *[x.ptr for x in ins], b[0].ptr, b[1].ptr, ws.ptr, s.s))
    s.sync()
    got = [x.download() for x in a]
    assert np.array_equal(got[0], b[0].download()) and np.array_equal(got[1], b[1].download())
    return ins, got


def test_noalloc_matches_evalmult_rescale_two_calls(G):
    assert G.lib.gf_noalloc_scratch_bytes() >= 318 * (1 << 16) * 8
    s = G.Stream()
    # first call: the largest supported shape, n = 2^16, L = 24, dnum = 4, K = 6
    P1 = make_params(1 << 16, 24, 4)
    assert len(P1.p) == 6
    ctx1 = G.Context(P1.n, P1.q, P1.p, P1.roots, 4)
    key1 = G.Key.uniform(ctx1, 4, 71, s)
    ins1, got1 = run_both(G, ctx1, key1, 24, 500, s)
    av, bv = key1.download(s)
    w0, w1 = P1.evalmult_rescale(24, *[x.download() for x in ins1], av, bv)
    assert np.array_equal(got1[0], w0) and np.array_equal(got1[1], w1)
    # second call, smaller: another context (n = 2^13, L = 6, dnum = 3) at level 4
    P2 = make_params(1 << 13, 6, 3)
    ctx2 = G.Context(P2.n, P2.q, P2.p, P2.roots, 3)
    key2 = G.Key.uniform(ctx2, 3, 72, s)
    ins2, got2 = run_both(G, ctx2, key2, 4, 600, s)
    av, bv = key2.download(s)
    w0, w1 = P2.evalmult_rescale(4, *[x.download() for x in ins2], av, bv)
    assert np.array_equal(got2[0], w0) and np.array_equal(got2[1], w1)
    # and the large context again, at a lower level, after the small call
    run_both(G, ctx1, key1, 13, 700, s)


def test_noalloc_rejects_contexts_above_the_maximum(G):
    s = G.Stream()
    P = make_params(1 << 16, 26, 4)            # 26 towers: the workspace exceeds 318 towers
    ctx = G.Context(P.n, P.q, P.p, P.roots, 4)
    assert ctx.ws_bytes > 318 * (1 << 16) * 8
    key = G.Key.uniform(ctx, 4, 73, s)
    d = [G.Dev(26 * P.n) for _ in range(6)]
    with pytest.raises(G.GpuFheError, match="static scratch maximum"):
        G.check(G.lib.gf_evalmult_rescale_noalloc(ctx.h, key.h, 26, *[x.ptr for x in d], s.s))
