"""The reference's EvalMult+Rescale launch sequence over the drop-in boundary
(bench/bench_evalmult_resident.cpp:96-100: LaunchTensor -> keyswitch_resident
-> LaunchCombine -> rescale_resident_raw on each component), launched directly
and captured as a graph per chain, as the reference benchmark runs it: every
result equals the oracle's EvalMult+Rescale, on every graph replay (with
unrelated eager work between the replays), and equals the batched plan."""
import numpy as np
import pytest

from common import make_params, parallel_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


@pytest.mark.parametrize("n,L,dnum,level", [(1 << 13, 5, 2, 5), (1 << 15, 4, 2, 4), (1 << 13, 6, 3, 4)])
def test_reference_sequence_direct_and_graph(G, n, L, dnum, level):
    P = make_params(n, L, dnum)
    s = G.Stream()
    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    key = G.Key.uniform(ctx, dnum, 31, s)
    av, bv = key.download(s)
    d_mods = G.Dev.of(np.asarray(P.q, np.uint64))
    chains = [G.RefSequence(ctx, key, level, d_mods, 100 * c + 1) for c in range(3)]
    s.sync()
    ins = [[x.download() for x in c.ins] for c in chains]
    want = parallel_map(lambda x: P.evalmult_rescale(level, *x, av, bv), ins)
    for c in chains:
        c.launch()
    G.device_sync()
    for c, (w0, w1) in zip(chains, want):
        r0, r1 = c.outputs()
        assert np.array_equal(r0, w0) and np.array_equal(r1, w1), "direct"
    for c in chains:
        c.capture()
    scratch = [G.Dev(2 * n) for _ in range(2)]
    for rep in range(3):
        for c in chains:
            for r in c.r:
                G.check(G.lib.gf_fill_uniform(ctx.h, r.ptr, level, 900 + rep, c.stream.s))
            c.graph.launch(c.stream)
        G.device_sync()
        # unrelated eager work between the replays
        G.check(G.lib.gf_add(ctx.h, 1, chains[0].ins[0].ptr, chains[0].ins[1].ptr, chains[1].ins[0].ptr,
                             chains[1].ins[1].ptr, scratch[0].ptr, scratch[1].ptr, s.s))
        s.sync()
        for c, (w0, w1) in zip(chains, want):
            r0, r1 = c.outputs()
            assert np.array_equal(r0, w0) and np.array_equal(r1, w1), ("graph replay", rep)
    # the batched plan over the same inputs gives the same residues
    outs = [[G.Dev((level - 1) * n) for _ in range(2)] for _ in chains]
    col = lambda j: [c.ins[j].ptr for c in chains]  # noqa: E731
    plan = G.plan_evalmult_rescale(ctx, key, [level] * len(chains), col(0), col(1), col(2), col(3),
                                   [o[0].ptr for o in outs], [o[1].ptr for o in outs], [c.ws.ptr for c in chains])
    plan.run_graph(s)
    s.sync()
    for o, (w0, w1) in zip(outs, want):
        assert np.array_equal(o[0].download(), w0) and np.array_equal(o[1].download(), w1), "plan"
