"""Serving loop (gf_server_*): continuous batching of mixed ring sizes, levels
and operations.  Every request's result in the mix must equal the same
operation run alone (and the oracle), bit for bit; admission must respect the
workspace budget."""
import numpy as np
import pytest

from common import make_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


def ring_setup(G, n, L, dnum, nrot, seed):
    P = make_params(n, L, dnum)
    rng = np.random.default_rng(seed)

    def key():
        return [np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
                for _ in range(2)]

    ctx = G.Context(n, P.q, P.p, P.roots, dnum)
    rav, rbv = key()
    rlk = G.Key(ctx, rav, rbv, P.dnum)
    rot_k = [pow(5, j + 1, 2 * n) for j in range(nrot)]
    rk, rkeys = [], []
    for _ in range(nrot):
        av, bv = key()
        rk.append(G.Key(ctx, av, bv, P.dnum))
        rkeys.append((av, bv))
    return dict(P=P, ctx=ctx, rlk=rlk, rlk_host=(rav, rbv), rot_k=rot_k, rk=rk, rkeys=rkeys, rng=rng)


def rand_ct(P, level, rng):
    return [np.concatenate([rng.integers(0, int(q), P.n, dtype=np.uint64) for q in P.q[:level]]) for _ in range(2)]


def build_trace(G, rings, count, seed, mean_gap_ms):
    rng = np.random.default_rng(seed)
    reqs, keep, meta = [], [], []
    t = 0.0
    for i in range(count):
        r = int(rng.integers(0, len(rings)))
        R = rings[r]
        P = R["P"]
        u = rng.random()
        op = G.OP_EVALMULT_RESCALE if u < 0.6 else (G.OP_ROTATE if u < 0.9 else G.OP_ADD)
        level = int(rng.integers(2, P.L + 1))
        rot = int(rng.integers(0, len(R["rk"])))
        a = rand_ct(P, level, rng)
        b = rand_ct(P, level, rng)
        da = [G.Dev.of(x) for x in a]
        db = [G.Dev.of(x) for x in b]
        out_l = level - 1 if op == G.OP_EVALMULT_RESCALE else level
        do = [G.Dev(out_l * P.n) for _ in range(2)]
        t += float(rng.exponential(mean_gap_ms))
        reqs.append(G.Request(r, op, level, rot, da[0].ptr, da[1].ptr, db[0].ptr, db[1].ptr, do[0].ptr, do[1].ptr, t))
        keep.append((da, db, do))
        meta.append(dict(ring=r, op=op, level=level, rot=rot, a=a, b=b, out_l=out_l))
    return reqs, keep, meta


def run_alone(G, R, m, da, db):
    P, ctx = R["P"], R["ctx"]
    s = G.Stream()
    ws = ctx.workspace()
    o = [G.Dev(m["out_l"] * P.n) for _ in range(2)]
    lv = m["level"]
    if m["op"] == G.OP_EVALMULT_RESCALE:
        G.check(G.lib.gf_evalmult_rescale(ctx.h, R["rlk"].h, lv, da[0].ptr, da[1].ptr, db[0].ptr, db[1].ptr,
                                          o[0].ptr, o[1].ptr, ws.ptr, s.s))
    elif m["op"] == G.OP_ROTATE:
        G.check(G.lib.gf_rotate(ctx.h, R["rk"][m["rot"]].h, lv, R["rot_k"][m["rot"]], da[0].ptr, da[1].ptr,
                                o[0].ptr, o[1].ptr, ws.ptr, s.s))
    else:
        G.check(G.lib.gf_add(ctx.h, lv, da[0].ptr, da[1].ptr, db[0].ptr, db[1].ptr, o[0].ptr, o[1].ptr, s.s))
    s.sync()
    return o[0].download(), o[1].download()


def test_serving_mix_equals_alone(G):
    rings = [ring_setup(G, 1 << 12, 6, 2, 3, 1), ring_setup(G, 1 << 13, 8, 4, 2, 2), ring_setup(G, 1 << 11, 4, 2, 2, 3)]
    srv = G.Server(ws_budget_bytes=2 << 30, lanes=4, max_batch=16)
    for R in rings:
        R["id"] = srv.add_ring(R["ctx"], R["rlk"], R["rk"], R["rot_k"])
    reqs, keep, meta = build_trace(G, rings, 96, 11, mean_gap_ms=0.02)
    comps = srv.run_trace(reqs)
    assert len(comps) == len(reqs)
    sizes = [c.batch_size for c in comps]
    assert max(sizes) > 1, "no batching happened"
    for c, m in zip(comps, meta):
        assert c.done_ms >= c.arrival_ms and c.launch_ms >= c.arrival_ms - 1e-6
        assert c.ring == m["ring"] and c.op == m["op"] and c.level == m["level"]
    st = srv.stats()
    assert st["admitted"] == len(reqs) and st["batches"] < len(reqs)
    for (da, db, do), m in zip(keep, meta):
        R = rings[m["ring"]]
        w0, w1 = run_alone(G, R, m, da, db)
        assert np.array_equal(do[0].download(), w0), m
        assert np.array_equal(do[1].download(), w1), m
    # the first few against the oracle as well
    for (da, db, do), m in list(zip(keep, meta))[:6]:
        R = rings[m["ring"]]
        P = R["P"]
        if m["op"] == G.OP_EVALMULT_RESCALE:
            w0, w1 = P.evalmult_rescale(m["level"], m["a"][0], m["a"][1], m["b"][0], m["b"][1], *R["rlk_host"])
        elif m["op"] == G.OP_ROTATE:
            av, bv = R["rkeys"][m["rot"]]
            w0, w1 = P.rotate(m["level"], m["a"][0], m["a"][1], R["rot_k"][m["rot"]], av, bv)
        else:
            continue
        assert np.array_equal(do[0].download(), w0) and np.array_equal(do[1].download(), w1), m


def test_serving_admission_by_footprint(G):
    R = ring_setup(G, 1 << 12, 6, 2, 1, 5)
    probe = G.Server(ws_budget_bytes=1 << 30, lanes=1, max_batch=4)
    rid = probe.add_ring(R["ctx"], R["rlk"], R["rk"], R["rot_k"])
    slot = probe.slot_bytes(rid)
    del probe
    srv = G.Server(ws_budget_bytes=2 * slot, lanes=2, max_batch=8)
    rid = srv.add_ring(R["ctx"], R["rlk"], R["rk"], R["rot_k"])
    P = R["P"]
    rng = np.random.default_rng(8)
    keep = []
    for i in range(10):
        a = [G.Dev.of(x) for x in rand_ct(P, P.L, rng)]
        b = [G.Dev.of(x) for x in rand_ct(P, P.L, rng)]
        keep.append((a, b))
        srv.submit(G.Request(rid, G.OP_EVALMULT_RESCALE, P.L, 0, a[0].ptr, a[1].ptr, b[0].ptr, b[1].ptr,
                             None, None, -1.0))
    srv.drain()
    comps = srv.poll()
    assert len(comps) == 10
    assert max(c.batch_size for c in comps) <= 2
    assert srv.stats()["slot_bytes"] <= 2 * slot


def test_serving_rejects_bad_requests(G):
    R = ring_setup(G, 1 << 11, 4, 2, 1, 6)
    srv = G.Server(ws_budget_bytes=1 << 30)
    rid = srv.add_ring(R["ctx"], R["rlk"], R["rk"], R["rot_k"])
    d = G.Dev(4 * R["P"].n)
    with pytest.raises(G.GpuFheError):
        srv.submit(G.Request(rid, G.OP_ROTATE, 4, 5, d.ptr, d.ptr, None, None, None, None, -1.0))
    with pytest.raises(G.GpuFheError):
        srv.submit(G.Request(rid + 1, G.OP_ADD, 4, 0, d.ptr, d.ptr, d.ptr, d.ptr, None, None, -1.0))
    with pytest.raises(G.GpuFheError):
        srv.submit(G.Request(rid, G.OP_EVALMULT_RESCALE, 9, 0, d.ptr, d.ptr, d.ptr, d.ptr, None, None, -1.0))


def test_serving_clock_alignment(G):
    """Latency subtracts a host-clock arrival from a device-event completion.
    The two clocks are aligned once, when the server is created; this bounds
    their skew over a few seconds: each request's device completion time
    falls between the host time just before it was submitted and the host
    time just after it was drained (plus 1 ms of slack for the event
    timestamp's resolution)."""
    import time
    R = ring_setup(G, 1 << 12, 4, 2, 1, 12)
    srv = G.Server(ws_budget_bytes=1 << 30, lanes=2, max_batch=4)
    rid = srv.add_ring(R["ctx"], R["rlk"], R["rk"], R["rot_k"])
    P = R["P"]
    a = [G.Dev.of(x) for x in rand_ct(P, P.L, R["rng"])]
    b = [G.Dev.of(x) for x in rand_ct(P, P.L, R["rng"])]
    for i in range(12):
        before = G.lib.gf_server_now_ms(srv.h)
        srv.submit(G.Request(rid, G.OP_EVALMULT_RESCALE, P.L, 0, a[0].ptr, a[1].ptr, b[0].ptr, b[1].ptr,
                             None, None, -1.0))
        srv.drain()
        after = G.lib.gf_server_now_ms(srv.h)
        (c,) = srv.poll()
        assert before <= c.arrival_ms <= c.launch_ms <= after, (i, before, c.arrival_ms, c.launch_ms, after)
        assert before - 1.0 <= c.done_ms <= after + 1.0, (i, before, c.done_ms, after)
        time.sleep(0.25)
