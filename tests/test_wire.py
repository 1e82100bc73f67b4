"""Wire format (gf_ct_pack / gf_key_pack): round trips, and refusal of
corrupted, truncated or mismatched data.  The ciphertext functions are host
code and run without a GPU."""
import numpy as np
import pytest

from common import make_params


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


def rand_ct(P, level, rng):
    return [np.concatenate([rng.integers(0, int(q), P.n, dtype=np.uint64) for q in P.q[:level]]) for _ in range(2)]


def test_ciphertext_round_trip(G):
    P = make_params(1 << 10, 5, 2)
    rng = np.random.default_rng(1)
    for level in (5, 3, 1):
        c0, c1 = rand_ct(P, level, rng)
        buf = G.pack_ciphertext(c0, c1, P.n, P.q[:level], scale=2.0 ** 40)
        assert len(buf) == G.lib.gf_ct_packed_bytes(P.n, level)
        r0, r1, n, mods, scale = G.unpack_ciphertext(buf)
        assert n == P.n and scale == 2.0 ** 40
        assert np.array_equal(mods, P.q[:level]) and np.array_equal(r0, c0) and np.array_equal(r1, c1)


def test_ciphertext_rejects_damage(G):
    P = make_params(1 << 10, 3, 1)
    c0, c1 = rand_ct(P, 3, np.random.default_rng(2))
    buf = bytearray(G.pack_ciphertext(c0, c1, P.n, P.q))
    flipped = bytearray(buf)
    flipped[32 + 8 * 3 + 17] ^= 0x04          # a bit of c0 (after the 32-byte header and 3 moduli)
    with pytest.raises(G.GpuFheError, match="checksum"):
        G.unpack_ciphertext(bytes(flipped))
    with pytest.raises(G.GpuFheError, match="length"):
        G.unpack_ciphertext(bytes(buf[:-8]))
    bad = bytearray(buf)
    bad[0:8] = b"NOTACTXT"
    with pytest.raises(G.GpuFheError, match="not a packed ciphertext"):
        G.unpack_ciphertext(bytes(bad))
    over = c0.copy()
    over[5] = P.q[0]                            # a residue equal to its modulus
    with pytest.raises(G.GpuFheError, match="residue"):
        G.pack_ciphertext(over, c1, P.n, P.q)


@pytest.mark.gpu
def test_key_round_trip_and_parameter_check(G):
    P = make_params(1 << 12, 4, 2)
    ctx = G.Context(P.n, P.q, P.p, P.roots, P.dnum)
    rng = np.random.default_rng(3)
    av = np.concatenate([rng.integers(0, int(m), P.n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), P.n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    key = G.Key(ctx, av, bv, P.dnum)
    buf = G.pack_key(key)
    back = G.unpack_key(ctx, buf)
    ra, rb = back.download()
    assert np.array_equal(ra, av) and np.array_equal(rb, bv)
    # the reloaded key switches exactly like the original
    d = np.concatenate([rng.integers(0, int(q), P.n, dtype=np.uint64) for q in P.q])
    s = G.Stream()
    ws = ctx.workspace()
    outs = []
    for k in (key, back):
        dd, b0, b1 = G.Dev.of(d), G.Dev(P.L * P.n), G.Dev(P.L * P.n)
        G.check(G.lib.gf_keyswitch(ctx.h, k.h, P.L, dd.ptr, b0.ptr, b1.ptr, ws.ptr, s.s))
        s.sync()
        outs.append((b0.download(), b1.download()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    other = make_params(1 << 12, 4, 2, qbits=49)
    ctx2 = G.Context(other.n, other.q, other.p, other.roots, other.dnum)
    with pytest.raises(G.GpuFheError, match="moduli"):
        G.unpack_key(ctx2, buf)
