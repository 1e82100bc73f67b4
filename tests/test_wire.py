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


def fnv1a(b):
    h = 0xcbf29ce484222325
    for x in b:
        h = ((h ^ x) * 0x100000001b3) & (2**64 - 1)
    return h


def test_ciphertext_header_bounds(G):
    """Header fields are bounded before any length is computed from them: a
    crafted header with huge n or towers (whose length would wrap) is refused,
    and gf_ct_packed_bytes is 0 outside the format's shapes."""
    import struct
    assert G.lib.gf_ct_packed_bytes(3 * 1024, 2) == 0          # n not a power of two
    assert G.lib.gf_ct_packed_bytes(1 << 18, 2) == 0           # n above 2^17
    assert G.lib.gf_ct_packed_bytes(1024, 0) == 0
    assert G.lib.gf_ct_packed_bytes(1024, 5000) == 0
    assert G.lib.gf_ct_packed_bytes(1 << 17, 4096) > 0
    for n, towers in ((1 << 31, 1 << 31), (1024, 0xFFFFFFFF), (1000, 2)):
        hdr = b"GFHECT01" + struct.pack("<IIdQ", n, towers, 1.0, 0)
        with pytest.raises(G.GpuFheError, match="out of range"):
            G.unpack_ciphertext(hdr + bytes(64))


@pytest.mark.gpu
def test_key_unpack_checks_residues(G):
    """A packed key whose checksum is consistent but which holds a residue not
    below its row's modulus (a Q row and a P row) is refused."""
    import struct
    P = make_params(1 << 12, 4, 2)
    ctx = G.Context(P.n, P.q, P.p, P.roots, P.dnum)
    rng = np.random.default_rng(4)
    av = np.concatenate([rng.integers(0, int(m), P.n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), P.n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    buf = bytearray(G.pack_key(G.Key(ctx, av, bv, P.dnum)))
    W = len(P.mods)
    body0 = 40 + 8 * W
    for row, which in ((1, 0), (P.L + 1, 1)):      # av row 1 (mod q_1); bv of the second P row
        bad = bytearray(buf)
        word = (which * P.dnum * W + W + row) * P.n + 7    # part 1
        bad[body0 + 8 * word: body0 + 8 * word + 8] = struct.pack("<Q", int(P.mods[row]))
        bad[32:40] = struct.pack("<Q", fnv1a(bytes(bad[40:])))
        with pytest.raises(G.GpuFheError, match="residue"):
            G.unpack_key(ctx, bytes(bad))


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
