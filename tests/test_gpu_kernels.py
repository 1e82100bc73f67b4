"""Bit-exact parity of the reference C-ABI launchers (include/gpufhe_kernels.h)
against the CPU oracle, on seeded inputs (seeds follow the reference's tests)."""
import numpy as np
import pytest

import oracle as O
from common import make_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpufhe
    return gpufhe


def rnd(seed, q, n):
    return np.random.default_rng(seed).integers(0, q, n, dtype=np.uint64)


NTT_CASES = [(10, 50), (12, 60), (13, 50), (14, 60), (15, 50), (16, 50), (16, 60), (17, 60)]


@pytest.mark.parametrize("logn,bits", NTT_CASES)
def test_ntt_intt(G, logn, bits):
    n = 1 << logn
    q = int(O.native_primes(1, bits, n)[0])
    psi = int(O.lib.or_native_root(n, q))
    fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
    x = rnd(20260725 + logn, q, n)
    d = G.Dev.of(x)
    tabs = [G.Dev.of(t) for t in (fr, fp, ir, ip)]
    G.lib.LaunchNTT_CT(d.ptr, tabs[0].ptr, tabs[1].ptr, n, q, None)
    G.device_sync()
    y = d.download()
    assert np.array_equal(y, O.ntt(x, q, fr, fp))
    G.lib.LaunchINTT_GS(d.ptr, tabs[2].ptr, tabs[3].ptr, n, q, ni, nip, None)
    G.device_sync()
    assert np.array_equal(d.download(), x)


def test_rns_mult_and_edges(G):
    n = 1 << 14
    for bits in (50, 60):
        q = int(O.native_primes(1, bits, n)[0])
        a = rnd(20260727, q, n)
        b = rnd(20260728, q, n)
        a[:3] = [0, q - 1, q - 1]
        b[:3] = [q - 1, q - 1, 1]
        da, db, dr = G.Dev.of(a), G.Dev.of(b), G.Dev(n)
        G.lib.LaunchRNSMultTower(da.ptr, db.ptr, dr.ptr, q, n, None)
        G.device_sync()
        want = np.zeros(n, np.uint64)
        O.lib.or_rns_mult(O.P(a), O.P(b), O.P(want), q, n)
        assert np.array_equal(dr.download(), want)


@pytest.mark.parametrize("bo,bn", [(55, 50), (50, 55), (60, 50)])
def test_switch_modulus(G, bo, bn):
    n = 4096
    qo = int(O.native_primes(1, bo, n)[0])
    qn = int(O.native_primes(1, bn, n, [qo])[0])
    v = rnd(20260726, qo, n)
    v[:3] = [qo // 2, qo // 2 + 1, qo - 1]
    d = G.Dev.of(v)
    G.lib.LaunchSwitchModulus(d.ptr, qo, qn, n, None)
    G.device_sync()
    w = v.copy()
    O.lib.or_switch_modulus(O.P(w), n, qo, qn)
    assert np.array_equal(d.download(), w)


def test_rescale_fuse_moddown_add_ptmulacc(G):
    n = 1 << 13
    q = int(O.native_primes(1, 50, n)[0])
    a, b, c, dd = (rnd(s, q, n) for s in (1, 2, 3, 4))
    s1, s2 = 123456789 % q, (q - 987654321) % q
    da, db = G.Dev.of(a), G.Dev.of(b)
    G.lib.LaunchRescaleFuse(da.ptr, db.ptr, s1, s2, q, n, None)
    G.device_sync()
    w = a.copy()
    O.lib.or_rescale_fuse(O.P(w), O.P(b), s1, s2, q, n)
    assert np.array_equal(da.download(), w)

    dans, da2 = G.Dev(n), G.Dev.of(a)
    G.lib.LaunchModDownCombine(dans.ptr, da2.ptr, db.ptr, s1, q, n, None)
    G.device_sync()
    w = np.zeros(n, np.uint64)
    O.lib.or_moddown_combine(O.P(w), O.P(a), O.P(b), s1, q, n)
    assert np.array_equal(dans.download(), w)

    da = G.Dev.of(a)
    G.lib.LaunchAddInto(da.ptr, db.ptr, q, n, None)
    G.device_sync()
    w = a.copy()
    O.lib.or_add_into(O.P(w), O.P(b), q, n)
    assert np.array_equal(da.download(), w)

    acc0, acc1 = G.Dev.of(a), G.Dev.of(b)
    dc, ddd, dmul = G.Dev.of(c), G.Dev.of(dd), G.Dev.of(a)   # kept alive across the launch
    G.lib.LaunchPtMulAcc(acc0.ptr, acc1.ptr, dc.ptr, ddd.ptr, dmul.ptr, q, n, None)
    G.device_sync()
    w0 = (a.astype(object) + c.astype(object) * a.astype(object)) % q
    w1 = (b.astype(object) + dd.astype(object) * a.astype(object)) % q
    assert np.array_equal(acc0.download(), w0.astype(np.uint64))
    assert np.array_equal(acc1.download(), w1.astype(np.uint64))


@pytest.mark.parametrize("off", [0, 1])
def test_ks_macc_single_and_batched(G, off):
    P = make_params(1 << 12, 4, 2)
    n, L, K = P.n, P.L, P.K
    W = L + K
    mods = P.mods
    rng = np.random.default_rng(20260729)
    dig = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for m in mods])
    key_rows = 9
    keymods = list(mods[:L]) + [3] * (key_rows - W) + list(mods[L:])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for m in keymods])
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for m in keymods])
    o0 = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for m in mods])
    o1 = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for m in mods])
    delta = key_rows - W
    rows = [i if i < L else i + delta for i in range(W)]
    w0, w1 = o0.copy(), o1.copy()
    for i in range(W):
        sl = slice(i * n, (i + 1) * n)
        ks = slice(rows[i] * n, (rows[i] + 1) * n)
        t0, t1 = w0[sl].copy(), w1[sl].copy()
        O.lib.or_ks_macc(O.P(t0), O.P(t1), O.P(dig[sl].copy()), O.P(bv[ks].copy()), O.P(av[ks].copy()),
                         int(mods[i]), n)
        w0[sl], w1[sl] = t0, t1
    d0, d1, ddig, dbv, dav = (Shifted(G, x, off=off) for x in (o0, o1, dig, bv, av))
    ptrs = G.Dev.of(np.array([ddig.ptr + 8 * i * n for i in range(W)], np.uint64))
    krow = np.array(rows, np.uint32)
    dkrow = G.Dev((W + 1) // 2)
    G.check(G.lib.gf_h2d(dkrow.ptr, krow.ctypes.data, krow.nbytes))
    dmods = G.Dev.of(mods)
    G.lib.LaunchKSMultAccBatched(d0.ptr, d1.ptr, ptrs.ptr, dbv.ptr, dav.ptr, dkrow.ptr, dmods.ptr, W, n, None)
    G.device_sync()
    assert np.array_equal(d0.download(), w0) and np.array_equal(d1.download(), w1)
    # per-tower form, one tower
    e0, e1 = G.Dev.of(o0[:n]), G.Dev.of(o1[:n])
    G.lib.LaunchKSMultAcc(e0.ptr, e1.ptr, ddig.ptr, dbv.ptr, dav.ptr, int(mods[0]), n, None)
    G.device_sync()
    assert np.array_equal(e0.download(), w0[:n]) and np.array_equal(e1.download(), w1[:n])


# (n, L, dnum, bits): 4 sources of <= 60 bits (column sums), 10 sources
# (generic 128-bit sums), 61-bit moduli (generic sums), and 9 sources of 61
# bits, whose sums can reach p 2^64 (the Barrett form with the caller's mu)
APPROX_CASES = [(1 << 15, 12, 3, (50, 60, 60)), (1 << 12, 10, 1, (50, 60, 60)), (1 << 12, 4, 1, (61, 61, 61)),
                (1 << 12, 9, 1, (61, 61, 61))]


@pytest.mark.parametrize("n,L,dnum,bits", APPROX_CASES)
def test_approx_switch_crt_basis(G, n, L, dnum, bits):
    P = make_params(n, L, dnum, qbits=bits[0], q0bits=bits[1], pbits=bits[2])
    k = O.ks_constants(P.q, P.p, dnum)
    W = P.L + P.K
    for part in range(dnum):
        a = k["alpha"]
        sp, sc = int(k["sizePart"][part]), int(k["sizeCompl"][part])
        qs = k["partSrcMod"][part * a: part * a + sp].copy()
        rng = np.random.default_rng(20260729 + part)
        src = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in qs])
        src[::97] = np.repeat(qs - 1, n)[::97]       # largest residues
        qhi = k["partQHatInv"][part * a: part * a + sp].copy()
        qhip = k["partQHatInvPrec"][part * a: part * a + sp].copy()
        qhmp = k["partQHatModp"][part * a * W: part * a * W + sp * sc].copy()
        pm = k["partComplMod"][part * W: part * W + sc].copy()
        mlo = k["partBMuLo"][part * W: part * W + sc].copy()
        mhi = k["partBMuHi"][part * W: part * W + sc].copy()
        want = np.zeros(sc * n, np.uint64)
        O.lib.or_approx_switch(O.P(src), O.P(want), O.P(qhi), O.P(qhip), O.P(qs), O.P(qhmp), O.P(pm), sp, sc, n)
        dst = G.Dev(sc * n)
        bufs = [G.Dev.of(x) for x in (src, qhi, qhip, qs, qhmp, pm, mlo, mhi)]
        G.lib.LaunchApproxSwitchCRTBasis(bufs[0].ptr, dst.ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, bufs[4].ptr,
                                         bufs[5].ptr, bufs[6].ptr, bufs[7].ptr, sp, sc, n, None)
        G.device_sync()
        assert np.array_equal(dst.download(), want)


class Shifted:
    """A device copy of `arr` starting `off` words into its allocation (off = 1:
    only 8-byte aligned, so the launchers take their one-word access path)."""

    def __init__(self, G, arr=None, words=None, off=0):
        w = arr.size if arr is not None else words
        self.buf = G.Dev(w + 1)
        self.ptr = self.buf.at(off)
        self.off, self.words = off, w
        if arr is not None:
            self.buf.upload(arr, off)

    def download(self):
        return self.buf.download(self.words, self.off)


# moduli: 60 + 50-bit, a 40-bit modulus, 61-bit moduli (the Montgomery
# constants each launcher derives on the device), both access widths
@pytest.mark.parametrize("bits,off", [((50, 60), 0), ((50, 60), 1), ((40, 40), 0), ((61, 61), 1)])
def test_tensor_combine(G, bits, off):
    P = make_params(1 << 12, 5, 2, qbits=bits[0], q0bits=bits[1], pbits=61)
    n, L = P.n, 5
    rng = np.random.default_rng(303)
    cts = [np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q]) for _ in range(6)]
    cts[0][::13] = np.repeat(P.q - 1, n)[::13]
    want = [np.zeros(L * n, np.uint64) for _ in range(3)]
    O.lib.or_tensor(*[O.P(c) for c in cts[:4]], *[O.P(w) for w in want], O.P(P.q), L, n)
    d = [Shifted(G, c, off=off) for c in cts]
    t = [Shifted(G, words=L * n, off=off) for _ in range(3)]
    dm = G.Dev.of(P.q)
    G.lib.LaunchTensor(*[x.ptr for x in d[:4]], *[x.ptr for x in t], dm.ptr, L, n, None)
    G.device_sync()
    for i in range(3):
        assert np.array_equal(t[i].download(), want[i])
    r = [Shifted(G, words=L * n, off=off), Shifted(G, words=L * n, off=off)]
    G.lib.LaunchCombine(r[0].ptr, r[1].ptr, t[0].ptr, t[1].ptr, d[4].ptr, d[5].ptr, dm.ptr, L, n, None)
    G.device_sync()
    w0, w1 = np.zeros(L * n, np.uint64), np.zeros(L * n, np.uint64)
    O.lib.or_combine(O.P(w0), O.P(w1), O.P(want[0]), O.P(want[1]), O.P(cts[4]), O.P(cts[5]), O.P(P.q), L, n)
    assert np.array_equal(r[0].download(), w0) and np.array_equal(r[1].download(), w1)


def test_automorph_coeff(G):
    n = 1 << 12
    q = int(O.native_primes(1, 50, n)[0])
    x = rnd(9, q, n)
    x[::7] = 0
    for k in (5, 25, 2 * n - 1):
        d, o = G.Dev.of(x), G.Dev(n)
        G.lib.LaunchAutomorphPermute(o.ptr, d.ptr, n, k, q, None)
        G.device_sync()
        w = np.zeros(n, np.uint64)
        O.lib.or_automorph_coeff(O.P(w), O.P(x), n, k, q)
        assert np.array_equal(o.download(), w)
