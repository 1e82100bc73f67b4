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


def test_ks_macc_single_and_batched(G):
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
    d0, d1, ddig, dbv, dav = (G.Dev.of(x) for x in (o0, o1, dig, bv, av))
    ptrs = G.Dev.of(np.array([ddig.at(i * n) for i in range(W)], np.uint64))
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


def test_approx_switch_crt_basis(G):
    P = make_params(1 << 15, 12, 3)
    n = P.n
    k = O.ks_constants(P.q, P.p, 3)
    W = P.L + P.K
    for part in range(3):
        a = k["alpha"]
        sp, sc = int(k["sizePart"][part]), int(k["sizeCompl"][part])
        qs = k["partSrcMod"][part * a: part * a + sp].copy()
        rng = np.random.default_rng(20260729 + part)
        src = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in qs])
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


def test_tensor_combine(G):
    P = make_params(1 << 12, 5, 2)
    n, L = P.n, 5
    rng = np.random.default_rng(303)
    cts = [np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q]) for _ in range(6)]
    want = [np.zeros(L * n, np.uint64) for _ in range(3)]
    O.lib.or_tensor(*[O.P(c) for c in cts[:4]], *[O.P(w) for w in want], O.P(P.q), L, n)
    d = [G.Dev.of(c) for c in cts]
    t = [G.Dev(L * n) for _ in range(3)]
    dm = G.Dev.of(P.q)
    G.lib.LaunchTensor(*[x.ptr for x in d[:4]], *[x.ptr for x in t], dm.ptr, L, n, None)
    G.device_sync()
    for i in range(3):
        assert np.array_equal(t[i].download(), want[i])
    r = [G.Dev(L * n), G.Dev(L * n)]
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
