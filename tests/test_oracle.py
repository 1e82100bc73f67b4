"""Pins the CPU oracle against known answers, the reference's own host code
(oracle/_ref), committed golden vectors and closed-form identities.  CPU only."""
import ctypes as C
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from common import (KeySet, centred, crt, decode, encode, make_params, mulmod_vec,
                    poly_mul_negacyclic_small, prod)

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))
# file check only: collecting this module (also under -m gpu) must not load the reference-built library
REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (reference sources absent)")


def brv(v, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (v & 1)
        v >>= 1
    return r


def test_oracle_import_maps_no_reference_code():
    """Importing the oracle loads liboracle.so only; oracle/_ref/libref.so maps
    on first use of oracle.ref (so GPU test processes never load it)."""
    code = ("import sys; sys.path.insert(0, %r); import oracle; "
            "m = open('/proc/self/maps').read(); assert 'liboracle.so' in m; assert 'libref.so' not in m"
            % os.path.dirname(os.path.dirname(O.__file__)))
    subprocess.check_call([__import__("sys").executable, "-c", code])


# ----------------------------------------------------------------- RNG ----
def test_mt64_known_answer():
    g = O.MT64()
    O.lib.or_mt64_seed(C.byref(g), 5489)
    v = 0
    for _ in range(10000):
        v = O.lib.or_mt64_next(C.byref(g))
    assert v == 9981545732273789042  # C++ [rand.predef] required behaviour


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_mt64_matches_libstdcxx():
    src = ("#include <random>\n#include <cstdio>\nint main(){std::mt19937_64 r(20260725);"
           "for(int i=0;i<700;++i)printf(\"%llu\\n\",(unsigned long long)r());}")
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "m.cpp"), "w").write(src)
        subprocess.check_call(["g++", "-O1", "-o", os.path.join(d, "m"), os.path.join(d, "m.cpp")])
        want = [int(x) for x in subprocess.check_output([os.path.join(d, "m")]).split()]
    g = O.MT64()
    O.lib.or_mt64_seed(C.byref(g), 20260725)
    got = [O.lib.or_mt64_next(C.byref(g)) for _ in range(700)]
    assert got == want


def test_gauss_sampler_moments():
    g = O.MT64()
    O.lib.or_mt64_seed(C.byref(g), 7)
    xs = np.array([O.lib.or_gauss_sample(C.byref(g), 3.19) for _ in range(200000)])
    assert abs(xs.mean()) < 0.05
    assert abs(xs.std() - 3.19) < 0.05
    assert np.abs(xs).max() <= int(np.ceil(3.19 * 12))


def test_ctr_draw_known_answers():
    """draw(seed, 0, i - 1) is the i-th output of sequential SplitMix64: the
    published outputs from seed 1234567; other domains are other counters."""
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [O.lib.or_ctr_draw(1234567, 0, i) for i in range(5)] == want
    # domain d, index i is counter d 2^32 + i + 1 of the same sequence
    g = 0x9E3779B97F4A7C15
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    for d, i in ((1, 0), (2, 17), (0x10000 + 3, 65535), (0x20002, 5)):
        assert O.lib.or_ctr_draw(99, d, i) == mix((99 + ((d << 32) + i + 1) * g) & M)
    assert len({O.lib.or_ctr_draw(7, d, 0) for d in range(64)}) == 64


def test_ctr_gauss_distribution():
    """The counter Gaussian has the table sampler's law: moments against sigma
    and a histogram matching or_gauss_sample's within sampling error."""
    N = 200000
    xs = np.array([O.lib.or_ctr_gauss(11, 2, i, 3.19) for i in range(N)])
    assert abs(xs.mean()) < 0.05
    assert abs(xs.std() - 3.19) < 0.05
    assert np.abs(xs).max() <= int(np.ceil(3.19 * 12))
    g = O.MT64()
    O.lib.or_mt64_seed(C.byref(g), 5)
    ys = np.array([O.lib.or_gauss_sample(C.byref(g), 3.19) for _ in range(N)])
    for k in range(-8, 9):
        p1, p2 = (xs == k).mean(), (ys == k).mean()
        assert abs(p1 - p2) < 5 * np.sqrt(max(p2, 1e-4) / N) + 1e-4, k
    # the discrete Gaussian's closed-form mass at 0: 1 / sum_k exp(-k^2 / 2 sigma^2)
    p0 = 1.0 / sum(np.exp(-k * k / (2 * 3.19 ** 2)) for k in range(-40, 41))
    assert abs((xs == 0).mean() - p0) < 5 * np.sqrt(p0 / N)


def test_keygen_ctr_identity_and_statistics():
    """Counter-sampled keys: s ternary with equal frequencies, a uniform,
    b + a s = ns e with e small, on every tower."""
    P = make_params(1 << 11, 4, 2)
    n = P.n
    s, a, b = O.keygen_ctr(n, P.mods, P.roots, ns=1, seed=42)
    for t in (0, 3, len(P.mods) - 1):
        q = int(P.mods[t])
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, int(P.roots[t]))
        sl = slice(t * n, (t + 1) * n)
        sc = np.array([centred(int(x), q) for x in O.intt(s[sl], q, ir, ip, ni, nip)])
        assert set(sc.tolist()) <= {-1, 0, 1}
        assert all(abs((sc == v).mean() - 1 / 3) < 0.05 for v in (-1, 0, 1))
        ph = (b[sl].astype(object) + mulmod_vec(a[sl], s[sl], q).astype(object)) % q
        e = [centred(int(x), q) for x in O.intt(O.U(list(ph)), q, ir, ip, ni, nip)]
        assert max(abs(x) for x in e) <= 40 and np.std(e) > 2.5
        assert abs(a[sl].astype(float).mean() / q - 0.5) < 0.03
    # the secret is one polynomial: the same coefficients on every tower
    s0 = [centred(int(x), int(P.mods[0])) for x in O.intt(s[:n], int(P.mods[0]), *O.root_tables(n, int(P.mods[0]), int(P.roots[0]))[2:])]
    q1 = int(P.mods[1])
    s1 = [centred(int(x), q1) for x in O.intt(s[n:2 * n], q1, *O.root_tables(n, q1, int(P.roots[1]))[2:])]
    assert s0 == s1


@pytest.mark.parametrize("level", [4, 2])
def test_evalkeygen_ctr_keyswitch_identity(level):
    """A relinearisation key from the counter sampler switches s^2 -> s:
    ba0 + ba1 s == d s^2 + small noise."""
    P = make_params(1 << 10, 4, 2)
    ks = KeySet.counter(P)
    assert _ks_identity(P, ks, level, level + 5) < (1 << 20)


# ---------------------------------------------------------- parameters ----
PRIME_CASES = [(4, 35, 1024), (3, 50, 1 << 15), (24, 50, 1 << 16), (6, 60, 1 << 16), (2, 36, 1024)]


@needs_ref
@pytest.mark.parametrize("count,bits,n", PRIME_CASES)
def test_primes_match_reference(count, bits, n):
    avoid = O.native_primes(1, bits, n)
    for av in ([], list(avoid)):
        mine = O.native_primes(count, bits, n, av)
        theirs = np.zeros(count, np.uint64)
        a = O.U(av or [0])
        assert O.ref.ref_native_primes(O.P(theirs), count, bits, n, O.P(a), len(av)) == 0
        assert list(mine) == list(theirs)


@pytest.mark.parametrize("count,bits,n", PRIME_CASES[:3])
def test_primes_closed_form(count, bits, n):
    ps = [int(x) for x in O.native_primes(count, bits, n)]
    M = 2 * n
    start = ((1 << bits) // M) * M + 1
    assert all(p % M == 1 and p < (1 << bits) for p in ps)
    # every skipped candidate above the last prime is composite (Fermat witness)
    got = set(ps)
    for c in range(start, ps[-1] - 1, -M):
        if c not in got:
            assert any(pow(b, c - 1, c) != 1 for b in (2, 3, 5, 7, 11, 13)) or c % 2 == 0 or \
                any(c % d == 0 for d in range(3, 2000, 2))
        else:
            assert all(pow(b, c - 1, c) == 1 for b in (2, 3, 5, 7, 11, 13))


@pytest.mark.parametrize("n,bits", [(1024, 35), (1 << 15, 50), (1 << 16, 60)])
def test_root_is_primitive_2n(n, bits):
    q = int(O.native_primes(1, bits, n)[0])
    r = int(O.lib.or_native_root(n, q))
    assert pow(r, n, q) == q - 1 and pow(r, 2 * n, q) == 1
    if O.ref is not None:
        assert r == O.ref.ref_native_root(n, q)
    # smallest-generator rule: no g < the winning base gives a primitive root
    M = 2 * n
    g = next(g for g in range(2, 10000) if pow(pow(g, (q - 1) // M, q), n, q) == q - 1)
    assert pow(g, (q - 1) // M, q) == r


@pytest.mark.parametrize("n,L", [(1024, 4), (1 << 16, 24)])
def test_rescale_consts(n, L):
    P = make_params(n, L, 4 if L > 4 else 2)
    mods = [int(x) for x in P.q]
    for last in (L - 1, L // 2):
        s1 = np.zeros(last, np.uint64)
        s2 = np.zeros(last, np.uint64)
        O.lib.or_rescale_consts(O.P(s1), O.P(s2), O.P(P.q), last)
        ql = mods[last]
        Ql = prod(mods[:last])
        D = (Ql * pow(Ql % ql, -1, ql)) // ql
        for i in range(last):
            assert int(s1[i]) == pow(ql, -1, mods[i])
            assert int(s2[i]) == D % mods[i]
            assert (int(s1[i]) + int(s2[i])) % mods[i] == 0     # s2 == -s1 identity
        if O.ref is not None:
            r1 = np.zeros(last, np.uint64)
            r2 = np.zeros(last, np.uint64)
            O.ref.ref_rescale_consts(O.P(r1), O.P(r2), O.P(P.q), L, last)
            assert np.array_equal(r1, s1) and np.array_equal(r2, s2)


KSC_CASES = [(1024, 4, 2), (1 << 15, 12, 3), (1 << 16, 24, 4), (1 << 16, 20, 4)]


@pytest.mark.parametrize("n,L,dnum", KSC_CASES)
def test_ks_constants_bignum(n, L, dnum):
    P = make_params(n, L, dnum)
    q = [int(x) for x in P.q]
    p = [int(x) for x in P.p]
    k = O.ks_constants(P.q, P.p, dnum)
    W = L + len(p)
    Pp = prod(p)
    assert [int(x) for x in k["pInvModq"]] == [pow(Pp % m, -1, m) for m in q]
    assert [int(x) for x in k["pHatInv"]] == [pow((Pp // pj) % pj, -1, pj) for pj in p]
    assert [int(x) for x in k["pHatModq"]] == [(Pp // pj) % m for pj in p for m in q]
    assert [int(x) for x in k["mdBMuLo"]] == [((1 << 128) - 1) // m % (1 << 64) for m in q]
    a = k["alpha"]
    for part in range(dnum):
        st, sp = int(k["startIdx"][part]), int(k["sizePart"][part])
        src = q[st:st + sp]
        Qp = prod(src)
        compl = [m for i, m in enumerate(q) if i < st or i >= st + sp] + p
        sc = int(k["sizeCompl"][part])
        assert sc == len(compl)
        assert [int(x) for x in k["partComplMod"][part * W: part * W + sc]] == compl
        assert [int(x) for x in k["partQHatInv"][part * a: part * a + sp]] == \
            [pow((Qp // m) % m, -1, m) for m in src]
        blk = k["partQHatModp"][part * a * W:]
        for i, m in enumerate(src):
            assert [int(x) for x in blk[i * sc:(i + 1) * sc]] == [(Qp // m) % c for c in compl]


@needs_ref
@pytest.mark.parametrize("n,L,dnum", KSC_CASES)
def test_ks_constants_match_reference(n, L, dnum):
    P = make_params(n, L, dnum)
    mine = O.ks_constants(P.q, P.p, dnum)
    theirs = O.ref_ks_constants(P.q, P.p, dnum)
    assert mine["alpha"] == theirs["alpha"]
    a, W = mine["alpha"], L + P.K
    for key in ("sizePart", "sizeCompl", "startIdx", "pHatInv", "pHatInvPrec", "pHatModq",
                "mdBMuLo", "mdBMuHi", "pInvModq"):
        assert np.array_equal(mine[key], theirs[key]), key
    for part in range(dnum):
        sp, sc = int(mine["sizePart"][part]), int(mine["sizeCompl"][part])
        for key, stride, cnt in (("partQHatInv", a, sp), ("partQHatInvPrec", a, sp),
                                 ("partSrcMod", a, sp), ("partQHatModp", a * W, sp * sc),
                                 ("partComplMod", W, sc), ("partBMuLo", W, sc),
                                 ("partBMuHi", W, sc)):
            o = part * stride
            assert np.array_equal(mine[key][o:o + cnt], theirs[key][o:o + cnt]), key


def test_ks_constants_golden():
    g = GOLD
    k = O.ks_constants(g["q"], g["p"], g["dnum"])
    W = len(g["q"]) + len(g["p"])
    assert [int(x) for x in k["pInvModq"]] == g["ksc"]["pInvModq"]
    assert [int(x) for x in k["pHatInv"]] == g["ksc"]["pHatInv"]
    for part, gp in enumerate(g["ksc"]["parts"]):
        sc = len(gp["compl"])
        assert [int(x) for x in k["partComplMod"][part * W: part * W + sc]] == gp["compl"]
        sp = gp["sizePart"]
        assert [int(x) for x in k["partQHatInv"][part * k["alpha"]: part * k["alpha"] + sp]] == gp["QHatInv"]


# ----------------------------------------------------------------- NTT ----
@pytest.mark.parametrize("n,bits", [(16, 30), (64, 40), (256, 50), (256, 60)])
def test_ntt_bruteforce_dft(n, bits):
    q = int(O.native_primes(1, bits, n)[0])
    psi = int(O.lib.or_native_root(n, q))
    fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
    rng = np.random.default_rng(n + bits)
    a = [int(x) for x in rng.integers(0, q, n, dtype=np.uint64)]
    got = O.ntt(a, q, fr, fp)
    lb = n.bit_length() - 1
    want = [sum(a[j] * pow(psi, (2 * brv(k, lb) + 1) * j, q) for j in range(n)) % q for k in range(n)]
    assert [int(x) for x in got] == want
    assert [int(x) for x in O.intt(got, q, ir, ip, ni, nip)] == a


def test_ntt_golden():
    g = GOLD
    n, q, psi = g["n"], g["q"][0], g["roots"][0]
    fr, fp, *_ = O.root_tables(n, q, psi)
    assert [int(x) for x in O.ntt(g["ntt_in"], q, fr, fp)] == g["ntt_out"]


@pytest.mark.parametrize("logn", [10, 12, 15, 16])
@pytest.mark.parametrize("bits", [50, 60])
def test_intt_roundtrip(logn, bits):
    n = 1 << logn
    q = int(O.native_primes(1, bits, n)[0])
    psi = int(O.lib.or_native_root(n, q))
    fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
    x = np.random.default_rng(logn).integers(0, q, n, dtype=np.uint64)
    y = O.ntt(x, q, fr, fp)
    assert not np.array_equal(x, y)
    assert np.array_equal(O.intt(y, q, ir, ip, ni, nip), x)
    # bit-reversed table layout: fr[i] = psi^brv(i)
    for i in (1, 2, 3, n // 2, n - 1):
        assert int(fr[i]) == pow(psi, brv(i, logn), q)


# ----------------------------------------------------- pointwise pieces ----
@pytest.mark.parametrize("bo,bn", [(55, 50), (50, 55), (60, 50), (50, 60)])
def test_switch_modulus_closed_form(bo, bn):
    n = 4096
    qo = int(O.native_primes(1, bo, n)[0])
    qn = int(O.native_primes(1, bn, n, [qo])[0])
    v = np.random.default_rng(20260726).integers(0, qo, n, dtype=np.uint64)
    v[:4] = [0, qo // 2, qo // 2 + 1, qo - 1]      # the strict '>' edge
    y = v.copy()
    O.lib.or_switch_modulus(O.P(y), n, qo, qn)
    want = [centred(int(x), qo) % qn for x in v]
    want[1] = (qo // 2) % qn                        # x == floor(q/2) is not lifted
    assert [int(x) for x in y] == want


def test_approx_switch_golden_and_bignum():
    g = GOLD
    part = g["ksc"]["parts"][0]
    n = g["n"]
    sp = part["sizePart"]
    qs = g["q"][:sp]
    k = O.ks_constants(g["q"], g["p"], g["dnum"])
    W = len(g["q"]) + len(g["p"])
    sc = len(part["compl"])
    src = O.U([x for col in g["conv_src"] for x in col])
    dst = np.zeros(sc * n, np.uint64)
    O.lib.or_approx_switch(O.P(src), O.P(dst), O.P(k["partQHatInv"]), O.P(k["partQHatInvPrec"]),
                           O.P(O.U(qs)), O.P(k["partQHatModp"][:sp * sc].copy()),
                           O.P(O.U(part["compl"])), sp, sc, n)
    assert [int(x) for x in dst] == [x for col in g["conv_out"] for x in col]
    _ = W


def test_approx_switch_bignum_large():
    P = make_params(1 << 12, 24, 4)
    n = P.n
    k = O.ks_constants(P.q, P.p, 4)
    sp, sc, a, W = int(k["sizePart"][1]), int(k["sizeCompl"][1]), k["alpha"], 24 + P.K
    qs = [int(x) for x in k["partSrcMod"][a:a + sp]]
    compl = [int(x) for x in k["partComplMod"][W:W + sc]]
    rng = np.random.default_rng(3)
    src = np.concatenate([rng.integers(0, q, n, dtype=np.uint64) for q in qs])
    dst = np.zeros(sc * n, np.uint64)
    O.lib.or_approx_switch(O.P(src), O.P(dst), O.P(k["partQHatInv"][a:].copy()),
                           O.P(k["partQHatInvPrec"][a:].copy()), O.P(O.U(qs)),
                           O.P(k["partQHatModp"][a * W:].copy()), O.P(O.U(compl)), sp, sc, n)
    Q = prod(qs)
    for r in range(0, n, 97):
        s = sum(((int(src[i * n + r]) * pow(Q // q, -1, q)) % q) * (Q // q) for i, q in enumerate(qs))
        for j, p in enumerate(compl):
            assert int(dst[j * n + r]) == s % p


def test_rescale_golden():
    g = GOLD
    P = O.Params(g["n"], g["q"], g["p"], g["roots"], g["dnum"])
    L = len(g["q"])
    c = O.U([x for col in g["rescale_in_eval"] for x in col])
    y = P.rescale(c, L)
    assert [int(x) for x in y] == [x for col in g["rescale_out_eval"] for x in col]


def test_rescale_crt_identity():
    P = make_params(1 << 10, 6, 2)
    n, L = P.n, 6
    mods = [int(x) for x in P.q]
    rng = np.random.default_rng(5)
    coef = np.concatenate([rng.integers(0, m, n, dtype=np.uint64) for m in mods])
    ev = coef.copy()
    for t in range(L):
        fr, fp, ir, ip, ni, nip = O.root_tables(n, mods[t], int(P.roots[t]))
        ev[t * n:(t + 1) * n] = O.ntt(coef[t * n:(t + 1) * n], mods[t], fr, fp)
    y = P.rescale(ev, L)
    ql = mods[L - 1]
    for t in range(L - 1):
        fr, fp, ir, ip, ni, nip = O.root_tables(n, mods[t], int(P.roots[t]))
        yc = O.intt(y[t * n:(t + 1) * n], mods[t], ir, ip, ni, nip)
        for r in range(0, n, 37):
            X = crt([int(coef[s * n + r]) for s in range(L)], mods)
            want = ((X - centred(X, ql)) // ql) % mods[t]
            assert int(yc[r]) == want


def test_automorph_eval_permutation():
    P = make_params(1 << 10, 3, 3)
    n = P.n
    logn = 10
    rng = np.random.default_rng(9)
    v = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q])
    for k in (5, 25, 2 * n - 1, 3):
        w = P.automorph(v, 3, k)
        for t in range(3):
            src = v[t * n:(t + 1) * n]
            perm = [brv((((2 * brv(i, logn) + 1) * k) % (2 * n) - 1) // 2, logn) for i in range(n)]
            assert np.array_equal(w[t * n:(t + 1) * n], src[perm])


# -------------------------------------------------------- keys, enc/dec ----
def test_keygen_identity():
    P = make_params(1 << 10, 4, 2)
    ks = KeySet(P)
    n = P.n
    q0 = int(P.q[0])
    fr, fp, ir, ip, ni, nip = O.root_tables(n, q0, int(P.roots[0]))
    b = ks.pkB[:n]
    a = ks.pkA[:n]
    s = ks.sQ[:n]
    ph = (b.astype(object) + mulmod_vec(a, s, q0).astype(object)) % q0
    e = [centred(int(x), q0) for x in O.intt(O.U(list(ph)), q0, ir, ip, ni, nip)]
    assert max(abs(x) for x in e) <= 40
    sc = [centred(int(x), q0) for x in O.intt(s, q0, ir, ip, ni, nip)]
    assert set(sc) <= {-1, 0, 1}


def test_encrypt_decrypt_roundtrip():
    P = make_params(1 << 10, 4, 2)
    ks = KeySet(P)
    n = P.n
    m = np.random.default_rng(1).integers(-1000, 1000, n).astype(np.int64) * (1 << 20)
    c0, c1 = O.encrypt(n, P.q, P.roots[:4], m, ks.pkA, ks.pkB)
    d = O.decrypt(n, int(P.q[0]), int(P.roots[0]), c0, c1, ks.sQ)
    assert np.max(np.abs(d - m)) < (1 << 12)
    assert np.array_equal(np.rint(d / (1 << 20)).astype(np.int64), m >> 20)


def _ks_identity(P, ks, level, seed):
    """KS_{s^2 -> s}(d) satisfies ba0 + ba1 s == d s^2 + small noise (tower 0)."""
    n = P.n
    av, bv = ks.rlk()
    rng = np.random.default_rng(seed)
    d = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q[:level]])
    ba0, ba1 = P.keyswitch(level, d, av, bv)
    s2 = ks.s2()
    worst = 0
    for t in (0, level - 1):
        q = int(P.q[t])
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, int(P.roots[t]))
        lhs = (ba0[t * n:(t + 1) * n].astype(object) +
               mulmod_vec(ba1[t * n:(t + 1) * n], ks.sQ[t * n:(t + 1) * n], q).astype(object)) % q
        rhs = mulmod_vec(d[t * n:(t + 1) * n], s2[t * n:(t + 1) * n], q).astype(object)
        diff = O.U([(x - y) % q for x, y in zip(lhs, rhs)])
        e = [centred(int(x), q) for x in O.intt(diff, q, ir, ip, ni, nip)]
        worst = max(worst, max(abs(x) for x in e))
    return worst


@pytest.mark.parametrize("level", [4, 3, 1])
def test_keyswitch_decrypt_identity(level):
    P = make_params(1 << 10, 4, 2)
    ks = KeySet(P)
    assert _ks_identity(P, ks, level, level) < (1 << 20)


def test_keyswitch_identity_cfg3_shape_small_ring():
    P = make_params(1 << 10, 24, 4)
    ks = KeySet(P)
    for level in (24, 13):
        assert _ks_identity(P, ks, level, 77) < (1 << 30)


def test_ks_digits_bignum():
    """Digit tower i of part p is d[i] inside the part (src/keyswitch_resident.cpp:191-194)
    and, outside it, NTT of the big-integer ModUp of the part's coefficient slice."""
    P = make_params(1 << 8, 6, 3)
    n, level = P.n, 5
    rng = np.random.default_rng(11)
    d = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q[:level]])
    dg = P.ks_digits(level, d)
    W = level + P.K
    mods = [int(x) for x in P.mods]
    rows = list(range(level)) + list(range(P.L, P.L + P.K))
    coef = []
    for t in range(level):
        fr, fp, ir, ip, ni, nip = O.root_tables(n, mods[t], int(P.roots[t]))
        coef.append(O.intt(d[t * n:(t + 1) * n], mods[t], ir, ip, ni, nip))
    for p in range(P.num_parts(level)):
        st = p * P.alpha
        part = list(range(st, min(st + P.alpha, level)))
        Qp = prod(mods[t] for t in part)
        for i in range(W):
            got = dg[(p * W + i) * n:(p * W + i + 1) * n]
            if i in part:
                assert np.array_equal(got, d[i * n:(i + 1) * n])
                continue
            m = mods[rows[i]]
            want = []
            for r in range(n):
                s_ = sum(((int(coef[t][r]) * pow(Qp // mods[t], -1, mods[t])) % mods[t]) * (Qp // mods[t])
                         for t in part)
                want.append(s_ % m)
            fr, fp, *_ = O.root_tables(n, m, int(P.roots[rows[i]]))
            assert np.array_equal(got, O.ntt(want, m, fr, fp))


def _exact_product_case(P, ks, level, D):
    n = P.n
    av, bv = ks.rlk()
    a1 = np.zeros(n, np.int64)
    a2 = np.zeros(n, np.int64)
    a1[0], a1[1] = 3, -2
    a2[0], a2[2] = 4, 5
    a1[n - 1] = 7
    qs, rs = P.q[:level], P.roots[:level]
    skA = ks.pkA[:level * n]
    skB = ks.pkB[:level * n]
    c0a, c1a = O.encrypt(n, qs, rs, a1 * D, skA, skB, seed=303)
    c0b, c1b = O.encrypt(n, qs, rs, a2 * D, skA, skB, seed=304)
    o0, o1 = P.evalmult_rescale(level, c0a, c1a, c0b, c1b, av, bv)
    dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0, o1, ks.sQ)
    ql = int(P.q[level - 1])
    got = np.array([round(int(x) * ql / (D * D)) for x in dec], dtype=object)
    want = poly_mul_negacyclic_small(a1, a2, n)
    return int(np.max(np.abs((got - want).astype(np.int64))))


def test_evalmult_rescale_exact_small():
    P = make_params(1 << 10, 4, 2)
    assert _exact_product_case(P, KeySet(P), 4, 1 << 40) == 0


@pytest.mark.slow
def test_evalmult_rescale_exact_cfg1_shape():
    """n=2^15, 60-bit first modulus + three 50-bit primes, dnum=2, 60-bit P."""
    P = make_params(1 << 15, 4, 2)
    assert P.K == 2
    assert _exact_product_case(P, KeySet(P), 4, 1 << 40) == 0


def test_evalmult_rescale_slots():
    P = make_params(1 << 11, 5, 2)
    ks = KeySet(P)
    n, level = P.n, 5
    av, bv = ks.rlk()
    rng = np.random.default_rng(2)
    z1 = rng.uniform(-1, 1, n // 2)
    z2 = rng.uniform(-1, 1, n // 2)
    delta = float(1 << 50)
    qs, rs = P.q[:level], P.roots[:level]
    c0a, c1a = O.encrypt(n, qs, rs, encode(z1, n, delta), ks.pkA[:level * n], ks.pkB[:level * n])
    c0b, c1b = O.encrypt(n, qs, rs, encode(z2, n, delta), ks.pkA[:level * n], ks.pkB[:level * n],
                         seed=304)
    o0, o1 = P.evalmult_rescale(level, c0a, c1a, c0b, c1b, av, bv)
    dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0, o1, ks.sQ)
    z = decode(dec, n, delta * delta / int(P.q[level - 1]))
    assert np.max(np.abs(z.real - z1 * z2)) < 1e-6


def test_rotation_slots():
    P = make_params(1 << 11, 4, 2)
    ks = KeySet(P)
    n, level = P.n, 4
    S = n // 2
    rng = np.random.default_rng(4)
    z = rng.uniform(-1, 1, S)
    delta = float(1 << 45)
    c0, c1 = O.encrypt(n, P.q, P.roots[:4], encode(z, n, delta), ks.pkA, ks.pkB, seed=505)
    for r in (1, 3):
        k = pow(5, r, 2 * n)
        av, bv = ks.rotkey(k)
        o0, o1 = P.rotate(level, c0, c1, k, av, bv)
        dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0, o1, ks.sQ)
        zz = decode(dec, n, delta).real
        assert np.max(np.abs(zz - np.roll(z, -r))) < 1e-6


def test_rotate_hoisted_identity_equals_keyswitch_path():
    """k = 1 is the identity automorphism: hoisted and unhoisted rotation must
    agree residue for residue at every level (pins the digit layout, the key
    rows and the ModDown of the hoisted path)."""
    P = make_params(1 << 10, 6, 3)
    rng = np.random.default_rng(12)
    n = P.n
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    for level in (6, 5, 3, 1):
        c0 = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q[:level]])
        c1 = np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q[:level]])
        (h0, h1), = P.rotate_hoisted(level, c0, c1, [1], [(av, bv)])
        w0, w1 = P.rotate(level, c0, c1, 1, av, bv)
        assert np.array_equal(h0, w0) and np.array_equal(h1, w1), level


def test_rotate_hoisted_slots():
    """Several rotations of one ciphertext from one decomposition decrypt to
    the rotated messages."""
    P = make_params(1 << 11, 4, 2)
    ks = KeySet(P)
    n, level = P.n, 4
    rng = np.random.default_rng(14)
    z = rng.uniform(-1, 1, n // 2)
    delta = float(1 << 45)
    c0, c1 = O.encrypt(n, P.q, P.roots[:4], encode(z, n, delta), ks.pkA, ks.pkB, seed=606)
    rs = (1, 2, 5)
    kk = [pow(5, r, 2 * n) for r in rs]
    outs = P.rotate_hoisted(level, c0, c1, kk, [ks.rotkey(k) for k in kk])
    for r, (o0, o1) in zip(rs, outs):
        dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0, o1, ks.sQ)
        assert np.max(np.abs(decode(dec, n, delta).real - np.roll(z, -r))) < 1e-6, r


def _pt_eval(P, coeffs, level):
    """Integer coefficient vector -> evaluation form over the first `level` towers."""
    out = []
    for t in range(level):
        q = int(P.q[t])
        fr, fp, *_ = O.root_tables(P.n, q, int(P.roots[t]))
        out.append(O.ntt(np.mod(coeffs, q).astype(np.uint64), q, fr, fp))
    return np.concatenate(out)


def test_linear_transform_unit_weights():
    """A single term with the all-ones plaintext is the hoisted rotation; the
    unrotated term with weight w is w (.) c residue for residue."""
    P = make_params(1 << 10, 5, 2)
    rng = np.random.default_rng(15)
    n, level = P.n, 4
    av = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    bv = np.concatenate([rng.integers(0, int(m), n, dtype=np.uint64) for _ in range(P.dnum) for m in P.mods])
    rnd = lambda: np.concatenate([rng.integers(0, int(q), n, dtype=np.uint64) for q in P.q[:level]])  # noqa: E731
    c0, c1 = rnd(), rnd()
    ones = np.ones(level * n, np.uint64)
    k = pow(5, 3, 2 * n)
    o0, o1 = P.linear_transform(level, c0, c1, [k], [(av, bv)], [ones])
    (h0, h1), = P.rotate_hoisted(level, c0, c1, [k], [(av, bv)])
    assert np.array_equal(o0, h0) and np.array_equal(o1, h1)
    w = rnd()
    o0, o1 = P.linear_transform(level, c0, c1, [1], [None], [w])
    for t in range(level):
        sl = slice(t * n, (t + 1) * n)
        assert np.array_equal(o0[sl], mulmod_vec(w[sl], c0[sl], P.q[t]))
        assert np.array_equal(o1[sl], mulmod_vec(w[sl], c1[sl], P.q[t]))


def test_linear_transform_slots():
    """Diagonal method: sum_j diag_j (.) rot_rj(z) decrypts from the weighted sum
    of rotations (the unrotated term included)."""
    P = make_params(1 << 11, 4, 2)
    ks = KeySet(P)
    n, level = P.n, 4
    rng = np.random.default_rng(16)
    z = rng.uniform(-1, 1, n // 2)
    d_ct, d_pt = float(1 << 30), float(1 << 20)
    c0, c1 = O.encrypt(n, P.q, P.roots[:4], encode(z, n, d_ct), ks.pkA, ks.pkB, seed=607)
    rs = (0, 1, 7)
    kk = [pow(5, r, 2 * n) for r in rs]
    diags = [rng.uniform(-1, 1, n // 2) for _ in rs]
    pts = [_pt_eval(P, encode(dg, n, d_pt), level) for dg in diags]
    keys = [None if r == 0 else ks.rotkey(k) for r, k in zip(rs, kk)]
    o0, o1 = P.linear_transform(level, c0, c1, kk, keys, pts)
    dec = O.decrypt(n, int(P.q[0]), int(P.roots[0]), o0, o1, ks.sQ)
    want = sum(dg * np.roll(z, -r) for dg, r in zip(diags, rs))
    assert np.max(np.abs(decode(dec, n, d_ct * d_pt).real - want)) < 1e-4


# -------------------------------------------------------------- encode ----
@needs_ref
def test_fast_encode_matches_reference_encode():
    n = 1024
    rng = np.random.default_rng(6)
    z = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    delta = float(1 << 40)
    mine = encode(z, n, delta)
    theirs = np.zeros(n, np.int64)
    re = np.ascontiguousarray(z.real)
    im = np.ascontiguousarray(z.imag)
    O.ref.ref_encode(O.P(theirs), O.P(re), O.P(im), n, delta)
    assert np.max(np.abs(mine - theirs)) <= 1          # the reference's own +-1 bar
    zr = np.zeros(n // 2)
    zi = np.zeros(n // 2)
    O.ref.ref_decode(O.P(zr), O.P(zi), O.P(theirs), n, delta)
    zz = decode(theirs, n, delta)
    assert np.max(np.abs(zz.real - zr)) < 1e-9 and np.max(np.abs(zz.imag - zi)) < 1e-9
