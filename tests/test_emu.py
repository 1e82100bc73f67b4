"""CPU tests of the device tile logic: tests/emu/emu.cpp compiles the same
headers as the HIP kernels (gpu-resident-library_amd/csrc/*.hpp) with g++ and
runs each workgroup sequentially; results must equal the oracle bit for bit
for every ring size, register radix and twiddle layout."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from common import make_params

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
u64p = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def emu():
    if shutil.which("g++") is None:
        pytest.skip("g++ absent")
    d = tempfile.mkdtemp()
    so = os.path.join(d, "libemu.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
                           "-I" + os.path.join(ROOT, "gpu-resident-library_amd", "csrc"),
                           "-o", so, os.path.join(HERE, "emu", "emu.cpp")])
    lib = C.CDLL(so)
    lib.emu_ntt.argtypes = [u64p, C.c_uint32, C.c_uint64, u64p, u64p, C.c_int, C.c_uint64, C.c_uint64,
                            C.c_int, C.c_uint32]
    for f in ("emu_mulmod", "emu_shoup", "emu_mulhi", "emu_shoup_lazy4", "emu_csub"):
        getattr(lib, f).restype = C.c_uint64
        getattr(lib, f).argtypes = [C.c_uint64] * (2 if f in ("emu_mulhi", "emu_csub") else 3)
    lib.emu_redc.restype = C.c_uint64
    lib.emu_redc.argtypes = [C.c_uint64] * 3
    lib.emu_dot30.restype = None
    lib.emu_dot30.argtypes = [u64p, u64p, C.c_uint32, u64p]
    lib.emu_ntt_ex.argtypes = [u64p, C.c_uint32, C.c_uint64, u64p, u64p, C.c_int, C.c_uint64, C.c_uint64,
                               C.c_int, C.c_uint32, C.c_int, u64p, u64p, C.c_uint64, C.c_uint64]
    lib.emu_lazy_red_of.restype = C.c_uint64
    lib.emu_lazy_red_of.argtypes = [C.c_uint64]
    lib.emu_ksi.restype = None
    lib.emu_ksi.argtypes = [C.c_uint32, C.c_uint64, u64p, u64p, C.c_uint32, C.c_uint32, C.c_uint32, u64p, u64p,
                            u64p, u64p, u64p, C.c_uint32, C.c_uint32, C.c_int, u64p, u64p, C.c_uint64, C.c_uint64,
                            u64p, u64p]
    lib.emu_tensor_ib.restype = None
    lib.emu_tensor_ib.argtypes = [C.c_uint32, C.c_uint64, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_int] + \
        [u64p] * 8
    lib.emu_intt_f64.restype = None
    lib.emu_intt_f64.argtypes = [u64p, C.c_uint32, C.c_uint64, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint32]
    lib.emu_inner_red.restype = C.c_uint64
    lib.emu_inner_red.argtypes = [u64p, u64p, C.c_uint32, C.c_uint64]
    lib.emu_reduce_red.restype = C.c_uint64
    lib.emu_reduce_red.argtypes = [C.c_uint64, C.c_uint64]
    pp = C.POINTER(u64p)
    lib.emu_moddown.restype = None
    lib.emu_moddown.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, pp, u64p, pp, pp, u64p, u64p, u64p, pp, pp,
                                u64p, u64p, u64p, u64p, C.c_uint64, C.c_uint64, u64p, u64p, C.c_int, C.c_int]
    lib.emu_modup.restype = None
    lib.emu_modup.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, pp, u64p, pp, pp, u64p, u64p, u64p, pp, pp, u64p,
                              u64p, C.c_int, C.c_uint32, C.c_int, C.c_int]
    yield lib
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("logn", list(range(1, 18)))
def test_ntt_engine_matches_oracle(emu, logn):
    n = 1 << logn
    for bits in (50, 60):
        q = int(O.native_primes(1, bits if logn > 3 else 30, n)[0])
        psi = int(O.lib.or_native_root(n, q))
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
        x = np.random.default_rng(logn + bits).integers(0, q, n, dtype=np.uint64)
        x[:2] = [q - 1, q - 1]
        want = O.ntt(x, q, fr, fp)
        for rmax in (2, 3, 4):
            for paired in (0, 1):
                y = x.copy()
                emu.emu_ntt(O.P(y), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax)
                assert np.array_equal(y, want), (bits, rmax, paired)
                z = want.copy()
                emu.emu_ntt(O.P(z), n, q, O.P(ir), O.P(ip), 0, ni, nip, paired, rmax)
                assert np.array_equal(z, x), (bits, rmax, paired)


def test_modarith_edges(emu):
    rng = np.random.default_rng(1)
    for bits in (30, 50, 59, 60):
        q = int(O.native_primes(1, bits, 1 << 10)[0])
        vals = [0, 1, q - 1, q - 2, q // 2] + [int(v) for v in rng.integers(0, q, 200, dtype=np.uint64)]
        for a in vals[:40]:
            for b in vals[:40]:
                assert emu.emu_mulmod(a, b, q) == (a * b) % q
                assert emu.emu_shoup(a, b, q) == (a * b) % q
        for _ in range(2000):
            a, b = (int(v) for v in rng.integers(0, 1 << 63, 2, dtype=np.uint64))
            a, b = a * 2 + 1, b
            assert emu.emu_mulhi(a, b) == (a * b) >> 64
        R = 1 << 64
        for _ in range(500):
            T = int(rng.integers(0, q, dtype=np.uint64)) * int(rng.integers(0, 1 << 63, dtype=np.uint64))
            T %= q * R
            assert emu.emu_redc(T % R, T >> 64, q) == (T * pow(R, -1, q)) % q


@pytest.mark.parametrize("bits", [30, 50, 59, 60])
def test_shoup_lazy4_range(emu, bits):
    """The transform butterflies' Shoup product: congruent to x w mod q and
    inside [0, 4q) for every 64-bit x, including the extremes that drive the
    quotient estimate furthest from the exact one (x0 = 2^32 - 1, x = 2^64 - 1);
    and csub's borrow select."""
    q = int(O.native_primes(1, bits, 1 << 10)[0])
    rng = np.random.default_rng(bits)
    ws = [0, 1, q - 1, q // 2] + [int(v) for v in rng.integers(0, q, 60, dtype=np.uint64)]
    xs = [0, 1, q - 1, q, 3 * q - 1, 8 * q - 1, (1 << 32) - 1, (1 << 64) - 1, ((1 << 64) - 1) ^ (1 << 31),
          (q << 2) | 0xFFFFFFFF] + [int(v) for v in rng.integers(0, 1 << 63, 200, dtype=np.uint64) * 2 + 1]
    hit = set()
    for w in ws:
        for x in xs:
            r = emu.emu_shoup_lazy4(x, w, q)
            assert r % q == (x * w) % q and r < 4 * q, (x, w)
            hit.add(r // q)
    assert {0, 1, 2} <= hit <= {0, 1, 2, 3}   # the slack is used, never exceeded
    for m in (q, 2 * q, 4 * q):
        for x in (0, 1, m - 1, m, m + 1, 2 * m - 1):
            assert emu.emu_csub(x, m) == (x - m if x >= m else x)


def test_dot30_column_sums(emu):
    """Conversion-kernel dot products: the 30-bit-split column sums equal the
    big-integer sum of products for 1..8 terms, at the largest operands
    (2^60 - 1, where every column is closest to overflow) and at random ones."""
    rng = np.random.default_rng(7)
    top = (1 << 60) - 1
    cases = [[top] * 8, [top, 0] * 4, [(1 << 30) - 1] * 8, [1 << 30] * 8]
    cases += [[int(v) for v in rng.integers(0, 1 << 60, 8, dtype=np.uint64)] for _ in range(300)]
    for a in cases:
        for b in (a[::-1], [top] * 8, [int(v) for v in rng.integers(0, 1 << 60, 8, dtype=np.uint64)]):
            for ns in range(1, 9):
                av, bv = np.asarray(a[:ns], np.uint64), np.asarray(b[:ns], np.uint64)
                out = np.zeros(2, np.uint64)
                emu.emu_dot30(O.P(av), O.P(bv), ns, O.P(out))
                want = sum(x * y for x, y in zip(a[:ns], b[:ns]))
                assert int(out[0]) | (int(out[1]) << 64) == want, (ns, a[:ns], b[:ns])


@pytest.mark.parametrize("bits", [33, 40, 50, 56])
def test_reduce_red(emu, bits):
    """The one-step reduction behind the forward transforms' lazy path: x mod q
    for every 64-bit x, at the edges of its quotient estimate (low word all
    ones, x just below a multiple of q, the largest x) and at random x."""
    q = int(O.native_primes(1, bits, 1 << 10)[0])
    assert emu.emu_lazy_red_of(q) == (1 << 64) // q
    rng = np.random.default_rng(bits)
    xs = [0, 1, q - 1, q, 2 * q - 1, 3 * q - 1, 69 * q - 1, 77 * q - 1, (1 << 64) - 1, (1 << 64) - q,
          ((1 << 64) - 1) // q * q - 1, (1 << 32) - 1, ((q >> 32) << 32) | 0xFFFFFFFF]
    xs += [int(v) for v in rng.integers(0, 1 << 63, 3000, dtype=np.uint64) * 2 + 1]
    for x in xs:
        assert emu.emu_reduce_red(x, q) == x % q, x


def test_lazy_red_range(emu):
    """Only moduli in (2^32, 2^57) take the lazy path: a 72q bound (80q with the
    epilogue) must fit 64 bits."""
    assert emu.emu_lazy_red_of((1 << 32) - 5) == 0
    assert emu.emu_lazy_red_of((1 << 32) + 15) != 0
    assert emu.emu_lazy_red_of((1 << 57) - 111) != 0
    assert emu.emu_lazy_red_of((1 << 57) + 1) == 0
    assert 72 * ((1 << 57) - 1) + 8 * ((1 << 57) - 1) < 1 << 64


@pytest.mark.parametrize("logn", [4, 7, 12, 13, 15, 16, 17])
def test_lazy_forward_transform(emu, logn):
    """Forward transforms without per-butterfly reductions (moduli below 2^57):
    reduced output equal to the oracle's, the unreduced output (lazy_out)
    congruent to it and below (4 + 4 logn) q, and the linear epilogue
    NTT(x) + e0 c0 + e1 c1 equal to its definition, for every radix and both
    twiddle layouts; inputs up to 4q - 1 (the unreduced outputs of the
    conversions that feed forward transforms) push the growth bound hardest."""
    n = 1 << logn
    for bits in (40, 56):
        q = int(O.native_primes(1, bits, n)[0])
        psi = int(O.lib.or_native_root(n, q))
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
        rng = np.random.default_rng(logn * 100 + bits)
        for x in (rng.integers(0, q, n, dtype=np.uint64), np.full(n, q - 1, np.uint64),
                  np.full(n, 4 * q - 1, np.uint64), rng.integers(0, 4 * q, n, dtype=np.uint64)):
            want = O.ntt(x % np.uint64(q), q, fr, fp)
            e0 = rng.integers(0, q, n, dtype=np.uint64)
            e1 = rng.integers(0, q, n, dtype=np.uint64)
            c0, c1 = int(rng.integers(0, q)), int(rng.integers(0, q))
            want_ep = np.array([(int(a) + int(b) * c0 + int(c) * c1) % q for a, b, c in zip(want, e0, e1)],
                               np.uint64)
            top = 0
            for rmax in (2, 3, 4):
                for paired in (0, 1):
                    y = x.copy()
                    emu.emu_ntt_ex(O.P(y), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax, 1, None, None, 0, 0)
                    assert np.array_equal(y, want), (bits, rmax, paired)
                    z = x.copy()
                    emu.emu_ntt_ex(O.P(z), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax, 2, None, None, 0, 0)
                    assert np.array_equal(z % np.uint64(q), want), (bits, rmax, paired)
                    assert int(z.max()) < (4 + 4 * logn) * q
                    top = max(top, int(z.max()))
                    e = x.copy()
                    emu.emu_ntt_ex(O.P(e), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax, 2, O.P(e0), O.P(e1),
                                   c0, c1)
                    assert np.array_equal(e, want_ep), (bits, rmax, paired)
                    e = x.copy()
                    emu.emu_ntt_ex(O.P(e), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax, 0, O.P(e0), O.P(e1),
                                   c0, c1)
                    assert np.array_equal(e, want_ep), (bits, rmax, paired)
            assert top >= 2 * q   # the unreduced values use their slack


@pytest.mark.parametrize("bits", [33, 50, 56])
def test_inner_product_unreduced_digits(emu, bits):
    """Key inner product over digits left unreduced by the lazy forward
    transforms (below 72q, up to 16 parts, keys in [0, q)): the 128-bit sum,
    REDC without its correction and the one-step reduction give
    sum * 2^-64 mod q, including at the largest digits and keys."""
    q = int(O.native_primes(1, bits, 1 << 10)[0])
    Rinv = pow(1 << 64, -1, q)
    rng = np.random.default_rng(bits + 5)
    cases = [([72 * q - 1] * 16, [q - 1] * 16), ([72 * q - 1] * 16, [0] * 16), ([0] * 16, [q - 1] * 16)]
    for _ in range(400):
        cases.append(([int(v) for v in rng.integers(0, 72 * q, 16, dtype=np.uint64)],
                      [int(v) for v in rng.integers(0, q, 16, dtype=np.uint64)]))
    for a, b in cases:
        for ns in (1, 4, 16):
            av, bv = np.asarray(a[:ns], np.uint64), np.asarray(b[:ns], np.uint64)
            want = sum(x * y for x, y in zip(a[:ns], b[:ns])) * Rinv % q
            assert emu.emu_inner_red(O.P(av), O.P(bv), ns, q) == want, (ns, a[:ns], b[:ns])


@pytest.mark.parametrize("logn,np_,xf", [(13, 4, 0b1110), (14, 3, 0b101), (16, 4, 0b0111), (16, 1, 0b1),
                                         (15, 7, 0b1011111)])
def test_fused_rowpass_inner_product(emu, logn, np_, xf):
    """The fused step of the keyswitch: each transformed part goes through the
    column pass and then the row pass inside the fused kernel body, the other
    parts are evaluation-form digits as given; the result equals
    sum_p digit_p * key_p * 2^-64 mod q with digit_p the oracle's forward
    transform, for moduli below 2^57 (unreduced digits) and above (reduced
    digits), with digits and keys below 2^60 (column sums) and above (128-bit
    sums), every register radix; and for the 50-bit modulus with the row
    passes in fp64 (digits below 2q, as the fp64 ModUp leaves them).  With at
    least two transformed parts, the step also runs as the P rows run it with
    the fused ModDown: the inverse row pass of both outputs inside it (the rows
    handed over in LDS), after which the column pass gives the oracle's
    inverse transform of each output."""
    n = 1 << logn
    # 50 bits: unreduced digits, column sums; 55: unreduced digits above 2^60
    # (128-bit sums); 59: reduced digits, column sums; 61: 128-bit sums
    for bits in (50, 55, 59, 61):
        q = int(O.native_primes(1, bits, n)[0])
        psi = int(O.lib.or_native_root(n, q))
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
        rng = np.random.default_rng(logn * 7 + bits + np_)
        src = rng.integers(0, q, (np_, n), dtype=np.uint64)
        src[0, :3] = q - 1
        kb = rng.integers(0, q, (np_, n), dtype=np.uint64)
        ka = rng.integers(0, q, (np_, n), dtype=np.uint64)
        dig = [O.ntt(src[p], q, fr, fp) if (xf >> p) & 1 else src[p] for p in range(np_)]
        Rinv = pow(1 << 64, -1, q)
        w0 = np.zeros(n, dtype=object)
        w1 = np.zeros(n, dtype=object)
        for p in range(np_):
            d = dig[p].astype(object)
            w0 += d * kb[p].astype(object)
            w1 += d * ka[p].astype(object)
        w0 = np.array([int(v) * Rinv % q for v in w0], np.uint64)
        w1 = np.array([int(v) * Rinv % q for v in w1], np.uint64)
        for rmax in (2, 3, 4):
            pairs = rmax % 2   # both access widths of the inner product
            for lazy_out in (0, 1):
                for f64 in ((0, 1) if bits == 50 else (0,)):
                    s = src.copy()
                    o0 = np.zeros(n, np.uint64)
                    o1 = np.zeros(n, np.uint64)
                    emu.emu_ksi(n, q, O.P(fr), O.P(fp), rmax, np_, xf, O.P(s), O.P(np.ascontiguousarray(kb)),
                                O.P(np.ascontiguousarray(ka)), O.P(o0), O.P(o1), pairs, lazy_out, f64,
                                None, None, 0, 0, None, None)
                    if lazy_out:   # unreduced, for the epilogue's lazy products: congruent, below 2^62
                        assert int(max(o0.max(), o1.max())) < 1 << 62
                        o0, o1 = o0 % np.uint64(q), o1 % np.uint64(q)
                    assert np.array_equal(o0, w0), (bits, rmax, pairs, lazy_out, f64)
                    assert np.array_equal(o1, w1), (bits, rmax, pairs, lazy_out, f64)
                    if lazy_out or f64 or bin(xf).count("1") < 2:
                        continue
                    s = src.copy()
                    x0, x1 = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
                    emu.emu_ksi(n, q, O.P(fr), O.P(fp), rmax, np_, xf, O.P(s), O.P(np.ascontiguousarray(kb)),
                                O.P(np.ascontiguousarray(ka)), O.P(o0), O.P(o1), pairs, 0, 0,
                                O.P(ir), O.P(ip), int(ni), int(nip), O.P(x0), O.P(x1))
                    assert np.array_equal(x0 % np.uint64(q), O.intt(w0, q, ir, ip, ni, nip)), ("rp", bits, rmax)
                    assert np.array_equal(x1 % np.uint64(q), O.intt(w1, q, ir, ip, ni, nip)), ("rp", bits, rmax)


@pytest.mark.parametrize("logn", [13, 15, 16, 17])
def test_tensor_fused_with_rowpass(emu, logn):
    """The tensor product fused with the inverse row pass of its t2 (t2 handed
    over in LDS per row tile, the row pass in fp64 or in integers, then the
    ordinary phase A) gives the oracle's t0 = a0 b0, t1 = a0 b1 + a1 b0,
    t2 = a1 b1 and its inverse transform of t2 times the final factor, for a
    50-bit modulus (both row-pass kinds) and a 60-bit one (integer), every
    register radix."""
    n = 1 << logn
    for bits, f64s in ((50, (0, 1)), (60, (0,))):
        q = int(O.native_primes(1, bits, n)[0])
        psi = int(O.lib.or_native_root(n, q))
        _, _, ir, ip, ni, nip = O.root_tables(n, q, psi)
        rng = np.random.default_rng(logn * 3 + bits)
        a = [rng.integers(0, q, n, dtype=np.uint64) for _ in range(4)]
        a[0][:3] = q - 1
        a[3][:3] = q - 1
        A = [x.astype(object) for x in a]
        w0 = np.array((A[0] * A[2]) % q, np.uint64)
        w1 = np.array((A[0] * A[3] + A[1] * A[2]) % q, np.uint64)
        w2 = np.array((A[1] * A[3]) % q, np.uint64)
        wx = O.intt(w2, q, ir, ip, ni, nip)
        for rmax in (2, 3, 4):
            for f64 in f64s:
                t = [np.zeros(n, np.uint64) for _ in range(3)]
                x = np.zeros(n, np.uint64)
                emu.emu_tensor_ib(n, q, O.P(ir), O.P(ip), int(ni), int(nip), rmax, f64, O.P(a[0]), O.P(a[1]),
                                  O.P(a[2]), O.P(a[3]), *[O.P(v) for v in t], O.P(x))
                assert np.array_equal(t[0], w0) and np.array_equal(t[1], w1) and np.array_equal(t[2], w2), (bits, rmax)
                assert np.array_equal(x % np.uint64(q), wx), (bits, rmax, f64)


@pytest.mark.parametrize("logn", [13, 14, 15, 16, 17])
def test_inverse_rowpass_f64(emu, logn):
    """An inverse transform whose phase B (the row pass) runs in fp64 and
    hands [0, 2q) to the integer phase A equals the oracle's inverse transform
    times the final factor, for inputs up to 4q - 1 (the lazy range phase B
    accepts), every register radix, 50-bit moduli."""
    n = 1 << logn
    for q in [int(v) for v in O.native_primes(2, 50, n)]:
        psi = int(O.lib.or_native_root(n, q))
        _, _, ir, ip, ni, nip = O.root_tables(n, q, psi)
        rng = np.random.default_rng(logn + q % 1000)
        x = rng.integers(0, 4 * q, n, dtype=np.uint64)
        x[:7] = 4 * q - 1
        want = O.intt(x % np.uint64(q), q, ir, ip, ni, nip)
        for rmax in (2, 3, 4):
            y = x.copy()
            emu.emu_intt_f64(O.P(y), n, q, O.P(ir), O.P(ip), int(ni), int(nip), rmax)
            assert np.array_equal(y, want), (q, rmax)


@pytest.mark.parametrize("logn,L,dnum,part", [(13, 6, 3, 0), (14, 5, 1, 0), (15, 12, 3, 2), (16, 24, 4, 0),
                                              (16, 24, 4, 3), (17, 3, 1, 0)])
def test_fused_modup_equals_unfused(emu, logn, L, dnum, part):
    """The fused ModUp (the sources' inverse column pass, the conversion to
    the part's complement and the digits' forward column pass in one
    workgroup per column tile) stores exactly what the unfused kernels store:
    each source's inverse column pass (standard tiles, radix 2^3 and 2^4), the
    conversion kernel's sums, each digit's forward column pass, over every
    ring split (2^13 .. 2^17), 1 to 6 sources, 50-bit (unreduced) and 60-bit
    (reduced butterflies) targets, inputs up to 4q - 1.  With the fp64 passes
    for moduli below 2^50 the 60-bit targets are still stored bit for bit
    (their sources' reduced values and the conversion sums are identical),
    and the fp64 targets hold the same residues in [0, 2q).  With the
    conversion's R = 2^60 column reduction (where the sources sum to at most
    1.2 2^60) every target holds the same residues, within the bounds the
    next passes take."""
    n = 1 << logn
    P = make_params(n, L, dnum)
    alpha = -(-L // dnum)
    st = alpha * part
    sp = min(alpha, L - st)
    srcm = [int(m) for m in P.q[st:st + sp]]
    tgt = [int(m) for i, m in enumerate(P.q) if not st <= i < st + sp] + [int(m) for m in P.p]
    sc = len(tgt)
    rng = np.random.default_rng(logn * 31 + part)
    R = 1 << 64
    src, sw, swp, post, postp = [], [], [], [], []
    for i, q in enumerate(srcm):
        x = rng.integers(0, 4 * q - 1, n, dtype=np.uint64)
        x[:5] = 4 * q - 1
        src.append(np.ascontiguousarray(x))
        _, _, ir, ip, ni, _ = O.root_tables(n, q, int(P.roots[st + i]))
        sw.append(ir)
        swp.append(ip)
        qhat = 1
        for k, m in enumerate(srcm):
            if k != i:
                qhat = qhat * m
        v = pow(qhat % q, -1, q) * int(ni) % q
        post.append(v)
        postp.append((v << 64) // q)
    tw, twp = [], []
    roots = [int(r) for r in P.roots]
    mods = [int(m) for m in P.mods]
    for t in tgt:
        fr, fp, _, _, _, _ = O.root_tables(n, t, roots[mods.index(t)])
        tw.append(fr)
        twp.append(fp)
    tab = np.zeros(sp * sc, np.uint64)
    for i in range(sp):
        qhat = 1
        for k, m in enumerate(srcm):
            if k != i:
                qhat *= m
        for j, t in enumerate(tgt):
            tab[i * sc + j] = (qhat % t) * R % t
    arr = lambda v: (u64p * len(v))(*[O.P(np.ascontiguousarray(a, dtype=np.uint64)) for a in v])  # noqa: E731
    keep = [src, sw, swp, tw, twp]
    outs = {}
    c60_ok = 5 * sum(srcm) <= 6 << 60
    for ref, rmax, f64, c60 in ((0, 3, 0, 0), (1, 3, 0, 0), (1, 4, 0, 0), (0, 3, 1, 0), (0, 3, 1, 1), (0, 3, 0, 1)):
        if c60 and not c60_ok:
            continue
        dst = np.zeros(sc * n, np.uint64)
        emu.emu_modup(n, sp, sc, arr(src), O.P(O.U(srcm)), arr(sw), arr(swp), O.P(O.U(post)), O.P(O.U(postp)),
                      O.P(O.U(tgt)), arr(tw), arr(twp), O.P(tab), O.P(dst), ref, rmax, f64, c60)
        outs[(ref, rmax, f64, c60)] = dst
    assert keep
    assert np.array_equal(outs[(0, 3, 0, 0)], outs[(1, 3, 0, 0)])
    assert np.array_equal(outs[(0, 3, 0, 0)], outs[(1, 4, 0, 0)])
    f = outs[(0, 3, 1, 0)].reshape(sc, n)
    r = outs[(1, 3, 0, 0)].reshape(sc, n)
    nf = 0
    for j, t in enumerate(tgt):
        if t < 1 << 50:
            nf += 1
            assert int(f[j].max()) < 2 * t, j
            assert np.array_equal(f[j] % np.uint64(t), r[j] % np.uint64(t)), j
        else:
            assert np.array_equal(f[j], r[j]), j
    assert nf == sum(t < 1 << 50 for t in tgt)
    if c60_ok:
        for key in ((0, 3, 1, 1), (0, 3, 0, 1)):
            c = outs[key].reshape(sc, n)
            for j, t in enumerate(tgt):
                assert np.array_equal(c[j] % np.uint64(t), r[j] % np.uint64(t)), (key, j)
                if key[2] and t < 1 << 50:
                    assert int(c[j].max()) < 2 * t, (key, j)
                elif t < 1 << 57:   # unreduced forward column pass from inputs below 2.2 t
                    assert int(c[j].max()) < 40 * t, (key, j)


@pytest.mark.parametrize("logn,L,dnum,level", [(13, 6, 3, 6), (14, 5, 1, 3), (15, 12, 3, 12), (16, 24, 4, 24),
                                               (16, 24, 4, 9), (17, 3, 1, 3)])
def test_fused_moddown_equals_unfused(emu, logn, L, dnum, level):
    """The fused ModDown + rescale prologue (the fused ModUp kernel's rescale
    mode: the P towers' inverse column pass, the dropped tower, the prologue
    value of every surviving tower and its forward column pass, in one
    workgroup per column tile) stores the residues the unfused order stores
    (its prologue folds the switch term into the conversion's Montgomery sum,
    so the lazy representatives differ), in [0, 2q) where the towers below
    2^50 take the fp64 passes; every ring
    split, 1 to 6 P towers, the dropped tower larger and smaller than the
    surviving ones."""
    n = 1 << logn
    P = make_params(n, L, dnum)
    K = len(P.p)
    if K > 6:
        pytest.skip("more P towers than the fused kernel takes")
    mods = [int(m) for m in P.mods]
    roots = [int(r) for r in P.roots]
    pm = [int(m) for m in P.p]
    tgt = [int(m) for m in P.q[:level - 1]]
    q_old = int(P.q[level - 1])
    sc = len(tgt)
    rng = np.random.default_rng(logn * 13 + level)
    src, sw, swp, post, postp = [], [], [], [], []
    for p in pm:
        x = rng.integers(0, 4 * p - 1, n, dtype=np.uint64)
        x[:3] = 4 * p - 1
        src.append(np.ascontiguousarray(x))
        _, _, ir, ip, _, _ = O.root_tables(n, p, roots[mods.index(p)])
        sw.append(ir)
        swp.append(ip)
        v = int(rng.integers(1, p))
        post.append(v)
        postp.append((v << 64) // p)
    tw, twp = [], []
    for t in tgt:
        fr, fp, _, _, _, _ = O.root_tables(n, t, roots[mods.index(t)])
        tw.append(fr)
        twp.append(fp)
    tab = np.array([int(rng.integers(0, t)) for _ in range(K) for t in tgt], np.uint64)
    tabd = np.array([int(rng.integers(0, q_old)) for _ in range(K)], np.uint64)
    s2 = np.array([int(rng.integers(0, t)) for t in tgt], np.uint64)
    dt = rng.integers(0, q_old, n, dtype=np.uint64)
    dr = rng.integers(0, q_old, n, dtype=np.uint64)
    dt[:2] = q_old - 1
    pinv = int(rng.integers(1, q_old))
    arr = lambda v: (u64p * len(v))(*[O.P(np.ascontiguousarray(a, dtype=np.uint64)) for a in v])  # noqa: E731
    outs = {}
    for ref, f64 in ((1, 0), (0, 0), (0, 1)):
        dst = np.zeros(sc * n, np.uint64)
        emu.emu_moddown(n, K, sc, arr(src), O.P(O.U(pm)), arr(sw), arr(swp), O.P(O.U(post)), O.P(O.U(postp)),
                        O.P(O.U(tgt)), arr(tw), arr(twp), O.P(tab), O.P(tabd), O.P(dt), O.P(dr), q_old, pinv,
                        O.P(s2), O.P(dst), ref, f64)
        outs[(ref, f64)] = dst.reshape(sc, n)
    r, a, f = outs[(1, 0)], outs[(0, 0)], outs[(0, 1)]
    for j, t in enumerate(tgt):
        assert np.array_equal(a[j] % np.uint64(t), r[j] % np.uint64(t)), j
        if t < 1 << 50:
            assert int(f[j].max()) < 2 * t, j
        else:
            assert np.array_equal(f[j], a[j]), j   # the integer passes either way
        assert np.array_equal(f[j] % np.uint64(t), r[j] % np.uint64(t)), j
