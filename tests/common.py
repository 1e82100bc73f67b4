"""Shared helpers for the tests: parameter sets, encode/decode, big-integer CRT."""
import functools

import numpy as np

import oracle as O

U = O.U


@functools.lru_cache(maxsize=None)
def make_params(n, L, dnum, qbits=50, q0bits=60, pbits=60, K=None):
    """Native parameter set the way the reference builds its OpenFHE-free
    tests (tests/test_native_e2e.cpp:28-34): descending NTT-friendly primes,
    first-found 2n-th roots.  K defaults to ceil(max part bits / pbits)."""
    q = list(O.native_primes(1, q0bits, n))
    if L > 1:
        q += list(O.native_primes(L - 1, qbits, n, q))
    alpha = -(-L // dnum)
    if K is None:
        part_bits = q0bits + (alpha - 1) * qbits
        K = -(-part_bits // pbits)
    p = list(O.native_primes(K, pbits, n, q))
    roots = [int(O.lib.or_native_root(n, int(m))) for m in q + p]
    return O.Params(n, q, p, roots, dnum)


def prod(xs):
    r = 1
    for x in xs:
        r *= int(x)
    return r


def crt(residues, mods):
    """CRT lift of per-modulus residues (lists of ints) -> int in [0, prod)."""
    M = prod(mods)
    x = 0
    for r, m in zip(residues, mods):
        Mi = M // int(m)
        x += int(r) * Mi * pow(Mi, -1, int(m))
    return x % M


def centred(x, m):
    x %= m
    return x - m if x > m // 2 else x


def rot_group(N):
    M = 2 * N
    r = np.zeros(N // 2, np.int64)
    rk = 1
    for k in range(N // 2):
        r[k] = rk
        rk = (rk * 5) % M
    return r


def encode(z, N, delta):
    """CKKS canonical embedding encode (same map as src/encode.cpp:13-28),
    computed with an FFT of length 2N instead of the O(N^2) sum."""
    r = rot_group(N)
    V = np.zeros(2 * N, np.complex128)
    V[r] = z
    V[(2 * N - r) % (2 * N)] = np.conj(z)
    F = np.fft.fft(V)[:N]
    return np.rint(F.real * (delta / N)).astype(np.int64)


def decode(m, N, delta):
    """Inverse of encode (src/encode.cpp:30-44) via an inverse FFT."""
    r = rot_group(N)
    x = np.zeros(2 * N, np.complex128)
    x[:N] = m.astype(np.float64)
    return np.fft.ifft(x)[r] * (2 * N / delta)


def poly_mul_negacyclic_small(a, b, n):
    out = np.zeros(n, dtype=object)
    for i in np.nonzero(a)[0]:
        for j in np.nonzero(b)[0]:
            k = int(i) + int(j)
            s = 1
            if k >= n:
                k -= n
                s = -1
            out[k] += s * int(a[i]) * int(b[j])
    return out


class KeySet:
    """Secret, public keys and relinearisation key for a parameter set, built
    by the oracle's keygen (tests/test_evalmult_rescale.cpp:34-50 pattern)."""

    def __init__(self, P, seed=101, ek_seed=202, hamming=0):
        self.P = P
        n = P.n
        self.sQP, self.pkAQP, self.pkBQP = O.keygen(n, P.mods, P.roots, seed=seed, hamming=hamming)
        L = P.L
        self.sQ = self.sQP[: L * n]
        self.pkA = self.pkAQP[: L * n]
        self.pkB = self.pkBQP[: L * n]
        self.ek_seed = ek_seed
        self._rlk = None

    @classmethod
    def counter(cls, P, seed=101, ek_seed=202):
        """The same key set drawn by the counter sampler (the library's device
        key generation: O.keygen_ctr / O.evalkeygen_ctr)."""
        ks = cls.__new__(cls)
        ks.P = P
        n, L = P.n, P.L
        ks.sQP, ks.pkAQP, ks.pkBQP = O.keygen_ctr(n, P.mods, P.roots, seed=seed)
        ks.sQ, ks.pkA, ks.pkB = ks.sQP[: L * n], ks.pkAQP[: L * n], ks.pkBQP[: L * n]
        ks.ek_seed = ek_seed
        ks._rlk = O.evalkeygen_ctr(n, P.L, P.K, P.alpha, P.dnum, P.mods, P.roots, ks.s2(), ks.pkAQP, ks.pkBQP,
                                   seed=ek_seed)
        return ks

    def s2(self):
        P = self.P
        n = P.n
        s = self.sQP.reshape(-1, n)
        out = np.empty_like(s)
        for t, q in enumerate(P.mods):
            out[t] = [(int(a) * int(a)) % int(q) for a in s[t]] if n <= 64 else \
                _mulmod_vec(s[t], s[t], int(q))
        return out.reshape(-1)

    def rlk(self):
        if self._rlk is None:
            P = self.P
            self._rlk = O.evalkeygen_sold(P.n, P.L, P.K, P.alpha, P.dnum, P.mods, P.roots,
                                          self.s2(), self.pkAQP, self.pkBQP, seed=self.ek_seed)
        return self._rlk

    def rotkey(self, k, seed=404):
        P = self.P
        s_auto = P.automorph(self.sQP, P.L + P.K, k, P.K)
        return O.evalkeygen_sold(P.n, P.L, P.K, P.alpha, P.dnum, P.mods, P.roots, s_auto,
                                 self.pkAQP, self.pkBQP, seed=seed)


def _mulmod_vec(a, b, q):
    """Exact elementwise (a*b) % q for uint64 arrays (splits into 32-bit halves)."""
    a = a.astype(object)
    b = b.astype(object)
    return np.array((a * b) % q, dtype=np.uint64)


def mulmod_vec(a, b, q):
    return _mulmod_vec(np.asarray(a, np.uint64), np.asarray(b, np.uint64), int(q))


def parallel_map(fn, items, workers=None):
    """fn over items on a thread pool (the oracle's ctypes calls release the
    GIL, so independent oracle operations run in parallel); results in order.
    At most 16 threads, the host share of one GPU's box."""
    import concurrent.futures as cf
    import os
    items = list(items)
    if workers is None:
        workers = max(1, min(16, len(os.sched_getaffinity(0)), len(items)))
    with cf.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(fn, items))
