"""ctypes view of the CPU oracle (liboracle.so) and of the reference's own
host-only code (oracle/_ref/libref.so, present only where the reference
sources were available to build it).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        if name == "liboracle.so":
            build()
        else:
            return None
    return C.CDLL(path)


def P(a):
    """numpy array -> ctypes pointer of the matching element type."""
    if a is None:
        return None
    if a.dtype == np.uint64:
        return a.ctypes.data_as(u64p)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(u32p)
    if a.dtype == np.int64:
        return a.ctypes.data_as(i64p)
    if a.dtype == np.float64:
        return a.ctypes.data_as(f64p)
    raise TypeError(a.dtype)


def U(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


class OrKsc(C.Structure):
    _fields_ = [("sizeQl", C.c_uint32), ("sizeP", C.c_uint32), ("numPart", C.c_uint32),
                ("alpha", C.c_uint32),
                ("sizePart", u32p), ("sizeCompl", u32p), ("startIdx", u32p),
                ("partQHatInv", u64p), ("partQHatInvPrec", u64p), ("partSrcMod", u64p),
                ("partQHatModp", u64p), ("partComplMod", u64p), ("partBMuLo", u64p),
                ("partBMuHi", u64p), ("pHatInv", u64p), ("pHatInvPrec", u64p),
                ("pHatModq", u64p), ("mdBMuLo", u64p), ("mdBMuHi", u64p), ("pInvModq", u64p)]


class OrParams(C.Structure):
    _fields_ = [("n", C.c_uint32), ("L", C.c_uint32), ("K", C.c_uint32), ("alpha", C.c_uint32),
                ("dnum", C.c_uint32), ("mod", u64p), ("root", u64p), ("fr", u64p), ("fp", u64p),
                ("ir", u64p), ("ip", u64p), ("ninv", u64p), ("ninvp", u64p)]


lib = _load("liboracle.so")
_ref = []   # oracle/_ref/libref.so, loaded on first use of `ref` (tests that compare with it)

_sig = {
    "or_mulmod": (C.c_uint64, [C.c_uint64] * 3),
    "or_powmod": (C.c_uint64, [C.c_uint64] * 3),
    "or_invmod": (C.c_uint64, [C.c_uint64] * 2),
    "or_shoup_precon": (C.c_uint64, [C.c_uint64] * 2),
    "or_shoup_mulmod": (C.c_uint64, [C.c_uint64] * 4),
    "or_is_prime": (C.c_int, [C.c_uint64]),
    "or_native_primes": (C.c_int, [u64p, C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_uint32]),
    "or_native_root": (C.c_uint64, [C.c_uint32, C.c_uint64]),
    "or_rescale_consts": (None, [u64p, u64p, u64p, C.c_uint32]),
    "or_root_tables": (None, [C.c_uint32, C.c_uint64, C.c_uint64, u64p, u64p, u64p, u64p, u64p]),
    "or_ntt": (None, [u64p, C.c_uint32, C.c_uint64, u64p, u64p]),
    "or_intt": (None, [u64p, C.c_uint32, C.c_uint64, u64p, u64p, C.c_uint64, C.c_uint64]),
    "or_rns_mult": (None, [u64p, u64p, u64p, C.c_uint64, C.c_uint32]),
    "or_switch_modulus": (None, [u64p, C.c_uint32, C.c_uint64, C.c_uint64]),
    "or_rescale_fuse": (None, [u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]),
    "or_approx_switch": (None, [u64p, u64p, u64p, u64p, u64p, u64p, u64p, C.c_uint32, C.c_uint32,
                                C.c_uint32]),
    "or_ks_macc": (None, [u64p, u64p, u64p, u64p, u64p, C.c_uint64, C.c_uint32]),
    "or_moddown_combine": (None, [u64p, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint32]),
    "or_tensor": (None, [u64p] * 8 + [C.c_uint32, C.c_uint32]),
    "or_combine": (None, [u64p] * 7 + [C.c_uint32, C.c_uint32]),
    "or_add_into": (None, [u64p, u64p, C.c_uint64, C.c_uint32]),
    "or_automorph_coeff": (None, [u64p, u64p, C.c_uint32, C.c_uint64, C.c_uint64]),
    "or_ks_constants": (C.POINTER(OrKsc), [u64p, C.c_uint32, u64p, C.c_uint32, C.c_uint32,
                                           C.c_uint32]),
    "or_ks_constants_free": (None, [C.POINTER(OrKsc)]),
    "or_params_create": (C.POINTER(OrParams), [C.c_uint32, u64p, C.c_uint32, u64p, C.c_uint32,
                                               u64p, C.c_uint32]),
    "or_params_free": (None, [C.POINTER(OrParams)]),
    "or_num_parts": (C.c_uint32, [C.POINTER(OrParams), C.c_uint32]),
    "or_keyswitch": (C.c_int, [C.POINTER(OrParams), C.c_uint32, u64p, u64p, u64p, u64p, u64p]),
    "or_ks_digits": (C.c_int, [C.POINTER(OrParams), C.c_uint32, u64p, u64p]),
    "or_rescale_raw": (C.c_int, [C.POINTER(OrParams), u64p, C.c_uint32]),
    "or_evalmult_rescale": (C.c_int, [C.POINTER(OrParams), C.c_uint32] + [u64p] * 8),
    "or_evalmult_norescale": (C.c_int, [C.POINTER(OrParams), C.c_uint32] + [u64p] * 8),
    "or_automorph_eval": (C.c_int, [C.POINTER(OrParams), u64p, C.c_uint32, C.c_uint32, C.c_int]),
    "or_rotate": (C.c_int, [C.POINTER(OrParams), C.c_uint32, u64p, u64p, C.c_uint32, u64p, u64p,
                            u64p, u64p]),
    "or_rotate_hoisted": (C.c_int, [C.POINTER(OrParams), C.c_uint32, u64p, u64p, C.c_uint32, u32p,
                                    C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u64p)]),
    "or_linear_transform": (C.c_int, [C.POINTER(OrParams), C.c_uint32, u64p, u64p, C.c_uint32, u32p,
                                      C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u64p), u64p, u64p]),
    "or_keygen": (None, [C.c_uint32, u64p, u64p, C.c_uint32, C.c_uint64, C.c_double, C.c_uint64,
                         C.c_uint32, u64p, u64p, u64p]),
    "or_evalkeygen_sold": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u64p,
                                  u64p, u64p, u64p, u64p, C.c_uint64, C.c_double, C.c_uint64,
                                  u64p, u64p]),
    "or_encrypt": (None, [C.c_uint32, u64p, u64p, C.c_uint32, i64p, u64p, u64p, C.c_uint64,
                          C.c_double, C.c_uint64, u64p, u64p]),
    "or_decrypt": (None, [C.c_uint32, C.c_uint64, C.c_uint64, u64p, u64p, u64p, i64p]),
    "or_ctr_draw": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32]),
    "or_ctr_gauss": (C.c_int64, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double]),
    "or_keygen_ctr": (None, [C.c_uint32, u64p, u64p, C.c_uint32, C.c_uint64, C.c_double, C.c_uint64,
                             u64p, u64p, u64p]),
    "or_evalkeygen_ctr": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u64p,
                                 u64p, u64p, u64p, u64p, C.c_uint64, C.c_double, C.c_uint64, u64p,
                                 u64p]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

def _load_ref():
    """The reference's own host code, built into oracle/_ref/libref.so from its
    sources, or None where they were absent.  Loaded only on first access of
    `oracle.ref`, so importing the oracle maps no reference-built code."""
    if _ref:
        return _ref[0]
    r = _load(os.path.join("_ref", "libref.so"))
    if r is not None:
        r.ref_native_primes.restype = C.c_int
        r.ref_native_primes.argtypes = [u64p, C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_uint32]
        r.ref_native_root.restype = C.c_uint64
        r.ref_native_root.argtypes = [C.c_uint32, C.c_uint64]
        r.ref_rescale_consts.restype = None
        r.ref_rescale_consts.argtypes = [u64p, u64p, u64p, C.c_uint32, C.c_uint32]
        r.ref_ks_constants.restype = C.c_int
        r.ref_ks_constants.argtypes = ([u64p, C.c_uint32, u64p, C.c_uint32, C.c_uint32, u32p, u32p,
                                        u32p, u32p] + [u64p] * 13)
        r.ref_encode.restype = None
        r.ref_encode.argtypes = [i64p, f64p, f64p, C.c_uint32, C.c_double]
        r.ref_decode.restype = None
        r.ref_decode.argtypes = [f64p, f64p, i64p, C.c_uint32, C.c_double]
    _ref.append(r)
    return r


def __getattr__(name):   # module attribute hook: `oracle.ref` loads the library lazily
    if name == "ref":
        return _load_ref()
    raise AttributeError(name)


class MT64(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("idx", C.c_int)]


lib.or_mt64_seed.restype = None
lib.or_mt64_seed.argtypes = [C.POINTER(MT64), C.c_uint64]
lib.or_mt64_next.restype = C.c_uint64
lib.or_mt64_next.argtypes = [C.POINTER(MT64)]
lib.or_gauss_sample.restype = C.c_int64
lib.or_gauss_sample.argtypes = [C.POINTER(MT64), C.c_double]


# ---------------------------------------------------------------- helpers --
def native_primes(count, bits, n, avoid=()):
    out = np.zeros(count, np.uint64)
    av = U(list(avoid) or [0])
    rc = lib.or_native_primes(P(out), count, bits, n, P(av), len(avoid))
    if rc:
        raise RuntimeError("prime search exhausted")
    return out


def root_tables(n, q, root):
    fr, fp, ir, ip = (np.zeros(n, np.uint64) for _ in range(4))
    pair = np.zeros(2, np.uint64)
    lib.or_root_tables(n, q, root, P(fr), P(fp), P(ir), P(ip), P(pair))
    return fr, fp, ir, ip, int(pair[0]), int(pair[1])


def ntt(x, q, fr, fp):
    y = U(x).copy()
    lib.or_ntt(P(y), len(y), q, P(fr), P(fp))
    return y


def intt(x, q, ir, ip, ninv, ninvp):
    y = U(x).copy()
    lib.or_intt(P(y), len(y), q, P(ir), P(ip), ninv, ninvp)
    return y


class Params:
    """A parameter set (Q moduli, P moduli, roots, dnum) with root tables."""

    def __init__(self, n, qmods, pmods, roots, dnum):
        self.n = n
        self.q = U(qmods)
        self.p = U(pmods)
        self.L = len(qmods)
        self.K = len(pmods)
        self.dnum = dnum
        self.mods = U(list(qmods) + list(pmods))
        self.roots = U(roots)
        self.alpha = (self.L + dnum - 1) // dnum
        self.h = lib.or_params_create(n, P(self.q), self.L, P(self.p), self.K, P(self.roots), dnum)

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_params_free(self.h)
            self.h = None

    def num_parts(self, level):
        return int(lib.or_num_parts(self.h, level))

    def keyswitch(self, level, d, av, bv):
        T = level * self.n
        ba0 = np.zeros(T, np.uint64)
        ba1 = np.zeros(T, np.uint64)
        rc = lib.or_keyswitch(self.h, level, P(U(d)), P(U(av)), P(U(bv)), P(ba0), P(ba1))
        if rc:
            raise RuntimeError("or_keyswitch failed")
        return ba0, ba1

    def ks_digits(self, level, d):
        W = level + self.K
        dg = np.zeros(self.num_parts(level) * W * self.n, np.uint64)
        if lib.or_ks_digits(self.h, level, P(U(d)), P(dg)):
            raise RuntimeError("or_ks_digits failed")
        return dg

    def rescale(self, c, towers):
        y = U(c).copy()
        lib.or_rescale_raw(self.h, P(y), towers)
        return y[: (towers - 1) * self.n]

    def evalmult_rescale(self, level, c0a, c1a, c0b, c1b, av, bv):
        T = (level - 1) * self.n
        o0 = np.zeros(T, np.uint64)
        o1 = np.zeros(T, np.uint64)
        rc = lib.or_evalmult_rescale(self.h, level, P(U(c0a)), P(U(c1a)), P(U(c0b)), P(U(c1b)),
                                     P(U(av)), P(U(bv)), P(o0), P(o1))
        if rc:
            raise RuntimeError("or_evalmult_rescale failed")
        return o0, o1

    def evalmult(self, level, c0a, c1a, c0b, c1b, av, bv):
        T = level * self.n
        o0 = np.zeros(T, np.uint64)
        o1 = np.zeros(T, np.uint64)
        rc = lib.or_evalmult_norescale(self.h, level, P(U(c0a)), P(U(c1a)), P(U(c0b)), P(U(c1b)),
                                       P(U(av)), P(U(bv)), P(o0), P(o1))
        if rc:
            raise RuntimeError("or_evalmult_norescale failed")
        return o0, o1

    def automorph(self, v, towers, k, p_towers=0):
        y = U(v).copy()
        lib.or_automorph_eval(self.h, P(y), towers, k, p_towers)
        return y

    def rotate(self, level, c0, c1, k, av, bv):
        T = level * self.n
        o0 = np.zeros(T, np.uint64)
        o1 = np.zeros(T, np.uint64)
        rc = lib.or_rotate(self.h, level, P(U(c0)), P(U(c1)), k, P(U(av)), P(U(bv)), P(o0), P(o1))
        if rc:
            raise RuntimeError("or_rotate failed")
        return o0, o1

    def rotate_hoisted(self, level, c0, c1, ks, keys):
        """m rotations of one ciphertext sharing one decomposition of c1;
        keys = [(av, bv), ...] one per automorphism index in ks."""
        m, T = len(ks), level * self.n
        outs = [(np.zeros(T, np.uint64), np.zeros(T, np.uint64)) for _ in range(m)]
        avs = [U(a) for a, _ in keys]
        bvs = [U(b) for _, b in keys]
        arr = lambda xs: (u64p * m)(*[P(x) for x in xs])  # noqa: E731
        kk = np.asarray(ks, np.uint32)
        rc = lib.or_rotate_hoisted(self.h, level, P(U(c0)), P(U(c1)), m, P(kk), arr(avs), arr(bvs),
                                   arr([o[0] for o in outs]), arr([o[1] for o in outs]))
        if rc:
            raise RuntimeError("or_rotate_hoisted failed")
        return outs

    def linear_transform(self, level, c0, c1, ks, keys, pts):
        """sum_j pts[j] * rot_ks[j](c0, c1); keys[j] = (av, bv), or None with
        ks[j] = 1 for the unrotated term.  pts[j]: level*n, evaluation form."""
        m, T = len(ks), level * self.n
        o0 = np.zeros(T, np.uint64)
        o1 = np.zeros(T, np.uint64)
        avs = [None if k is None else U(k[0]) for k in keys]
        bvs = [None if k is None else U(k[1]) for k in keys]
        ptv = [U(p) for p in pts]
        arr = lambda xs: (u64p * m)(*[None if x is None else P(x) for x in xs])  # noqa: E731
        kk = np.asarray(ks, np.uint32)
        rc = lib.or_linear_transform(self.h, level, P(U(c0)), P(U(c1)), m, P(kk), arr(avs), arr(bvs),
                                     arr(ptv), P(o0), P(o1))
        if rc:
            raise RuntimeError("or_linear_transform failed")
        return o0, o1


def keygen(n, mods, roots, ns=1, sigma=3.19, seed=101, hamming=0):
    mods = U(mods)
    roots = U(roots)
    T = len(mods) * n
    s, a, b = (np.zeros(T, np.uint64) for _ in range(3))
    lib.or_keygen(n, P(mods), P(roots), len(mods), ns, sigma, seed, hamming, P(s), P(a), P(b))
    return s, a, b


def evalkeygen_sold(n, sizeQ, sizeP, alpha, numPart, modQP, rootQP, s_old, pkA, pkB, ns=1,
                    sigma=3.19, seed=202):
    W = sizeQ + sizeP
    av = np.zeros(numPart * W * n, np.uint64)
    bv = np.zeros(numPart * W * n, np.uint64)
    lib.or_evalkeygen_sold(n, sizeQ, sizeP, alpha, numPart, P(U(modQP)), P(U(rootQP)),
                           P(U(s_old)), P(U(pkA)), P(U(pkB)), ns, sigma, seed, P(av), P(bv))
    return av, bv


def keygen_ctr(n, mods, roots, ns=1, sigma=3.19, seed=101):
    """or_keygen with the counter sampler (the library's device key generation)."""
    mods = U(mods)
    roots = U(roots)
    T = len(mods) * n
    s, a, b = (np.zeros(T, np.uint64) for _ in range(3))
    lib.or_keygen_ctr(n, P(mods), P(roots), len(mods), ns, sigma, seed, P(s), P(a), P(b))
    return s, a, b


def evalkeygen_ctr(n, sizeQ, sizeP, alpha, numPart, modQP, rootQP, s_old, pkA, pkB, ns=1, sigma=3.19,
                   seed=202):
    W = sizeQ + sizeP
    av = np.zeros(numPart * W * n, np.uint64)
    bv = np.zeros(numPart * W * n, np.uint64)
    lib.or_evalkeygen_ctr(n, sizeQ, sizeP, alpha, numPart, P(U(modQP)), P(U(rootQP)), P(U(s_old)),
                          P(U(pkA)), P(U(pkB)), ns, sigma, seed, P(av), P(bv))
    return av, bv


def encrypt(n, mods, roots, m, pkA, pkB, ns=1, sigma=3.19, seed=303):
    mods = U(mods)
    T = len(mods) * n
    c0 = np.zeros(T, np.uint64)
    c1 = np.zeros(T, np.uint64)
    mm = np.ascontiguousarray(np.asarray(m, dtype=np.int64))
    lib.or_encrypt(n, P(mods), P(U(roots)), len(mods), P(mm), P(U(pkA)), P(U(pkB)), ns, sigma,
                   seed, P(c0), P(c1))
    return c0, c1


def decrypt(n, q0, root0, c0, c1, s):
    out = np.zeros(n, np.int64)
    lib.or_decrypt(n, q0, root0, P(U(c0)), P(U(c1)), P(U(s)), P(out))
    return out


def ks_constants(qmods, pmods, numPart, alpha=0):
    """or_ks_constants as a dict of numpy arrays (reference flat layout)."""
    q = U(qmods)
    p = U(pmods)
    h = lib.or_ks_constants(P(q), len(q), P(p), len(p), numPart, alpha)
    if not h:
        raise ValueError("invalid part layout")
    k = h.contents
    W = len(q) + len(p)
    a, npart = k.alpha, k.numPart

    def arr(ptr, cnt, dt=np.uint64):
        return np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt).copy()

    d = dict(alpha=a, numPart=npart,
             sizePart=arr(k.sizePart, npart, np.uint32), sizeCompl=arr(k.sizeCompl, npart, np.uint32),
             startIdx=arr(k.startIdx, npart, np.uint32),
             partQHatInv=arr(k.partQHatInv, npart * a), partQHatInvPrec=arr(k.partQHatInvPrec, npart * a),
             partSrcMod=arr(k.partSrcMod, npart * a), partQHatModp=arr(k.partQHatModp, npart * a * W),
             partComplMod=arr(k.partComplMod, npart * W), partBMuLo=arr(k.partBMuLo, npart * W),
             partBMuHi=arr(k.partBMuHi, npart * W), pHatInv=arr(k.pHatInv, len(p)),
             pHatInvPrec=arr(k.pHatInvPrec, len(p)), pHatModq=arr(k.pHatModq, len(p) * len(q)),
             mdBMuLo=arr(k.mdBMuLo, len(q)), mdBMuHi=arr(k.mdBMuHi, len(q)),
             pInvModq=arr(k.pInvModq, len(q)))
    lib.or_ks_constants_free(h)
    return d


def ref_ks_constants(qmods, pmods, numPart):
    """The reference's compute_keyswitch_constants (oracle/_ref) in the same layout."""
    q = U(qmods)
    p = U(pmods)
    W = len(q) + len(p)
    a = -(-len(q) // numPart)
    alpha = np.zeros(1, np.uint32)
    sp, sc, st = (np.zeros(numPart, np.uint32) for _ in range(3))
    qhi, qhip, src = (np.zeros(numPart * a, np.uint64) for _ in range(3))
    qhmp = np.zeros(numPart * a * W, np.uint64)
    cm, mlo, mhi = (np.zeros(numPart * W, np.uint64) for _ in range(3))
    phi, phip = np.zeros(len(p), np.uint64), np.zeros(len(p), np.uint64)
    phmq = np.zeros(len(p) * len(q), np.uint64)
    dlo, dhi, pinv = (np.zeros(len(q), np.uint64) for _ in range(3))
    _load_ref().ref_ks_constants(P(q), len(q), P(p), len(p), numPart, P(alpha), P(sp), P(sc), P(st), P(qhi),
                         P(qhip), P(src), P(qhmp), P(cm), P(mlo), P(mhi), P(phi), P(phip), P(phmq),
                         P(dlo), P(dhi), P(pinv))
    return dict(alpha=int(alpha[0]), numPart=numPart, sizePart=sp, sizeCompl=sc, startIdx=st,
                partQHatInv=qhi, partQHatInvPrec=qhip, partSrcMod=src, partQHatModp=qhmp,
                partComplMod=cm, partBMuLo=mlo, partBMuHi=mhi, pHatInv=phi, pHatInvPrec=phip,
                pHatModq=phmq, mdBMuLo=dlo, mdBMuHi=dhi, pInvModq=pinv)
