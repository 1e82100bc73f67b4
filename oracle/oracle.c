/*
 * CPU oracle (test infrastructure only; see oracle.h for the pinning map).
 * Reference paths are relative to the reference repository root.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* scalar arithmetic                                                   */
/* ------------------------------------------------------------------ */

/* (a*b) mod q -- the reference's plain 128-bit remainder
 * (kernels/tensor.cu:7-8, kernels/ks_fastcore.cu:12-14). */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

/* src/params.cpp:16 */
uint64_t or_powmod(uint64_t b, uint64_t e, uint64_t q) {
    u128 r = 1, bb = b % q;
    while (e) {
        if (e & 1) r = (r * bb) % q;
        bb = (bb * bb) % q;
        e >>= 1;
    }
    return (uint64_t)r;
}

/* src/ks_constants.cpp:17 (q prime) */
uint64_t or_invmod(uint64_t a, uint64_t q) { return or_powmod(a % q, q - 2, q); }

/* floor(w * 2^64 / q): src/ks_constants.cpp:19 */
uint64_t or_shoup_precon(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

/* x*w mod q with one correction: kernels/ntt_ct.cu:11-16 */
uint64_t or_shoup_mulmod(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t hi = (uint64_t)(((u128)x * wp) >> 64);
    uint64_t r = x * w - hi * q;
    return r >= q ? r - q : r;
}

static inline uint64_t addm(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}
static inline uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
/* signed small integer -> [0,q): src/keygen.cpp:93 */
static inline uint64_t embed(int64_t v, uint64_t q) {
    int64_t r = v % (int64_t)q;
    return (uint64_t)(r < 0 ? r + (int64_t)q : r);
}

/* ------------------------------------------------------------------ */
/* mt19937_64 -- the standard engine (C++ [rand.predef]); the reference */
/* seeds one per sampling call (src/keygen.cpp:60).                     */
/* ------------------------------------------------------------------ */
void or_mt64_seed(or_mt64* g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}
uint64_t or_mt64_next(or_mt64* g) {
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[(i + 1) % 312] & LM);
            uint64_t xa = x >> 1;
            if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/*
 * Discrete Gaussian, rounded-normal shape with standard deviation sigma.
 * The reference draws std::lround(std::normal_distribution(0,sigma)(rng))
 * (src/keygen.cpp:61,83), whose output is implementation-defined; this build
 * specifies its own sampler so the library and the oracle draw identical
 * noise:  weights w_0 = 1, w_k = 2 exp(-k^2 / (2 sigma^2)) for 1 <= k <= 12 sigma;
 * T[k] = (uint64)(cum_k / total * 2^63);  r = next() >> 1; k = first index
 * with r < T[k] (last index if none); if k > 0 the sign is next() & 1.
 */
#define GMAX 256
/* the table of the Gaussian sampler for sigma: K and T[0..K] */
static void gauss_table(double sigma, uint64_t* T, int* Kout) {
    int K = (int)ceil(sigma * 12.0);
    if (K > GMAX) K = GMAX;
    double w[GMAX + 1], total = 0.0, cum = 0.0;
    for (int k = 0; k <= K; ++k) {
        w[k] = (k == 0) ? 1.0 : 2.0 * exp(-(double)k * (double)k / (2.0 * sigma * sigma));
        total += w[k];
    }
    for (int k = 0; k <= K; ++k) {
        cum += w[k];
        T[k] = (uint64_t)ldexp(cum / total, 63);
    }
    *Kout = K;
}
int64_t or_gauss_sample(or_mt64* g, double sigma) {
    static double cached_sigma = -1.0;
    static uint64_t T[GMAX + 1];
    static int K = 0;
    if (sigma != cached_sigma) {
        gauss_table(sigma, T, &K);
        cached_sigma = sigma;
    }
    uint64_t r = or_mt64_next(g) >> 1;
    int k = 0;
    while (k < K && r >= T[k]) ++k;
    if (k == 0) return 0;
    return (or_mt64_next(g) & 1) ? -(int64_t)k : (int64_t)k;
}

/* ------------------------------------------------------------------ */
/* parameters                                                          */
/* ------------------------------------------------------------------ */

/* Miller-Rabin over the first 12 prime bases: src/params.cpp:17-25 */
int or_is_prime(uint64_t x) {
    static const uint64_t B[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (x < 2) return 0;
    for (int i = 0; i < 12; ++i)
        if (x % B[i] == 0) return x == B[i];
    uint64_t d = x - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (int i = 0; i < 12; ++i) {
        uint64_t v = or_powmod(B[i], d, x);
        if (v == 1 || v == x - 1) continue;
        int ok = 0;
        for (int r = 1; r < s; ++r) {
            v = or_mulmod(v, v, x);
            if (v == x - 1) { ok = 1; break; }
        }
        if (!ok) return 0;
    }
    return 1;
}

/* First `count` primes q = 1 mod 2n descending from floor(2^bits/2n)*2n+1,
 * skipping `avoid`: src/params.cpp:29-39.  Returns -1 when exhausted. */
int or_native_primes(uint64_t* out, uint32_t count, uint32_t bits, uint32_t n,
                     const uint64_t* avoid, uint32_t navoid) {
    const uint64_t M = 2ULL * n;
    uint64_t q = ((1ULL << bits) / M) * M + 1;
    uint32_t got = 0;
    while (got < count) {
        if (q < (1ULL << (bits - 1))) return -1;
        int skip = 0;
        for (uint32_t a = 0; a < navoid; ++a)
            if (avoid[a] == q) skip = 1;
        if (!skip && or_is_prime(q)) out[got++] = q;
        q -= M;
    }
    return 0;
}

/* first g^((q-1)/2n), g = 2,3,..., whose n-th power is -1: src/params.cpp:42-48 */
uint64_t or_native_root(uint32_t n, uint64_t q) {
    const uint64_t M = 2ULL * n;
    for (uint64_t g = 2; g < q; ++g) {
        uint64_t c = or_powmod(g, (q - 1) / M, q);
        if (or_powmod(c, n, q) == q - 1) return c;
    }
    return 0;
}

/* (a*b) mod m for m < 2^127 by double-and-add */
static u128 mulmod128(u128 a, u128 b, u128 m) {
    a %= m;
    u128 r = 0;
    while (b) {
        if (b & 1) { r += a; if (r >= m) r -= m; }
        a <<= 1; if (a >= m) a -= m;
        b >>= 1;
    }
    return r;
}

/* s1[i] = q_last^-1 mod q_i;  s2[i] = floor(Q_l [Q_l^-1]_{q_last} / q_last) mod q_i
 * with Q_l = prod_{t<last} q_t:  src/params.cpp:54-75 (same folding). */
void or_rescale_consts(uint64_t* s1, uint64_t* s2, const uint64_t* mod, uint32_t last) {
    uint64_t ql = mod[last];
    uint64_t Qmod = 1 % ql;
    for (uint32_t t = 0; t < last; ++t) Qmod = or_mulmod(Qmod, mod[t] % ql, ql);
    uint64_t u = or_powmod(Qmod, ql - 2, ql);
    for (uint32_t i = 0; i < last; ++i) {
        uint64_t qi = mod[i];
        s1[i] = or_powmod(ql % qi, qi - 2, qi);
        u128 m2 = (u128)ql * qi;
        u128 Q2 = 1 % m2;
        for (uint32_t t = 0; t < last; ++t) Q2 = mulmod128(Q2, mod[t], m2);
        u128 prod = mulmod128(Q2, u, m2);
        u128 pm1 = (prod == 0) ? m2 - 1 : prod - 1;
        s2[i] = (uint64_t)((pm1 / ql) % qi);
    }
}

static uint32_t brv(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t k = 0; k < bits; ++k) { r = (r << 1) | (v & 1); v >>= 1; }
    return r;
}
static uint32_t ilog2u(uint32_t n) {
    uint32_t l = 0;
    while ((1u << l) < n) ++l;
    return l;
}

/* fr[i] = psi^brv(i), ir[i] = psi^-brv(i), Shoup precons, n^-1:
 * src/keyswitch.cpp:22-35 */
void or_root_tables(uint32_t n, uint64_t q, uint64_t root, uint64_t* fr, uint64_t* fp,
                    uint64_t* ir, uint64_t* ip, uint64_t* ninv_pair) {
    uint32_t logn = ilog2u(n);
    uint64_t* pw = (uint64_t*)malloc((size_t)n * 8);
    uint64_t bases[2] = {root, or_invmod(root, q)};
    uint64_t* tr[2] = {fr, ir};
    uint64_t* tp[2] = {fp, ip};
    for (int s = 0; s < 2; ++s) {
        pw[0] = 1;
        for (uint32_t i = 1; i < n; ++i) pw[i] = or_mulmod(pw[i - 1], bases[s], q);
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t r = pw[brv(i, logn)];
            tr[s][i] = r;
            tp[s][i] = or_shoup_precon(r, q);
        }
    }
    free(pw);
    ninv_pair[0] = or_invmod(n, q);
    ninv_pair[1] = or_shoup_precon(ninv_pair[0], q);
}

/* ------------------------------------------------------------------ */
/* per-tower primitives                                                */
/* ------------------------------------------------------------------ */

/* Cooley-Tukey, standard order in, bit-reversed out, twiddle roots[m+i]:
 * kernels/ntt_ct.cu:19-39 (stage), :92-112 (stage sequence). */
void or_ntt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* fr, const uint64_t* fp) {
    for (uint32_t m = 1, t = n >> 1; m < n; m <<= 1, t >>= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            uint64_t w = fr[m + i], wp = fp[m + i];
            uint32_t j1 = 2 * i * t;
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t u = x[j];
                uint64_t v = or_shoup_mulmod(x[j + t], w, wp, q);
                x[j] = addm(u, v, q);
                x[j + t] = subm(u, v, q);
            }
        }
    }
}

/* Gentleman-Sande, bit-reversed in, standard out, then x n^-1:
 * kernels/intt_ct.cu:21-42 (stage), :86-90 (scale), :94-121 (sequence). */
void or_intt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* ir, const uint64_t* ip,
             uint64_t ninv, uint64_t ninv_p) {
    for (uint32_t m = n >> 1, t = 1; m >= 1; m >>= 1, t <<= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            uint64_t w = ir[m + i], wp = ip[m + i];
            uint32_t j1 = 2 * i * t;
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t u = x[j], v = x[j + t];
                x[j] = addm(u, v, q);
                x[j + t] = or_shoup_mulmod(subm(u, v, q), w, wp, q);
            }
        }
    }
    for (uint32_t j = 0; j < n; ++j) x[j] = or_shoup_mulmod(x[j], ninv, ninv_p, q);
}

/* r = a*b mod q: kernels/rns_arith.cu:21-33 (Montgomery, output == (a*b)%q) */
void or_rns_mult(const uint64_t* a, const uint64_t* b, uint64_t* r, uint64_t q, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) r[i] = or_mulmod(a[i], b[i], q);
}

/* centred switch q_old -> q_new, strict '>' half test: kernels/switch_modulus.cu:11-42 */
void or_switch_modulus(uint64_t* v, uint32_t n, uint64_t q_old, uint64_t q_new) {
    uint64_t half = q_old >> 1;
    if (q_new > q_old) {
        uint64_t diff = q_new - q_old;
        for (uint32_t i = 0; i < n; ++i) v[i] = v[i] + (v[i] > half ? diff : 0);
    } else {
        uint64_t diff = (q_old - q_new) % q_new;
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t d = v[i] > half ? diff : 0;
            uint64_t xm = v[i] >= q_new ? v[i] % q_new : v[i];
            v[i] = xm >= d ? xm - d : xm + q_new - d;
        }
    }
}

/* a = a*s1 + b*s2 mod q: kernels/rescale.cu:10-21 */
void or_rescale_fuse(uint64_t* a, const uint64_t* b, uint64_t s1, uint64_t s2, uint64_t q,
                     uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) a[i] = addm(or_mulmod(a[i], s1, q), or_mulmod(b[i], s2, q), q);
}

/* dst_j = (sum_i [src_i QHatInv_i]_{q_i} QHatModp[i][j]) mod p_j, exact
 * (Barrett-128 with mu = floor((2^128-1)/p) returns the exact remainder):
 * kernels/basis_convert.cu:24-55 (Barrett), :61-101 (kernel). */
void or_approx_switch(const uint64_t* src, uint64_t* dst, const uint64_t* QHatInv,
                      const uint64_t* QHatInvPre, const uint64_t* q, const uint64_t* QHatModp,
                      const uint64_t* p, uint32_t sizeQ, uint32_t sizeP, uint32_t n) {
    uint64_t* xi = (uint64_t*)malloc((size_t)sizeQ * 8);
    for (uint32_t r = 0; r < n; ++r) {
        for (uint32_t i = 0; i < sizeQ; ++i)
            xi[i] = or_shoup_mulmod(src[(size_t)i * n + r], QHatInv[i], QHatInvPre[i], q[i]);
        for (uint32_t j = 0; j < sizeP; ++j) {
            u128 sum = 0;
            for (uint32_t i = 0; i < sizeQ; ++i) sum += (u128)xi[i] * QHatModp[(size_t)i * sizeP + j];
            dst[(size_t)j * n + r] = (uint64_t)(sum % p[j]);
        }
    }
    free(xi);
}

/* out0 += digit*b, out1 += digit*a (mod q): kernels/ks_fastcore.cu:18-31 */
void or_ks_macc(uint64_t* out0, uint64_t* out1, const uint64_t* digit, const uint64_t* b,
                const uint64_t* a, uint64_t q, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) {
        out0[k] = addm(out0[k], or_mulmod(digit[k], b[k], q), q);
        out1[k] = addm(out1[k], or_mulmod(digit[k], a[k], q), q);
    }
}

/* ans = (a - b) * s mod q: kernels/moddown.cu:10-22 */
void or_moddown_combine(uint64_t* ans, const uint64_t* a, const uint64_t* b, uint64_t s,
                        uint64_t q, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) ans[k] = or_mulmod(subm(a[k], b[k], q), s, q);
}

/* t0 = a0 b0, t1 = a0 b1 + a1 b0, t2 = a1 b1: kernels/tensor.cu:11-27 */
void or_tensor(const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
               const uint64_t* c1b, uint64_t* t0, uint64_t* t1, uint64_t* t2,
               const uint64_t* mods, uint32_t sizeQ, uint32_t n) {
    for (uint32_t t = 0; t < sizeQ; ++t) {
        uint64_t q = mods[t];
        for (uint32_t k = 0; k < n; ++k) {
            size_t x = (size_t)t * n + k;
            t0[x] = or_mulmod(c0a[x], c0b[x], q);
            t1[x] = addm(or_mulmod(c0a[x], c1b[x], q), or_mulmod(c1a[x], c0b[x], q), q);
            t2[x] = or_mulmod(c1a[x], c1b[x], q);
        }
    }
}

/* r0 = t0 + ba0, r1 = t1 + ba1: kernels/tensor.cu:39-52 */
void or_combine(uint64_t* r0, uint64_t* r1, const uint64_t* t0, const uint64_t* t1,
                const uint64_t* ba0, const uint64_t* ba1, const uint64_t* mods,
                uint32_t sizeQ, uint32_t n) {
    for (uint32_t t = 0; t < sizeQ; ++t) {
        uint64_t q = mods[t];
        for (uint32_t k = 0; k < n; ++k) {
            size_t x = (size_t)t * n + k;
            r0[x] = addm(t0[x], ba0[x], q);
            r1[x] = addm(t1[x], ba1[x], q);
        }
    }
}

/* a += b mod q: kernels/stage_acc.cu:41-49 */
void or_add_into(uint64_t* a, const uint64_t* b, uint64_t q, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) a[k] = addm(a[k], b[k], q);
}

/* coefficient-domain sigma_k: out[(j k) mod 2n] = +-in[j]: kernels/automorph.cu:9-20
 * (out is fully overwritten since j -> j k mod 2n is a bijection onto residues
 * mod n when k is odd). */
void or_automorph_coeff(uint64_t* out, const uint64_t* in, uint32_t n, uint64_t k, uint64_t q) {
    uint64_t M = 2ULL * n;
    for (uint32_t j = 0; j < n; ++j) {
        uint64_t idx = ((uint64_t)j * k) % M;
        uint64_t v = in[j];
        if (idx < n) out[idx] = v;
        else out[idx - n] = v ? q - v : 0;
    }
}

/* ------------------------------------------------------------------ */
/* keyswitch constants: src/ks_constants.cpp:31-102                    */
/* ------------------------------------------------------------------ */
static uint64_t prod_mod(const uint64_t* mods, uint32_t cnt, uint64_t q, int skip) {
    uint64_t r = 1 % q;
    for (uint32_t k = 0; k < cnt; ++k) {
        if ((int)k == skip) continue;
        r = or_mulmod(r, mods[k] % q, q);
    }
    return r;
}

or_ksc* or_ks_constants(const uint64_t* moduliQ, uint32_t sizeQ, const uint64_t* moduliP,
                        uint32_t sizeP, uint32_t numPart, uint32_t alpha) {
    if (alpha == 0) alpha = (sizeQ + numPart - 1) / numPart;   /* ceil(sizeQ/numPart) :39 */
    if (numPart == 0) numPart = (sizeQ + alpha - 1) / alpha;
    if (alpha * (numPart - 1) >= sizeQ) return NULL;           /* empty trailing part */
    const uint32_t W = sizeQ + sizeP;
    or_ksc* K = (or_ksc*)calloc(1, sizeof(or_ksc));
    K->sizeQl = sizeQ; K->sizeP = sizeP; K->numPart = numPart; K->alpha = alpha;
    K->sizePart = (uint32_t*)calloc(numPart, 4);
    K->sizeCompl = (uint32_t*)calloc(numPart, 4);
    K->startIdx = (uint32_t*)calloc(numPart, 4);
    K->partQHatInv = (uint64_t*)calloc((size_t)numPart * alpha, 8);
    K->partQHatInvPrec = (uint64_t*)calloc((size_t)numPart * alpha, 8);
    K->partSrcMod = (uint64_t*)calloc((size_t)numPart * alpha, 8);
    K->partQHatModp = (uint64_t*)calloc((size_t)numPart * alpha * W, 8);
    K->partComplMod = (uint64_t*)calloc((size_t)numPart * W, 8);
    K->partBMuLo = (uint64_t*)calloc((size_t)numPart * W, 8);
    K->partBMuHi = (uint64_t*)calloc((size_t)numPart * W, 8);
    K->pHatInv = (uint64_t*)calloc(sizeP, 8);
    K->pHatInvPrec = (uint64_t*)calloc(sizeP, 8);
    K->pHatModq = (uint64_t*)calloc((size_t)sizeP * sizeQ, 8);
    K->mdBMuLo = (uint64_t*)calloc(sizeQ, 8);
    K->mdBMuHi = (uint64_t*)calloc(sizeQ, 8);
    K->pInvModq = (uint64_t*)calloc(sizeQ, 8);

    /* ModDown family :45-63 */
    for (uint32_t i = 0; i < sizeQ; ++i)
        K->pInvModq[i] = or_invmod(prod_mod(moduliP, sizeP, moduliQ[i], -1), moduliQ[i]);
    for (uint32_t j = 0; j < sizeP; ++j) {
        uint64_t pj = moduliP[j];
        K->pHatInv[j] = or_invmod(prod_mod(moduliP, sizeP, pj, (int)j), pj);
        K->pHatInvPrec[j] = or_shoup_precon(K->pHatInv[j], pj);
        for (uint32_t i = 0; i < sizeQ; ++i)
            K->pHatModq[(size_t)j * sizeQ + i] = prod_mod(moduliP, sizeP, moduliQ[i], (int)j);
    }
    for (uint32_t i = 0; i < sizeQ; ++i) {
        u128 mu = (~(u128)0) / moduliQ[i];
        K->mdBMuLo[i] = (uint64_t)mu; K->mdBMuHi[i] = (uint64_t)(mu >> 64);
    }
    /* decompose families :65-101 */
    for (uint32_t part = 0; part < numPart; ++part) {
        uint32_t st = alpha * part;
        uint32_t sp = (sizeQ > st + alpha) ? alpha : (sizeQ - st);
        K->startIdx[part] = st; K->sizePart[part] = sp;
        const uint64_t* src = moduliQ + st;
        for (uint32_t i = 0; i < sp; ++i) {
            uint64_t q = src[i];
            uint64_t v = or_invmod(prod_mod(src, sp, q, (int)i), q);
            K->partSrcMod[(size_t)part * alpha + i] = q;
            K->partQHatInv[(size_t)part * alpha + i] = v;
            K->partQHatInvPrec[(size_t)part * alpha + i] = or_shoup_precon(v, q);
        }
        /* complement: non-part Q towers in index order, then P (:85-92) */
        uint32_t sc = 0;
        uint64_t* cm = K->partComplMod + (size_t)part * W;
        for (uint32_t i = 0; i < sizeQ; ++i)
            if (i < st || i >= st + sp) cm[sc++] = moduliQ[i];
        for (uint32_t j = 0; j < sizeP; ++j) cm[sc++] = moduliP[j];
        K->sizeCompl[part] = sc;
        /* [sizePart x sizeCompl] packed inside the part's alpha*W block (:93-96) */
        for (uint32_t i = 0; i < sp; ++i)
            for (uint32_t j = 0; j < sc; ++j)
                K->partQHatModp[(size_t)part * alpha * W + (size_t)i * sc + j] = prod_mod(src, sp, cm[j], (int)i);
        for (uint32_t j = 0; j < sc; ++j) {
            u128 mu = (~(u128)0) / cm[j];
            K->partBMuLo[(size_t)part * W + j] = (uint64_t)mu;
            K->partBMuHi[(size_t)part * W + j] = (uint64_t)(mu >> 64);
        }
    }
    return K;
}

void or_ks_constants_free(or_ksc* K) {
    if (!K) return;
    free(K->sizePart); free(K->sizeCompl); free(K->startIdx);
    free(K->partQHatInv); free(K->partQHatInvPrec); free(K->partSrcMod); free(K->partQHatModp);
    free(K->partComplMod); free(K->partBMuLo); free(K->partBMuHi);
    free(K->pHatInv); free(K->pHatInvPrec); free(K->pHatModq);
    free(K->mdBMuLo); free(K->mdBMuHi); free(K->pInvModq);
    free(K);
}

/* ------------------------------------------------------------------ */
/* parameter set                                                       */
/* ------------------------------------------------------------------ */
or_params* or_params_create(uint32_t n, const uint64_t* qmods, uint32_t L,
                            const uint64_t* pmods, uint32_t K, const uint64_t* roots,
                            uint32_t dnum) {
    or_params* P = (or_params*)calloc(1, sizeof(or_params));
    uint32_t M = L + K;
    P->n = n; P->L = L; P->K = K; P->dnum = dnum;
    P->alpha = (L + dnum - 1) / dnum;
    P->mod = (uint64_t*)malloc((size_t)M * 8);
    P->root = (uint64_t*)malloc((size_t)M * 8);
    memcpy(P->mod, qmods, (size_t)L * 8);
    memcpy(P->mod + L, pmods, (size_t)K * 8);
    memcpy(P->root, roots, (size_t)M * 8);
    P->fr = (uint64_t*)malloc((size_t)M * n * 8);
    P->fp = (uint64_t*)malloc((size_t)M * n * 8);
    P->ir = (uint64_t*)malloc((size_t)M * n * 8);
    P->ip = (uint64_t*)malloc((size_t)M * n * 8);
    P->ninv = (uint64_t*)malloc((size_t)M * 8);
    P->ninvp = (uint64_t*)malloc((size_t)M * 8);
    for (uint32_t r = 0; r < M; ++r) {
        uint64_t pair[2];
        size_t o = (size_t)r * n;
        or_root_tables(n, P->mod[r], P->root[r], P->fr + o, P->fp + o, P->ir + o, P->ip + o, pair);
        P->ninv[r] = pair[0]; P->ninvp[r] = pair[1];
    }
    return P;
}

void or_params_free(or_params* P) {
    if (!P) return;
    free(P->mod); free(P->root); free(P->fr); free(P->fp); free(P->ir); free(P->ip);
    free(P->ninv); free(P->ninvp);
    free(P);
}

uint32_t or_num_parts(const or_params* P, uint32_t sizeQl) {
    return (sizeQl + P->alpha - 1) / P->alpha;
}

static void ntt_row(const or_params* P, uint32_t row, uint64_t* x) {
    size_t o = (size_t)row * P->n;
    or_ntt(x, P->n, P->mod[row], P->fr + o, P->fp + o);
}
static void intt_row(const or_params* P, uint32_t row, uint64_t* x) {
    size_t o = (size_t)row * P->n;
    or_intt(x, P->n, P->mod[row], P->ir + o, P->ip + o, P->ninv[row], P->ninvp[row]);
}
/* row of a modulus in the full Q|P list (row_of: src/keyswitch_resident.cpp:134-137) */
static uint32_t row_of(const or_params* P, uint64_t q) {
    for (uint32_t r = 0; r < P->L + P->K; ++r)
        if (P->mod[r] == q) return r;
    return 0xFFFFFFFFu;
}

/*
 * Digit decomposition (ModUp) for all parts at this level:
 * digits[p][i] for i < sizeQl + K, eval form.  Tower i of part p is d[i]
 * inside the part's range, otherwise the NTT of the basis-converted
 * complement tower (src/keyswitch_resident.cpp:149-200, :191-194).
 */
int or_ks_digits(const or_params* P, uint32_t sizeQl, const uint64_t* d, uint64_t* digits) {
    const uint32_t n = P->n, K = P->K, W = sizeQl + K;
    const uint32_t np = or_num_parts(P, sizeQl);
    or_ksc* C = or_ks_constants(P->mod, sizeQl, P->mod + P->L, K, np, P->alpha);
    if (!C) return -1;
    uint64_t* part = (uint64_t*)malloc((size_t)P->alpha * n * 8);
    uint64_t* compl = (uint64_t*)malloc((size_t)W * n * 8);
    for (uint32_t p = 0; p < np; ++p) {
        uint32_t sp = C->sizePart[p], sc = C->sizeCompl[p], st = C->startIdx[p], en = st + sp;
        memcpy(part, d + (size_t)st * n, (size_t)sp * n * 8);
        for (uint32_t i = 0; i < sp; ++i) intt_row(P, st + i, part + (size_t)i * n);
        or_approx_switch(part, compl, C->partQHatInv + (size_t)p * P->alpha,
                         C->partQHatInvPrec + (size_t)p * P->alpha,
                         C->partSrcMod + (size_t)p * P->alpha,
                         C->partQHatModp + (size_t)p * P->alpha * W, C->partComplMod + (size_t)p * W,
                         sp, sc, n);
        for (uint32_t j = 0; j < sc; ++j)
            ntt_row(P, row_of(P, C->partComplMod[(size_t)p * W + j]), compl + (size_t)j * n);
        uint64_t* dg = digits + (size_t)p * W * n;
        for (uint32_t i = 0; i < W; ++i) {
            const uint64_t* src = (i >= st && i < en) ? d + (size_t)i * n
                                : (i < st) ? compl + (size_t)i * n
                                           : compl + (size_t)(i - sp) * n;
            memcpy(dg + (size_t)i * n, src, (size_t)n * 8);
        }
    }
    free(part); free(compl);
    or_ks_constants_free(C);
    return 0;
}

/*
 * Hybrid keyswitch of d (sizeQl towers, eval) to (ba0, ba1) over Q_l:
 * src/keyswitch_resident.cpp:139-225 -- decompose, inner product with the
 * key rows (row = i for Q towers, i + (L - sizeQl) for P towers, :93-96,181),
 * ApproxModDown of both accumulators (:203-224).  av/bv: [dnum][L+K][n].
 */
/* inner product of the digits with the key and ApproxModDown of both
 * accumulators: the second half of or_keyswitch (:171-224 of the file above) */
static int ks_from_digits(const or_params* P, uint32_t sizeQl, const uint64_t* digits,
                          const uint64_t* av, const uint64_t* bv, uint64_t* ba0, uint64_t* ba1);

int or_keyswitch(const or_params* P, uint32_t sizeQl, const uint64_t* d, const uint64_t* av,
                 const uint64_t* bv, uint64_t* ba0, uint64_t* ba1) {
    const uint32_t n = P->n, L = P->L, K = P->K, W = sizeQl + K;
    const uint32_t np = or_num_parts(P, sizeQl);
    if (sizeQl == 0 || sizeQl > L || np > P->dnum) return -1;
    uint64_t* digits = (uint64_t*)malloc((size_t)np * W * n * 8);
    if (or_ks_digits(P, sizeQl, d, digits)) { free(digits); return -1; }
    int rc = ks_from_digits(P, sizeQl, digits, av, bv, ba0, ba1);
    free(digits);
    return rc;
}

static int ks_from_digits(const or_params* P, uint32_t sizeQl, const uint64_t* digits,
                          const uint64_t* av, const uint64_t* bv, uint64_t* ba0, uint64_t* ba1) {
    const uint32_t n = P->n, L = P->L, K = P->K, W = sizeQl + K;
    const uint32_t np = or_num_parts(P, sizeQl);
    const uint32_t delta = L - sizeQl;
    uint64_t* res0 = (uint64_t*)calloc((size_t)W * n, 8);
    uint64_t* res1 = (uint64_t*)calloc((size_t)W * n, 8);
    const size_t keyStride = (size_t)(L + K) * n;
    for (uint32_t p = 0; p < np; ++p)
        for (uint32_t i = 0; i < W; ++i) {
            uint32_t row = (i >= sizeQl) ? i + delta : i;
            or_ks_macc(res0 + (size_t)i * n, res1 + (size_t)i * n,
                       digits + ((size_t)p * W + i) * n, bv + p * keyStride + (size_t)row * n,
                       av + p * keyStride + (size_t)row * n, P->mod[row], n);
        }
    or_ksc* C = or_ks_constants(P->mod, sizeQl, P->mod + L, K, np, P->alpha);
    uint64_t* pw = (uint64_t*)malloc((size_t)K * n * 8);
    uint64_t* qsw = (uint64_t*)malloc((size_t)sizeQl * n * 8);
    uint64_t* res[2] = {res0, res1};
    uint64_t* out[2] = {ba0, ba1};
    for (int c = 0; c < 2; ++c) {
        memcpy(pw, res[c] + (size_t)sizeQl * n, (size_t)K * n * 8);
        for (uint32_t j = 0; j < K; ++j) intt_row(P, L + j, pw + (size_t)j * n);
        or_approx_switch(pw, qsw, C->pHatInv, C->pHatInvPrec, P->mod + L, C->pHatModq, P->mod,
                         K, sizeQl, n);
        for (uint32_t i = 0; i < sizeQl; ++i) {
            ntt_row(P, i, qsw + (size_t)i * n);
            or_moddown_combine(out[c] + (size_t)i * n, res[c] + (size_t)i * n, qsw + (size_t)i * n,
                               C->pInvModq[i], P->mod[i], n);
        }
    }
    free(pw); free(qsw); free(res0); free(res1);
    or_ks_constants_free(C);
    return 0;
}

/* drop the last tower into the survivors: src/keyswitch_resident.cpp:227-244 */
int or_rescale_raw(const or_params* P, uint64_t* c, uint32_t towers) {
    const uint32_t n = P->n;
    if (towers < 2) return 0;
    const uint32_t last = towers - 1;
    uint64_t* s1 = (uint64_t*)malloc((size_t)last * 8);
    uint64_t* s2 = (uint64_t*)malloc((size_t)last * 8);
    or_rescale_consts(s1, s2, P->mod, last);
    uint64_t* drop = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* scr = (uint64_t*)malloc((size_t)n * 8);
    memcpy(drop, c + (size_t)last * n, (size_t)n * 8);
    intt_row(P, last, drop);
    for (uint32_t t = 0; t < last; ++t) {
        memcpy(scr, drop, (size_t)n * 8);
        or_switch_modulus(scr, n, P->mod[last], P->mod[t]);
        ntt_row(P, t, scr);
        or_rescale_fuse(c + (size_t)t * n, scr, s1[t], s2[t], P->mod[t], n);
    }
    free(s1); free(s2); free(drop); free(scr);
    return 0;
}

/* tensor -> keyswitch(t2) -> combine: tests/test_evalmult_rescale.cpp:69-71 */
int or_evalmult_norescale(const or_params* P, uint32_t sizeQl, const uint64_t* c0a,
                          const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                          const uint64_t* av, const uint64_t* bv, uint64_t* r0,
                          uint64_t* r1) {
    const size_t T = (size_t)sizeQl * P->n;
    uint64_t* t0 = (uint64_t*)malloc(T * 8);
    uint64_t* t1 = (uint64_t*)malloc(T * 8);
    uint64_t* t2 = (uint64_t*)malloc(T * 8);
    uint64_t* b0 = (uint64_t*)malloc(T * 8);
    uint64_t* b1 = (uint64_t*)malloc(T * 8);
    or_tensor(c0a, c1a, c0b, c1b, t0, t1, t2, P->mod, sizeQl, P->n);
    int rc = or_keyswitch(P, sizeQl, t2, av, bv, b0, b1);
    if (rc == 0) or_combine(r0, r1, t0, t1, b0, b1, P->mod, sizeQl, P->n);
    free(t0); free(t1); free(t2); free(b0); free(b1);
    return rc;
}

/* ... -> rescale both components: tests/test_evalmult_rescale.cpp:69-73;
 * out0/out1 receive sizeQl-1 towers. */
int or_evalmult_rescale(const or_params* P, uint32_t sizeQl, const uint64_t* c0a,
                        const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                        const uint64_t* av, const uint64_t* bv, uint64_t* out0,
                        uint64_t* out1) {
    const size_t T = (size_t)sizeQl * P->n;
    uint64_t* r0 = (uint64_t*)malloc(T * 8);
    uint64_t* r1 = (uint64_t*)malloc(T * 8);
    int rc = or_evalmult_norescale(P, sizeQl, c0a, c1a, c0b, c1b, av, bv, r0, r1);
    if (rc == 0) {
        or_rescale_raw(P, r0, sizeQl);
        or_rescale_raw(P, r1, sizeQl);
        memcpy(out0, r0, (size_t)(sizeQl - 1) * P->n * 8);
        memcpy(out1, r1, (size_t)(sizeQl - 1) * P->n * 8);
    }
    free(r0); free(r1);
    return rc;
}

/* INTT -> coefficient permutation -> NTT per tower: src/rotate.cpp:70-84.
 * The last p_towers towers use the P rows of the parameter set. */
int or_automorph_eval(const or_params* P, uint64_t* v, uint32_t towers, uint32_t k,
                      int p_towers) {
    const uint32_t n = P->n;
    uint64_t* tmp = (uint64_t*)malloc((size_t)n * 8);
    uint32_t nq = towers - (uint32_t)p_towers;
    for (uint32_t t = 0; t < towers; ++t) {
        uint32_t row = (t < nq) ? t : P->L + (t - nq);
        uint64_t* col = v + (size_t)t * n;
        intt_row(P, row, col);
        or_automorph_coeff(tmp, col, n, k, P->mod[row]);
        memcpy(col, tmp, (size_t)n * 8);
        ntt_row(P, row, col);
    }
    free(tmp);
    return 0;
}

/*
 * Hoisted rotations of one ciphertext by m automorphisms.  The digits of c1
 * (or_ks_digits) are computed once; rotation j applies sigma_kj to every
 * digit tower (over Q_l P) and to c0, then runs the inner product with
 * rotation key j and ApproxModDown:  out_j = (sigma(c0) + ba0, ba1).
 * The reference has no hoisting (it decomposes sigma_k(c1) for every rotation,
 * src/rotate.cpp:118-132); because ModUp is approximate, hoisted and
 * unhoisted results decrypt to the same message with different noise, and
 * coincide exactly for k = 1.  av[j], bv[j]: [dnum][L+K][n].
 */
int or_rotate_hoisted(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
                      uint32_t m, const uint32_t* ks, const uint64_t* const* av,
                      const uint64_t* const* bv, uint64_t* const* out0, uint64_t* const* out1) {
    const uint32_t n = P->n, K = P->K, W = sizeQl + K;
    const uint32_t np = or_num_parts(P, sizeQl);
    if (sizeQl == 0 || sizeQl > P->L || np > P->dnum) return -1;
    const size_t T = (size_t)sizeQl * n, D = (size_t)np * W * n;
    uint64_t* digits = (uint64_t*)malloc(D * 8);
    if (or_ks_digits(P, sizeQl, c1, digits)) { free(digits); return -1; }
    uint64_t* pd = (uint64_t*)malloc(D * 8);
    uint64_t* a0 = (uint64_t*)malloc(T * 8);
    uint64_t* b0 = (uint64_t*)malloc(T * 8);
    int rc = 0;
    for (uint32_t j = 0; j < m && rc == 0; ++j) {
        memcpy(pd, digits, D * 8);
        for (uint32_t p = 0; p < np; ++p) or_automorph_eval(P, pd + (size_t)p * W * n, W, ks[j], (int)K);
        memcpy(a0, c0, T * 8);
        or_automorph_eval(P, a0, sizeQl, ks[j], 0);
        rc = ks_from_digits(P, sizeQl, pd, av[j], bv[j], b0, out1[j]);
        for (uint32_t t = 0; t < sizeQl && rc == 0; ++t)
            or_add_into(a0 + (size_t)t * n, b0 + (size_t)t * n, P->mod[t], n);
        if (rc == 0) memcpy(out0[j], a0, T * 8);
    }
    free(digits); free(pd); free(a0); free(b0);
    return rc;
}

/*
 * Plaintext-weighted sum of rotations (diagonal method of a linear transform):
 * out = sum_j pt_j * rot_kj(c0, c1).  Terms with a key use the hoisted
 * rotations above; a term with av[j] == NULL must have ks[j] == 1 and is the
 * ciphertext itself.  The weighting is the pointwise multiply-accumulate of
 * kernels/stage_acc.cu:12-28 applied to every term.  pts[j]: sizeQl x n, eval.
 */
int or_linear_transform(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
                        uint32_t m, const uint32_t* ks, const uint64_t* const* av,
                        const uint64_t* const* bv, const uint64_t* const* pts, uint64_t* out0,
                        uint64_t* out1) {
    const uint32_t n = P->n;
    const size_t T = (size_t)sizeQl * n;
    uint32_t* rk = (uint32_t*)malloc((size_t)m * 4);
    const uint64_t** ra = (const uint64_t**)malloc((size_t)m * sizeof(void*));
    const uint64_t** rb = (const uint64_t**)malloc((size_t)m * sizeof(void*));
    uint64_t** r0 = (uint64_t**)malloc((size_t)m * sizeof(void*));
    uint64_t** r1 = (uint64_t**)malloc((size_t)m * sizeof(void*));
    const uint64_t** x0 = (const uint64_t**)malloc((size_t)m * sizeof(void*));
    const uint64_t** x1 = (const uint64_t**)malloc((size_t)m * sizeof(void*));
    uint32_t nr = 0;
    int rc = 0;
    for (uint32_t j = 0; j < m; ++j) {
        if (!av[j]) {
            if (ks[j] != 1) rc = -1;
            x0[j] = c0;
            x1[j] = c1;
            continue;
        }
        rk[nr] = ks[j];
        ra[nr] = av[j];
        rb[nr] = bv[j];
        r0[nr] = (uint64_t*)malloc(T * 8);
        r1[nr] = (uint64_t*)malloc(T * 8);
        x0[j] = r0[nr];
        x1[j] = r1[nr];
        ++nr;
    }
    if (rc == 0 && nr) rc = or_rotate_hoisted(P, sizeQl, c0, c1, nr, rk, ra, rb, r0, r1);
    if (rc == 0)
        for (uint32_t t = 0; t < sizeQl; ++t) {
            const uint64_t q = P->mod[t];
            for (uint32_t k = 0; k < n; ++k) {
                const size_t x = (size_t)t * n + k;
                uint64_t s0 = 0, s1 = 0;
                for (uint32_t j = 0; j < m; ++j) {
                    s0 = addm(s0, or_mulmod(pts[j][x], x0[j][x], q), q);
                    s1 = addm(s1, or_mulmod(pts[j][x], x1[j][x], q), q);
                }
                out0[x] = s0;
                out1[x] = s1;
            }
        }
    for (uint32_t j = 0; j < nr; ++j) { free(r0[j]); free(r1[j]); }
    free(rk); free(ra); free(rb); free(r0); free(r1); free(x0); free(x1);
    return rc;
}

/* (sigma_k(c0) + ba0, ba1) with (ba0,ba1) = KS(sigma_k(c1)): src/rotate.cpp:118-132 */
int or_rotate(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
              uint32_t k, const uint64_t* av, const uint64_t* bv, uint64_t* out0,
              uint64_t* out1) {
    const size_t T = (size_t)sizeQl * P->n;
    uint64_t* a0 = (uint64_t*)malloc(T * 8);
    uint64_t* a1 = (uint64_t*)malloc(T * 8);
    uint64_t* b0 = (uint64_t*)malloc(T * 8);
    memcpy(a0, c0, T * 8);
    memcpy(a1, c1, T * 8);
    or_automorph_eval(P, a0, sizeQl, k, 0);
    or_automorph_eval(P, a1, sizeQl, k, 0);
    int rc = or_keyswitch(P, sizeQl, a1, av, bv, b0, out1);
    if (rc == 0) {
        for (uint32_t t = 0; t < sizeQl; ++t) or_add_into(a0 + (size_t)t * P->n, b0 + (size_t)t * P->n, P->mod[t], P->n);
        memcpy(out0, a0, T * 8);
    }
    free(a0); free(a1); free(b0);
    return rc;
}

/* ------------------------------------------------------------------ */
/* keys and encryption                                                 */
/* ------------------------------------------------------------------ */
static void ntt_fresh(uint64_t* x, uint32_t n, uint64_t q, uint64_t root) {
    uint64_t* fr = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* fp = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* ir = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* ip = (uint64_t*)malloc((size_t)n * 8);
    uint64_t pair[2];
    or_root_tables(n, q, root, fr, fp, ir, ip, pair);
    or_ntt(x, n, q, fr, fp);
    free(fr); free(fp); free(ir); free(ip);
}

/* s ternary (or sparse with Hamming weight h), e Gaussian, a uniform (taken
 * as eval form), b = ns e - a s, all eval: src/keygen.cpp:55-112 with the
 * Gaussian replaced by or_gauss_sample. */
void or_keygen(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
               uint64_t ns, double sigma, uint64_t seed, uint32_t hamming, uint64_t* s,
               uint64_t* pkA, uint64_t* pkB) {
    or_mt64 g;
    or_mt64_seed(&g, seed);
    int64_t* sc = (int64_t*)calloc(n, 8);
    int64_t* ec = (int64_t*)calloc(n, 8);
    if (hamming == 0) {
        for (uint32_t k = 0; k < n; ++k) sc[k] = (int64_t)(or_mt64_next(&g) % 3) - 1;
    } else {
        uint32_t* idx = (uint32_t*)malloc((size_t)n * 4);
        for (uint32_t k = 0; k < n; ++k) idx[k] = k;
        for (uint32_t k = n - 1; k > 0; --k) {
            uint32_t j = (uint32_t)(or_mt64_next(&g) % (k + 1));
            uint32_t t = idx[k]; idx[k] = idx[j]; idx[j] = t;
        }
        uint32_t h = hamming < n ? hamming : n;
        for (uint32_t k = 0; k < h; ++k) sc[idx[k]] = (or_mt64_next(&g) & 1) ? 1 : -1;
        free(idx);
    }
    for (uint32_t k = 0; k < n; ++k) ec[k] = or_gauss_sample(&g, sigma);
    uint64_t* eT = (uint64_t*)malloc((size_t)n * 8);
    for (uint32_t t = 0; t < count; ++t) {
        uint64_t q = mods[t];
        uint64_t* sT = s + (size_t)t * n;
        uint64_t* aT = pkA + (size_t)t * n;
        uint64_t* bT = pkB + (size_t)t * n;
        for (uint32_t k = 0; k < n; ++k) {
            sT[k] = embed(sc[k], q);
            eT[k] = embed(ec[k], q);
            aT[k] = or_mt64_next(&g) % q;
        }
        ntt_fresh(sT, n, q, roots[t]);
        ntt_fresh(eT, n, q, roots[t]);
        for (uint32_t k = 0; k < n; ++k)
            bT[k] = subm(or_mulmod(ns % q, eT[k], q), or_mulmod(aT[k], sT[k], q), q);
    }
    free(eT); free(sc); free(ec);
}

/* Eval key over QP, per part p:  av = pkA u + ns e1,  bv = pkB u + ns e0
 * (+ (P mod q_t) sOld inside the part's Q range): src/evalkeygen.cpp:42-96.
 * av/bv: [numPart][sizeQ+sizeP][n]. */
void or_evalkeygen_sold(uint32_t n, uint32_t sizeQ, uint32_t sizeP, uint32_t alpha,
                        uint32_t numPart, const uint64_t* modQP, const uint64_t* rootQP,
                        const uint64_t* sOld, const uint64_t* pkA, const uint64_t* pkB,
                        uint64_t ns, double sigma, uint64_t seed, uint64_t* av,
                        uint64_t* bv) {
    const uint32_t W = sizeQ + sizeP;
    or_mt64 g;
    or_mt64_seed(&g, seed);
    int64_t* uc = (int64_t*)malloc((size_t)n * 8);
    int64_t* e0c = (int64_t*)malloc((size_t)n * 8);
    int64_t* e1c = (int64_t*)malloc((size_t)n * 8);
    uint64_t* uT = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e0T = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e1T = (uint64_t*)malloc((size_t)n * 8);
    for (uint32_t part = 0; part < numPart; ++part) {
        uint32_t st = alpha * part;
        uint32_t sp = (sizeQ > st + alpha) ? alpha : (sizeQ - st);
        uint32_t en = st + sp;
        for (uint32_t k = 0; k < n; ++k) {
            uc[k] = (int64_t)(or_mt64_next(&g) % 3) - 1;
            e0c[k] = or_gauss_sample(&g, sigma);
            e1c[k] = or_gauss_sample(&g, sigma);
        }
        for (uint32_t t = 0; t < W; ++t) {
            uint64_t q = modQP[t];
            for (uint32_t k = 0; k < n; ++k) {
                uT[k] = embed(uc[k], q); e0T[k] = embed(e0c[k], q); e1T[k] = embed(e1c[k], q);
            }
            ntt_fresh(uT, n, q, rootQP[t]);
            ntt_fresh(e0T, n, q, rootQP[t]);
            ntt_fresh(e1T, n, q, rootQP[t]);
            uint64_t Pq = prod_mod(modQP + sizeQ, sizeP, q, -1);
            size_t o = ((size_t)part * W + t) * n;
            for (uint32_t k = 0; k < n; ++k) {
                uint64_t a = addm(or_mulmod(pkA[(size_t)t * n + k], uT[k], q), or_mulmod(ns % q, e1T[k], q), q);
                uint64_t b = addm(or_mulmod(pkB[(size_t)t * n + k], uT[k], q), or_mulmod(ns % q, e0T[k], q), q);
                if (t >= st && t < en) b = addm(b, or_mulmod(Pq, sOld[(size_t)t * n + k], q), q);
                av[o + k] = a;
                bv[o + k] = b;
            }
        }
    }
    free(uc); free(e0c); free(e1c); free(uT); free(e0T); free(e1T);
}

/* ---- counter sampler (the library's device key generation) ----
 * draw(seed, domain, i) = SplitMix64's output function of
 * seed + (domain 2^32 + i + 1) 0x9E3779B97F4A7C15 (so draw(seed, 0, i - 1) is
 * the i-th output of sequential SplitMix64 from `seed`); ternary = draw mod 3
 * - 1; uniform = draw mod q; Gaussian = #{j < K : (draw >> 1) >= T[j]} of the
 * or_gauss_sample table, negated when draw & 1.  Domains: secret 1, error 2,
 * uniform a of tower t 0x10000 + t; switching key part p: u 0x20000 + 3p,
 * e0 0x20001 + 3p, e1 0x20002 + 3p. */
uint64_t or_ctr_draw(uint64_t seed, uint32_t domain, uint32_t index) {
    uint64_t c = ((uint64_t)domain << 32) + (uint64_t)index + 1;
    uint64_t z = seed + c * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
int64_t or_ctr_gauss(uint64_t seed, uint32_t domain, uint32_t index, double sigma) {
    uint64_t T[GMAX + 1];
    int K;
    gauss_table(sigma, T, &K);
    uint64_t r = or_ctr_draw(seed, domain, index);
    int64_t k = 0;
    for (int j = 0; j < K; ++j) k += (r >> 1) >= T[j];
    return (r & 1) ? -k : k;
}
static void ctr_gauss_poly(uint64_t seed, uint32_t domain, uint32_t n, double sigma, int64_t* out) {
    uint64_t T[GMAX + 1];
    int K;
    gauss_table(sigma, T, &K);
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t r = or_ctr_draw(seed, domain, i);
        int64_t k = 0;
        for (int j = 0; j < K; ++j) k += (r >> 1) >= T[j];
        out[i] = (r & 1) ? -k : k;
    }
}

/* or_keygen with the counter sampler: s ternary (domain 1), e Gaussian
 * (domain 2), a of tower t uniform (domain 0x10000 + t, taken as eval form),
 * b = ns e - a s: src/keygen.cpp:55-112. */
void or_keygen_ctr(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
                   uint64_t ns, double sigma, uint64_t seed, uint64_t* s, uint64_t* pkA,
                   uint64_t* pkB) {
    int64_t* sc = (int64_t*)malloc((size_t)n * 8);
    int64_t* ec = (int64_t*)malloc((size_t)n * 8);
    uint64_t* eT = (uint64_t*)malloc((size_t)n * 8);
    for (uint32_t k = 0; k < n; ++k) sc[k] = (int64_t)(or_ctr_draw(seed, 1, k) % 3) - 1;
    ctr_gauss_poly(seed, 2, n, sigma, ec);
    for (uint32_t t = 0; t < count; ++t) {
        uint64_t q = mods[t];
        uint64_t* sT = s + (size_t)t * n;
        uint64_t* aT = pkA + (size_t)t * n;
        uint64_t* bT = pkB + (size_t)t * n;
        for (uint32_t k = 0; k < n; ++k) {
            sT[k] = embed(sc[k], q);
            eT[k] = embed(ec[k], q);
            aT[k] = or_ctr_draw(seed, 0x10000 + t, k) % q;
        }
        ntt_fresh(sT, n, q, roots[t]);
        ntt_fresh(eT, n, q, roots[t]);
        for (uint32_t k = 0; k < n; ++k)
            bT[k] = subm(or_mulmod(ns % q, eT[k], q), or_mulmod(aT[k], sT[k], q), q);
    }
    free(eT); free(sc); free(ec);
}

/* or_evalkeygen_sold with the counter sampler: part p draws u (domain
 * 0x20000 + 3p, ternary), e0 (+1) and e1 (+2, Gaussian):
 * src/evalkeygen.cpp:42-96. */
void or_evalkeygen_ctr(uint32_t n, uint32_t sizeQ, uint32_t sizeP, uint32_t alpha,
                       uint32_t numPart, const uint64_t* modQP, const uint64_t* rootQP,
                       const uint64_t* sOld, const uint64_t* pkA, const uint64_t* pkB,
                       uint64_t ns, double sigma, uint64_t seed, uint64_t* av,
                       uint64_t* bv) {
    const uint32_t W = sizeQ + sizeP;
    int64_t* uc = (int64_t*)malloc((size_t)n * 8);
    int64_t* e0c = (int64_t*)malloc((size_t)n * 8);
    int64_t* e1c = (int64_t*)malloc((size_t)n * 8);
    uint64_t* uT = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e0T = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e1T = (uint64_t*)malloc((size_t)n * 8);
    for (uint32_t part = 0; part < numPart; ++part) {
        uint32_t st = alpha * part;
        uint32_t sp = (sizeQ > st + alpha) ? alpha : (sizeQ - st);
        uint32_t en = st + sp;
        for (uint32_t k = 0; k < n; ++k) uc[k] = (int64_t)(or_ctr_draw(seed, 0x20000 + 3 * part, k) % 3) - 1;
        ctr_gauss_poly(seed, 0x20001 + 3 * part, n, sigma, e0c);
        ctr_gauss_poly(seed, 0x20002 + 3 * part, n, sigma, e1c);
        for (uint32_t t = 0; t < W; ++t) {
            uint64_t q = modQP[t];
            for (uint32_t k = 0; k < n; ++k) {
                uT[k] = embed(uc[k], q); e0T[k] = embed(e0c[k], q); e1T[k] = embed(e1c[k], q);
            }
            ntt_fresh(uT, n, q, rootQP[t]);
            ntt_fresh(e0T, n, q, rootQP[t]);
            ntt_fresh(e1T, n, q, rootQP[t]);
            uint64_t Pq = prod_mod(modQP + sizeQ, sizeP, q, -1);
            size_t o = ((size_t)part * W + t) * n;
            for (uint32_t k = 0; k < n; ++k) {
                uint64_t a = addm(or_mulmod(pkA[(size_t)t * n + k], uT[k], q), or_mulmod(ns % q, e1T[k], q), q);
                uint64_t b = addm(or_mulmod(pkB[(size_t)t * n + k], uT[k], q), or_mulmod(ns % q, e0T[k], q), q);
                if (t >= st && t < en) b = addm(b, or_mulmod(Pq, sOld[(size_t)t * n + k], q), q);
                av[o + k] = a;
                bv[o + k] = b;
            }
        }
    }
    free(uc); free(e0c); free(e1c); free(uT); free(e0T); free(e1T);
}

/* c0 = pkB v + ns e0 + m, c1 = pkA v + ns e1 (eval): src/encdec.cpp:41-64 */
void or_encrypt(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
                const int64_t* m, const uint64_t* pkA, const uint64_t* pkB, uint64_t ns,
                double sigma, uint64_t seed, uint64_t* c0, uint64_t* c1) {
    or_mt64 g;
    or_mt64_seed(&g, seed);
    int64_t* vc = (int64_t*)malloc((size_t)n * 8);
    int64_t* e0c = (int64_t*)malloc((size_t)n * 8);
    int64_t* e1c = (int64_t*)malloc((size_t)n * 8);
    for (uint32_t k = 0; k < n; ++k) {
        vc[k] = (int64_t)(or_mt64_next(&g) % 3) - 1;
        e0c[k] = or_gauss_sample(&g, sigma);
        e1c[k] = or_gauss_sample(&g, sigma);
    }
    uint64_t* vT = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e0T = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* e1T = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* mT = (uint64_t*)malloc((size_t)n * 8);
    for (uint32_t t = 0; t < count; ++t) {
        uint64_t q = mods[t];
        for (uint32_t k = 0; k < n; ++k) {
            vT[k] = embed(vc[k], q); e0T[k] = embed(e0c[k], q);
            e1T[k] = embed(e1c[k], q); mT[k] = embed(m[k], q);
        }
        ntt_fresh(vT, n, q, roots[t]); ntt_fresh(e0T, n, q, roots[t]);
        ntt_fresh(e1T, n, q, roots[t]); ntt_fresh(mT, n, q, roots[t]);
        for (uint32_t k = 0; k < n; ++k) {
            size_t x = (size_t)t * n + k;
            uint64_t b = addm(or_mulmod(pkB[x], vT[k], q), or_mulmod(ns % q, e0T[k], q), q);
            uint64_t a = addm(or_mulmod(pkA[x], vT[k], q), or_mulmod(ns % q, e1T[k], q), q);
            c0[x] = addm(b, mT[k], q);
            c1[x] = a;
        }
    }
    free(vc); free(e0c); free(e1c); free(vT); free(e0T); free(e1T); free(mT);
}

/* tower 0 only: INTT(c0 + c1 s), centred: src/encdec.cpp:68-81 */
void or_decrypt(uint32_t n, uint64_t q0, uint64_t root0, const uint64_t* c0,
                const uint64_t* c1, const uint64_t* s, int64_t* mout) {
    uint64_t* ph = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* fr = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* fp = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* ir = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* ip = (uint64_t*)malloc((size_t)n * 8);
    uint64_t pair[2];
    or_root_tables(n, q0, root0, fr, fp, ir, ip, pair);
    for (uint32_t k = 0; k < n; ++k) ph[k] = addm(c0[k], or_mulmod(c1[k], s[k], q0), q0);
    or_intt(ph, n, q0, ir, ip, pair[0], pair[1]);
    for (uint32_t k = 0; k < n; ++k) {
        uint64_t c = ph[k];
        mout[k] = (c > q0 / 2) ? (int64_t)c - (int64_t)q0 : (int64_t)c;
    }
    free(ph); free(fr); free(fp); free(ir); free(ip);
}
