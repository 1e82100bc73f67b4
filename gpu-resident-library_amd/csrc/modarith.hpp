// Exact 64-bit modular arithmetic for gfx950, built from 32-bit multiplies.
//
// gfx950 has no 64x64-bit integer multiply: every 64-bit product below is
// assembled from v_mad_u64_u32 / v_mul_lo_u32 / v_mul_hi_u32 pieces.  The
// helpers are also compiled for the host (GF_HD) so the CPU test suite can run
// the same index and arithmetic logic against the oracle.
//
// Moduli: q < 2^61 (the library uses <= 60-bit primes), all residues in [0,q).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GF_HD __host__ __device__ __forceinline__
#else
#define GF_HD inline
#endif

namespace gf {

// (a * b + c) with 32-bit a, b and 64-bit c: one v_mad_u64_u32
GF_HD uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

struct U128 {
    uint64_t lo, hi;
};

// full 64x64 -> 128 product from four 32x32 pieces
GF_HD U128 mul128(uint64_t a, uint64_t b) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;
    uint64_t t = mad32(a0, b1, p00 >> 32);
    uint64_t u = mad32(a1, b0, (uint32_t)t);
    U128 r;
    r.hi = mad32(a1, b1, (t >> 32) + (u >> 32));
    r.lo = (u << 32) | (uint32_t)p00;
    return r;
}

// high 64 bits of a*b
GF_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t t = mad32(a0, b1, ((uint64_t)a0 * b0) >> 32);
    uint64_t u = mad32(a1, b0, (uint32_t)t);
    return mad32(a1, b1, (t >> 32) + (u >> 32));
}

GF_HD U128 add128(U128 a, U128 b) {
    U128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
    return r;
}

GF_HD uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}
GF_HD uint64_t submod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

// Shoup: x*w mod q with wp = floor(w 2^64 / q); exact for any x < 2^64.
GF_HD uint64_t shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t r = x * w - mulhi64(x, wp) * q;
    return r >= q ? r - q : r;
}
// Shoup without the final correction: x*w mod q + {0 or q}, in [0, 2q), any x < 2^64
GF_HD uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    return x * w - mulhi64(x, wp) * q;
}
// x in [0, 2m) -> [0, m)
GF_HD uint64_t csub(uint64_t x, uint64_t m) { return x >= m ? x - m : x; }

GF_HD uint32_t umulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
// floor(a b / 2^64) minus 0, 1 or 2: the a0*b0 product and the low halves of
// the cross products are dropped (their sum is < 3 * 2^64 / 2^64)
GF_HD uint64_t mulhi64_approx(uint64_t a, uint64_t b) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    return mad32(a1, b1, (uint64_t)umulhi32(a0, b1) + umulhi32(a1, b0));
}
// Shoup with the approximate quotient: x*w mod q + {0, q, 2q, 3q}, in [0, 4q), any x < 2^64
GF_HD uint64_t shoup_lazy4(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    return x * w - mulhi64_approx(x, wp) * q;
}

// Montgomery reduction, R = 2^64, qneg = -q^{-1} mod 2^64.
// For T < q * 2^64 returns T * R^{-1} mod q, in [0, q).
GF_HD uint64_t redc(U128 T, uint64_t q, uint64_t qneg) {
    uint64_t m = T.lo * qneg;
    uint64_t t = T.hi + mulhi64(m, q) + (T.lo != 0 ? 1u : 0u);
    return t >= q ? t - q : t;
}

// a*b mod q for arbitrary residues: REDC(REDC(a b) * R^2).  r2 = 2^128 mod q.
GF_HD uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q, uint64_t qneg, uint64_t r2) {
    uint64_t t = redc(mul128(a, b), q, qneg);
    return redc(mul128(t, r2), q, qneg);
}

// ---- host-side constant helpers (setup only) ----
inline uint64_t host_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((unsigned __int128)a * b) % q);
}
inline uint64_t host_powmod(uint64_t b, uint64_t e, uint64_t q) {
    unsigned __int128 r = 1, bb = b % q;
    while (e) {
        if (e & 1) r = (r * bb) % q;
        bb = (bb * bb) % q;
        e >>= 1;
    }
    return (uint64_t)r;
}
inline uint64_t host_invmod(uint64_t a, uint64_t q) { return host_powmod(a % q, q - 2, q); }
inline uint64_t host_precon(uint64_t w, uint64_t q) {
    return (uint64_t)(((unsigned __int128)w << 64) / q);
}
inline uint64_t host_qneg(uint64_t q) {  // -q^{-1} mod 2^64 by Newton iteration
    uint64_t x = q;
    for (int i = 0; i < 6; ++i) x *= 2 - q * x;
    return (uint64_t)0 - x;
}
inline uint64_t host_r1(uint64_t q) { return (uint64_t)((((unsigned __int128)1) << 64) % q); }
inline uint64_t host_r2(uint64_t q) {
    uint64_t r = host_r1(q);
    return host_mulmod(r, r, q);
}
// x * 2^64 mod q (Montgomery form of x)
inline uint64_t host_to_mont(uint64_t x, uint64_t q) { return host_mulmod(x, host_r1(q), q); }

}  // namespace gf
