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

GF_HD uint32_t umulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

// The value itself, made opaque to the optimiser on the device: a 64-bit
// addend assembled from two 32-bit halves then feeds v_mad_u64_u32 as one
// register pair instead of being split back into a separate 64-bit add.
// Emits no instruction.
GF_HD uint64_t opaque64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(v));
#endif
    return v;
}

// The same for a 32-bit value.  Inside a loop it keeps a loop-invariant
// 32-bit operand from being widened to 64 bits once, before the loop, and
// held in a register pair.
GF_HD uint32_t opaque32(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(v));
#endif
    return v;
}

// a + b of two 32-bit words as a 33-bit value (carry from the add itself)
GF_HD uint64_t add33(uint32_t a, uint32_t b) {
    uint32_t s;
    const uint32_t cy = __builtin_add_overflow(a, b, &s) ? 1u : 0u;
    return opaque64(((uint64_t)cy << 32) | s);
}

// full 64x64 -> 128 product from four 32x32 pieces
GF_HD U128 mul128(uint64_t a, uint64_t b) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;
    uint64_t t = mad32(a0, b1, p00 >> 32);
    uint64_t u = mad32(a1, b0, (uint32_t)t);
    U128 r;
    r.hi = mad32(a1, b1, add33((uint32_t)(t >> 32), (uint32_t)(u >> 32)));
    r.lo = (u << 32) | (uint32_t)p00;
    return r;
}

// high 64 bits of a*b
GF_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t t = mad32(a0, b1, umulhi32(a0, b0));
    uint64_t u = mad32(a1, b0, (uint32_t)t);
    return mad32(a1, b1, add33((uint32_t)(t >> 32), (uint32_t)(u >> 32)));
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
// x in [0, 2m) -> [0, m); the select keys on the subtraction's own borrow
GF_HD uint64_t csub(uint64_t x, uint64_t m) {
    uint64_t d;
    return __builtin_sub_overflow(x, m, &d) ? x : d;
}

// Shoup product for the transform butterflies:
//   quotient  qh = x1 wp1 + floor(x0 wp1 / 2^32) + floor(x1 wp0 / 2^32)
//             (x = x1 2^32 + x0, wp likewise), which is floor(x wp / 2^64)
//             minus 0, 1 or 2: the terms left out (x0 wp0 / 2^64 and the two
//             fractional parts) sum to less than 3;
//   remainder x w - qh q  computed mod 2^64 as lo64(x w) + lo64(qh nq),
//             nq = 2^64 - q, in one multiply-add chain.
// With wp = floor(w 2^64 / q) the exact quotient leaves [0, 2q) for any
// x < 2^64, so the result is x w mod q + {0, q, 2q, 3q}, in [0, 4q).
GF_HD uint64_t shoup_lazy4(uint64_t x, uint64_t w, uint64_t wp, uint64_t nq) {
    const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
    const uint32_t p0 = (uint32_t)wp, p1 = (uint32_t)(wp >> 32);
    const uint64_t qh = mad32(x1, p1, add33(umulhi32(x0, p1), umulhi32(x1, p0)));
    const uint32_t h0 = (uint32_t)qh, h1 = (uint32_t)(qh >> 32);
    const uint32_t n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
    const uint64_t p = (uint64_t)x0 * w0;
    const uint32_t hi = (uint32_t)(p >> 32) + x0 * w1 + x1 * w0 + h0 * n1 + h1 * n0;
    return mad32(h0, n0, opaque64(((uint64_t)hi << 32) | (uint32_t)p));
}

// x mod q for any 64-bit x, given 2^32 < q < 2^62 and m = floor(2^64 / q)
// (so m < 2^32).  With x = xh 2^32 + xl, the estimate e = floor(xh m / 2^32)
// satisfies x/q - 3 < e <= x/q: xh m / 2^32 lies within xh / 2^32 < 1 below
// xh 2^32 / q, which lies within xl / q < 1 below x / q.  So x - e q is in
// [0, 3q) and two conditional subtractions finish.
GF_HD uint64_t reduce_red(uint64_t x, uint64_t q, uint32_t m) {
    const uint32_t e = umulhi32((uint32_t)(x >> 32), m);
    const uint64_t r = x - (uint64_t)e * q;
    return csub(csub(r, 2 * q), q);
}

// Exact sum of NS <= 8 products a_s b_s of values below 2^60, as 128 bits.
// Each operand is split at bit 30 (a = a1 2^30 + a0, a0 < 2^30, a1 < 2^30),
// and the partial products go to three 64-bit columns with one
// v_mad_u64_u32 each:
//   lo = sum a0 b0 < 8 2^60,  mid = sum a0 b1 + a1 b0 < 8 2^61,  hi = sum a1 b1 < 8 2^60,
// none of which can overflow; the total is lo + mid 2^30 + hi 2^60.  The
// sources come pre-split (alo, ahi); b(s) returns the s-th constant.
enum : uint32_t { SPLIT30_MASK = (1u << 30) - 1 };
GF_HD uint32_t split30_lo(uint64_t v) { return (uint32_t)v & SPLIT30_MASK; }
GF_HD uint32_t split30_hi(uint64_t v) { return (uint32_t)(v >> 30); }
template <uint32_t NS, class B>
GF_HD U128 dot30(const uint32_t* alo, const uint32_t* ahi, B b) {
    static_assert(NS >= 1 && NS <= 8, "column sums are exact for at most 8 terms");
    uint64_t lo = 0, mid = 0, hi = 0;
#pragma unroll
    for (uint32_t s = 0; s < NS; ++s) {
        const uint64_t v = b(s);
        const uint32_t b0 = split30_lo(v), b1 = split30_hi(v);
        lo = mad32(alo[s], b0, lo);
        mid = mad32(alo[s], b1, mid);
        mid = mad32(ahi[s], b0, mid);
        hi = mad32(ahi[s], b1, hi);
    }
    U128 T;
    T.lo = lo + (mid << 30);
    uint64_t c = T.lo < lo ? 1u : 0u;
    const uint64_t h = hi << 60;
    T.lo += h;
    c += T.lo < h ? 1u : 0u;
    T.hi = (mid >> 34) + (hi >> 4) + c;
    return T;
}

// The column sums of dot30 themselves (lo + mid 2^30 + hi 2^60), for a
// reduction that takes them as they are (redc60_lazy)
struct Cols30 {
    uint64_t lo, mid, hi;
};
template <uint32_t NS, class B>
GF_HD Cols30 dot30_cols(const uint32_t* alo, const uint32_t* ahi, B b) {
    static_assert(NS >= 1 && NS <= 8, "column sums are exact for at most 8 terms");
    Cols30 c{0, 0, 0};
#pragma unroll
    for (uint32_t s = 0; s < NS; ++s) {
        const uint64_t v = b(s);
        const uint32_t b0 = split30_lo(v), b1 = split30_hi(v);
        c.lo = mad32(alo[s], b0, c.lo);
        c.mid = mad32(alo[s], b1, c.mid);
        c.mid = mad32(ahi[s], b0, c.mid);
        c.hi = mad32(ahi[s], b1, c.hi);
    }
    return c;
}
// Montgomery reduction with R = 2^60 straight from the column sums
// T = lo + mid 2^30 + hi 2^60 of dot30_cols over at most 6 terms (two
// values below 2^60 each): two steps of one 30-bit digit each,
// m = (low 30 bits) (-q^-1 mod 2^30) mod 2^30, adding m q split at bit 30
// into the two columns it touches and carrying the (now zero) low 30 bits
// away.  Returns (T + M q) / 2^60 for some M < 2^60: congruent to T 2^-60,
// below T / 2^60 + q.  Needs q < 2^60 (qneg = -q^-1 mod 2^64).  With NS
// terms the columns stay below (NS + 1) 2^60 (lo) and (2 NS + 2) 2^60 + 2^35
// (mid after both steps), so below 2^64 for NS <= 6.
GF_HD uint64_t redc60_lazy(Cols30 c, uint64_t q, uint64_t qneg) {
    const uint32_t qi = (uint32_t)qneg & SPLIT30_MASK, q0 = (uint32_t)q & SPLIT30_MASK, q1 = (uint32_t)(q >> 30);
    const uint32_t m1 = ((uint32_t)c.lo * qi) & SPLIT30_MASK;
    const uint64_t lo = mad32(m1, q0, c.lo);
    const uint64_t mid = mad32(m1, q1, c.mid) + (lo >> 30);
    const uint32_t m2 = ((uint32_t)mid * qi) & SPLIT30_MASK;
    const uint64_t mid2 = mad32(m2, q0, mid);
    return mad32(m2, q1, c.hi) + (mid2 >> 30);
}

// Montgomery reduction, R = 2^64, qneg = -q^{-1} mod 2^64, without the final
// correction: T R^{-1} mod q plus a multiple of q, below T / 2^64 + q + 1
// (any T with T / 2^64 + q + 1 <= 2^64).
GF_HD uint64_t redc_lazy(U128 T, uint64_t q, uint64_t qneg) {
    uint64_t m = T.lo * qneg;
    return T.hi + mulhi64(m, q) + (T.lo != 0 ? 1u : 0u);
}
// For T < q * 2^64 returns T * R^{-1} mod q, in [0, q).
GF_HD uint64_t redc(U128 T, uint64_t q, uint64_t qneg) {
    const uint64_t t = redc_lazy(T, q, qneg);
    return t >= q ? t - q : t;
}
// T * R^{-1} mod q in [0, q) for any T with T / 2^64 + q + 1 <= 2^64, given
// 2^32 < q < 2^62 and m = floor(2^64 / q) (sums of products of unreduced values)
GF_HD uint64_t redc_red(U128 T, uint64_t q, uint64_t qneg, uint32_t m) {
    return reduce_red(redc_lazy(T, q, qneg), q, m);
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
