// Counter-based sampler for key generation on the device.
//
// Every coefficient is drawn on its own from (seed, domain, index), so a
// kernel samples all coefficients of a polynomial at once:
//     c = domain * 2^32 + index + 1,   z = seed + c * G   (mod 2^64),
//     draw = mix(z)
// with G = 0x9E3779B97F4A7C15 and mix the output function of SplitMix64
// (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014): z ^= z >> 30; z *= 0xBF58476D1CE4E5B9; z ^= z >> 27;
// z *= 0x94D049BB133111EB; z ^= z >> 31.  (draw(seed, 0, i - 1) is the i-th
// output of sequential SplitMix64 seeded with `seed`.)
//   ternary:   (int64)(draw mod 3) - 1
//   uniform:   draw mod q
//   Gaussian:  the cumulative table of GaussSampler (thresholds T[0..K]):
//              k = #{j < K : (draw >> 1) >= T[j]}, negated when draw & 1
// Domains: secret 1, error 2, uniform a of tower t 0x10000 + t; switching
// key part p: u 0x20000 + 3p, e0 0x20001 + 3p, e1 0x20002 + 3p.
#pragma once
#include <stdint.h>

#include "modarith.hpp"

namespace gf {

enum : uint32_t {
    CTR_SECRET = 1,
    CTR_ERROR = 2,
    CTR_UNIFORM = 0x10000,   // + tower
    CTR_KEY = 0x20000        // + 3 part (+ 0: u, 1: e0, 2: e1)
};

GF_HD uint64_t ctr_draw(uint64_t seed, uint32_t domain, uint32_t index) {
    const uint64_t c = ((uint64_t)domain << 32) + (uint64_t)index + 1;
    uint64_t z = seed + c * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
GF_HD int64_t ctr_ternary(uint64_t r) { return (int64_t)(r % 3) - 1; }
// T: K thresholds, read through f(j) (a table in any address space)
template <class F>
GF_HD int64_t ctr_gauss(uint64_t r, uint32_t K, F T) {
    const uint64_t m = r >> 1;
    int64_t k = 0;
    for (uint32_t j = 0; j < K; ++j) k += m >= T(j) ? 1 : 0;
    return (r & 1) ? -k : k;
}

}  // namespace gf
