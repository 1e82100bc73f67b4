// Native parameter generation (OpenFHE-free): NTT-friendly primes, 2n-th
// roots, and the rescale constant pair.  Same rules as the reference's
// src/params.cpp:29-75 so that parameter sets agree bit for bit.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "modarith.hpp"

namespace gpufhe {
using gf::host_mulmod;
using gf::host_powmod;

static bool is_prime_u64(uint64_t x) {
    static const uint64_t B[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (x < 2) return false;
    for (uint64_t b : B)
        if (x % b == 0) return x == b;
    uint64_t d = x - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (uint64_t b : B) {   // deterministic for 64-bit x with these bases
        uint64_t v = host_powmod(b, d, x);
        if (v == 1 || v == x - 1) continue;
        bool ok = false;
        for (int r = 1; r < s && !ok; ++r) {
            v = host_mulmod(v, v, x);
            ok = (v == x - 1);
        }
        if (!ok) return false;
    }
    return true;
}

// the first `count` primes q = 1 mod 2n below 2^bits, scanning downward from
// floor(2^bits / 2n) 2n + 1 and skipping `avoid` (src/params.cpp:29-39)
void native_primes(std::vector<uint64_t>& out, uint32_t count, uint32_t bits, uint32_t n,
                   const std::vector<uint64_t>& avoid) {
    if (bits < 2 || bits > 62) throw std::runtime_error("native_primes: bits out of range");
    const uint64_t M = 2ull * n;
    for (uint64_t q = ((1ull << bits) / M) * M + 1; out.size() < count; q -= M) {
        if (q < (1ull << (bits - 1))) throw std::runtime_error("prime search exhausted");
        bool skip = false;
        for (uint64_t a : avoid) skip |= (a == q);
        if (!skip && is_prime_u64(q)) out.push_back(q);
    }
}

// g^((q-1)/2n) for the smallest g = 2, 3, ... whose n-th power is -1 (src/params.cpp:42-48)
uint64_t native_root(uint32_t n, uint64_t q) {
    const uint64_t M = 2ull * n;
    for (uint64_t g = 2; g < q; ++g) {
        const uint64_t c = host_powmod(g, (q - 1) / M, q);
        if (host_powmod(c, n, q) == q - 1) return c;
    }
    throw std::runtime_error("no 2n-th root");
}

// s1[i] = q_last^-1 mod q_i; s2[i] = floor(Q_l [Q_l^-1]_{q_last} / q_last) mod q_i,
// Q_l = prod_{t<last} q_t (src/params.cpp:50-75).  Since Q_l [Q_l^-1] = q_last D + 1
// and q_i | Q_l, D = -q_last^-1 (mod q_i): s2 is computed as q_i - s1.
void native_rescale_consts(std::vector<uint64_t>& s1, std::vector<uint64_t>& s2,
                           const std::vector<uint64_t>& mod, uint32_t last) {
    s1.resize(last);
    s2.resize(last);
    const uint64_t ql = mod[last];
    for (uint32_t i = 0; i < last; ++i) {
        const uint64_t qi = mod[i];
        s1[i] = host_powmod(ql % qi, qi - 2, qi);
        s2[i] = s1[i] ? qi - s1[i] : 0;
    }
}

}  // namespace gpufhe
