// Host-side noise sampler shared by key generation and encryption.
//
// Discrete Gaussian of standard deviation sigma on [-12 sigma, 12 sigma]:
// weights w_0 = 1, w_k = 2 exp(-k^2 / (2 sigma^2)); thresholds
// T[k] = (uint64)(cumulative_k / total * 2^63); draw r = rng() >> 1, take the
// first k with r < T[k] (the last index if none), then a sign bit rng() & 1
// when k > 0.  The reference rounds std::normal_distribution draws
// (src/keygen.cpp:61,83), whose values are implementation-defined; this fixed
// table sampler makes keys and ciphertexts reproducible everywhere.
#pragma once
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace gf {

class GaussSampler {
public:
    explicit GaussSampler(double sigma) {
        K_ = (int)std::ceil(sigma * 12.0);
        if (K_ > 256) K_ = 256;
        std::vector<double> w(K_ + 1);
        double total = 0.0, cum = 0.0;
        for (int k = 0; k <= K_; ++k) {
            w[k] = k == 0 ? 1.0 : 2.0 * std::exp(-(double)k * (double)k / (2.0 * sigma * sigma));
            total += w[k];
        }
        T_.resize(K_ + 1);
        for (int k = 0; k <= K_; ++k) {
            cum += w[k];
            T_[k] = (uint64_t)std::ldexp(cum / total, 63);
        }
    }
    int64_t operator()(std::mt19937_64& g) const {
        const uint64_t r = g() >> 1;
        int k = 0;
        while (k < K_ && r >= T_[k]) ++k;
        if (k == 0) return 0;
        return (g() & 1) ? -(int64_t)k : (int64_t)k;
    }

    // the thresholds T[0..K] (the counter sampler uses T[0..K-1])
    const std::vector<uint64_t>& table() const { return T_; }
    int K() const { return K_; }

private:
    int K_ = 0;
    std::vector<uint64_t> T_;
};

// the per-part draws of a switching key (src/evalkeygen.cpp:62-66): for each
// coefficient u (ternary), e0, e1 (Gaussian), parts in order
inline void sample_switch_key(uint32_t n, uint32_t parts, double sigma, uint64_t seed, std::vector<int64_t>& u,
                              std::vector<int64_t>& e0, std::vector<int64_t>& e1) {
    std::mt19937_64 g(seed);
    const GaussSampler gauss(sigma);
    u.resize((size_t)parts * n);
    e0.resize((size_t)parts * n);
    e1.resize((size_t)parts * n);
    for (uint32_t p = 0; p < parts; ++p)
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)p * n + k;
            u[x] = (int64_t)(g() % 3) - 1;
            e0[x] = gauss(g);
            e1[x] = gauss(g);
        }
}

// the draws of one encryption (src/encdec.cpp:47-53): for each coefficient
// v (ternary), e0, e1 (Gaussian)
inline void sample_encrypt(uint32_t n, double sigma, uint64_t seed, std::vector<int64_t>& v,
                           std::vector<int64_t>& e0, std::vector<int64_t>& e1) {
    sample_switch_key(n, 1, sigma, seed, v, e0, e1);
}

}  // namespace gf
