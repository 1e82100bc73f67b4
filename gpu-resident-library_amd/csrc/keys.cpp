// Keys, encryption, host-vector rotation and CKKS encoding for the C++ API.
//
// Randomness: std::mt19937_64 seeded by the caller, drawn in the order of the
// reference (src/keygen.cpp:55-112, src/evalkeygen.cpp:42-96,
// src/encdec.cpp:41-64).  The reference rounds std::normal_distribution
// draws, whose values are implementation-defined; the Gaussian here is a
// fixed cumulative-table sampler (GaussSampler, sampler.hpp), so keys and
// ciphertexts are reproducible on every platform.
//
// Transforms run on the device (gf_ntt_towers); the pointwise key algebra is
// host work on 128-bit products, as in the reference (setup only).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <random>
#include <stdexcept>
#include <string>

#include "../../include/gpufhe.hpp"
#include "../../include/gpufhe_capi.h"
#include "sampler.hpp"

namespace gpufhe {

namespace {
using u128 = unsigned __int128;
uint32_t g_hamming = 0;

void ck(int rc, const char* what) {
    if (rc) throw std::runtime_error(std::string("gpufhe: ") + what + ": " + gf_last_error());
}
uint64_t mulm(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
uint64_t addm(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;
    return s >= q ? s - q : s;
}
uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
uint64_t embed(int64_t v, uint64_t q) {
    const int64_t r = v % (int64_t)q;
    return r < 0 ? (uint64_t)(r + (int64_t)q) : (uint64_t)r;
}

using Gauss = gf::GaussSampler;

// in-place transforms of host towers on the device: v is towers x n
void xform(std::vector<uint64_t>& v, uint32_t towers, uint32_t n, const uint64_t* mods, const uint64_t* roots,
           bool inverse) {
    const size_t bytes = (size_t)towers * n * 8;
    if (!bytes) return;
    void* d = nullptr;
    ck(gf_malloc(&d, bytes), "transform buffer");
    int rc = gf_h2d(d, v.data(), bytes);
    if (!rc) rc = gf_ntt_towers(n, (uint64_t*)d, towers, mods, roots, inverse ? 1 : 0, nullptr);
    if (!rc) rc = gf_device_sync();
    if (!rc) rc = gf_d2h(v.data(), d, bytes);
    (void)gf_free(d);
    ck(rc, "tower transform");
}
// the same small signed polynomial embedded in every tower, then transformed
std::vector<uint64_t> embed_eval(const std::vector<int64_t>& c, uint32_t towers, uint32_t n, const uint64_t* mods,
                                 const uint64_t* roots) {
    std::vector<uint64_t> v((size_t)towers * n);
    for (uint32_t t = 0; t < towers; ++t)
        for (uint32_t k = 0; k < n; ++k) v[(size_t)t * n + k] = embed(c[k], mods[t]);
    xform(v, towers, n, mods, roots, false);
    return v;
}
}  // namespace

void set_secret_hamming_weight(uint32_t h) { g_hamming = h; }

// src/keygen.cpp:55-112: s ternary (or exactly h nonzero +-1 at positions of a
// Fisher-Yates shuffle), e Gaussian, a uniform per tower taken as evaluation
// form, b = ns e - a s.
KeyPairHost keygen_host(uint32_t n, const std::vector<uint64_t>& moduli, const std::vector<uint64_t>& roots,
                        uint64_t ns, double sigma, uint64_t seed) {
    const uint32_t L = (uint32_t)moduli.size();
    if (roots.size() < L) throw std::runtime_error("keygen_host: one root per modulus needed");
    std::mt19937_64 g(seed);
    const Gauss gauss(sigma);
    std::vector<int64_t> sc(n, 0), ec(n, 0);
    if (g_hamming == 0) {
        for (uint32_t k = 0; k < n; ++k) sc[k] = (int64_t)(g() % 3) - 1;
    } else {
        std::vector<uint32_t> idx(n);
        for (uint32_t k = 0; k < n; ++k) idx[k] = k;
        for (uint32_t k = n - 1; k > 0; --k) std::swap(idx[k], idx[(uint32_t)(g() % (k + 1))]);
        const uint32_t h = g_hamming < n ? g_hamming : n;
        for (uint32_t k = 0; k < h; ++k) sc[idx[k]] = (g() & 1) ? 1 : -1;
    }
    for (uint32_t k = 0; k < n; ++k) ec[k] = gauss(g);
    KeyPairHost kp;
    kp.pkA.resize((size_t)L * n);
    for (uint32_t t = 0; t < L; ++t)
        for (uint32_t k = 0; k < n; ++k) kp.pkA[(size_t)t * n + k] = g() % moduli[t];
    kp.s = embed_eval(sc, L, n, moduli.data(), roots.data());
    const std::vector<uint64_t> e = embed_eval(ec, L, n, moduli.data(), roots.data());
    kp.pkB.resize((size_t)L * n);
    for (uint32_t t = 0; t < L; ++t) {
        const uint64_t q = moduli[t], nsq = ns % q;
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)t * n + k;
            kp.pkB[x] = subm(mulm(nsq, e[x], q), mulm(kp.pkA[x], kp.s[x], q), q);
        }
    }
    return kp;
}

// src/evalkeygen.cpp:42-96: per part, fresh u (ternary), e0, e1 (Gaussian);
// av = pkA u + ns e1, bv = pkB u + ns e0, plus (P mod q_t) sOld on the part's
// own Q towers.  Writes K.av/K.bv [numPart][|QP| x n] and evalKeyTowers = |QP|.
void evalkeygen_host_sold(KeySwitchConstants& K, const std::vector<uint64_t>& sQP, const std::vector<uint64_t>& sOld,
                          const std::vector<uint64_t>& pkA_QP, const std::vector<uint64_t>& pkB_QP,
                          const std::vector<uint64_t>& PModq_QP, const std::vector<uint64_t>& modQP,
                          const std::vector<uint64_t>& rootQP, uint64_t ns, double sigma, uint64_t seed) {
    (void)sQP;
    const uint32_t n = K.n, W = (uint32_t)modQP.size(), sizeQl = K.sizeQl, alpha = K.alpha, parts = K.numPart;
    const size_t T = (size_t)W * n;
    if (!n || !alpha || rootQP.size() < W || PModq_QP.size() < W || sOld.size() < (size_t)sizeQl * n ||
        pkA_QP.size() < T || pkB_QP.size() < T)
        throw std::runtime_error("evalkeygen_host_sold: inputs shorter than the QP basis");
    K.evalKeyTowers = W;
    K.av.assign(parts, {});
    K.bv.assign(parts, {});
    std::mt19937_64 g(seed);
    const Gauss gauss(sigma);
    std::vector<int64_t> uc(n), e0c(n), e1c(n);
    for (uint32_t p = 0; p < parts; ++p) {
        const uint32_t st = alpha * p, en = st + (sizeQl > st + alpha ? alpha : sizeQl - st);
        for (uint32_t k = 0; k < n; ++k) {
            uc[k] = (int64_t)(g() % 3) - 1;
            e0c[k] = gauss(g);
            e1c[k] = gauss(g);
        }
        const std::vector<uint64_t> u = embed_eval(uc, W, n, modQP.data(), rootQP.data());
        const std::vector<uint64_t> e0 = embed_eval(e0c, W, n, modQP.data(), rootQP.data());
        const std::vector<uint64_t> e1 = embed_eval(e1c, W, n, modQP.data(), rootQP.data());
        std::vector<uint64_t>& av = K.av[p];
        std::vector<uint64_t>& bv = K.bv[p];
        av.resize(T);
        bv.resize(T);
        for (uint32_t t = 0; t < W; ++t) {
            const uint64_t q = modQP[t], nsq = ns % q, Pq = PModq_QP[t] % q;
            const bool own = t >= st && t < en;
            for (uint32_t k = 0; k < n; ++k) {
                const size_t x = (size_t)t * n + k;
                av[x] = addm(mulm(pkA_QP[x], u[x], q), mulm(nsq, e1[x], q), q);
                uint64_t b = addm(mulm(pkB_QP[x], u[x], q), mulm(nsq, e0[x], q), q);
                if (own) b = addm(b, mulm(Pq, sOld[x], q), q);
                bv[x] = b;
            }
        }
    }
}

// src/evalkeygen.cpp:99-113: relinearisation key, sOld = s^2
void evalkeygen_host(KeySwitchConstants& K, const std::vector<uint64_t>& sQP, const std::vector<uint64_t>& pkA_QP,
                     const std::vector<uint64_t>& pkB_QP, const std::vector<uint64_t>& PModq_QP,
                     const std::vector<uint64_t>& modQP, const std::vector<uint64_t>& rootQP, uint64_t ns, double sigma,
                     uint64_t seed) {
    const uint32_t n = K.n, W = (uint32_t)modQP.size();
    if (sQP.size() < (size_t)W * n) throw std::runtime_error("evalkeygen_host: secret shorter than the QP basis");
    std::vector<uint64_t> s2((size_t)W * n);
    for (uint32_t t = 0; t < W; ++t)
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)t * n + k;
            s2[x] = mulm(sQP[x], sQP[x], modQP[t]);
        }
    evalkeygen_host_sold(K, sQP, s2, pkA_QP, pkB_QP, PModq_QP, modQP, rootQP, ns, sigma, seed);
}

// src/encdec.cpp:41-64: c0 = pkB v + ns e0 + m, c1 = pkA v + ns e1 (eval form)
void encrypt_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, const std::vector<int64_t>& m,
                  const std::vector<uint64_t>& pkA, const std::vector<uint64_t>& pkB, uint32_t n,
                  const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root, uint64_t ns, double sigma,
                  uint64_t seed) {
    const uint32_t L = (uint32_t)mod.size();
    const size_t T = (size_t)L * n;
    if (m.size() < n || pkA.size() < T || pkB.size() < T || root.size() < L)
        throw std::runtime_error("encrypt_host: inputs shorter than towers x n");
    std::vector<int64_t> vc, e0c, e1c;
    gf::sample_encrypt(n, sigma, seed, vc, e0c, e1c);
    const std::vector<uint64_t> v = embed_eval(vc, L, n, mod.data(), root.data());
    const std::vector<uint64_t> e0 = embed_eval(e0c, L, n, mod.data(), root.data());
    const std::vector<uint64_t> e1 = embed_eval(e1c, L, n, mod.data(), root.data());
    const std::vector<uint64_t> mt = embed_eval(std::vector<int64_t>(m.begin(), m.begin() + n), L, n, mod.data(),
                                                root.data());
    c0.resize(T);
    c1.resize(T);
    for (uint32_t t = 0; t < L; ++t) {
        const uint64_t q = mod[t], nsq = ns % q;
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)t * n + k;
            c0[x] = addm(addm(mulm(pkB[x], v[x], q), mulm(nsq, e0[x], q), q), mt[x], q);
            c1[x] = addm(mulm(pkA[x], v[x], q), mulm(nsq, e1[x], q), q);
        }
    }
}

// src/encdec.cpp:68-81: tower 0 only, INTT(c0 + c1 s), centred
void decrypt_host(std::vector<int64_t>& mout, const std::vector<uint64_t>& c0, const std::vector<uint64_t>& c1,
                  const std::vector<uint64_t>& s, uint32_t n, const std::vector<uint64_t>& mod,
                  const std::vector<uint64_t>& root) {
    if (mod.empty() || root.empty() || c0.size() < n || c1.size() < n || s.size() < n)
        throw std::runtime_error("decrypt_host: inputs shorter than one tower");
    const uint64_t q = mod[0];
    std::vector<uint64_t> ph(n);
    for (uint32_t k = 0; k < n; ++k) ph[k] = addm(c0[k], mulm(c1[k], s[k], q), q);
    xform(ph, 1, n, mod.data(), root.data(), true);
    mout.resize(n);
    for (uint32_t k = 0; k < n; ++k) mout[k] = ph[k] > q / 2 ? (int64_t)ph[k] - (int64_t)q : (int64_t)ph[k];
}

// ------------------------------------------------------------ rotation --
// src/rotate.cpp:70-84 in place, with an n-word scratch tower: the
// automorphism is one permutation of the evaluation points per tower
void automorphism_eval_device(uint64_t* d_v, uint32_t towers, uint32_t n, uint32_t k,
                              const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root,
                              uint64_t* d_scratch, hipStream_t s) {
    (void)mod;
    (void)root;
    for (uint32_t t = 0; t < towers; ++t) {
        uint64_t* col = d_v + (size_t)t * n;
        ck(gf_automorph(n, col, d_scratch, 1, k, s), "automorphism");
        if (hipMemcpyAsync(col, d_scratch, (size_t)n * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
            throw std::runtime_error("gpufhe: automorphism copy failed");
    }
}

// src/rotate.cpp:118-132: (sigma_k(c0) + ba0, ba1), ba = KS(sigma_k(c1)), in
// place on the device pointers; one batched plan (permutation, keyswitch,
// add) instead of per-tower launches, so scratch/b0/b1 are not needed
void rotate_ct_device(uint64_t* d_c0, uint64_t* d_c1, uint32_t k, const DeviceKSContext& C, DeviceKSWork& W,
                      uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod,
                      const std::vector<uint64_t>& root, uint64_t* d_scratch, uint64_t* d_b0, uint64_t* d_b1,
                      hipStream_t s) {
    (void)mod;
    (void)root;
    (void)d_scratch;
    (void)d_b0;
    (void)d_b1;
    if (n != C.n || towers == 0 || towers > C.sizeQl) throw std::runtime_error("rotate_ct_device: level outside the context");
    if (!W.d_ws) throw std::runtime_error("rotate_ct_device: DeviceKSWork not created");
    ck(gf_rotate(C.engine(), C.key(), towers, k, d_c0, d_c1, d_c0, d_c1, W.d_ws, s), "rotate_ct_device");
}

// src/rotate.cpp:140-172: host vectors in and out
void rotate_ct_resident(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, uint32_t k,
                        const KeySwitchConstants& Krot, uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod,
                        const std::vector<uint64_t>& root) {
    const size_t T = (size_t)towers * n;
    if (c0.size() < T || c1.size() < T) throw std::runtime_error("rotate_ct_resident: ciphertext shorter than towers x n");
    DeviceKSContext C = ks_context_create(Krot);
    DeviceKSWork W = ks_work_create(C);
    DeviceCiphertext ct(Scheme::CKKS, n, towers);
    try {
        ct.upload(std::vector<uint64_t>(c0.begin(), c0.begin() + T), std::vector<uint64_t>(c1.begin(), c1.begin() + T));
        rotate_ct_device(ct.c0(), ct.c1(), k, C, W, towers, n, mod, root, nullptr, nullptr, nullptr, nullptr);
        ck(gf_device_sync(), "sync");
        ct.to_host(c0, c1);
    } catch (...) {
        ks_work_destroy(W);
        throw;
    }
    ks_work_destroy(W);
}
// src/rotate.cpp:176-188 (the reference's host-orchestrated variant; same result)
void rotate_ct_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, uint32_t k, const KeySwitchConstants& Krot,
                    uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root) {
    rotate_ct_resident(c0, c1, k, Krot, towers, n, mod, root);
}
// src/rotate.cpp:87-101
void automorphism_eval_host(std::vector<uint64_t>& v, uint32_t towers, uint32_t n, uint32_t k,
                            const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root) {
    (void)mod;
    (void)root;
    const size_t bytes = (size_t)towers * n * 8;
    if (v.size() * 8 < bytes) throw std::runtime_error("automorphism_eval_host: input shorter than towers x n");
    void *a = nullptr, *b = nullptr;
    ck(gf_malloc(&a, bytes), "buffer");
    int rc = gf_malloc(&b, bytes);
    if (!rc) rc = gf_h2d(a, v.data(), bytes);
    if (!rc) rc = gf_automorph(n, (const uint64_t*)a, (uint64_t*)b, towers, k, nullptr);
    if (!rc) rc = gf_device_sync();
    if (!rc) rc = gf_d2h(v.data(), b, bytes);
    (void)gf_free(a);
    (void)gf_free(b);
    ck(rc, "automorphism_eval_host");
}
// src/rotate.cpp:104-113
void pt_to_eval_host(std::vector<uint64_t>& out, const std::vector<int64_t>& m, uint32_t towers, uint32_t n,
                     const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root) {
    if (m.size() < n || mod.size() < towers || root.size() < towers)
        throw std::runtime_error("pt_to_eval_host: inputs shorter than towers x n");
    out = embed_eval(std::vector<int64_t>(m.begin(), m.begin() + n), towers, n, mod.data(), root.data());
}
// src/ptops.cpp:13-19
void ct_add_ct_host(std::vector<uint64_t>& a0, std::vector<uint64_t>& a1, const std::vector<uint64_t>& b0,
                    const std::vector<uint64_t>& b1, uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod) {
    for (uint32_t t = 0; t < towers; ++t)
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)t * n + k;
            a0[x] = addm(a0[x], b0[x], mod[t]);
            a1[x] = addm(a1[x], b1[x], mod[t]);
        }
}
// src/ptops.cpp:20-26
void ct_mul_pt_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, const std::vector<uint64_t>& ptEval,
                    uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod) {
    const size_t T = (size_t)towers * n;
    if (c0.size() < T || c1.size() < T || ptEval.size() < T || mod.size() < towers)
        throw std::runtime_error("ct_mul_pt_host: inputs shorter than towers x n");
    for (uint32_t t = 0; t < towers; ++t)
        for (uint32_t k = 0; k < n; ++k) {
            const size_t x = (size_t)t * n + k;
            c0[x] = mulm(c0[x], ptEval[x], mod[t]);
            c1[x] = mulm(c1[x], ptEval[x], mod[t]);
        }
}

// ------------------------------------------------------------ encoding --
// Slots z_k sit at the evaluation points zeta^(r_k), zeta = exp(i pi / N),
// r_k = 5^k mod 2N (src/encode.cpp:1-5).  Both directions are one 2N-point
// complex FFT: encode puts z_k at index r_k of a length-2N array A and reads
// m_j = round((2 Delta / N) Re(sum_t A[t] zeta^(-j t))) for j < N; decode
// sums m_j zeta^(j t) and reads index r_k.  O(N log N) against the
// reference's O(N^2) sums; results agree to double rounding.
namespace {
using cd = std::complex<double>;
// in-place iterative radix-2 DFT, X[f] = sum_t x[t] exp(sign 2 pi i f t / M)
void fft(std::vector<cd>& a, int sign) {
    const size_t M = a.size();
    for (size_t i = 1, j = 0; i < M; ++i) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    std::vector<cd> w(M / 2);
    for (size_t k = 0; k < M / 2; ++k) w[k] = std::polar(1.0, sign * 2.0 * M_PI * (double)k / (double)M);
    for (size_t len = 2; len <= M; len <<= 1) {
        const size_t step = M / len;
        for (size_t i = 0; i < M; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const cd u = a[i + k], v = a[i + k + len / 2] * w[k * step];
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}
std::vector<uint64_t> orbit(uint32_t N) {
    const uint64_t M = 2ull * N;
    std::vector<uint64_t> r(N / 2);
    uint64_t x = 1;
    for (uint32_t k = 0; k < N / 2; ++k, x = x * 5 % M) r[k] = x;
    return r;
}
}  // namespace

void encode_host(std::vector<int64_t>& m, const std::vector<std::complex<double>>& z, uint32_t N, double Delta) {
    if (N < 2 || (N & (N - 1)) || z.size() < N / 2) throw std::runtime_error("encode_host: N must be a power of two, N/2 slots");
    const std::vector<uint64_t> r = orbit(N);
    std::vector<cd> A(2 * (size_t)N, cd(0, 0));
    for (uint32_t k = 0; k < N / 2; ++k) A[r[k]] = z[k];
    fft(A, -1);
    m.assign(N, 0);
    const double sc = 2.0 * Delta / (double)N;
    for (uint32_t j = 0; j < N; ++j) m[j] = (int64_t)std::llround(sc * A[j].real());
}

void decode_host(std::vector<std::complex<double>>& z, const std::vector<int64_t>& m, uint32_t N, double Delta) {
    if (N < 2 || (N & (N - 1)) || m.size() < N) throw std::runtime_error("decode_host: N must be a power of two, N coefficients");
    const std::vector<uint64_t> r = orbit(N);
    std::vector<cd> B(2 * (size_t)N, cd(0, 0));
    for (uint32_t j = 0; j < N; ++j) B[j] = cd((double)m[j], 0.0);
    fft(B, +1);
    z.assign(N / 2, cd(0, 0));
    for (uint32_t k = 0; k < N / 2; ++k) z[k] = B[r[k]] / Delta;
}

}  // namespace gpufhe
