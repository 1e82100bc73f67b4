// C++ API tests, written the way the reference's own tests call its API
// (tests/test_ks_resident.cpp, tests/test_evalmult_rescale.cpp,
// tests/test_rot_resident.cpp, tests/test_native_e2e.cpp,
// tests/test_keyreuse.cpp, tests/test_keygen.cpp, tests/test_encdec.cpp),
// with the oracle (oracle/liboracle.so) as the checker.
//
//   test_api cpu   -- host-only functions (constants, encoding, parameters)
//   test_api gpu   -- everything that runs kernels
//
// Prints one "PASS <name>" or "FAIL <name>: <why>" line per case; exits 1 if
// any case failed.
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpufhe.hpp"
#include "../../include/gpufhe_capi.h"
#include "../../include/gpufhe_kernels.h"
#include "../../oracle/oracle.h"

using namespace gpufhe;
using u64 = uint64_t;
using V = std::vector<u64>;

static int g_fail = 0;
static void run(const char* name, const std::function<void()>& f) {
    std::printf("case %s\n", name);
    std::fflush(stdout);
    try {
        f();
        std::printf("PASS %s\n", name);
    } catch (const std::exception& e) {
        std::printf("FAIL %s: %s\n", name, e.what());
        ++g_fail;
    }
    std::fflush(stdout);
}
static void expect(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error(what);
}
static void same(const V& a, const V& b, const std::string& what) {
    if (a.size() != b.size()) throw std::runtime_error(what + ": size " + std::to_string(a.size()) + " vs " + std::to_string(b.size()));
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i] != b[i]) throw std::runtime_error(what + ": first difference at " + std::to_string(i));
}
static void hip_ok(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

// ---- parameter set: 60-bit q0, 50-bit q1.., 60-bit P ----
struct Params {
    uint32_t n, L, K, dnum, alpha;
    V q, p, qp, roots;
    or_params* P = nullptr;
    Params(uint32_t n_, uint32_t L_, uint32_t dnum_, uint32_t K_) : n(n_), L(L_), K(K_), dnum(dnum_) {
        alpha = (L + dnum - 1) / dnum;
        q.resize(L);
        expect(or_native_primes(q.data(), 1, 60, n, nullptr, 0) == 0, "q0");
        if (L > 1) expect(or_native_primes(q.data() + 1, L - 1, 50, n, q.data(), 1) == 0, "q");
        p.resize(K);
        expect(or_native_primes(p.data(), K, 60, n, q.data(), L) == 0, "p");
        qp = q;
        qp.insert(qp.end(), p.begin(), p.end());
        for (u64 m : qp) roots.push_back(or_native_root(n, m));
        P = or_params_create(n, q.data(), L, p.data(), K, roots.data(), dnum);
    }
    ~Params() { or_params_free(P); }
    KeySwitchConstants consts(uint32_t sizeQl) const {
        KeySwitchConstants Kc;
        Kc.n = n;
        const uint32_t parts = (sizeQl + alpha - 1) / alpha;
        compute_keyswitch_constants(Kc, V(q.begin(), q.begin() + sizeQl), p, parts);
        Kc.rootModList = qp;
        Kc.rootValList = roots;
        return Kc;
    }
    V rand_towers(uint32_t towers, std::mt19937_64& g, bool with_p = false) const {
        V v((size_t)towers * n);
        for (uint32_t t = 0; t < towers; ++t)
            for (uint32_t k = 0; k < n; ++k) {
                const u64 m = with_p ? qp[t] : q[t];
                v[(size_t)t * n + k] = g() % m;
            }
        return v;
    }
    V rand_key(std::mt19937_64& g) const {   // [dnum][L+K][n], uniform per row modulus
        V v((size_t)dnum * (L + K) * n);
        for (uint32_t d = 0; d < dnum; ++d)
            for (uint32_t t = 0; t < L + K; ++t)
                for (uint32_t k = 0; k < n; ++k) v[((size_t)d * (L + K) + t) * n + k] = g() % qp[t];
        return v;
    }
};
static void set_key(KeySwitchConstants& Kc, const Params& P, const V& av, const V& bv) {
    const size_t per = (size_t)(P.L + P.K) * P.n;
    Kc.av.assign(Kc.numPart, {});
    Kc.bv.assign(Kc.numPart, {});
    for (uint32_t d = 0; d < Kc.numPart; ++d) {
        Kc.av[d].assign(av.begin() + d * per, av.begin() + (d + 1) * per);
        Kc.bv[d].assign(bv.begin() + d * per, bv.begin() + (d + 1) * per);
    }
    Kc.evalKeyTowers = P.L + P.K;
    Kc.fullQ = P.L;
}
static V P_mod_qp(const Params& P) {
    V r;
    for (u64 m : P.qp) {
        unsigned __int128 x = 1;
        for (u64 pj : P.p) x = x * (pj % m) % m;
        r.push_back((u64)x);
    }
    return r;
}
struct Dev {
    u64* p = nullptr;
    size_t words = 0;
    explicit Dev(size_t w) : words(w) { hip_ok(hipMalloc((void**)&p, w * 8)); }
    explicit Dev(const V& v) : Dev(v.size()) { hip_ok(hipMemcpy(p, v.data(), v.size() * 8, hipMemcpyHostToDevice)); }
    ~Dev() { (void)hipFree(p); }
    V get(size_t w = 0) const {
        V v(w ? w : words);
        hip_ok(hipMemcpy(v.data(), p, v.size() * 8, hipMemcpyDeviceToHost));
        return v;
    }
};
static std::vector<int64_t> negacyclic(const std::vector<int64_t>& a, const std::vector<int64_t>& b, uint32_t n) {
    std::vector<int64_t> r(n, 0);
    for (uint32_t i = 0; i < n; ++i)
        if (a[i])
            for (uint32_t j = 0; j < n; ++j)
                if (b[j]) {
                    const uint32_t k = i + j;
                    if (k < n) r[k] += a[i] * b[j];
                    else r[k - n] -= a[i] * b[j];
                }
    return r;
}

// ------------------------------------------------------------ cpu cases --
static void cpu_cases() {
    run("compute_keyswitch_constants_vs_oracle", [] {
        const uint32_t n = 1 << 12;
        for (auto shape : std::vector<std::pair<uint32_t, uint32_t>>{{4, 2}, {6, 3}, {12, 3}, {24, 4}, {5, 2}, {7, 7}}) {
            const uint32_t L = shape.first, dnum = shape.second;
            const uint32_t alpha = (L + dnum - 1) / dnum, K = (60 + (alpha - 1) * 50 + 59) / 60;
            Params P(n, L, dnum, K);
            KeySwitchConstants Kc;
            Kc.n = n;
            compute_keyswitch_constants(Kc, P.q, P.p, dnum);
            or_ksc* O = or_ks_constants(P.q.data(), L, P.p.data(), K, dnum, 0);
            expect(O != nullptr, "oracle constants");
            const uint32_t W = L + K;
            expect(Kc.alpha == O->alpha && Kc.numPart == dnum && Kc.fullQ == L, "scalars");
            same(Kc.pInvModq, V(O->pInvModq, O->pInvModq + L), "pInvModq");
            same(Kc.pHatInv, V(O->pHatInv, O->pHatInv + K), "pHatInv");
            same(Kc.pHatInvPrec, V(O->pHatInvPrec, O->pHatInvPrec + K), "pHatInvPrec");
            same(Kc.pHatModq, V(O->pHatModq, O->pHatModq + (size_t)K * L), "pHatModq");
            same(Kc.mdBMuLo, V(O->mdBMuLo, O->mdBMuLo + L), "mdBMuLo");
            same(Kc.mdBMuHi, V(O->mdBMuHi, O->mdBMuHi + L), "mdBMuHi");
            for (uint32_t pt = 0; pt < dnum; ++pt) {
                const uint32_t sp = O->sizePart[pt], sc = O->sizeCompl[pt];
                expect(Kc.sizePart[pt] == sp && Kc.sizeCompl[pt] == sc && Kc.startIdx[pt] == O->startIdx[pt], "part sizes");
                const size_t a = (size_t)pt * O->alpha;
                same(Kc.partQHatInv[pt], V(O->partQHatInv + a, O->partQHatInv + a + sp), "partQHatInv");
                same(Kc.partQHatInvPrec[pt], V(O->partQHatInvPrec + a, O->partQHatInvPrec + a + sp), "partQHatInvPrec");
                same(Kc.partSrcMod[pt], V(O->partSrcMod + a, O->partSrcMod + a + sp), "partSrcMod");
                const size_t b = (size_t)pt * O->alpha * W;
                same(Kc.partQHatModp[pt], V(O->partQHatModp + b, O->partQHatModp + b + (size_t)sp * sc), "partQHatModp");
                const size_t c = (size_t)pt * W;
                same(Kc.partComplMod[pt], V(O->partComplMod + c, O->partComplMod + c + sc), "partComplMod");
                same(Kc.partBMuLo[pt], V(O->partBMuLo + c, O->partBMuLo + c + sc), "partBMuLo");
                same(Kc.partBMuHi[pt], V(O->partBMuHi + c, O->partBMuHi + c + sc), "partBMuHi");
            }
            or_ks_constants_free(O);
        }
    });
    run("compute_keyswitch_constants_rejects_empty_part", [] {
        Params P(1 << 10, 9, 3, 2);
        KeySwitchConstants Kc;
        bool threw = false;
        try {
            compute_keyswitch_constants(Kc, P.q, P.p, 4);   // alpha 3 covers 9 towers in 3 parts
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, "no error for an empty trailing part");
    });
    run("native_params_vs_oracle", [] {
        for (uint32_t n : {1u << 10, 1u << 15, 1u << 16}) {
            V a, b(8);
            native_primes(a, 8, 50, n, {});
            expect(or_native_primes(b.data(), 8, 50, n, nullptr, 0) == 0, "oracle primes");
            same(a, b, "native_primes");
            for (u64 m : a) expect(native_root(n, m) == or_native_root(n, m), "native_root");
            V s1, s2, o1(7), o2(7);
            native_rescale_consts(s1, s2, a, 7);
            or_rescale_consts(o1.data(), o2.data(), a.data(), 7);
            same(V(s1.begin(), s1.begin() + 7), o1, "s1");
            same(V(s2.begin(), s2.begin() + 7), o2, "s2");
        }
    });
    run("encode_decode_vs_direct_sums", [] {
        // the direct O(N^2) sums of src/encode.cpp:13-44 restated here at small N
        for (uint32_t N : {16u, 256u, 2048u}) {
            const double Delta = std::ldexp(1.0, 40);
            std::mt19937_64 g(N);
            std::uniform_real_distribution<double> U(-1, 1);
            std::vector<std::complex<double>> z(N / 2);
            for (auto& x : z) x = {U(g), U(g)};
            std::vector<int64_t> m;
            encode_host(m, z, N, Delta);
            const uint64_t M = 2ull * N;
            std::vector<uint64_t> r(N / 2);
            for (uint64_t k = 0, x = 1; k < N / 2; ++k, x = x * 5 % M) r[k] = x;
            int64_t worst = 0;
            for (uint32_t j = 0; j < N; ++j) {
                double acc = 0;
                for (uint32_t k = 0; k < N / 2; ++k) {
                    const double th = -M_PI * (double)((j * r[k]) % M) / N;
                    acc += z[k].real() * std::cos(th) - z[k].imag() * std::sin(th);
                }
                const int64_t want = std::llround(2.0 * Delta / N * acc);
                worst = std::max<int64_t>(worst, std::llabs(want - m[j]));
            }
            // both sides round a double sum of N/2 terms of size ~Delta: agree to 2 units
            expect(worst <= 2, "encode differs from the direct sum by " + std::to_string(worst));
            std::vector<std::complex<double>> y;
            decode_host(y, m, N, Delta);
            double err = 0;
            for (uint32_t k = 0; k < N / 2; ++k) err = std::max(err, std::abs(y[k] - z[k]));
            expect(err < 1e-9, "decode(encode(z)) error " + std::to_string(err));
        }
    });
    run("ct_add_ct_host", [] {
        Params P(1 << 10, 3, 3, 1);
        std::mt19937_64 g(5);
        V a0 = P.rand_towers(3, g), a1 = P.rand_towers(3, g), b0 = P.rand_towers(3, g), b1 = P.rand_towers(3, g);
        V w0 = a0, w1 = a1;
        for (uint32_t t = 0; t < 3; ++t) {
            or_add_into(w0.data() + t * P.n, b0.data() + t * P.n, P.q[t], P.n);
            or_add_into(w1.data() + t * P.n, b1.data() + t * P.n, P.q[t], P.n);
        }
        ct_add_ct_host(a0, a1, b0, b1, 3, P.n, P.q);
        same(a0, w0, "c0");
        same(a1, w1, "c1");
    });
    run("ct_mul_pt_host", [] {
        Params P(1 << 10, 3, 3, 1);
        std::mt19937_64 g(6);
        V a0 = P.rand_towers(3, g), a1 = P.rand_towers(3, g), pt = P.rand_towers(3, g);
        V w0(a0.size()), w1(a1.size());
        for (uint32_t t = 0; t < 3; ++t) {
            or_rns_mult(a0.data() + t * P.n, pt.data() + t * P.n, w0.data() + t * P.n, P.q[t], P.n);
            or_rns_mult(a1.data() + t * P.n, pt.data() + t * P.n, w1.data() + t * P.n, P.q[t], P.n);
        }
        ct_mul_pt_host(a0, a1, pt, 3, P.n, P.q);
        same(a0, w0, "c0");
        same(a1, w1, "c1");
    });
    run("ks_context_create_rejects_bad_parts", [] {
        Params P(1 << 10, 4, 2, 2);
        KeySwitchConstants Kc = P.consts(4);
        Kc.numPart = 3;
        bool threw = false;
        try {
            (void)ks_context_create(Kc);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, "no error for numPart inconsistent with alpha");
    });
}

// ------------------------------------------------------------ gpu cases --
static void gpu_cases() {
    run("device_ciphertext_roundtrip_and_move", [] {
        Params P(1 << 10, 3, 3, 1);
        std::mt19937_64 g(1);
        V c0 = P.rand_towers(3, g), c1 = P.rand_towers(3, g);
        DeviceCiphertext ct(Scheme::CKKS, P.n, 3);
        ct.upload(c0, c1);
        DeviceCiphertext moved(std::move(ct));
        expect(ct.c0() == nullptr && moved.num_towers() == 3, "move");
        V h0, h1;
        moved.to_host(h0, h1);
        same(h0, c0, "c0");
        same(h1, c1, "c1");
        moved.drop_tower();
        moved.to_host(h0, h1);
        expect(h0.size() == 2 * (size_t)P.n, "drop_tower count");
        bool threw = false;
        try {
            moved.upload(c0, c1);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, "upload size mismatch not reported");
    });

    run("keygen_encrypt_decrypt_vs_oracle", [] {
        Params P(1 << 12, 4, 2, 2);
        for (uint32_t h : {0u, 64u}) {
            set_secret_hamming_weight(h);
            KeyPairHost kp = keygen_host(P.n, P.qp, P.roots, 1, 3.19, 777);
            const size_t T = P.qp.size() * (size_t)P.n;
            V s(T), a(T), b(T);
            or_keygen(P.n, P.qp.data(), P.roots.data(), (uint32_t)P.qp.size(), 1, 3.19, 777, h, s.data(), a.data(), b.data());
            same(kp.s, s, "s");
            same(kp.pkA, a, "pkA");
            same(kp.pkB, b, "pkB");
            std::vector<int64_t> m(P.n, 0);
            for (uint32_t k = 0; k < P.n; k += 7) m[k] = (int64_t)(k % 1000) - 500;
            const V pa(a.begin(), a.begin() + (size_t)P.L * P.n), pb(b.begin(), b.begin() + (size_t)P.L * P.n);
            V c0, c1;
            encrypt_host(c0, c1, m, pa, pb, P.n, P.q, P.roots, 1, 3.19, 909);
            V o0((size_t)P.L * P.n), o1((size_t)P.L * P.n);
            or_encrypt(P.n, P.q.data(), P.roots.data(), P.L, m.data(), pa.data(), pb.data(), 1, 3.19, 909, o0.data(), o1.data());
            same(c0, o0, "c0");
            same(c1, o1, "c1");
            std::vector<int64_t> dec, odec(P.n);
            decrypt_host(dec, c0, c1, kp.s, P.n, P.q, P.roots);
            or_decrypt(P.n, P.q[0], P.roots[0], c0.data(), c1.data(), kp.s.data(), odec.data());
            expect(dec == odec, "decrypt vs oracle");
            int64_t noise = 0;
            for (uint32_t k = 0; k < P.n; ++k) noise = std::max<int64_t>(noise, std::llabs(dec[k] - m[k]));
            expect(noise < 2000, "decryption noise " + std::to_string(noise));
        }
        set_secret_hamming_weight(0);
    });

    run("evalkeygen_vs_oracle", [] {
        Params P(1 << 11, 4, 2, 2);
        KeyPairHost kp = keygen_host(P.n, P.qp, P.roots, 1, 3.19, 31);
        KeySwitchConstants Kc = P.consts(P.L);
        evalkeygen_host(Kc, kp.s, kp.pkA, kp.pkB, P_mod_qp(P), P.qp, P.roots, 1, 3.19, 41);
        const uint32_t W = P.L + P.K;
        V s2((size_t)W * P.n);
        for (uint32_t t = 0; t < W; ++t)
            for (uint32_t k = 0; k < P.n; ++k) s2[(size_t)t * P.n + k] = or_mulmod(kp.s[t * P.n + k], kp.s[t * P.n + k], P.qp[t]);
        V av((size_t)P.dnum * W * P.n), bv(av.size());
        or_evalkeygen_sold(P.n, P.L, P.K, P.alpha, P.dnum, P.qp.data(), P.roots.data(), s2.data(), kp.pkA.data(),
                           kp.pkB.data(), 1, 3.19, 41, av.data(), bv.data());
        expect(Kc.evalKeyTowers == W && Kc.av.size() == P.dnum, "key shape");
        for (uint32_t d = 0; d < P.dnum; ++d) {
            same(Kc.av[d], V(av.begin() + (size_t)d * W * P.n, av.begin() + (size_t)(d + 1) * W * P.n), "av");
            same(Kc.bv[d], V(bv.begin() + (size_t)d * W * P.n, bv.begin() + (size_t)(d + 1) * W * P.n), "bv");
        }
    });

    run("keyswitch_resident_vs_core_and_oracle", [] {
        // tests/test_ks_resident.cpp: resident KS == host-orchestrated KS, and both == oracle
        for (auto shape : std::vector<std::array<uint32_t, 4>>{{1u << 12, 4, 2, 2}, {1u << 13, 6, 3, 2}}) {
            Params P(shape[0], shape[1], shape[2], shape[3]);
            std::mt19937_64 g(shape[1]);
            const V av = P.rand_key(g), bv = P.rand_key(g);
            KeySwitchConstants Kc = P.consts(P.L);
            set_key(Kc, P, av, bv);
            const V d = P.rand_towers(P.L, g);
            auto C = ks_context_create(Kc);
            auto W = ks_work_create(C);
            Dev dd(d), b0(d.size()), b1(d.size());
            keyswitch_resident(dd.p, b0.p, b1.p, C, W, nullptr);
            hip_ok(hipDeviceSynchronize());
            KeySwitchResult R = keyswitch_core_resident(d, Kc);
            V w0(d.size()), w1(d.size());
            expect(or_keyswitch(P.P, P.L, d.data(), av.data(), bv.data(), w0.data(), w1.data()) == 0, "oracle");
            same(b0.get(), w0, "ba0 resident");
            same(b1.get(), w1, "ba1 resident");
            same(R.ba0, w0, "ba0 core");
            same(R.ba1, w1, "ba1 core");
            ks_work_destroy(W);
            ks_context_destroy(C);
        }
    });

    run("plaintext_ops_vs_oracle", [] {
        // src/ptops.cpp gates on the device: (c0 pt, c1 pt) and (c0 + pt, c1)
        Params P(1 << 12, 5, 5, 1);
        const uint32_t n = P.n, l = 4;
        std::mt19937_64 g(21);
        KeySwitchConstants Kc = P.consts(P.L);
        auto C = ks_context_create(Kc);
        const V c0 = P.rand_towers(l, g), c1 = P.rand_towers(l, g), pt = P.rand_towers(l, g);
        Dev dpt(pt);
        DeviceCiphertext a(Scheme::CKKS, n, l), out(Scheme::CKKS, n, P.L);
        a.upload(c0, c1);
        EvalMultPlain(a, dpt.p, out, C, nullptr);
        hip_ok(hipDeviceSynchronize());
        expect(out.num_towers() == l, "EvalMultPlain level");
        V h0, h1, w0(c0.size()), w1(c1.size());
        out.to_host(h0, h1);
        for (uint32_t t = 0; t < l; ++t) {
            or_rns_mult(c0.data() + t * n, pt.data() + t * n, w0.data() + t * n, P.q[t], n);
            or_rns_mult(c1.data() + t * n, pt.data() + t * n, w1.data() + t * n, P.q[t], n);
        }
        same(h0, w0, "EvalMultPlain c0");
        same(h1, w1, "EvalMultPlain c1");
        EvalAddPlain(a, dpt.p, a, C, nullptr);   // in place
        hip_ok(hipDeviceSynchronize());
        a.to_host(h0, h1);
        V s0 = c0;
        for (uint32_t t = 0; t < l; ++t) or_add_into(s0.data() + t * n, pt.data() + t * n, P.q[t], n);
        same(h0, s0, "EvalAddPlain c0");
        same(h1, c1, "EvalAddPlain c1");
        ks_context_destroy(C);
    });

    run("rotate_hoisted_and_linear_transform_vs_oracle", [] {
        // hoisted rotations and the diagonal-method sum through the C++ API,
        // residue for residue against the oracle's restatements
        Params P(1 << 12, 6, 3, 2);
        const uint32_t n = P.n, l = 5;
        std::mt19937_64 g(31);
        const uint32_t ks[2] = {5, 125};
        V av[2], bv[2];
        std::vector<KeySwitchConstants> Kc(2);
        std::vector<DeviceKSContext> C;
        for (int j = 0; j < 2; ++j) {
            av[j] = P.rand_key(g);
            bv[j] = P.rand_key(g);
            Kc[j] = P.consts(P.L);
            set_key(Kc[j], P, av[j], bv[j]);
            C.push_back(ks_context_create(Kc[j]));
        }
        const V c0 = P.rand_towers(l, g), c1 = P.rand_towers(l, g);
        DeviceCiphertext ct(Scheme::CKKS, n, l);
        ct.upload(c0, c1);
        auto W = ks_work_create(C[0]);
        std::vector<DeviceKSWork> Wr{ks_work_create(C[0]), ks_work_create(C[1])};
        std::vector<DeviceCiphertext> outs;
        outs.emplace_back(Scheme::CKKS, n, l);
        outs.emplace_back(Scheme::CKKS, n, l);
        RotateHoisted(ct, {ks[0], ks[1]}, {&C[0], &C[1]}, outs, W, Wr, nullptr);
        hip_ok(hipDeviceSynchronize());
        std::vector<V> w0(2, V((size_t)l * n)), w1(2, V((size_t)l * n));
        const u64* oa[2] = {av[0].data(), av[1].data()};
        const u64* ob[2] = {bv[0].data(), bv[1].data()};
        u64* p0[2] = {w0[0].data(), w0[1].data()};
        u64* p1[2] = {w1[0].data(), w1[1].data()};
        expect(or_rotate_hoisted(P.P, l, c0.data(), c1.data(), 2, ks, oa, ob, p0, p1) == 0, "oracle hoisted");
        for (int j = 0; j < 2; ++j) {
            V h0, h1;
            outs[j].to_host(h0, h1);
            same(h0, w0[j], "RotateHoisted c0 " + std::to_string(j));
            same(h1, w1[j], "RotateHoisted c1 " + std::to_string(j));
        }
        // three terms, the first unrotated (no context, no workspace)
        const V pt0 = P.rand_towers(l, g), pt1 = P.rand_towers(l, g), pt2 = P.rand_towers(l, g);
        Dev d0(pt0), d1(pt1), d2(pt2);
        std::vector<DeviceKSWork> Wl{DeviceKSWork{}, Wr[0], Wr[1]};
        DeviceCiphertext out(Scheme::CKKS, n, l);
        LinearTransform(ct, {1, ks[0], ks[1]}, {nullptr, &C[0], &C[1]}, {d0.p, d1.p, d2.p}, out, W, Wl, nullptr);
        hip_ok(hipDeviceSynchronize());
        const uint32_t lk[3] = {1, ks[0], ks[1]};
        const u64* la[3] = {nullptr, av[0].data(), av[1].data()};
        const u64* lb[3] = {nullptr, bv[0].data(), bv[1].data()};
        const u64* lp[3] = {pt0.data(), pt1.data(), pt2.data()};
        V x0((size_t)l * n), x1((size_t)l * n);
        expect(or_linear_transform(P.P, l, c0.data(), c1.data(), 3, lk, la, lb, lp, x0.data(), x1.data()) == 0,
               "oracle linear transform");
        V h0, h1;
        out.to_host(h0, h1);
        same(h0, x0, "LinearTransform c0");
        same(h1, x1, "LinearTransform c1");
        ks_work_destroy(W);
        for (auto& w : Wr) ks_work_destroy(w);
        for (auto& c : C) ks_context_destroy(c);
    });

    run("keyreuse_lower_level_context", [] {
        // tests/test_keyreuse.cpp: a context built for a lower level reuses the
        // full-level key (P rows at fullQ + k)
        Params P(1 << 12, 6, 3, 2);
        std::mt19937_64 g(9);
        const V av = P.rand_key(g), bv = P.rand_key(g);
        for (uint32_t l : {5u, 4u, 2u}) {
            KeySwitchConstants Kc = P.consts(l);
            set_key(Kc, P, av, bv);
            Kc.fullQ = P.L;
            const V d = P.rand_towers(l, g);
            auto C = ks_context_create(Kc);
            auto W = ks_work_create(C);
            Dev dd(d), b0(d.size()), b1(d.size());
            keyswitch_resident(dd.p, b0.p, b1.p, C, W, nullptr);
            hip_ok(hipDeviceSynchronize());
            V w0(d.size()), w1(d.size());
            expect(or_keyswitch(P.P, l, d.data(), av.data(), bv.data(), w0.data(), w1.data()) == 0, "oracle");
            same(b0.get(), w0, "ba0 level " + std::to_string(l));
            same(b1.get(), w1, "ba1 level " + std::to_string(l));
            ks_work_destroy(W);
        }
    });

    run("evalmult_rescale_reference_sequence_exact_product", [] {
        // tests/test_evalmult_rescale.cpp:52-93: tensor -> keyswitch_resident ->
        // combine -> rescale_resident_raw x2, then an exact small-integer product
        Params P(1 << 12, 4, 2, 2);
        const uint32_t n = P.n, L = P.L;
        KeyPairHost kp = keygen_host(n, P.qp, P.roots, 1, 3.19, 101);
        KeySwitchConstants Kc = P.consts(L);
        evalkeygen_host(Kc, kp.s, kp.pkA, kp.pkB, P_mod_qp(P), P.qp, P.roots, 1, 3.19, 202);
        const int64_t D = (int64_t)1 << 40;
        std::vector<int64_t> a1(n, 0), a2(n, 0), m1(n, 0), m2(n, 0);
        a1[0] = 3; a1[1] = -2; a1[n - 1] = 7;
        a2[0] = 4; a2[2] = 5;
        for (uint32_t k = 0; k < n; ++k) { m1[k] = a1[k] * D; m2[k] = a2[k] * D; }
        const V pa(kp.pkA.begin(), kp.pkA.begin() + (size_t)L * n), pb(kp.pkB.begin(), kp.pkB.begin() + (size_t)L * n);
        V c0a, c1a, c0b, c1b;
        encrypt_host(c0a, c1a, m1, pa, pb, n, P.q, P.roots, 1, 3.19, 303);
        encrypt_host(c0b, c1b, m2, pa, pb, n, P.q, P.roots, 1, 3.19, 304);
        auto C = ks_context_create(Kc);
        auto W = ks_work_create(C);
        const size_t T = (size_t)L * n;
        Dev A0(c0a), A1(c1a), B0(c0b), B1(c1b), t0(T), t1(T), t2(T), ba0(T), ba1(T), r0(T), r1(T), mods(P.q), scr(n), drp(n);
        LaunchTensor(A0.p, A1.p, B0.p, B1.p, t0.p, t1.p, t2.p, mods.p, L, n, nullptr);
        keyswitch_resident(t2.p, ba0.p, ba1.p, C, W, nullptr);
        LaunchCombine(r0.p, r1.p, t0.p, t1.p, ba0.p, ba1.p, mods.p, L, n, nullptr);
        V s1, s2;
        native_rescale_consts(s1, s2, P.q, L - 1);
        rescale_resident_raw(r0.p, L, C, s1, s2, scr.p, drp.p, nullptr);
        rescale_resident_raw(r1.p, L, C, s1, s2, scr.p, drp.p, nullptr);
        hip_ok(hipDeviceSynchronize());
        const V g0 = r0.get((size_t)(L - 1) * n), g1 = r1.get((size_t)(L - 1) * n);
        V w0((size_t)(L - 1) * n), w1(w0.size());
        const V av = [&] { V v; for (auto& x : Kc.av) v.insert(v.end(), x.begin(), x.end()); return v; }();
        const V bv = [&] { V v; for (auto& x : Kc.bv) v.insert(v.end(), x.begin(), x.end()); return v; }();
        expect(or_evalmult_rescale(P.P, L, c0a.data(), c1a.data(), c0b.data(), c1b.data(), av.data(), bv.data(),
                                   w0.data(), w1.data()) == 0, "oracle");
        same(g0, w0, "out0 vs oracle");
        same(g1, w1, "out1 vs oracle");
        // first-class EvalMultRescale gives the same residues
        DeviceCiphertext a(Scheme::CKKS, n, L), b(Scheme::CKKS, n, L), o(Scheme::CKKS, n, L);
        a.upload(c0a, c1a);
        b.upload(c0b, c1b);
        EvalMultRescale(a, b, o, C, W, nullptr);
        hip_ok(hipDeviceSynchronize());
        expect(o.num_towers() == L - 1, "EvalMultRescale level");
        V h0, h1;
        o.to_host(h0, h1);
        same(h0, w0, "EvalMultRescale out0");
        same(h1, w1, "EvalMultRescale out1");
        // EvalMult then Rescale (first-class) once more
        DeviceCiphertext o2(Scheme::CKKS, n, L);
        EvalMult(a, b, o2, C, W, nullptr);
        Rescale(o2, C, W, nullptr);
        hip_ok(hipDeviceSynchronize());
        o2.to_host(h0, h1);
        same(h0, w0, "EvalMult+Rescale out0");
        same(h1, w1, "EvalMult+Rescale out1");
        std::vector<int64_t> dec;
        decrypt_host(dec, g0, g1, kp.s, n, P.q, P.roots);
        const std::vector<int64_t> want = negacyclic(a1, a2, n);
        const double ql = (double)P.q[L - 1];
        for (uint32_t k = 0; k < n; ++k) {
            const int64_t got = std::llround((double)dec[k] * ql / ((double)D * (double)D));
            expect(got == want[k], "product coefficient " + std::to_string(k));
        }
        ks_work_destroy(W);
    });

    run("rotate_vs_oracle_and_slots", [] {
        // tests/test_rot_resident.cpp device/host gates plus a decode check
        Params P(1 << 12, 4, 2, 2);
        const uint32_t n = P.n, L = P.L;
        KeyPairHost kp = keygen_host(n, P.qp, P.roots, 1, 3.19, 55);
        const uint32_t r = 3;
        uint64_t k = 1;
        for (uint32_t i = 0; i < r; ++i) k = k * 5 % (2ull * n);
        V s_auto = kp.s;
        expect(or_automorph_eval(P.P, s_auto.data(), L + P.K, (uint32_t)k, P.K) == 0, "oracle automorphism");
        // the library's own automorphism of the secret agrees
        V s_lib = kp.s;
        automorphism_eval_host(s_lib, L + P.K, n, (uint32_t)k, P.qp, P.roots);
        same(s_lib, s_auto, "automorphism_eval_host");
        KeySwitchConstants Kr = P.consts(L);
        evalkeygen_host_sold(Kr, kp.s, s_auto, kp.pkA, kp.pkB, P_mod_qp(P), P.qp, P.roots, 1, 3.19, 66);
        std::mt19937_64 g(77);
        std::uniform_real_distribution<double> U(-1, 1);
        std::vector<std::complex<double>> z(n / 2);
        for (auto& x : z) x = {U(g), 0.0};
        const double Delta = std::ldexp(1.0, 45);
        std::vector<int64_t> m;
        encode_host(m, z, n, Delta);
        const V pa(kp.pkA.begin(), kp.pkA.begin() + (size_t)L * n), pb(kp.pkB.begin(), kp.pkB.begin() + (size_t)L * n);
        V c0, c1;
        encrypt_host(c0, c1, m, pa, pb, n, P.q, P.roots, 1, 3.19, 88);
        V w0(c0.size()), w1(c0.size());
        const V av = [&] { V v; for (auto& x : Kr.av) v.insert(v.end(), x.begin(), x.end()); return v; }();
        const V bv = [&] { V v; for (auto& x : Kr.bv) v.insert(v.end(), x.begin(), x.end()); return v; }();
        expect(or_rotate(P.P, L, c0.data(), c1.data(), (uint32_t)k, av.data(), bv.data(), w0.data(), w1.data()) == 0, "oracle rotate");
        auto C = ks_context_create(Kr);
        auto W = ks_work_create(C);
        DeviceCiphertext ct(Scheme::CKKS, n, L);
        ct.upload(c0, c1);
        Rotate(ct, (uint32_t)k, C, W, nullptr);
        hip_ok(hipDeviceSynchronize());
        V h0, h1;
        ct.to_host(h0, h1);
        same(h0, w0, "Rotate c0");
        same(h1, w1, "Rotate c1");
        Dev d0(c0), d1(c1), scr(n), t0(c0.size()), t1(c0.size());
        rotate_ct_device(d0.p, d1.p, (uint32_t)k, C, W, L, n, P.q, P.roots, scr.p, t0.p, t1.p, nullptr);
        hip_ok(hipDeviceSynchronize());
        same(d0.get(), w0, "rotate_ct_device c0");
        same(d1.get(), w1, "rotate_ct_device c1");
        V r0 = c0, r1 = c1;
        rotate_ct_resident(r0, r1, (uint32_t)k, Kr, L, n, P.q, P.roots);
        same(r0, w0, "rotate_ct_resident c0");
        same(r1, w1, "rotate_ct_resident c1");
        std::vector<int64_t> dec;
        decrypt_host(dec, h0, h1, kp.s, n, P.q, P.roots);
        std::vector<std::complex<double>> y;
        decode_host(y, dec, n, Delta);
        double err = 0;
        for (uint32_t i = 0; i < n / 2; ++i) err = std::max(err, std::abs(y[i].real() - z[(i + r) % (n / 2)].real()));
        expect(err < 1e-6, "rotated slots error " + std::to_string(err));
        // automorphism_eval_device in place agrees with the oracle too
        Dev v(c0);
        automorphism_eval_device(v.p, L, n, (uint32_t)k, P.q, P.roots, scr.p, nullptr);
        hip_ok(hipDeviceSynchronize());
        V want = c0;
        expect(or_automorph_eval(P.P, want.data(), L, (uint32_t)k, 0) == 0, "oracle");
        same(v.get(), want, "automorphism_eval_device");
        ks_work_destroy(W);
    });

    run("relin_free_chain_tables", [] {
        // src/ckks_ops.cpp: mul_resident then rescale_resident vs the oracle
        Params P(1 << 12, 5, 5, 1);
        const uint32_t n = P.n, L = P.L;
        std::mt19937_64 g(12);
        V a0 = P.rand_towers(L, g), a1 = P.rand_towers(L, g), b0 = P.rand_towers(L, g), b1 = P.rand_towers(L, g);
        DeviceTables T;
        build_device_tables(T, n, P.q);
        for (uint32_t t = 0; t < L; ++t) expect(T.root[t] == P.roots[t], "table root");
        DeviceCiphertext a(Scheme::CKKS, n, L), b(Scheme::CKKS, n, L);
        a.upload(a0, a1);
        b.upload(b0, b1);
        mul_resident(a, b, T, nullptr);
        V s1, s2;
        native_rescale_consts(s1, s2, P.q, L - 1);
        Dev scr(n), drp(n);
        rescale_resident(a, T, s1, s2, nullptr, scr.p, drp.p);
        hip_ok(hipDeviceSynchronize());
        V h0, h1;
        a.to_host(h0, h1);
        V w0(a0.size()), w1(a1.size());
        for (uint32_t t = 0; t < L; ++t) {
            or_rns_mult(a0.data() + t * n, b0.data() + t * n, w0.data() + t * n, P.q[t], n);
            or_rns_mult(a1.data() + t * n, b1.data() + t * n, w1.data() + t * n, P.q[t], n);
        }
        expect(or_rescale_raw(P.P, w0.data(), L) == 0 && or_rescale_raw(P.P, w1.data(), L) == 0, "oracle rescale");
        w0.resize((size_t)(L - 1) * n);
        w1.resize((size_t)(L - 1) * n);
        same(h0, w0, "c0");
        same(h1, w1, "c1");
    });

    run("rescale_resident_raw_two_streams_vs_oracle", [] {
        // the context's one rescale scratch shared by calls alternating over two
        // streams (no allocation per call; each call on a new stream waits for
        // the previous one on the device): every result equals the oracle
        Params P(1 << 12, 6, 3, 2);
        std::mt19937_64 g(17);
        KeySwitchConstants Kc = P.consts(P.L);
        auto C = ks_context_create(Kc);
        hipStream_t st[2];
        hip_ok(hipStreamCreate(&st[0]));
        hip_ok(hipStreamCreate(&st[1]));
        std::vector<V> in;
        std::vector<std::unique_ptr<Dev>> d;
        const uint32_t lv[6] = {6, 5, 6, 3, 2, 6};
        for (uint32_t i = 0; i < 6; ++i) {
            in.push_back(P.rand_towers(lv[i], g));
            d.emplace_back(new Dev(in.back()));
        }
        for (uint32_t i = 0; i < 6; ++i) {
            V s1, s2;
            native_rescale_consts(s1, s2, P.q, lv[i] - 1);
            rescale_resident_raw(d[i]->p, lv[i], C, s1, s2, nullptr, nullptr, st[i % 2]);
        }
        hip_ok(hipDeviceSynchronize());
        for (uint32_t i = 0; i < 6; ++i) {
            V w = in[i];
            expect(or_rescale_raw(P.P, w.data(), lv[i]) == 0, "oracle");
            w.resize((size_t)(lv[i] - 1) * P.n);
            same(d[i]->get((size_t)(lv[i] - 1) * P.n), w, "rescale vs oracle");
        }
        hip_ok(hipStreamDestroy(st[0]));
        hip_ok(hipStreamDestroy(st[1]));
    });

    run("rescale_resident_raw_rejects_foreign_constants", [] {
        Params P(1 << 10, 4, 2, 2);
        std::mt19937_64 g(3);
        KeySwitchConstants Kc = P.consts(P.L);
        auto C = ks_context_create(Kc);
        V s1, s2;
        native_rescale_consts(s1, s2, P.q, P.L - 1);
        s1[0] ^= 1;
        Dev c(P.rand_towers(P.L, g));
        bool threw = false;
        try {
            rescale_resident_raw(c.p, P.L, C, s1, s2, nullptr, nullptr, nullptr);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, "no error for constants of other moduli");
    });
}

int main(int argc, char** argv) {
    const std::string which = argc > 1 ? argv[1] : "cpu";
    if (which == "cpu" || which == "all") cpu_cases();
    if (which == "gpu" || which == "all") gpu_cases();
    std::printf("%s: %d failed\n", which.c_str(), g_fail);
    return g_fail ? 1 : 0;
}
