// C-ABI wrappers around the reference's host-only functions, so the tests can
// call the reference itself (compiled from its own sources by oracle/Makefile)
// next to the oracle.  Declarations follow the reference's own tests, which
// declare these functions the same way (tests/test_native_e2e.cpp:13-20).
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include "keyswitch.h"

namespace gpufhe {
void native_primes(std::vector<uint64_t>&, uint32_t, uint32_t, uint32_t, const std::vector<uint64_t>&);
uint64_t native_root(uint32_t, uint64_t);
void native_rescale_consts(std::vector<uint64_t>&, std::vector<uint64_t>&, const std::vector<uint64_t>&, uint32_t);
void encode_host(std::vector<int64_t>&, const std::vector<std::complex<double>>&, uint32_t, double);
void decode_host(std::vector<std::complex<double>>&, const std::vector<int64_t>&, uint32_t, double);
}

extern "C" {
int ref_native_primes(uint64_t* out, uint32_t count, uint32_t bits, uint32_t n,
                      const uint64_t* avoid, uint32_t navoid) {
    std::vector<uint64_t> o, a(avoid, avoid + navoid);
    try { gpufhe::native_primes(o, count, bits, n, a); } catch (...) { return -1; }
    std::memcpy(out, o.data(), o.size() * 8);
    return 0;
}
uint64_t ref_native_root(uint32_t n, uint64_t q) { return gpufhe::native_root(n, q); }
void ref_rescale_consts(uint64_t* s1, uint64_t* s2, const uint64_t* mod, uint32_t count, uint32_t last) {
    std::vector<uint64_t> a, b, m(mod, mod + count);
    gpufhe::native_rescale_consts(a, b, m, last);
    std::memcpy(s1, a.data(), a.size() * 8);
    std::memcpy(s2, b.data(), b.size() * 8);
}
// Flattened KeySwitchConstants with the same layout as the oracle's or_ksc:
// per part p: partQHatInv/Prec/SrcMod at p*alpha, partQHatModp packed at
// p*alpha*W (stride sizeCompl), complement arrays at p*W; W = sizeQ + sizeP.
int ref_ks_constants(const uint64_t* q, uint32_t sizeQ, const uint64_t* p, uint32_t sizeP,
                     uint32_t numPart, uint32_t* alpha_out, uint32_t* sizePart, uint32_t* sizeCompl,
                     uint32_t* startIdx, uint64_t* qhi, uint64_t* qhip, uint64_t* srcMod,
                     uint64_t* qhmp, uint64_t* complMod, uint64_t* muLo, uint64_t* muHi,
                     uint64_t* pHatInv, uint64_t* pHatInvPrec, uint64_t* pHatModq,
                     uint64_t* mdMuLo, uint64_t* mdMuHi, uint64_t* pInvModq) {
    gpufhe::KeySwitchConstants K;
    std::vector<uint64_t> mq(q, q + sizeQ), mp(p, p + sizeP);
    gpufhe::compute_keyswitch_constants(K, mq, mp, numPart);
    const uint32_t a = K.alpha, W = sizeQ + sizeP;
    *alpha_out = a;
    for (uint32_t k = 0; k < numPart; ++k) {
        sizePart[k] = K.sizePart[k]; sizeCompl[k] = K.sizeCompl[k]; startIdx[k] = K.startIdx[k];
        std::memcpy(qhi + (size_t)k * a, K.partQHatInv[k].data(), K.partQHatInv[k].size() * 8);
        std::memcpy(qhip + (size_t)k * a, K.partQHatInvPrec[k].data(), K.partQHatInvPrec[k].size() * 8);
        std::memcpy(srcMod + (size_t)k * a, K.partSrcMod[k].data(), K.partSrcMod[k].size() * 8);
        std::memcpy(qhmp + (size_t)k * a * W, K.partQHatModp[k].data(), K.partQHatModp[k].size() * 8);
        std::memcpy(complMod + (size_t)k * W, K.partComplMod[k].data(), K.partComplMod[k].size() * 8);
        std::memcpy(muLo + (size_t)k * W, K.partBMuLo[k].data(), K.partBMuLo[k].size() * 8);
        std::memcpy(muHi + (size_t)k * W, K.partBMuHi[k].data(), K.partBMuHi[k].size() * 8);
    }
    std::memcpy(pHatInv, K.pHatInv.data(), sizeP * 8);
    std::memcpy(pHatInvPrec, K.pHatInvPrec.data(), sizeP * 8);
    std::memcpy(pHatModq, K.pHatModq.data(), (size_t)sizeP * sizeQ * 8);
    std::memcpy(mdMuLo, K.mdBMuLo.data(), sizeQ * 8);
    std::memcpy(mdMuHi, K.mdBMuHi.data(), sizeQ * 8);
    std::memcpy(pInvModq, K.pInvModq.data(), sizeQ * 8);
    return 0;
}
void ref_encode(int64_t* m, const double* re, const double* im, uint32_t N, double delta) {
    std::vector<std::complex<double>> z(N / 2);
    for (uint32_t k = 0; k < N / 2; ++k) z[k] = {re[k], im[k]};
    std::vector<int64_t> o;
    gpufhe::encode_host(o, z, N, delta);
    std::memcpy(m, o.data(), (size_t)N * 8);
}
void ref_decode(double* re, double* im, const int64_t* m, uint32_t N, double delta) {
    std::vector<int64_t> mm(m, m + N);
    std::vector<std::complex<double>> z;
    gpufhe::decode_host(z, mm, N, delta);
    for (uint32_t k = 0; k < N / 2; ++k) { re[k] = z[k].real(); im[k] = z[k].imag(); }
}
}
