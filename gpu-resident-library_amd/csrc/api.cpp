// C++ host API (include/gpufhe.hpp) over the extern "C" layer.
//
// Ciphertext ownership, the reference's resident keyswitch context / work set
// and its first-class operations, the relin-free chain helpers and the
// keyswitch constant families.  Key generation, encryption and encoding live
// in keys.cpp.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/gpufhe.hpp"
#include "../../include/gpufhe_capi.h"
#include "../../include/gpufhe_kernels.h"

namespace gpufhe {

namespace {
using u128 = unsigned __int128;

void ck(int rc, const char* what) {
    if (rc) throw std::runtime_error(std::string("gpufhe: ") + what + ": " + gf_last_error());
}
void hk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("gpufhe: ") + what + ": " + hipGetErrorString(e));
}
uint64_t mulm(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
uint64_t powm(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    for (; e; e >>= 1, b = mulm(b, b, q))
        if (e & 1) r = mulm(r, b, q);
    return r;
}
uint64_t invm(uint64_t a, uint64_t q) { return powm(a, q - 2, q); }   // q prime
uint64_t shoup_pre(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
// prod of mods[0..cnt) except index `skip`, mod q
uint64_t prod_except(const uint64_t* mods, size_t cnt, uint64_t q, long skip) {
    uint64_t r = 1 % q;
    for (size_t i = 0; i < cnt; ++i)
        if ((long)i != skip) r = mulm(r, mods[i] % q, q);
    return r;
}
uint64_t root_for(const KeySwitchConstants& K, uint64_t q) {
    for (size_t i = 0; i < K.rootModList.size() && i < K.rootValList.size(); ++i)
        if (K.rootModList[i] == q) return K.rootValList[i];
    return native_root(K.n, q);
}
}  // namespace

// ----------------------------------------------------------- ciphertext --
// include/device_ciphertext.h:15-43, src/device_ciphertext.cpp:12-61
DeviceCiphertext::DeviceCiphertext(Scheme scheme, uint32_t n, uint32_t num_towers)
    : scheme_(scheme), n_(n), towers_(num_towers), cap_(num_towers) {
    const size_t bytes = (size_t)n * num_towers * 8;
    if (!bytes) return;
    hk(hipMalloc((void**)&d_c0_, bytes), "hipMalloc c0");
    hipError_t e = hipMalloc((void**)&d_c1_, bytes);
    if (e != hipSuccess) {
        (void)hipFree(d_c0_);
        d_c0_ = nullptr;
        hk(e, "hipMalloc c1");
    }
}
DeviceCiphertext::~DeviceCiphertext() {
    if (d_c0_) (void)hipFree(d_c0_);
    if (d_c1_) (void)hipFree(d_c1_);
}
DeviceCiphertext::DeviceCiphertext(DeviceCiphertext&& o) noexcept
    : scheme_(o.scheme_), format_(o.format_), n_(o.n_), towers_(o.towers_), cap_(o.cap_), d_c0_(o.d_c0_),
      d_c1_(o.d_c1_) {
    o.d_c0_ = o.d_c1_ = nullptr;
    o.towers_ = o.cap_ = 0;
}
DeviceCiphertext& DeviceCiphertext::operator=(DeviceCiphertext&& o) noexcept {
    if (this != &o) {
        if (d_c0_) (void)hipFree(d_c0_);
        if (d_c1_) (void)hipFree(d_c1_);
        scheme_ = o.scheme_;
        format_ = o.format_;
        n_ = o.n_;
        towers_ = o.towers_;
        cap_ = o.cap_;
        d_c0_ = o.d_c0_;
        d_c1_ = o.d_c1_;
        o.d_c0_ = o.d_c1_ = nullptr;
        o.towers_ = o.cap_ = 0;
    }
    return *this;
}
void DeviceCiphertext::upload(const std::vector<uint64_t>& c0, const std::vector<uint64_t>& c1) {
    const size_t need = (size_t)n_ * towers_;
    if (c0.size() != need || c1.size() != need)
        throw std::runtime_error("DeviceCiphertext::upload size mismatch");
    hk(hipMemcpy(d_c0_, c0.data(), need * 8, hipMemcpyHostToDevice), "upload c0");
    hk(hipMemcpy(d_c1_, c1.data(), need * 8, hipMemcpyHostToDevice), "upload c1");
}
void DeviceCiphertext::to_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1) const {
    const size_t need = (size_t)n_ * towers_;
    c0.resize(need);
    c1.resize(need);
    hk(hipMemcpy(c0.data(), d_c0_, need * 8, hipMemcpyDeviceToHost), "to_host c0");
    hk(hipMemcpy(c1.data(), d_c1_, need * 8, hipMemcpyDeviceToHost), "to_host c1");
}

// ------------------------------------------------------ relin-free chain --
DeviceTables::~DeviceTables() {}

void build_device_tables(DeviceTables& T, uint32_t n, const std::vector<uint64_t>& moduli) {
    T.n = n;
    T.q = moduli;
    const size_t L = moduli.size();
    T.d_froots.assign(L, nullptr);
    T.d_fprecon.assign(L, nullptr);
    T.d_iroots.assign(L, nullptr);
    T.d_iprecon.assign(L, nullptr);
    T.root.assign(L, 0);
    T.ninv.assign(L, 0);
    T.ninv_precon.assign(L, 0);
    for (size_t t = 0; t < L; ++t) {
        T.root[t] = native_root(n, moduli[t]);
        ck(gf_root_tables(n, moduli[t], T.root[t], &T.d_froots[t], &T.d_fprecon[t], &T.d_iroots[t],
                          &T.d_iprecon[t], &T.ninv[t], &T.ninv_precon[t]),
           "root tables");
    }
}

void mul_resident(DeviceCiphertext& a, const DeviceCiphertext& b, const DeviceTables& T, hipStream_t stream) {
    const uint32_t n = a.n(), towers = a.num_towers();
    if (b.num_towers() < towers || T.q.size() < towers) throw std::runtime_error("mul_resident: tower count mismatch");
    for (uint32_t t = 0; t < towers; ++t) {
        const size_t o = (size_t)t * n;
        LaunchRNSMultTower(a.c0() + o, b.c0() + o, a.c0() + o, T.q[t], n, stream);
        LaunchRNSMultTower(a.c1() + o, b.c1() + o, a.c1() + o, T.q[t], n, stream);
    }
    hk(hipGetLastError(), "mul_resident launch");
}

void rescale_resident(DeviceCiphertext& ct, const DeviceTables& T, const std::vector<uint64_t>& s1,
                      const std::vector<uint64_t>& s2, hipStream_t stream, uint64_t* scratch, uint64_t* dropCoeff) {
    const uint32_t n = ct.n(), towers = ct.num_towers();
    if (towers < 2) return;
    const uint32_t last = towers - 1;
    if (T.q.size() < towers || s1.size() < last || s2.size() < last)
        throw std::runtime_error("rescale_resident: tables or constants shorter than the ciphertext");
    uint64_t* comps[2] = {ct.c0(), ct.c1()};
    for (uint64_t* base : comps) {
        hk(hipMemcpyAsync(dropCoeff, base + (size_t)last * n, (size_t)n * 8, hipMemcpyDeviceToDevice, stream),
           "copy dropped tower");
        LaunchINTT_GS(dropCoeff, T.d_iroots[last], T.d_iprecon[last], n, T.q[last], T.ninv[last],
                      T.ninv_precon[last], stream);
        for (uint32_t t = 0; t < last; ++t) {
            hk(hipMemcpyAsync(scratch, dropCoeff, (size_t)n * 8, hipMemcpyDeviceToDevice, stream), "copy");
            LaunchSwitchModulus(scratch, T.q[last], T.q[t], n, stream);
            LaunchNTT_CT(scratch, T.d_froots[t], T.d_fprecon[t], n, T.q[t], stream);
            LaunchRescaleFuse(base + (size_t)t * n, scratch, s1[t], s2[t], T.q[t], n, stream);
        }
    }
    hk(hipGetLastError(), "rescale_resident launch");
    ct.drop_tower();
}

// ----------------------------------------------------- keyswitch constants --
// src/ks_constants.cpp:31-102: ModDown family (P -> Q), then per part the
// decompose family over the part's complement (other Q towers in index order,
// then P).  Everything is a definitional residue, computed with 128-bit
// products modulo each target.
void compute_keyswitch_constants(KeySwitchConstants& K, const std::vector<uint64_t>& moduliQ,
                                 const std::vector<uint64_t>& moduliP, uint32_t numPart) {
    const uint32_t sizeQ = (uint32_t)moduliQ.size(), sizeP = (uint32_t)moduliP.size();
    if (numPart == 0 || sizeQ == 0) throw std::runtime_error("compute_keyswitch_constants: empty basis or no parts");
    const uint32_t alpha = (sizeQ + numPart - 1) / numPart;
    if (alpha * (numPart - 1) >= sizeQ)
        throw std::runtime_error("compute_keyswitch_constants: numPart leaves an empty digit part");
    K.sizeQl = sizeQ;
    K.sizeP = sizeP;
    K.numPart = numPart;
    K.alpha = alpha;
    K.fullQ = sizeQ;
    K.qMod = moduliQ;
    K.pMod = moduliP;
    const uint64_t* Pm = moduliP.data();

    K.pInvModq.resize(sizeQ);
    for (uint32_t i = 0; i < sizeQ; ++i) K.pInvModq[i] = invm(prod_except(Pm, sizeP, moduliQ[i], -1), moduliQ[i]);
    K.pHatInv.resize(sizeP);
    K.pHatInvPrec.resize(sizeP);
    K.pHatModq.resize((size_t)sizeP * sizeQ);
    for (uint32_t j = 0; j < sizeP; ++j) {
        const uint64_t pj = moduliP[j];
        K.pHatInv[j] = invm(prod_except(Pm, sizeP, pj, j), pj);
        K.pHatInvPrec[j] = shoup_pre(K.pHatInv[j], pj);
        for (uint32_t i = 0; i < sizeQ; ++i) K.pHatModq[(size_t)j * sizeQ + i] = prod_except(Pm, sizeP, moduliQ[i], j);
    }
    K.mdBMuLo.resize(sizeQ);
    K.mdBMuHi.resize(sizeQ);
    for (uint32_t i = 0; i < sizeQ; ++i) {
        const u128 mu = ~(u128)0 / moduliQ[i];
        K.mdBMuLo[i] = (uint64_t)mu;
        K.mdBMuHi[i] = (uint64_t)(mu >> 64);
    }

    K.sizePart.assign(numPart, 0);
    K.sizeCompl.assign(numPart, 0);
    K.startIdx.assign(numPart, 0);
    for (auto* v : {&K.partQHatInv, &K.partQHatInvPrec, &K.partSrcMod, &K.partQHatModp, &K.partComplMod,
                    &K.partBMuLo, &K.partBMuHi})
        v->assign(numPart, {});
    for (uint32_t p = 0; p < numPart; ++p) {
        const uint32_t st = alpha * p, sp = std::min(alpha, sizeQ - st);
        K.startIdx[p] = st;
        K.sizePart[p] = sp;
        const uint64_t* src = moduliQ.data() + st;
        for (uint32_t i = 0; i < sp; ++i) {
            const uint64_t v = invm(prod_except(src, sp, src[i], i), src[i]);
            K.partSrcMod[p].push_back(src[i]);
            K.partQHatInv[p].push_back(v);
            K.partQHatInvPrec[p].push_back(shoup_pre(v, src[i]));
        }
        std::vector<uint64_t>& cm = K.partComplMod[p];
        for (uint32_t i = 0; i < sizeQ; ++i)
            if (i < st || i >= st + sp) cm.push_back(moduliQ[i]);
        cm.insert(cm.end(), moduliP.begin(), moduliP.end());
        const uint32_t sc = (uint32_t)cm.size();
        K.sizeCompl[p] = sc;
        K.partQHatModp[p].resize((size_t)sp * sc);
        for (uint32_t i = 0; i < sp; ++i)
            for (uint32_t j = 0; j < sc; ++j) K.partQHatModp[p][(size_t)i * sc + j] = prod_except(src, sp, cm[j], i);
        for (uint32_t j = 0; j < sc; ++j) {
            const u128 mu = ~(u128)0 / cm[j];
            K.partBMuLo[p].push_back((uint64_t)mu);
            K.partBMuHi[p].push_back((uint64_t)(mu >> 64));
        }
    }
}

// ------------------------------------------------------ resident context --
namespace detail {
struct KSImpl {
    gf_ctx* ctx = nullptr;
    gf_key* key = nullptr;
    std::vector<uint64_t> s1, s2;       // rescale constants of the context's top level
    // rescale_resident_raw's scratch (gf_rescale_ws_bytes, L towers each): one
    // allocated with the context and one more with every ks_work_create, kept
    // in a pool; a stream takes one from the pool on its first call and keeps
    // it, so calls on different streams never share scratch and callers that
    // create one work set per stream never allocate in a call
    std::mutex mu;
    std::vector<void*> rs_all, rs_pool;
    std::map<hipStream_t, void*> rs_of;
    ~KSImpl() {
        for (void* p : rs_all) (void)gf_free(p);
        if (key) gf_key_destroy(key);
        if (ctx) gf_ctx_destroy(ctx);
    }
    void add_rescale_scratch() {
        void* p = nullptr;
        ck(gf_malloc(&p, gf_rescale_ws_bytes(ctx)), "rescale scratch");
        rs_all.push_back(p);
        rs_pool.push_back(p);
    }
    uint64_t* rescale_scratch(hipStream_t s) {
        auto it = rs_of.find(s);
        if (it != rs_of.end()) return (uint64_t*)it->second;
        if (rs_pool.empty()) add_rescale_scratch();   // more streams than work sets
        void* p = rs_pool.back();
        rs_pool.pop_back();
        rs_of.emplace(s, p);
        return (uint64_t*)p;
    }
};
}  // namespace detail

gf_ctx* DeviceKSContext::engine() const {
    if (!impl) throw std::runtime_error("gpufhe: keyswitch context is empty or destroyed");
    return impl->ctx;
}
gf_key* DeviceKSContext::key() const {
    if (!impl || !impl->key) throw std::runtime_error("gpufhe: keyswitch context holds no key");
    return impl->key;
}

// src/keyswitch_resident.cpp:32-101.  One engine context per KeySwitchConstants
// (towers = sizeQl, part width alpha); root tables come from the process-wide
// cache; the eval key is uploaded once in the reference's row layout.
DeviceKSContext ks_context_create(const KeySwitchConstants& K) {
    if (K.n == 0 || K.sizeQl == 0 || K.qMod.size() < K.sizeQl || K.pMod.size() < K.sizeP)
        throw std::runtime_error("ks_context_create: incomplete KeySwitchConstants");
    const uint32_t alpha = K.alpha ? K.alpha : (K.sizeQl + K.numPart - 1) / K.numPart;
    if ((K.sizeQl + alpha - 1) / alpha != K.numPart)
        throw std::runtime_error("ks_context_create: numPart does not match sizeQl and alpha");
    DeviceKSContext C;
    C.n = K.n;
    C.sizeQl = K.sizeQl;
    C.sizeP = K.sizeP;
    C.sizeQlP = K.sizeQl + K.sizeP;
    C.numPart = K.numPart;
    C.fullQ = K.fullQ ? K.fullQ : K.sizeQl;
    C.alpha = alpha;
    C.sizePart = K.sizePart;
    C.sizeCompl = K.sizeCompl;
    C.startIdx = K.startIdx;
    C.qModHost.assign(K.qMod.begin(), K.qMod.begin() + K.sizeQl);
    C.pModHost.assign(K.pMod.begin(), K.pMod.begin() + K.sizeP);
    C.modList = C.qModHost;
    C.modList.insert(C.modList.end(), C.pModHost.begin(), C.pModHost.end());
    for (uint64_t q : C.modList) {
        const uint64_t r = root_for(K, q);
        const uint64_t *fr, *fp, *ir, *ip;
        uint64_t ni, nip;
        ck(gf_root_tables(K.n, q, r, &fr, &fp, &ir, &ip, &ni, &nip), "root tables");
        C.rootList.push_back(r);
        C.d_fr.push_back(fr);
        C.d_fp.push_back(fp);
        C.d_ir.push_back(ir);
        C.d_ip.push_back(ip);
        C.ninv.push_back(ni);
        C.ninv_p.push_back(nip);
    }
    C.pInvModqHost.resize(K.sizeQl);
    for (uint32_t i = 0; i < K.sizeQl; ++i)
        C.pInvModqHost[i] = invm(prod_except(C.pModHost.data(), K.sizeP, C.qModHost[i], -1), C.qModHost[i]);

    auto impl = std::make_shared<detail::KSImpl>();
    impl->ctx = gf_ctx_create_alpha(K.n, C.qModHost.data(), K.sizeQl, C.pModHost.data(), K.sizeP,
                                    C.rootList.data(), alpha);
    if (!impl->ctx) ck(-1, "engine context");
    impl->add_rescale_scratch();
    if (!K.av.empty()) {
        const uint32_t rows = K.evalKeyTowers ? K.evalKeyTowers : C.fullQ + K.sizeP;
        if (K.av.size() < K.numPart || K.bv.size() < K.numPart)
            throw std::runtime_error("ks_context_create: eval key has fewer parts than numPart");
        const size_t per = (size_t)rows * K.n;
        std::vector<uint64_t> a(per * K.numPart), b(per * K.numPart);
        for (uint32_t p = 0; p < K.numPart; ++p) {
            if (K.av[p].size() < per || K.bv[p].size() < per)
                throw std::runtime_error("ks_context_create: eval key part shorter than evalKeyTowers x n");
            std::copy(K.av[p].begin(), K.av[p].begin() + per, a.begin() + p * per);
            std::copy(K.bv[p].begin(), K.bv[p].begin() + per, b.begin() + p * per);
        }
        impl->key = gf_key_create_rows(impl->ctx, a.data(), b.data(), K.numPart, rows, C.fullQ, nullptr);
        if (!impl->key) ck(-1, "eval key upload");
    }
    C.impl = impl;
    return C;
}

DeviceKSWork ks_work_create(const DeviceKSContext& C) {
    if (!C.impl) throw std::runtime_error("gpufhe: keyswitch context is empty or destroyed");
    {
        std::lock_guard<std::mutex> lk(C.impl->mu);
        C.impl->add_rescale_scratch();   // for the stream this work set will serve
    }
    DeviceKSWork W;
    W.bytes = gf_ws_bytes(C.engine());
    void* p = nullptr;
    ck(gf_malloc(&p, W.bytes), "work set");
    W.d_ws = (uint64_t*)p;
    return W;
}
void ks_context_destroy(DeviceKSContext& C) {
    C.impl.reset();
    C.d_fr.clear();
    C.d_fp.clear();
    C.d_ir.clear();
    C.d_ip.clear();
}
void ks_work_destroy(DeviceKSWork& W) {
    if (W.d_ws) (void)gf_free(W.d_ws);
    W.d_ws = nullptr;
    W.bytes = 0;
}

static void need_ws(const DeviceKSWork& W) {
    if (!W.d_ws) throw std::runtime_error("gpufhe: DeviceKSWork not created");
}

void keyswitch_resident(const uint64_t* d_a, uint64_t* d_ba0, uint64_t* d_ba1, const DeviceKSContext& C,
                        DeviceKSWork& W, hipStream_t s) {
    need_ws(W);
    ck(gf_keyswitch(C.engine(), C.key(), C.sizeQl, d_a, d_ba0, d_ba1, W.d_ws, s), "keyswitch_resident");
}

void rescale_resident_raw(uint64_t* d_c, uint32_t towers, const DeviceKSContext& C,
                          const std::vector<uint64_t>& s1, const std::vector<uint64_t>& s2, uint64_t* d_scratch,
                          uint64_t* d_drop, hipStream_t s) {
    (void)d_scratch;
    (void)d_drop;
    if (towers < 2) return;
    if (towers > C.sizeQl) throw std::runtime_error("rescale_resident_raw: more towers than the context");
    // the engine folds its own constants of these moduli; they must be the caller's
    std::vector<uint64_t> e1, e2;
    native_rescale_consts(e1, e2, C.qModHost, towers - 1);
    if (s1.size() < towers - 1 || s2.size() < towers - 1 ||
        !std::equal(e1.begin(), e1.begin() + (towers - 1), s1.begin()) ||
        !std::equal(e2.begin(), e2.begin() + (towers - 1), s2.begin()))
        throw std::runtime_error("rescale_resident_raw: s1/s2 are not the rescale constants of the context's moduli");
    uint64_t* scratch;
    {
        std::lock_guard<std::mutex> lk(C.impl->mu);
        scratch = C.impl->rescale_scratch(s);
    }
    ck(gf_rescale_comp(C.engine(), towers, d_c, d_c, scratch, s), "rescale_resident_raw");
}

KeySwitchResult keyswitch_core_resident(const std::vector<uint64_t>& aTowers, const KeySwitchConstants& K) {
    const size_t T = (size_t)K.sizeQl * K.n;
    if (aTowers.size() < T) throw std::runtime_error("keyswitch_core_resident: input shorter than sizeQl x n");
    DeviceKSContext C = ks_context_create(K);
    DeviceKSWork W = ks_work_create(C);
    void *a = nullptr, *b0 = nullptr, *b1 = nullptr;
    KeySwitchResult R;
    try {
        ck(gf_malloc(&a, T * 8), "input");
        ck(gf_malloc(&b0, T * 8), "output");
        ck(gf_malloc(&b1, T * 8), "output");
        ck(gf_h2d(a, aTowers.data(), T * 8), "upload");
        keyswitch_resident((const uint64_t*)a, (uint64_t*)b0, (uint64_t*)b1, C, W, nullptr);
        ck(gf_device_sync(), "sync");
        R.ba0.resize(T);
        R.ba1.resize(T);
        ck(gf_d2h(R.ba0.data(), b0, T * 8), "download");
        ck(gf_d2h(R.ba1.data(), b1, T * 8), "download");
    } catch (...) {
        (void)gf_free(a);
        (void)gf_free(b0);
        (void)gf_free(b1);
        ks_work_destroy(W);
        throw;
    }
    (void)gf_free(a);
    (void)gf_free(b0);
    (void)gf_free(b1);
    ks_work_destroy(W);
    return R;
}

// ----------------------------------------------------- first-class ops --
static void same_level(const DeviceCiphertext& a, const DeviceCiphertext& b, const DeviceKSContext& C) {
    if (a.n() != C.n || b.n() != C.n) throw std::runtime_error("gpufhe: ring dimension differs from the context");
    if (a.num_towers() != b.num_towers()) throw std::runtime_error("gpufhe: operands at different levels");
    if (a.num_towers() == 0 || a.num_towers() > C.sizeQl)
        throw std::runtime_error("gpufhe: operand level outside the context");
}
// the result's tower count must not exceed out's current count (drop_tower only lowers it)
static void check_out(const DeviceCiphertext& out, uint32_t towers, uint32_t n, const char* op) {
    if (out.n() != n || out.num_towers() < towers)
        throw std::runtime_error(std::string(op) + ": output ciphertext has fewer towers than the result");
}

void EvalMult(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out, const DeviceKSContext& C,
              DeviceKSWork& W, hipStream_t s) {
    same_level(a, b, C);
    need_ws(W);
    const uint32_t l = a.num_towers();
    check_out(out, l, C.n, "EvalMult");
    ck(gf_evalmult(C.engine(), C.key(), l, a.c0(), a.c1(), b.c0(), b.c1(), out.c0(), out.c1(), W.d_ws, s),
       "EvalMult");
    while (out.num_towers() > l) out.drop_tower();
}
void EvalMultRescale(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out,
                     const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s) {
    same_level(a, b, C);
    need_ws(W);
    const uint32_t l = a.num_towers();
    if (l < 2) throw std::runtime_error("EvalMultRescale: needs at least two towers");
    check_out(out, l - 1, C.n, "EvalMultRescale");
    ck(gf_evalmult_rescale(C.engine(), C.key(), l, a.c0(), a.c1(), b.c0(), b.c1(), out.c0(), out.c1(), W.d_ws, s),
       "EvalMultRescale");
    while (out.num_towers() > l - 1) out.drop_tower();
}
void Rescale(DeviceCiphertext& ct, const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s) {
    need_ws(W);
    const uint32_t l = ct.num_towers();
    if (ct.n() != C.n || l < 2 || l > C.sizeQl) throw std::runtime_error("Rescale: level outside the context");
    ck(gf_rescale(C.engine(), l, ct.c0(), ct.c1(), ct.c0(), ct.c1(), W.d_ws, s), "Rescale");
    ct.drop_tower();
}
void Rotate(DeviceCiphertext& ct, uint32_t k, const DeviceKSContext& rotC, DeviceKSWork& W, hipStream_t s) {
    need_ws(W);
    const uint32_t l = ct.num_towers();
    if (ct.n() != rotC.n || l == 0 || l > rotC.sizeQl) throw std::runtime_error("Rotate: level outside the context");
    ck(gf_rotate(rotC.engine(), rotC.key(), l, k, ct.c0(), ct.c1(), ct.c0(), ct.c1(), W.d_ws, s), "Rotate");
}
void KeySwitch(const uint64_t* d_a, uint32_t towers, uint64_t* d_ba0, uint64_t* d_ba1, const DeviceKSContext& C,
               DeviceKSWork& W, hipStream_t s) {
    need_ws(W);
    if (towers == 0 || towers > C.sizeQl) throw std::runtime_error("KeySwitch: level outside the context");
    ck(gf_keyswitch(C.engine(), C.key(), towers, d_a, d_ba0, d_ba1, W.d_ws, s), "KeySwitch");
}
void EvalAdd(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out, const DeviceKSContext& C,
             hipStream_t s) {
    same_level(a, b, C);
    const uint32_t l = a.num_towers();
    check_out(out, l, C.n, "EvalAdd");
    ck(gf_add(C.engine(), l, a.c0(), a.c1(), b.c0(), b.c1(), out.c0(), out.c1(), s), "EvalAdd");
    while (out.num_towers() > l) out.drop_tower();
}
void EvalMultPlain(const DeviceCiphertext& a, const uint64_t* d_pt, DeviceCiphertext& out, const DeviceKSContext& C,
                   hipStream_t s) {
    same_level(a, a, C);
    const uint32_t l = a.num_towers();
    check_out(out, l, C.n, "EvalMultPlain");
    ck(gf_mul_plain(C.engine(), l, a.c0(), a.c1(), d_pt, out.c0(), out.c1(), s), "EvalMultPlain");
    while (out.num_towers() > l) out.drop_tower();
}
void EvalAddPlain(const DeviceCiphertext& a, const uint64_t* d_pt, DeviceCiphertext& out, const DeviceKSContext& C,
                  hipStream_t s) {
    same_level(a, a, C);
    const uint32_t l = a.num_towers();
    check_out(out, l, C.n, "EvalAddPlain");
    ck(gf_add_plain(C.engine(), l, a.c0(), a.c1(), d_pt, out.c0(), out.c1(), s), "EvalAddPlain");
    while (out.num_towers() > l) out.drop_tower();
}

// The rotation contexts of one call must share the parameter set: their keys
// then have one layout and run on the first context's engine.
static const DeviceKSContext& rotation_engine(const DeviceCiphertext& ct,
                                              const std::vector<const DeviceKSContext*>& rotC,
                                              std::vector<gf_key*>& keys, const char* op) {
    const DeviceKSContext* E = nullptr;
    for (const DeviceKSContext* r : rotC)
        if (r) {
            E = r;
            break;
        }
    if (!E) throw std::runtime_error(std::string(op) + ": needs at least one rotation context");
    for (const DeviceKSContext* r : rotC)
        if (r && (r->modList != E->modList || r->numPart != E->numPart || r->sizeQl != E->sizeQl))
            throw std::runtime_error(std::string(op) + ": rotation contexts of different parameter sets");
    const uint32_t l = ct.num_towers();
    if (ct.n() != E->n || l == 0 || l > E->sizeQl)
        throw std::runtime_error(std::string(op) + ": level outside the context");
    keys.clear();
    for (const DeviceKSContext* r : rotC) keys.push_back(r ? r->key() : nullptr);
    return *E;
}
// Wrot[j] serves term j; a term without a key (the unrotated one) uses none,
// so its entry may be empty
static std::vector<uint64_t*> work_ptrs(std::vector<DeviceKSWork>& Wrot, const std::vector<gf_key*>& keys,
                                        const char* op) {
    if (Wrot.size() < keys.size()) throw std::runtime_error(std::string(op) + ": one workspace per term needed");
    std::vector<uint64_t*> p;
    for (size_t j = 0; j < keys.size(); ++j) {
        if (keys[j]) need_ws(Wrot[j]);
        p.push_back(Wrot[j].d_ws);
    }
    return p;
}

void RotateHoisted(const DeviceCiphertext& ct, const std::vector<uint32_t>& ks,
                   const std::vector<const DeviceKSContext*>& rotC, std::vector<DeviceCiphertext>& outs,
                   DeviceKSWork& W, std::vector<DeviceKSWork>& Wrot, hipStream_t s) {
    const size_t m = ks.size();
    if (m == 0 || rotC.size() != m || outs.size() != m)
        throw std::runtime_error("RotateHoisted: ks, rotC and outs differ in length");
    for (const DeviceKSContext* r : rotC)
        if (!r) throw std::runtime_error("RotateHoisted: every index needs a rotation context");
    need_ws(W);
    std::vector<gf_key*> keys;
    const DeviceKSContext& E = rotation_engine(ct, rotC, keys, "RotateHoisted");
    const uint32_t l = ct.num_towers();
    std::vector<uint64_t*> o0, o1;
    for (DeviceCiphertext& o : outs) {
        check_out(o, l, E.n, "RotateHoisted");
        o0.push_back(o.c0());
        o1.push_back(o.c1());
    }
    std::vector<uint64_t*> wr = work_ptrs(Wrot, keys, "RotateHoisted");
    std::vector<uint64_t> k64(ks.begin(), ks.end());
    ck(gf_rotate_hoisted(E.engine(), l, ct.c0(), ct.c1(), (uint32_t)m, keys.data(), k64.data(), o0.data(), o1.data(),
                         W.d_ws, wr.data(), s),
       "RotateHoisted");
    for (DeviceCiphertext& o : outs)
        while (o.num_towers() > l) o.drop_tower();
}

void LinearTransform(const DeviceCiphertext& ct, const std::vector<uint32_t>& ks,
                     const std::vector<const DeviceKSContext*>& rotC, const std::vector<const uint64_t*>& d_pts,
                     DeviceCiphertext& out, DeviceKSWork& W, std::vector<DeviceKSWork>& Wrot, hipStream_t s) {
    const size_t m = ks.size();
    if (m == 0 || rotC.size() != m || d_pts.size() != m)
        throw std::runtime_error("LinearTransform: ks, rotC and d_pts differ in length");
    need_ws(W);
    std::vector<gf_key*> keys;
    const DeviceKSContext& E = rotation_engine(ct, rotC, keys, "LinearTransform");
    const uint32_t l = ct.num_towers();
    check_out(out, l, E.n, "LinearTransform");
    std::vector<uint64_t*> wr = work_ptrs(Wrot, keys, "LinearTransform");
    std::vector<uint64_t> k64(ks.begin(), ks.end());
    ck(gf_linear_transform(E.engine(), l, ct.c0(), ct.c1(), (uint32_t)m, keys.data(), k64.data(), d_pts.data(),
                           out.c0(), out.c1(), W.d_ws, wr.data(), s),
       "LinearTransform");
    while (out.num_towers() > l) out.drop_tower();
}

}  // namespace gpufhe
