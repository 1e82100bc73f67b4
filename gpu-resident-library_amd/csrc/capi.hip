// extern "C" handle API (include/gpufhe_capi.h) over the engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpufhe_capi.h"
#include "handles.hpp"

namespace gpufhe {
void native_primes(std::vector<uint64_t>&, uint32_t, uint32_t, uint32_t, const std::vector<uint64_t>&);
uint64_t native_root(uint32_t, uint64_t);
void native_rescale_consts(std::vector<uint64_t>&, std::vector<uint64_t>&, const std::vector<uint64_t>&,
                           uint32_t);
}  // namespace gpufhe

using namespace gf;

static thread_local std::string g_err;
void gf::set_last_error(const char* msg) { g_err = msg; }

// single-op plans are cached by (op, arguments), keys by serial number;
// building one uploads its job arrays synchronously, which a stream capture
// cannot contain.  The caller holds a reference while it launches, so
// eviction by another thread never frees a plan in use.  `keys`: the serials
// of the keys the plan reads (gf_key_destroy evicts those plans).
static std::shared_ptr<Plan> cached_plan(gf_ctx* c, std::vector<uint64_t> key, std::vector<uint64_t> keys,
                                         hipStream_t s, const std::function<void(PlanBuilder&)>& build) {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->cache.find(key);
    if (it != c->cache.end()) return it->second.plan;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
        throw std::runtime_error("gpufhe: first call of an operation with new arguments inside a stream capture; "
                                 "run it once outside the capture first");
    if (c->cache.size() > 4096) c->cache.clear();
    PlanBuilder b(*c->C);
    build(b);
    std::shared_ptr<Plan> P(b.finish());
    c->cache.emplace(std::move(key), gf_cached_plan{P, std::move(keys)});
    return P;
}

// ---- allocation-free EvalMult+Rescale (gf_evalmult_rescale_noalloc) ----
// The workspace and the plan's job arrays live in one file-scope device array.
// Sized for one operation at n = 2^16, L = 24, K = 6, dnum = 4 (a workspace of
// 318 towers), plus 1 MiB of job arrays; each GPU holds its copy until the
// process ends.  Host code never passes `scratch` itself to a kernel: each
// device's address of it comes from hipGetSymbolAddress, resolved once per
// device (with that device current) when a context is created there.
namespace {
constexpr size_t NOALLOC_WS_WORDS = (size_t)318 << 16;
constexpr size_t NOALLOC_PLAN_BYTES = (size_t)1 << 20;
constexpr size_t NOALLOC_WORDS = NOALLOC_WS_WORDS + NOALLOC_PLAN_BYTES / 8;
}  // namespace
static __device__ __align__(256) uint64_t scratch[NOALLOC_WORDS];

static uint64_t* scratch_of_current_device(int* dev_out) {
    static std::mutex mu;
    static std::map<int, uint64_t*> addr;
    int dev = -1;
    hip_check(hipGetDevice(&dev), "current device");
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    uint64_t* p = nullptr;
    hip_check(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)), "address of the static scratch array");
    addr.emplace(dev, p);
    return p;
}

static void ctx_setup(gf_ctx* c) {
    c->scratch = scratch_of_current_device(&c->device);
    hip_check(hipHostMalloc(&c->stage, NOALLOC_PLAN_BYTES, hipHostMallocDefault), "plan staging");
    hip_check(hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming), "plan staging event");
}

gf_ctx::~gf_ctx() {
    if (stage_done) {
        (void)hipEventSynchronize(stage_done);
        (void)hipEventDestroy(stage_done);
    }
    if (stage) (void)hipHostFree(stage);
}

uint64_t gf_key::next_serial() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

#define U(x) ((uint64_t)(uintptr_t)(x))

extern "C" {

const char* gf_last_error(void) { return g_err.c_str(); }

int gf_malloc(void** p, size_t bytes) { return guard([&] { hip_check(hipMalloc(p, bytes), "hipMalloc"); }); }
int gf_free(void* p) { return guard([&] { hip_check(hipFree(p), "hipFree"); }); }
int gf_h2d(void* d, const void* h, size_t b) {
    return guard([&] { hip_check(hipMemcpy(d, h, b, hipMemcpyHostToDevice), "H2D"); });
}
int gf_d2h(void* h, const void* d, size_t b) {
    return guard([&] { hip_check(hipMemcpy(h, d, b, hipMemcpyDeviceToHost), "D2H"); });
}
int gf_d2d(void* d, const void* s, size_t b) {
    return guard([&] { hip_check(hipMemcpy(d, s, b, hipMemcpyDeviceToDevice), "D2D"); });
}
int gf_device_sync(void) { return guard([&] { hip_check(hipDeviceSynchronize(), "sync"); }); }
int gf_stream_create(hipStream_t* s) {
    return guard([&] { hip_check(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "stream"); });
}
int gf_stream_destroy(hipStream_t s) { return guard([&] { hip_check(hipStreamDestroy(s), "stream"); }); }
int gf_stream_sync(hipStream_t s) { return guard([&] { hip_check(hipStreamSynchronize(s), "stream sync"); }); }

int gf_native_primes(uint64_t* out, uint32_t count, uint32_t bits, uint32_t n, const uint64_t* avoid,
                     uint32_t navoid) {
    return guard([&] {
        std::vector<uint64_t> o, a(avoid, avoid + navoid);
        gpufhe::native_primes(o, count, bits, n, a);
        std::memcpy(out, o.data(), o.size() * 8);
    });
}
uint64_t gf_native_root(uint32_t n, uint64_t q) {
    uint64_t r = 0;
    guard([&] { r = gpufhe::native_root(n, q); });
    return r;
}
void gf_native_rescale_consts(uint64_t* s1, uint64_t* s2, const uint64_t* mod, uint32_t last) {
    std::vector<uint64_t> a, b, m(mod, mod + last + 1);
    gpufhe::native_rescale_consts(a, b, m, last);
    std::memcpy(s1, a.data(), last * 8);
    std::memcpy(s2, b.data(), last * 8);
}

gf_ctx* gf_ctx_create(uint32_t n, const uint64_t* q, uint32_t L, const uint64_t* p, uint32_t K,
                      const uint64_t* roots, uint32_t dnum) {
    return guard_ptr<gf_ctx>([&] {
        if (dnum == 0 || dnum > L) throw std::runtime_error("gpufhe: dnum must be in [1, L]");
        auto c = std::make_unique<gf_ctx>();
        c->C = std::make_unique<Ctx>(n, std::vector<uint64_t>(q, q + L), std::vector<uint64_t>(p, p + K),
                                     std::vector<uint64_t>(roots, roots + L + K), (L + dnum - 1) / dnum);
        ctx_setup(c.get());
        return c.release();
    });
}
void gf_ctx_destroy(gf_ctx* c) { delete c; }
size_t gf_ws_bytes(const gf_ctx* c) { return c->C->ws_words() * 8; }
size_t gf_rescale_ws_bytes(const gf_ctx* c) { return c->C->rescale_ws_words() * 8; }

gf_key* gf_key_create(gf_ctx* c, const uint64_t* av, const uint64_t* bv, uint32_t nparts, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        auto k = new gf_key;
        k->ctx = c;
        k->K = upload_key(*c->C, av, bv, nparts, c->C->L + c->C->K, c->C->L, s);
        return k;
    });
}
void gf_key_destroy(gf_key* k) {
    if (!k) return;
    {
        // plans that read this key hold its device pointers: drop them with it
        std::lock_guard<std::mutex> lk(k->ctx->mu);
        auto& cache = k->ctx->cache;
        for (auto it = cache.begin(); it != cache.end();) {
            const auto& ks = it->second.keys;
            if (std::find(ks.begin(), ks.end(), k->serial) != ks.end()) it = cache.erase(it);
            else ++it;
        }
    }
    delete k;
}
uint32_t gf_ctx_cached_plans(gf_ctx* c) {
    std::lock_guard<std::mutex> lk(c->mu);
    return (uint32_t)c->cache.size();
}

gf_ctx* gf_ctx_create_alpha(uint32_t n, const uint64_t* q, uint32_t L, const uint64_t* p, uint32_t K,
                            const uint64_t* roots, uint32_t alpha) {
    return guard_ptr<gf_ctx>([&] {
        if (alpha == 0 || alpha > L) throw std::runtime_error("gpufhe: part width alpha must be in [1, L]");
        auto c = std::make_unique<gf_ctx>();
        c->C = std::make_unique<Ctx>(n, std::vector<uint64_t>(q, q + L), std::vector<uint64_t>(p, p + K),
                                     std::vector<uint64_t>(roots, roots + L + K), alpha);
        ctx_setup(c.get());
        return c.release();
    });
}
gf_key* gf_key_create_rows(gf_ctx* c, const uint64_t* av, const uint64_t* bv, uint32_t nparts, uint32_t rows,
                           uint32_t fullQ, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        auto k = new gf_key;
        k->ctx = c;
        k->K = upload_key(*c->C, av, bv, nparts, rows, fullQ, s);
        return k;
    });
}
uint32_t gf_ctx_towers(const gf_ctx* c) { return c->C->L; }

gf_key* gf_key_generate(gf_ctx* c, const uint64_t* d_sold, const uint64_t* d_pka, const uint64_t* d_pkb,
                        uint64_t ns, double sigma, uint64_t seed, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        if (!(sigma > 0.0) || sigma > 20.0) throw std::runtime_error("gpufhe: sigma must be in (0, 20]");
        auto k = new gf_key;
        k->ctx = c;
        k->K = generate_key(*c->C, d_sold, d_pka, d_pkb, ns, sigma, seed, s);
        return k;
    });
}
gf_key* gf_key_generate_ctr(gf_ctx* c, const uint64_t* d_sold, const uint64_t* d_pka, const uint64_t* d_pkb,
                            uint64_t ns, double sigma, uint64_t seed, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        if (!(sigma > 0.0) || sigma > 20.0) throw std::runtime_error("gpufhe: sigma must be in (0, 20]");
        auto k = std::make_unique<gf_key>();
        k->ctx = c;
        k->K = generate_key(*c->C, d_sold, d_pka, d_pkb, ns, sigma, seed, s, true);
        return k.release();
    });
}
int gf_keygen(gf_ctx* c, uint64_t seed, double sigma, uint64_t ns, uint64_t* d_s, uint64_t* d_pka, uint64_t* d_pkb,
              hipStream_t s) {
    return guard([&] {
        if (!(sigma > 0.0) || sigma > 20.0) throw std::runtime_error("gpufhe: sigma must be in (0, 20]");
        keygen_device(*c->C, seed, sigma, ns, d_s, d_pka, d_pkb, s);
    });
}
int gf_key_download(const gf_key* k, uint64_t* h_av, uint64_t* h_bv, hipStream_t s) {
    return guard([&] { download_key(*k->ctx->C, *k->K, h_av, h_bv, s); });
}
int gf_encrypt(gf_ctx* c, uint32_t level, const uint64_t* d_pka, const uint64_t* d_pkb, const int64_t* h_m,
               uint64_t ns, double sigma, uint64_t seed, uint64_t* d_c0, uint64_t* d_c1, hipStream_t s) {
    return guard([&] {
        if (!(sigma > 0.0) || sigma > 20.0) throw std::runtime_error("gpufhe: sigma must be in (0, 20]");
        encrypt_device(*c->C, level, d_pka, d_pkb, h_m, ns, sigma, seed, d_c0, d_c1, s);
    });
}
int gf_decrypt(gf_ctx* c, const uint64_t* d_c0, const uint64_t* d_c1, const uint64_t* d_s, int64_t* h_m,
               hipStream_t s) {
    return guard([&] { decrypt_device(*c->C, d_c0, d_c1, d_s, h_m, s); });
}
void gf_key_shape(const gf_key* k, uint32_t* nparts, uint32_t* rows, uint32_t* fullQ) {
    *nparts = k->K->nparts;
    *rows = k->K->rows;
    *fullQ = k->K->fullQ;
}
gf_key* gf_key_create_uniform(gf_ctx* c, uint32_t nparts, uint64_t seed, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        const Ctx& C = *c->C;
        if (nparts < C.dnum) throw std::runtime_error("gpufhe: key has fewer parts than the context needs");
        auto K = std::make_unique<DevKey>();
        K->nparts = nparts;
        K->rows = C.L + C.K;
        K->fullQ = C.L;
        const size_t per = (size_t)K->rows * C.n, words = (size_t)nparts * per;
        K->a.alloc(words * 8);
        K->b.alloc(words * 8);
        DevBuf dm(C.mods.size() * 8);
        hip_check(hipMemcpy(dm.p, C.mods.data(), dm.bytes, hipMemcpyHostToDevice), "upload moduli");
        // uniform residues are uniform in Montgomery form too: no conversion needed
        for (uint32_t p = 0; p < nparts; ++p) {
            launch_fill(K->a.as<uint64_t>() + p * per, dm.as<uint64_t>(), K->rows, C.n, seed * 1000003 + 2 * p, s);
            launch_fill(K->b.as<uint64_t>() + p * per, dm.as<uint64_t>(), K->rows, C.n, seed * 1000003 + 2 * p + 1, s);
        }
        hip_check(hipStreamSynchronize(s), "key fill");
        auto k = new gf_key;
        k->ctx = c;
        k->K = std::move(K);
        return k;
    });
}

int gf_root_tables(uint32_t n, uint64_t q, uint64_t root, const uint64_t** fr, const uint64_t** fp,
                   const uint64_t** ir, const uint64_t** ip, uint64_t* ninv, uint64_t* ninv_precon) {
    return guard([&] {
        if (n < 2 || (n & (n - 1)) || n > (1u << 17)) throw std::runtime_error("gpufhe: ring dimension must be a power of two in [2, 2^17]");
        const RootTable t = root_table(n, q, root);
        *fr = t.fr;
        *fp = t.fp;
        *ir = t.ir;
        *ip = t.ip;
        *ninv = t.ninv;
        *ninv_precon = t.ninvp;
    });
}
int gf_ntt_towers(uint32_t n, uint64_t* d_x, uint32_t towers, const uint64_t* mods, const uint64_t* roots,
                  int inverse, hipStream_t s) {
    return guard([&] {
        if (n < 2 || (n & (n - 1)) || n > (1u << 17)) throw std::runtime_error("gpufhe: ring dimension must be a power of two in [2, 2^17]");
        const NttGeom g = make_geom(n);
        for (uint32_t t = 0; t < towers; ++t) {
            const RootTable r = root_table(n, mods[t], roots[t]);
            uint64_t* x = d_x + (size_t)t * n;
            NttJob j = inverse ? NttJob{x, x, r.ir, r.ip, mods[t], r.ninv, r.ninvp}
                               : NttJob{x, x, r.fr, r.fp, mods[t], 0, 0};
            launch_ntt_one(j, g, !inverse, s);
        }
        hip_check(hipGetLastError(), "NTT launch");
    });
}
int gf_automorph(uint32_t n, const uint64_t* d_in, uint64_t* d_out, uint32_t towers, uint64_t k, hipStream_t s) {
    return guard([&] {
        if (n < 2 || (n & (n - 1))) throw std::runtime_error("gpufhe: ring dimension must be a power of two");
        if (!(k & 1) || k >= 2ull * n) throw std::runtime_error("gpufhe: automorphism index must be odd and < 2n");
        if (d_in == d_out) throw std::runtime_error("gpufhe: the automorphism cannot run in place");
        uint32_t logn = 0;
        while ((1u << logn) < n) ++logn;
        for (uint32_t t = 0; t < towers; ++t)
            launch_perm_one(PermJob{d_in + (size_t)t * n, d_out + (size_t)t * n, k}, n, logn, s);
        hip_check(hipGetLastError(), "automorphism launch");
    });
}
int gf_rescale_comp(gf_ctx* c, uint32_t l, const uint64_t* d_c, uint64_t* d_o, uint64_t* ws, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {7, l, U(d_c), U(d_o), U(ws)}, {}, s,
                              [&](PlanBuilder& b) { b.rescale_one(l, d_c, d_o, ws); });
        P->run(s);
    });
}

int gf_evalmult_rescale(gf_ctx* c, gf_key* k, uint32_t l, const uint64_t* c0a, const uint64_t* c1a,
                        const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0, uint64_t* o1, uint64_t* ws,
                        hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {1, l, k->serial, U(c0a), U(c1a), U(c0b), U(c1b), U(o0), U(o1), U(ws)}, {k->serial}, s,
                              [&](PlanBuilder& b) { b.evalmult_rescale(l, c0a, c1a, c0b, c1b, o0, o1, *k->K, ws); });
        P->run(s);
    });
}
// gf_evalmult_rescale with its workspace and job arrays in the static scratch
// array: no device or pinned allocation in the call.  The plan is rebuilt on
// the host and its job arrays staged through the context's pinned buffer (after
// that buffer's previous copy has completed) on every call.
int gf_evalmult_rescale_noalloc(gf_ctx* c, gf_key* k, uint32_t l, const uint64_t* c0a, const uint64_t* c1a,
                                const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0, uint64_t* o1,
                                hipStream_t s) {
    return guard([&] {
        const Ctx& C = *c->C;
        if (C.ws_words() > NOALLOC_WS_WORDS)
            throw std::runtime_error("gpufhe: the context's workspace exceeds the static scratch maximum "
                                     "(318 towers of 2^16 words: n = 2^16, L = 24, K = 6, dnum = 4)");
        int dev = -1;
        hip_check(hipGetDevice(&dev), "current device");
        if (dev != c->device) throw std::runtime_error("gpufhe: context created on another device than the current one");
        PlanBuilder b(C);
        b.evalmult_rescale(l, c0a, c1a, c0b, c1b, o0, o1, *k->K, c->scratch);
        std::vector<char> blob;
        std::unique_ptr<Plan> P = b.finish_host(blob);
        if (blob.size() > NOALLOC_PLAN_BYTES) throw std::runtime_error("gpufhe: plan larger than the static job area");
        uint64_t* jobs = c->scratch + NOALLOC_WS_WORDS;
        std::lock_guard<std::mutex> lk(c->mu);
        hip_check(hipEventSynchronize(c->stage_done), "plan staging");
        if (!blob.empty()) {
            std::memcpy(c->stage, blob.data(), blob.size());
            hip_check(hipMemcpyAsync(jobs, c->stage, blob.size(), hipMemcpyHostToDevice, s), "stage plan");
        }
        hip_check(hipEventRecord(c->stage_done, s), "plan staging");
        P->base = jobs;
        P->run(s);
        hip_check(hipGetLastError(), "launch");
    });
}
size_t gf_noalloc_scratch_bytes(void) { return NOALLOC_WORDS * 8; }

int gf_evalmult(gf_ctx* c, gf_key* k, uint32_t l, const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
                const uint64_t* c1b, uint64_t* o0, uint64_t* o1, uint64_t* ws, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {2, l, k->serial, U(c0a), U(c1a), U(c0b), U(c1b), U(o0), U(o1), U(ws)}, {k->serial}, s,
                              [&](PlanBuilder& b) { b.evalmult(l, c0a, c1a, c0b, c1b, o0, o1, *k->K, ws); });
        P->run(s);
    });
}
int gf_keyswitch(gf_ctx* c, gf_key* k, uint32_t l, const uint64_t* d, uint64_t* ba0, uint64_t* ba1, uint64_t* ws,
                 hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {3, l, k->serial, U(d), U(ba0), U(ba1), U(ws)}, {k->serial}, s,
                              [&](PlanBuilder& b) { b.keyswitch(l, d, ba0, ba1, *k->K, ws); });
        P->run(s);
    });
}
int gf_rescale(gf_ctx* c, uint32_t l, const uint64_t* c0, const uint64_t* c1, uint64_t* o0, uint64_t* o1,
               uint64_t* ws, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {4, l, U(c0), U(c1), U(o0), U(o1), U(ws)}, {}, s,
                              [&](PlanBuilder& b) { b.rescale(l, c0, c1, o0, o1, ws); });
        P->run(s);
    });
}
int gf_rotate(gf_ctx* c, gf_key* k, uint32_t l, uint64_t kk, const uint64_t* c0, const uint64_t* c1, uint64_t* o0,
              uint64_t* o1, uint64_t* ws, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {5, l, k->serial, kk, U(c0), U(c1), U(o0), U(o1), U(ws)}, {k->serial}, s,
                              [&](PlanBuilder& b) { b.rotate(l, kk, c0, c1, o0, o1, *k->K, ws); });
        P->run(s);
    });
}
int gf_add(gf_ctx* c, uint32_t l, const uint64_t* a0, const uint64_t* a1, const uint64_t* b0, const uint64_t* b1,
           uint64_t* o0, uint64_t* o1, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {6, l, U(a0), U(a1), U(b0), U(b1), U(o0), U(o1)}, {}, s,
                              [&](PlanBuilder& b) { b.add(l, a0, a1, b0, b1, o0, o1); });
        P->run(s);
    });
}

gf_plan* gf_plan_evalmult_rescale(gf_ctx* c, gf_key* k, uint32_t count, const uint32_t* levels,
                                  const uint64_t* const* c0a, const uint64_t* const* c1a,
                                  const uint64_t* const* c0b, const uint64_t* const* c1b, uint64_t* const* o0,
                                  uint64_t* const* o1, uint64_t* const* ws) {
    return guard_ptr<gf_plan>([&] {
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i)
            b.evalmult_rescale(levels[i], c0a[i], c1a[i], c0b[i], c1b[i], o0[i], o1[i], *k->K, ws[i]);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
gf_plan* gf_plan_rotate(gf_ctx* c, gf_key* const* rks, uint32_t count, const uint32_t* levels, const uint64_t* ks,
                        const uint64_t* const* c0, const uint64_t* const* c1, uint64_t* const* o0,
                        uint64_t* const* o1, uint64_t* const* ws) {
    return guard_ptr<gf_plan>([&] {
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i) b.rotate(levels[i], ks[i], c0[i], c1[i], o0[i], o1[i], *rks[i]->K, ws[i]);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
gf_plan* gf_plan_rotate_hoisted(gf_ctx* c, uint32_t count, const uint32_t* levels, const uint64_t* const* c0,
                                const uint64_t* const* c1, uint32_t m, gf_key* const* rks, const uint64_t* ks,
                                uint64_t* const* o0, uint64_t* const* o1, uint64_t* const* ws_base,
                                uint64_t* const* ws_rot) {
    return guard_ptr<gf_plan>([&] {
        if (m == 0) throw std::runtime_error("gpufhe: hoisted rotation needs at least one index");
        std::vector<const DevKey*> keys(m);
        for (uint32_t j = 0; j < m; ++j) keys[j] = rks[j]->K.get();
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i)
            b.rotate_hoisted(levels[i], m, ks, c0[i], c1[i], o0 + (size_t)i * m, o1 + (size_t)i * m, keys.data(),
                             ws_base[i], ws_rot + (size_t)i * m);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
gf_plan* gf_plan_linear_transform(gf_ctx* c, uint32_t count, const uint32_t* levels, const uint64_t* const* c0,
                                  const uint64_t* const* c1, uint32_t m, gf_key* const* rks, const uint64_t* ks,
                                  const uint64_t* const* pts, uint64_t* const* o0, uint64_t* const* o1,
                                  uint64_t* const* ws_base, uint64_t* const* ws_rot) {
    return guard_ptr<gf_plan>([&] {
        std::vector<const DevKey*> keys(m);
        for (uint32_t j = 0; j < m; ++j) keys[j] = rks[j] ? rks[j]->K.get() : nullptr;
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i)
            b.linear_transform(levels[i], m, ks, c0[i], c1[i], pts, o0[i], o1[i], keys.data(), ws_base[i],
                               ws_rot + (size_t)i * m);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
int gf_mul_plain(gf_ctx* c, uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                 uint64_t* o1, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {10, l, U(c0), U(c1), U(pt), U(o0), U(o1)}, {}, s,
                             [&](PlanBuilder& b) { b.mul_plain(l, c0, c1, pt, o0, o1); });
        P->run(s);
    });
}
int gf_add_plain(gf_ctx* c, uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                 uint64_t* o1, hipStream_t s) {
    return guard([&] {
        auto P = cached_plan(c, {11, l, U(c0), U(c1), U(pt), U(o0), U(o1)}, {}, s,
                             [&](PlanBuilder& b) { b.add_plain(l, c0, c1, pt, o0, o1); });
        P->run(s);
    });
}
int gf_rotate_hoisted(gf_ctx* c, uint32_t l, const uint64_t* c0, const uint64_t* c1, uint32_t m, gf_key* const* rks,
                      const uint64_t* ks, uint64_t* const* o0, uint64_t* const* o1, uint64_t* ws_base,
                      uint64_t* const* ws_rot, hipStream_t s) {
    return guard([&] {
        if (m == 0) throw std::runtime_error("gpufhe: hoisted rotation needs at least one index");
        std::vector<uint64_t> key{12, l, m, U(c0), U(c1), U(ws_base)}, serials;
        std::vector<const DevKey*> keys(m);
        for (uint32_t j = 0; j < m; ++j) {
            if (!rks[j]) throw std::runtime_error("gpufhe: hoisted rotation needs a key per index");
            keys[j] = rks[j]->K.get();
            key.insert(key.end(), {rks[j]->serial, ks[j], U(o0[j]), U(o1[j]), U(ws_rot[j])});
            serials.push_back(rks[j]->serial);
        }
        auto P = cached_plan(c, std::move(key), std::move(serials), s, [&](PlanBuilder& b) {
            b.rotate_hoisted(l, m, ks, c0, c1, o0, o1, keys.data(), ws_base, ws_rot);
        });
        P->run(s);
    });
}
int gf_linear_transform(gf_ctx* c, uint32_t l, const uint64_t* c0, const uint64_t* c1, uint32_t m,
                        gf_key* const* rks, const uint64_t* ks, const uint64_t* const* pts, uint64_t* o0,
                        uint64_t* o1, uint64_t* ws_base, uint64_t* const* ws_rot, hipStream_t s) {
    return guard([&] {
        std::vector<uint64_t> key{13, l, m, U(c0), U(c1), U(o0), U(o1), U(ws_base)}, serials;
        std::vector<const DevKey*> keys(m);
        for (uint32_t j = 0; j < m; ++j) {
            keys[j] = rks[j] ? rks[j]->K.get() : nullptr;
            const uint64_t ser = rks[j] ? rks[j]->serial : 0;
            key.insert(key.end(), {ser, ks[j], U(pts[j]), U(ws_rot[j])});
            if (rks[j]) serials.push_back(ser);
        }
        auto P = cached_plan(c, std::move(key), std::move(serials), s, [&](PlanBuilder& b) {
            b.linear_transform(l, m, ks, c0, c1, pts, o0, o1, keys.data(), ws_base, ws_rot);
        });
        P->run(s);
    });
}
gf_plan* gf_plan_mul_plain(gf_ctx* c, uint32_t count, const uint32_t* levels, const uint64_t* const* c0,
                           const uint64_t* const* c1, const uint64_t* const* pt, uint64_t* const* o0,
                           uint64_t* const* o1) {
    return guard_ptr<gf_plan>([&] {
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i) b.mul_plain(levels[i], c0[i], c1[i], pt[i], o0[i], o1[i]);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
gf_plan* gf_plan_add_plain(gf_ctx* c, uint32_t count, const uint32_t* levels, const uint64_t* const* c0,
                           const uint64_t* const* c1, const uint64_t* const* pt, uint64_t* const* o0,
                           uint64_t* const* o1) {
    return guard_ptr<gf_plan>([&] {
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i) b.add_plain(levels[i], c0[i], c1[i], pt[i], o0[i], o1[i]);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
gf_plan* gf_plan_add(gf_ctx* c, uint32_t count, const uint32_t* levels, const uint64_t* const* a0,
                     const uint64_t* const* a1, const uint64_t* const* b0, const uint64_t* const* b1,
                     uint64_t* const* o0, uint64_t* const* o1) {
    return guard_ptr<gf_plan>([&] {
        PlanBuilder b(*c->C);
        for (uint32_t i = 0; i < count; ++i) b.add(levels[i], a0[i], a1[i], b0[i], b1[i], o0[i], o1[i]);
        auto p = new gf_plan;
        p->P = b.finish();
        return p;
    });
}
int gf_plan_run(gf_plan* p, hipStream_t s) { return guard([&] { p->P->run(s); }); }
int gf_plan_run_per_job(gf_plan* p, hipStream_t s) { return guard([&] { p->P->run_per_job(s); }); }
int gf_plan_run_graph(gf_plan* p, hipStream_t s) {
    return guard([&] {
        if (!s) throw std::runtime_error("gpufhe: graph capture needs a non-null stream");
        p->P->run_graph(s);
    });
}
uint32_t gf_plan_launches(const gf_plan* p) {
    uint32_t n = 0;
    for (const Step& s : p->P->steps)
        n += (s.kind == Step::NTT_F || s.kind == Step::NTT_I) ? (p->P->geom.single ? 1 : 2) : 1;
    return n;
}
void gf_plan_destroy(gf_plan* p) { delete p; }

uint32_t gf_plan_num_steps(const gf_plan* p) { return (uint32_t)p->P->steps.size(); }
int gf_plan_step_info(const gf_plan* p, uint32_t i, uint32_t* kind, uint32_t* jobs) {
    return guard([&] {
        if (i >= p->P->steps.size()) throw std::runtime_error("gpufhe: step index out of range");
        *kind = (uint32_t)p->P->steps[i].kind;
        *jobs = p->P->steps[i].count;
    });
}
int gf_plan_run_step(gf_plan* p, uint32_t i, hipStream_t s) {
    return guard([&] {
        if (i >= p->P->steps.size()) throw std::runtime_error("gpufhe: step index out of range");
        p->P->run_step(i, s);
    });
}
uint32_t gf_plan_step_kernels(const gf_plan* p, uint32_t i) {
    return i < p->P->steps.size() ? p->P->step_kernels(i) : 0u;
}
int gf_plan_run_step_kernel(gf_plan* p, uint32_t i, uint32_t which, hipStream_t s) {
    return guard([&] {
        if (i >= p->P->steps.size()) throw std::runtime_error("gpufhe: step index out of range");
        if (which < 1 || which > p->P->step_kernels(i))
            throw std::runtime_error("gpufhe: kernel index out of range for this step");
        p->P->run_step(i, s, which);
    });
}
int gf_ctx_set_ntt_radix(gf_ctx* c, uint32_t r) {
    return guard([&] {
        if (r < 2 || r > 4) throw std::runtime_error("gpufhe: NTT radix log2 must be 2, 3 or 4");
        std::lock_guard<std::mutex> lk(c->mu);
        c->C->set_radix(r);
        c->cache.clear();   // cached single-op plans carry the old geometry
    });
}
uint32_t gf_ctx_get_ntt_radix(const gf_ctx* c) { return c->C->geom.rmax; }
int gf_ctx_set_fusion(gf_ctx* c, int on) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(c->mu);
        c->C->fused = on != 0;
        c->cache.clear();
    });
}
int gf_ctx_autotune(gf_ctx* c, hipStream_t s, uint32_t* r, float* ms) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(c->mu);
        float m = 0;
        const uint32_t best = c->C->autotune(s, &m);
        c->cache.clear();
        if (r) *r = best;
        if (ms) *ms = m;
    });
}
int gf_ctx_autotune_plan(gf_ctx* c, uint32_t batch, hipStream_t s) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(c->mu);
        c->C->autotune_plan(s, batch);
        c->cache.clear();
    });
}
const char* gf_ctx_tuning_report(const gf_ctx* c) { return c->C->tune_report.c_str(); }
int gf_ctx_set_knob(gf_ctx* c, const char* name, uint32_t value) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(c->mu);
        c->C->set_knob(name, value);
        c->cache.clear();
    });
}
int gf_ctx_get_knob(const gf_ctx* c, const char* name, uint32_t* value) {
    return guard([&] { *value = c->C->knob(name); });
}
int gf_fill_uniform(gf_ctx* c, uint64_t* d, uint32_t level, uint64_t seed, hipStream_t s) {
    return guard([&] {
        if (level < 1 || level > c->C->L) throw std::runtime_error("gpufhe: bad level");
        launch_fill(d, c->C->lv[c->C->L].d_tqMD, level, c->C->n, seed, s);
    });
}
int gf_event_create(hipEvent_t* e) { return guard([&] { hip_check(hipEventCreate(e), "event"); }); }
int gf_event_destroy(hipEvent_t e) { return guard([&] { hip_check(hipEventDestroy(e), "event"); }); }
int gf_event_record(hipEvent_t e, hipStream_t s) { return guard([&] { hip_check(hipEventRecord(e, s), "event"); }); }
int gf_event_elapsed_ms(float* ms, hipEvent_t a, hipEvent_t b) {
    return guard([&] {
        hip_check(hipEventSynchronize(b), "event sync");
        hip_check(hipEventElapsedTime(ms, a, b), "event elapsed");
    });
}

int gf_capture_begin(hipStream_t s) {
    return guard([&] { hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"); });
}
int gf_capture_end(hipStream_t s, hipGraphExec_t* out) {
    return guard([&] {
        hipGraph_t g = nullptr;
        hip_check(hipStreamEndCapture(s, &g), "end capture");
        const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "instantiate graph");
    });
}
int gf_graph_launch(hipGraphExec_t g, hipStream_t s) {
    return guard([&] { hip_check(hipGraphLaunch(g, s), "graph launch"); });
}
int gf_graph_destroy(hipGraphExec_t g) {
    return guard([&] { hip_check(hipGraphExecDestroy(g), "graph destroy"); });
}

}  // extern "C"
