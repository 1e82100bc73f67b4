// Definitions behind the opaque handles of include/gpufhe_capi.h, shared by
// the C ABI translation units, and the exception-to-return-code guards.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

// A cached single-op plan embeds the device pointers of the keys it reads;
// `keys` lists those keys' serial numbers so that destroying a key evicts it.
struct gf_cached_plan {
    std::shared_ptr<gf::Plan> plan;
    std::vector<uint64_t> keys;
};
struct gf_ctx {
    std::unique_ptr<gf::Ctx> C;
    std::mutex mu;
    // single-op plans by (op, arguments); keys appear by serial, never by address
    std::map<std::vector<uint64_t>, gf_cached_plan> cache;
    // set up with the context for gf_evalmult_rescale_noalloc: the device the
    // context was created on, that device's address of the library's static
    // scratch array, and a pinned staging area (with the event of its last
    // copy) for one plan's job arrays
    int device = -1;
    uint64_t* scratch = nullptr;
    void* stage = nullptr;
    hipEvent_t stage_done = nullptr;
    ~gf_ctx();
};
struct gf_key {
    gf_ctx* ctx;
    std::unique_ptr<gf::DevKey> K;
    // process-wide unique and never reused (a new key may reuse a freed
    // handle's address, never its serial)
    uint64_t serial = next_serial();
    static uint64_t next_serial();
};
struct gf_plan {
    std::unique_ptr<gf::Plan> P;
};

namespace gf {
// message returned by gf_last_error() on this thread
void set_last_error(const char* msg);

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return -1;
    }
}
template <class T, class F>
T* guard_ptr(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
}  // namespace gf
