// Definitions behind the opaque handles of include/gpufhe_capi.h, shared by
// the C ABI translation units, and the exception-to-return-code guards.
#pragma once
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

struct gf_ctx {
    std::unique_ptr<gf::Ctx> C;
    std::mutex mu;
    std::map<std::vector<uint64_t>, std::shared_ptr<gf::Plan>> cache;  // single-op plans
};
struct gf_key {
    gf_ctx* ctx;
    std::unique_ptr<gf::DevKey> K;
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
