// Host emulation of the library's device tile logic (test infrastructure).
// Compiles the same headers the HIP kernels use with a plain C++ compiler and
// runs each workgroup sequentially, so index and arithmetic mistakes show up
// on the CPU against the oracle before any GPU run.
#include <cstdint>
#include <vector>
#include "ntt_engine.hpp"

using namespace gf;

namespace {
struct NoSync { void operator()() const {} };
}

extern "C" void emu_ntt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* tw, const uint64_t* twp,
                        int fwd, uint64_t post, uint64_t postp) {
    NttGeom g = make_geom(n);
    NttJob j{x, x, tw, twp, q, post, postp};
    std::vector<uint64_t> lds(NTT_TILE > n ? NTT_TILE : n);
    if (g.single) {
        if (fwd) single_body<true>(j, g, lds.data(), 0, 1, NoSync{});
        else single_body<false>(j, g, lds.data(), 0, 1, NoSync{});
        return;
    }
    const uint32_t tilesA = (1u << g.logN2) / g.C, tilesB = (1u << g.logN1) / g.R;
    if (fwd) {
        for (uint32_t t = 0; t < tilesA; ++t) phaseA_body<true>(j, g, t, lds.data(), 0, 1, NoSync{});
        for (uint32_t t = 0; t < tilesB; ++t) phaseB_body<true>(j, g, t, lds.data(), 0, 1, NoSync{});
    } else {
        for (uint32_t t = 0; t < tilesB; ++t) phaseB_body<false>(j, g, t, lds.data(), 0, 1, NoSync{});
        for (uint32_t t = 0; t < tilesA; ++t) phaseA_body<false>(j, g, t, lds.data(), 0, 1, NoSync{});
    }
}

extern "C" uint64_t emu_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return mulmod(a, b, q, host_qneg(q), host_r2(q));
}
extern "C" uint64_t emu_shoup(uint64_t x, uint64_t w, uint64_t q) {
    return shoup(x, w, host_precon(w, q), q);
}
extern "C" uint64_t emu_mulhi(uint64_t a, uint64_t b) { return mulhi64(a, b); }
