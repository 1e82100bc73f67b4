// The fused ModUp kernel (modup.hpp) and its launch.
#include "modup.hpp"
#include "ntt_kernels.hpp"

namespace gf {

// One workgroup per (part job, column tile).  Source count SP fixed at compile
// time, so every thread's unit of the sources stays in registers.
// sources I .. SP - 1, one after the other (compile-time source index, so the
// unit arrays are indexed by constants and stay in registers); source I is
// phase I and takes LDS tile I mod 2 and twiddle tile I mod 2; its twiddles
// are requested with its first pass's data
template <uint32_t S1, uint32_t SP, uint32_t I>
__device__ __forceinline__ void modup_sources(const ModupJob& j, const NttGeom& g, uint64_t* lds,
                                              uint64_t (*twl)[2u << S1], uint32_t tid, uint32_t tile,
                                              ModupUnit<SP>& x) {
    if constexpr (I < SP) {
        constexpr uint32_t TWP = 1u << S1, TWPER = TWP > MODUP_THREADS ? TWP / MODUP_THREADS : 1;
        Words2 tp[TWPER];
#pragma unroll
        for (uint32_t k = 0; k < TWPER; ++k)
            if (tid + k * MODUP_THREADS < TWP) tp[k] = modup_src_tw_pair(j, I, tid + k * MODUP_THREADS);
        uint64_t* tl = twl[I & 1];
        uint64_t v[MODUP_UNIT];
        modup_source_pass<S1, false>(j, I, g, tile, lds + (I & 1) * MODUP_TILE, tid, MODUP_THREADS, DevSync{}, v, tl,
                                     [&] {
#pragma unroll
                                         for (uint32_t k = 0; k < TWPER; ++k)
                                             if (tid + k * MODUP_THREADS < TWP)
                                                 modup_tw_put(tl, tid + k * MODUP_THREADS, tp[k]);
                                     });
        modup_take_unit<SP, S1>(j, I, v, x);
        modup_sources<S1, SP, I + 1>(j, g, lds, twl, tid, tile, x);
    }
}

// RS: the rescale mode (the fused ModDown + rescale prologue; ModupJob);
// C60: the conversion's R = 2^60 column reduction (ModupJob::c60)
template <uint32_t S1, uint32_t SP, bool RS, bool C60>
__global__ __launch_bounds__(MODUP_THREADS) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_modup(const ModupJob* __restrict__ jobs, uint32_t off, NttGeom g) {
    __shared__ uint64_t lds[2 * MODUP_TILE];   // phases alternate between the two tiles
    // the phases' column-pass twiddles (modup_src_tw_pair, modup_tw_pair), by
    // phase parity as the tiles: phase p + 2 rewrites phase p's copy only
    // after a barrier of phase p + 1, which every thread reaches after its
    // last read of it
    constexpr uint32_t TWP = 1u << S1, TWPER = TWP > MODUP_THREADS ? TWP / MODUP_THREADS : 1;
    __shared__ uint64_t twl[2][2 * TWP];
    const ModupJob& j = jobs[off + blockIdx.y];
    const uint32_t tid = threadIdx.x;
    // Workgroups go to the 8 XCDs in turn (linear id mod 8, and gridDim.x is
    // a multiple of 8 when remapped), each with its own L2.  A tile's rows are
    // 2^(10 - S1) words, shorter than a 128-byte L2 line at every ring size,
    // so neighbouring tiles share lines: give each XCD a run of consecutive
    // tiles of the tower instead of every eighth one.
    const uint32_t nt = gridDim.x;
    const uint32_t tile = nt % 8 ? blockIdx.x : (blockIdx.x % 8) * (nt / 8) + blockIdx.x / 8;
    ModupUnit<SP> x;
    modup_sources<S1, SP, 0>(j, g, lds, twl, tid, tile, x);
    // target t is phase SP + t; its twiddles are requested before its
    // conversion and stored to LDS after it
    uint64_t cv[MODUP_UNIT];
    auto target = [&](uint32_t t, auto&& convert) {
        Words2 tp[TWPER];
#pragma unroll
        for (uint32_t k = 0; k < TWPER; ++k)
            if (tid + k * MODUP_THREADS < TWP) tp[k] = modup_tw_pair(j, t, tid + k * MODUP_THREADS);
        convert();
        uint64_t* tl = twl[(SP + t) & 1];
#pragma unroll
        for (uint32_t k = 0; k < TWPER; ++k)
            if (tid + k * MODUP_THREADS < TWP) modup_tw_put(tl, tid + k * MODUP_THREADS, tp[k]);
        modup_target_pass<S1, false>(j, t, g, tile, lds + ((SP + t) & 1) * MODUP_TILE, tid, MODUP_THREADS, DevSync{},
                                     cv, tl);
    };
    if constexpr (RS) {
        uint64_t dv[MODUP_UNIT];
        modup_drop_unit<SP, S1>(j, g, tile, tid, x, dv);
        for (uint32_t t = 0; t < j.sc; ++t) target(t, [&] { modup_rs_unit<SP, S1>(j, t, x, dv, cv); });
    } else {
        for (uint32_t t = 0; t < j.sc; ++t) target(t, [&] { modup_convert_unit<SP, S1, C60>(j, t, x, cv); });
    }
}

// jobs of one launch share their source count sp (and c60)
template <bool RS, bool C60>
static void launch_modup_t(const ModupJob* d, uint32_t nj, const NttGeom& g, uint32_t sp, hipStream_t s) {
    if (g.single || g.n % MODUP_TILE || sp < 1 || sp > MODUP_MAXSRC)
        throw std::runtime_error("gpufhe: fused ModUp / ModDown launch shape out of range");
    const bool ok = with_split(g.logN1, g.logN2, [&](auto A, auto) {
        constexpr uint32_t S1 = decltype(A)::value;
        for (uint32_t off = 0; off < nj; off += MAXY) {
            const uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;
            const dim3 grid(g.n / MODUP_TILE, cnt);
            switch (sp) {
#define GF_SP(k) case k: k_modup<S1, k, RS, C60><<<grid, MODUP_THREADS, 0, s>>>(d, off, g); break;
                GF_SP(1) GF_SP(2) GF_SP(3) GF_SP(4) GF_SP(5) GF_SP(6)
#undef GF_SP
            }
        }
    });
    if (!ok) throw std::runtime_error("gpufhe: unsupported ring dimension for the fused ModUp");
}
// arg: sp | c60 << 8
void launch_modup(const ModupJob* d, uint32_t nj, const NttGeom& g, uint32_t arg, hipStream_t s) {
    if (arg >> 8) launch_modup_t<false, true>(d, nj, g, arg & 0xff, s);
    else launch_modup_t<false, false>(d, nj, g, arg & 0xff, s);
}
void launch_moddown(const ModupJob* d, uint32_t nj, const NttGeom& g, uint32_t sp, hipStream_t s) {
    launch_modup_t<true, false>(d, nj, g, sp, s);
}

}  // namespace gf
