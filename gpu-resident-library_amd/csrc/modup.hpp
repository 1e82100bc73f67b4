// Fused ModUp of one digit part on one column tile: the last pass of the
// inverse transforms of the part's towers, the basis conversion to the
// part's complement, and the first pass of the complement digits' forward
// transforms (the steps of src/keyswitch_resident.cpp:176-194 that run
// between the inverse transforms' row pass and the digits' row pass).
//
// The tile is MODUP_TILE elements: C = MODUP_TILE >> S1 columns of all
// N1 = 2^S1 rows.  With radix-4 register passes a column transform's first
// pass (chunk 0, stages 0..1) has exactly one unit per thread (256 threads):
// column s = unit & (C - 1), rows l + u TL (l = unit >> log C,
// TL = 2^(S1 - 2), u < 4), LDS slot s + (l + u TL) C.  The inverse
// transforms' last pass is the same chunk, so after it each thread keeps its
// unit of every source in registers (reduced, times n^-1 QHatInv, split into
// the 30-bit halves the conversion sums take: 8 VGPRs per source); for each
// target it forms the conversion of its
// 4 elements, stores them at their slots, and the target's forward column
// pass runs from LDS to HBM.  The coefficient-form sources and digits never
// travel through HBM.  Every butterfly, product and reduction is the one the
// unfused kernels apply, so the stored digit column
// passes equal theirs bit for bit.
#pragma once
#include "ntt_engine.hpp"

namespace gf {

// per-target and conversion tables are read at the same index by every lane:
// through the constant address space on the device (scalar loads)
#if defined(__HIP_DEVICE_COMPILE__)
#define GF_UNIFORM __attribute__((address_space(4)))
#else
#define GF_UNIFORM
#endif
template <class T>
GF_HD T uniform_at(const T* p, uint32_t i) {
    return ((const GF_UNIFORM T*)p)[i];
}

enum {
    MODUP_LOG_TILE = 10,                           // 1024 elements: 4 workgroups of 4 waves per CU
    MODUP_TILE = 1 << MODUP_LOG_TILE,
    MODUP_RM = 2,                                  // register radix 2^2: 4 elements per unit
    MODUP_UNIT = 1 << MODUP_RM,
    MODUP_THREADS = MODUP_TILE / MODUP_UNIT,       // one chunk-0 unit per thread
    MODUP_MAXSRC = 6
};

struct ModupJob {
    const uint64_t* src[MODUP_MAXSRC];   // the inverse transforms' row-pass output, tower st + i
    const uint64_t* stw[MODUP_MAXSRC];   // their paired inverse twiddle tables
    uint64_t sq[MODUP_MAXSRC], post[MODUP_MAXSRC], postp[MODUP_MAXSRC];   // modulus, n^-1 QHatInv
    uint64_t* dst;                        // complement digit j at dst + j n: its column-pass output
    const uint64_t* tab;                  // [sp][sc] conversion constants (Montgomery form)
    const uint64_t* tq;                   // [sc] target moduli
    const uint64_t* tqneg;                // [sc]
    const uint64_t* ttw;                  // [sc] addresses of the targets' paired forward tables
    const uint64_t* tred;                 // [sc] lazy_red_of(target modulus)
    uint32_t sp, sc;
};

// one thread's unit of every source: element u of source i (reduced, below
// 2^60) as its low and high 30 bits
template <uint32_t SP>
struct ModupUnit {
    uint32_t lo[MODUP_UNIT][SP], hi[MODUP_UNIT][SP];
};

template <uint32_t S1>
GF_HD uint32_t modup_slot(uint32_t unit, uint32_t u) {
    static_assert(chunk_size(S1, MODUP_RM, 0) == MODUP_RM, "chunk 0 must be a full radix-4 pass");
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC, TL = 1u << (S1 - MODUP_RM);
    return (unit & (C - 1)) + (((unit >> LOGC) + u * TL) << LOGC);
}

// source i's inverse column pass over the tile (all threads), its last pass
// into LDS; then each thread takes its unit (after the block barrier)
template <uint32_t S1, class Sync>
GF_HD void modup_source_pass(const ModupJob& j, uint32_t i, const NttGeom& g, uint32_t tile, uint64_t* lds,
                             uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC;
    const TileDesc d{LOGC, C, 1, 1, 0, true, false};
    const ColIn gin{gcast(j.src[i]), g.logN2, tile * C};
    tile_chunks<false, false, S1, MODUP_RM, 0>(lds, d, TwPaired{gcast(j.stw[i])}, TwPaired{gcast(j.stw[i])}, j.sq[i],
                                               tid, nthr, sync, gin, LdsOut{lds});
}
template <uint32_t SP, uint32_t S1>
GF_HD void modup_take_unit(const ModupJob& j, uint32_t i, const uint64_t* lds, uint32_t unit, ModupUnit<SP>& x) {
    const uint64_t post = j.post[i], postp = j.postp[i], q = j.sq[i];
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
        // the inverse transform's final store: times n^-1 QHatInv, reduced
        const uint64_t v = shoup(lds[modup_slot<S1>(unit, u)], post, postp, q);
        x.lo[u][i] = split30_lo(v);
        x.hi[u][i] = split30_hi(v);
    }
}
// target t's conversion of this thread's unit into LDS: the exact column sums
// over the part's sources, REDC without its correction ([0, 2 q_t)), as the
// conversion kernel leaves them for the forward transforms
template <uint32_t SP, uint32_t S1>
GF_HD void modup_convert_unit(const ModupJob& j, uint32_t t, uint64_t* lds, uint32_t unit, ModupUnit<SP>& x) {
    const uint64_t q = uniform_at(j.tq, t), qn = uniform_at(j.tqneg, t);
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
#pragma unroll
        for (uint32_t i = 0; i < SP; ++i) {
            // in place, so that each target's sums take the 32-bit halves as
            // they are (no 64-bit copies made once before the target loop)
            x.lo[u][i] = opaque32(x.lo[u][i]);
            x.hi[u][i] = opaque32(x.hi[u][i]);
        }
        const U128 acc = dot30<SP>(x.lo[u], x.hi[u], [&](uint32_t s) { return uniform_at(j.tab, s * j.sc + t); });
        lds[modup_slot<S1>(unit, u)] = redc_lazy(acc, q, qn);
    }
}
// target t's forward column pass from LDS, stored to HBM unreduced (the
// digits' row pass continues from these values)
template <uint32_t S1, class Sync>
GF_HD void modup_target_pass(const ModupJob& j, uint32_t t, const NttGeom& g, uint32_t tile, uint64_t* lds,
                             uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC;
    const TileDesc d{LOGC, C, 1, 1, 0, true, false};
    NttJob jt{};
    jt.q = uniform_at(j.tq, t);
    const ColOut<true> gout{gcast(j.dst + (size_t)t * g.n), g.logN2, tile * C, jt};
    const TwPaired tw{gcast((const uint64_t*)(uintptr_t)uniform_at(j.ttw, t))};
    if (uniform_at(j.tred, t))
        tile_chunks<true, true, S1, MODUP_RM, 0>(lds, d, tw, tw, jt.q, tid, nthr, sync, LdsIn{lds}, gout);
    else
        tile_chunks<true, false, S1, MODUP_RM, 0>(lds, d, tw, tw, jt.q, tid, nthr, sync, LdsIn{lds}, gout);
}

}  // namespace gf
