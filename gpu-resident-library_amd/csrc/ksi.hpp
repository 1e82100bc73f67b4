// Fused last pass of the ModUp digits' forward transform and the key inner
// product (the inner product of src/keyswitch_resident.cpp:171-200 over
// digits produced by the forward NTT of :184-194).
//
// One workgroup owns KSI_TILE consecutive positions of tower row i (ROWS
// transform rows of N2) for every part p of the decomposition.  Wave w runs
// the forward row pass (phase B) of the transformed parts w, w + nwaves, ...
// (in part order) on LDS region w, w + nwaves, ..., reading the phase-A output
// straight from HBM, and keeps the evaluation-form digits there; after one
// block barrier every thread takes KSI_TILE / nthr positions and accumulates
// sum_p digit_p * key_p for both key components as 128-bit sums.  A launch
// gives its jobs one wave per transformed part (up to KSI_WAVES), so no wave
// idles through the row passes.  The transformed digits therefore never
// travel through HBM.  Parts whose tower i is the input's own (no transform)
// are read from HBM at the inner product.
#pragma once
#include "ntt_engine.hpp"

namespace gf {

enum {
    KSI_MAXP = 7,                                    // parts per job (KSI_MAXP regions fit 64 KiB of LDS)
    KSI_LOG_TILE = 10,
    KSI_TILE = 1 << KSI_LOG_TILE,                    // positions per workgroup
    KSI_PART_WORDS = KSI_TILE + KSI_TILE / 16,       // one part's padded LDS region
    KSI_WAVES = 4,                                   // largest workgroup, in waves
    KSI_THREADS = 64 * KSI_WAVES
};

struct KsiJob {
    // part p's tower of row i: the phase-A output of its forward transform
    // (bit p of xf set) or an evaluation-form digit read as it is
    const uint64_t* src[KSI_MAXP];
    const uint64_t* kb[KSI_MAXP];   // key rows (Montgomery form)
    const uint64_t* ka[KSI_MAXP];
    uint64_t* o0;
    uint64_t* o1;
    const uint64_t* tw;     // paired forward twiddles of row i's modulus
    const uint64_t* twr;    // and their phase-B row table
    uint64_t q, qneg;
    uint64_t red;           // lazy_red_of(q): the transforms skip their reductions
    uint32_t np, xf;
    uint32_t pairs;         // every tower pointer 16-byte aligned: two positions per access
    // 1: the outputs only feed the rescale epilogue's lazy Shoup products
    // (any 64-bit input): REDC without its correction, no final reduction
    uint32_t lazy_out;
};

// V consecutive words from word i of a tower (V = 2: one 16-byte access; the
// address is 16-byte aligned), addressed as base + 32-bit byte offset (gat)
struct alignas(16) Words2 {
    uint64_t w[2];
};
template <uint32_t V>
GF_HD void load_words(const GF_GLOBAL uint64_t* base, uint32_t i, uint64_t* out) {
    if constexpr (V == 2) {
        const Words2 x = gat((const GF_GLOBAL Words2*)base, i / 2);
        out[0] = x.w[0];
        out[1] = x.w[1];
    } else {
        out[0] = gat(base, i);
    }
}
template <uint32_t V>
GF_HD void store_words(GF_GLOBAL uint64_t* base, uint32_t i, const uint64_t* in) {
    if constexpr (V == 2) {
        gat((GF_GLOBAL Words2*)base, i / 2) = Words2{{in[0], in[1]}};
    } else {
        gat(base, i) = in[0];
    }
}

// LZ = (j.red != 0).  The wave-level sync wsync separates the row passes of
// one wave's region; bsync is the workgroup barrier before the inner product.
template <uint32_t S2, uint32_t RM, bool LZ, class WSync, class BSync>
GF_HD void ksi_body(const KsiJob& j, uint32_t N1, uint32_t tile, uint64_t* lds, uint32_t wave, uint32_t nwaves,
                    uint32_t lane, uint32_t wlanes, uint32_t tid, uint32_t nthr, WSync wsync, BSync bsync) {
    constexpr uint32_t ROWS = KSI_TILE >> S2, LOGR = KSI_LOG_TILE - S2;
    const uint32_t a0 = tile * ROWS, off = a0 << S2;
    NttJob t{};
    t.tw = j.tw;
    t.twr = j.twr;
    t.q = j.q;
    t.red = j.red;
    const TileDesc d{LOGR, 1, padded(1u << S2), N1 + a0, 1, false, true};
    // transformed part k (in part order) -> LDS region k
    uint32_t k = 0;
    for (uint32_t p = 0; p < j.np; ++p) {
        if (!((j.xf >> p) & 1)) continue;
        if (k % nwaves == wave)
            row_tile<true, LZ, true, S2, RM, ROWS, 2, true, WSync, true>(
                t, d, gcast(j.src[p]), nullptr, off, lds + k * KSI_PART_WORDS, lane, wlanes, wsync);
        ++k;
    }
    bsync();
    // positions idx .. idx + V - 1 per step (V = 2: one 16-byte access per operand)
    auto inner = [&](auto Vc) {
        constexpr uint32_t V = decltype(Vc)::value;
        const uint64_t q = j.q, qneg = j.qneg;
        for (uint32_t idx = tid * V; idx < KSI_TILE; idx += nthr * V) {
            const uint32_t gi = off + idx;
            U128 s0[V], s1[V];
            for (uint32_t v = 0; v < V; ++v) s0[v] = s1[v] = U128{0, 0};
            for (uint32_t p = 0; p < j.np; ++p) {
                uint64_t dv[V], kb[V], ka[V];
                if ((j.xf >> p) & 1) {
                    const uint32_t k = __builtin_popcount(j.xf & ((1u << p) - 1));
                    for (uint32_t v = 0; v < V; ++v) {
                        dv[v] = lds[k * KSI_PART_WORDS + padded(idx + v)];
                        // LZ: below 72q (the sums stay below 2^125); else [0, 8q) -> [0, q)
                        if constexpr (!LZ) dv[v] = ntt_final<true>(dv[v], t);
                    }
                } else {
                    load_words<V>(gcast(j.src[p]), gi, dv);
                }
                load_words<V>(gcast(j.kb[p]), gi, kb);
                load_words<V>(gcast(j.ka[p]), gi, ka);
                for (uint32_t v = 0; v < V; ++v) {
                    s0[v] = add128(s0[v], mul128(dv[v], kb[v]));
                    s1[v] = add128(s1[v], mul128(dv[v], ka[v]));
                }
            }
            uint64_t r0[V], r1[V];
            for (uint32_t v = 0; v < V; ++v) {
                if (j.lazy_out) {
                    r0[v] = redc_lazy(s0[v], q, qneg);
                    r1[v] = redc_lazy(s1[v], q, qneg);
                } else if constexpr (LZ) {
                    r0[v] = redc_red(s0[v], q, qneg, (uint32_t)j.red);
                    r1[v] = redc_red(s1[v], q, qneg, (uint32_t)j.red);
                } else {
                    r0[v] = redc(s0[v], q, qneg);
                    r1[v] = redc(s1[v], q, qneg);
                }
            }
            store_words<V>(gcast(j.o0), gi, r0);
            store_words<V>(gcast(j.o1), gi, r1);
        }
    };
    if (j.pairs) inner(std::integral_constant<uint32_t, 2>{});
    else inner(std::integral_constant<uint32_t, 1>{});
}

}  // namespace gf
