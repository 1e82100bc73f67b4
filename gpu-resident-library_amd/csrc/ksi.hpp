// Fused last pass of the ModUp digits' forward transform and the key inner
// product (the inner product of src/keyswitch_resident.cpp:171-200 over
// digits produced by the forward NTT of :184-194).
//
// One workgroup owns KSI_TILE consecutive positions of tower row i (ROWS
// transform rows of N2) for every part p of the decomposition.  Wave w runs
// the forward row pass (phase B) of the transformed parts w, w + nwaves, ...
// (in part order) on LDS region w, w + nwaves, ..., reading the phase-A output
// straight from HBM, and keeps the evaluation-form digits there; after one
// block barrier every thread (including an optional extra wave that runs no
// row pass) takes KSI_TILE / nthr positions and accumulates
// sum_p digit_p * key_p for both key components (30-bit column sums where
// every value is below 2^60 and there are at most 4 parts, else 128-bit sums).  A launch
// gives its jobs one wave per transformed part (up to KSI_WAVES), so no wave
// idles through the row passes.  The transformed digits therefore never
// travel through HBM.  Parts whose tower i is the input's own (no transform)
// are read from HBM at the inner product.
#pragma once
#include "ntt_engine.hpp"
#include "ntt_f64.hpp"

namespace gf {

enum {
    KSI_MAXP = 7,                                    // parts per job (KSI_MAXP regions fit 64 KiB of LDS)
    KSI_LOG_TILE = 10,
    KSI_TILE = 1 << KSI_LOG_TILE,                    // positions per workgroup
    KSI_PART_WORDS = KSI_TILE + KSI_TILE / 16,       // one part's padded LDS region
    KSI_WAVES = 5,                                   // largest workgroup, in waves
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
    // 1: every digit and key value of the row is below 2^60 (ksi_d30), so the
    // inner product runs as 30-bit column sums
    uint32_t d30;
    // digits from the fp64 ModUp (below 2q) of a modulus below 2^50: the row
    // passes in fp64 (ntt_f64.hpp) with these pairs and their row table, q and
    // fl(1/q); null: the integer passes
    const uint64_t* twd;
    const uint64_t* twrd;
    double fq, fqinv;
    // rp_on: the row's two outputs are not stored; instead the inverse
    // transforms' row pass rp[c] (its input tower o0 / o1) runs on them after
    // the inner product, from LDS (at least two transformed parts, so two LDS
    // regions)
    uint32_t rp_on;
    NttJob rp[2];
};

// V consecutive words from word i of a tower (V = 2: one 16-byte access; the
// address is 16-byte aligned), addressed as base + 32-bit byte offset (gat)
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

// Digits and keys of a row fit the 30-bit column sums when both are below
// 2^60: keys are residues (< q); digits are residues too, or with red (the
// forward transform's unreduced values) below 72 q.
inline uint32_t ksi_d30(uint64_t q, uint64_t red) {
    const unsigned __int128 dmax = red ? (unsigned __int128)72 * q : (unsigned __int128)q;
    return dmax <= ((unsigned __int128)1 << 60) ? 1u : 0u;
}

// The inner product's outputs at positions idx .. idx + V - 1 (word gi of the
// row): to HBM, or (j.rp_on) into LDS regions 0 and 1 at those positions'
// slots for the row pass after it; each thread writes only its own
// positions, whose digits it has already read
template <uint32_t V>
GF_HD void ksi_store(const KsiJob& j, uint64_t* lds, uint32_t idx, uint32_t gi, const uint64_t* r0,
                     const uint64_t* r1) {
    if (j.rp_on) {
#pragma unroll
        for (uint32_t v = 0; v < V; ++v) {
            lds[padded(idx + v)] = r0[v];
            lds[KSI_PART_WORDS + padded(idx + v)] = r1[v];
        }
        return;
    }
    store_words<V>(gcast(j.o0), gi, r0);
    store_words<V>(gcast(j.o1), gi, r1);
}

// hi:lo of lo + mid 2^30 + hi 2^60 (the column sums of dot30's layout)
GF_HD U128 cols_to_u128(uint64_t lo, uint64_t mid, uint64_t hi) {
    U128 T;
    T.lo = lo + (mid << 30);
    uint64_t c = T.lo < lo ? 1u : 0u;
    const uint64_t h = hi << 60;
    T.lo += h;
    c += T.lo < h ? 1u : 0u;
    T.hi = (mid >> 34) + (hi >> 4) + c;
    return T;
}

// The inner product of NP parts (NP known at compile time): positions
// idx .. idx + V - 1 per step; all operands of a step (each part's digit, from
// LDS or HBM, and both key rows) are loaded before any arithmetic, and the
// sums are 30-bit column sums (j.d30): per part and component four 32 x 32
// multiply-adds with no carries.  NP <= 7 keeps every column below 2^64
// (lo, hi < 7 2^60; mid < 14 2^60).
template <uint32_t NP, uint32_t V, bool LZ>
GF_HD void ksi_inner_cols(const KsiJob& j, uint64_t* lds, uint32_t off, uint32_t tid, uint32_t nthr,
                          const NttJob& t) {
    const GF_GLOBAL uint64_t *src[NP], *kb[NP], *ka[NP];
    uint32_t reg[NP];
    bool xf[NP];
#pragma unroll
    for (uint32_t p = 0; p < NP; ++p) {
        src[p] = gcast(j.src[p]);
        kb[p] = gcast(j.kb[p]);
        ka[p] = gcast(j.ka[p]);
        xf[p] = (j.xf >> p) & 1;
        reg[p] = (uint32_t)__builtin_popcount(j.xf & ((1u << p) - 1)) * KSI_PART_WORDS;
    }
    const uint64_t q = j.q, qneg = j.qneg;
    for (uint32_t idx = tid * V; idx < KSI_TILE; idx += nthr * V) {
        const uint32_t gi = off + idx;
        uint64_t dv[NP][V], bv[NP][V], av[NP][V];
#pragma unroll
        for (uint32_t p = 0; p < NP; ++p) {
            if (xf[p]) {
#pragma unroll
                for (uint32_t v = 0; v < V; ++v) dv[p][v] = lds[reg[p] + padded(idx + v)];
            } else {
                load_words<V>(src[p], gi, dv[p]);
            }
            load_words<V>(kb[p], gi, bv[p]);
            load_words<V>(ka[p], gi, av[p]);
        }
        uint64_t r0[V], r1[V];
#pragma unroll
        for (uint32_t v = 0; v < V; ++v) {
            uint64_t l0 = 0, m0 = 0, h0 = 0, l1 = 0, m1 = 0, h1 = 0;
#pragma unroll
            for (uint32_t p = 0; p < NP; ++p) {
                uint64_t d = dv[p][v];
                // LZ: below 72q < 2^60 as is (j.d30); else [0, 8q) -> [0, q)
                if (!LZ && xf[p]) d = ntt_final<true>(d, t);
                const uint32_t d0 = split30_lo(d), d1 = split30_hi(d);
                const uint64_t b = bv[p][v], a = av[p][v];
                const uint32_t b0 = split30_lo(b), b1 = split30_hi(b), a0 = split30_lo(a), a1 = split30_hi(a);
                l0 = mad32(d0, b0, l0);
                m0 = mad32(d0, b1, m0);
                m0 = mad32(d1, b0, m0);
                h0 = mad32(d1, b1, h0);
                l1 = mad32(d0, a0, l1);
                m1 = mad32(d0, a1, m1);
                m1 = mad32(d1, a0, m1);
                h1 = mad32(d1, a1, h1);
            }
            const U128 s0 = cols_to_u128(l0, m0, h0), s1 = cols_to_u128(l1, m1, h1);
            if (j.lazy_out) {
                r0[v] = redc_lazy(s0, q, qneg);
                r1[v] = redc_lazy(s1, q, qneg);
            } else if constexpr (LZ) {
                r0[v] = redc_red(s0, q, qneg, (uint32_t)j.red);
                r1[v] = redc_red(s1, q, qneg, (uint32_t)j.red);
            } else {
                r0[v] = redc(s0, q, qneg);
                r1[v] = redc(s1, q, qneg);
            }
        }
        ksi_store<V>(j, lds, idx, gi, r0, r1);
    }
}

// ksi_inner_cols for the job's part count, when it has an instantiation
template <uint32_t V, bool LZ>
GF_HD bool ksi_inner_cols_np(const KsiJob& j, uint64_t* lds, uint32_t off, uint32_t tid, uint32_t nthr,
                             const NttJob& t) {
    switch (j.np) {
        case 1: ksi_inner_cols<1, V, LZ>(j, lds, off, tid, nthr, t); return true;
        case 2: ksi_inner_cols<2, V, LZ>(j, lds, off, tid, nthr, t); return true;
        case 3: ksi_inner_cols<3, V, LZ>(j, lds, off, tid, nthr, t); return true;
        case 4: ksi_inner_cols<4, V, LZ>(j, lds, off, tid, nthr, t); return true;
        default: return false;
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
        if (k % nwaves == wave) {
            if (j.twd)
                row_tile_f64<true, S2, RM, ROWS, true>(j.twd, j.twrd, j.fq, j.fqinv, d, gcast(j.src[p]), nullptr, off,
                                                       lds + k * KSI_PART_WORDS, lane, wlanes, wsync);
            else
                row_tile<true, LZ, true, S2, RM, ROWS, 2, true, WSync, true>(
                    t, d, gcast(j.src[p]), nullptr, off, lds + k * KSI_PART_WORDS, lane, wlanes, wsync);
        }
        ++k;
    }
    bsync();
    // up to 4 parts with every value below 2^60: column sums, loads first
    bool done = false;
    if (j.d30)
        done = j.pairs ? ksi_inner_cols_np<2, LZ>(j, lds, off, tid, nthr, t)
                       : ksi_inner_cols_np<1, LZ>(j, lds, off, tid, nthr, t);
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
            ksi_store<V>(j, lds, idx, gi, r0, r1);
        }
    };
    if (!done) {
        if (j.pairs) inner(std::integral_constant<uint32_t, 2>{});
        else inner(std::integral_constant<uint32_t, 1>{});
    }
    // the inverse transforms' row pass of the two output rows (inputs in
    // [0, q), lazy outputs for the column pass after it), component c by
    // wave c
    if (j.rp_on) {
        bsync();
        for (uint32_t c = wave; c < 2; c += nwaves) {
            const NttJob& r = j.rp[c];
            row_tile<false, false, true, S2, RM, ROWS, 0, true, WSync, false, true>(
                r, d, nullptr, gcast(r.out), off, lds + c * KSI_PART_WORDS, lane, wlanes, wsync);
        }
    }
}

}  // namespace gf
