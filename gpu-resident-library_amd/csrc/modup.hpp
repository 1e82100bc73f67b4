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
// TL = 2^(S1 - 2), u < 4).  The inverse transforms' last pass is the same
// chunk, so it leaves each thread's unit of a source in its registers, where
// the thread keeps it for the whole tile (reduced, times n^-1 QHatInv, split
// into the 30-bit halves the conversion sums take: 8 VGPRs per source); for
// each target it forms the conversion of its 4 elements and runs the target's
// first forward pass on them in registers, the later passes through LDS, the
// last to HBM.  The coefficient-form sources and digits never travel through
// HBM, and no pass of a source or target needs LDS at its register end.
//
// Each source and target (a phase) takes the next of two LDS tiles in turn,
// and likewise the next of two LDS twiddle tiles (the phase's column-pass
// twiddles, requested when the phase starts): a phase's first LDS write can
// only overlap the previous phase's reads of the other tile, and every phase
// of two or more passes has a barrier that all threads reach only after
// finishing the phase before it, so the phases need no barriers between
// them.  Every butterfly, product and reduction is the one the unfused
// kernels apply, so the stored digit column passes equal theirs bit for bit.
//
// Sources and targets whose modulus is below 2^50 run their column passes in
// double precision (ntt_f64.hpp) when the job gives their fp64 tables; the
// values stored for the digits' row pass are then other representatives of
// the same residues (in [0, 2q) instead of the integer pass's lazy range),
// and everything from the sources' reduced values to the conversion sums is
// identical.
#pragma once
#include <type_traits>

#include "ntt_engine.hpp"
#include "ntt_f64.hpp"

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

// per surviving tower constants of the fused ModDown + rescale prologue
struct RsTarget {
    uint64_t q, qneg, diff, new_larger, s2, s2p;
    uint64_t red;   // floor(2^64 / q): Shoup reduction x mod q = shoup(x, 1, red, q)
};
// The rescale prologue's value for one surviving tower t and one position:
// the centred switch of the dropped tower's value dv (hi: dv above q_old / 2)
// to q_t, times s2 (lazy Shoup, below 2q), plus the ModDown conversion sum acc
// after REDC without its correction (below 1.5q + 1 for at most 8 sources
// below 2^60): below 4q, which the forward transform after it takes.
GF_HD uint64_t rs_lazy_value(uint64_t dv, bool hi, const RsTarget& tg, const U128& acc) {
    uint64_t sw;
    if (tg.new_larger) {
        sw = dv + (hi ? tg.diff : 0);
    } else {
        const uint64_t d = hi ? tg.diff : 0;
        const uint64_t xm = shoup(dv, 1, tg.red, tg.q);   // dv mod q_t, exact for any dv < 2^64
        sw = xm >= d ? xm - d : xm + tg.q - d;
    }
    return shoup_lazy(sw, tg.s2, tg.s2p, tg.q) + redc_lazy(acc, tg.q, tg.qneg);
}

// The fused ModDown's prologue folds the switch term into the conversion's
// Montgomery sum: with x~ = dv - q_old [dv > q_old / 2] the centred dropped
// value, x~ s2 = dv s2 + [dv > q_old / 2] (-q_old s2) mod q_t, so dv and the
// 0/1 flag join the sources as two more terms with these constants
// (Montgomery form, R = 2^64): cs = s2 2^64, ch = -q_old s2 2^64 mod q.
inline void rs_sw_consts(uint64_t q, uint64_t q_old, uint64_t s2, uint64_t& cs, uint64_t& ch) {
    cs = (uint64_t)(((unsigned __int128)(s2 % q) << 64) % q);
    ch = (uint64_t)(((unsigned __int128)((q - q_old % q) % q) * cs) % q);
}

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
    // fp64 passes (moduli below 2^50).  Source i: its inverse pairs (TwF64
    // layout; null: the integer passes), q, fl(1/q), and n^-1 QHatInv centred
    // with its fl(./q).  Target t: ttwd[t] the address of its forward pairs
    // (0: the integer passes), tqf[2t], tqf[2t+1] the bit patterns of q and
    // fl(1/q).
    const uint64_t* stwd[MODUP_MAXSRC];
    double sqd[MODUP_MAXSRC], sqinv[MODUP_MAXSRC], spost[MODUP_MAXSRC], spostq[MODUP_MAXSRC];
    const uint64_t* ttwd;
    const uint64_t* tqf;
    // 1: tab holds the constants times 2^60 (not 2^64) and the conversion
    // reduces its column sums with redc60_lazy; the host sets it where the
    // part's source moduli sum to at most 1.2 2^60, so the targets stay
    // below 2.2 q_t (the forward passes take below 4q; the fp64 ones below
    // 1.25q after the shift by q).  Jobs of one launch share it.
    uint32_t c60;
    // Rescale mode (the fused ModDown + rescale prologue of one component):
    // the sources are the P towers (times n^-1 PHatInv), the targets the
    // surviving Q towers.  Per position the dropped tower
    //   dv = dt + (dr - REDC(sum_k x_k tabd[k strided])) pinv   (mod q_old)
    // then per target rs_lazy_value(dv, dv > half, tg[t], sum_k x_k tab[k sc + t]),
    // as the prologue kernel computes them (RsJob)
    const uint64_t* dt;
    const uint64_t* dr;
    const uint64_t* tabd;
    const RsTarget* tg;
    // [2 sc] per target: s2 2^64 and -q_old s2 2^64 mod q_t (rs_sw_consts)
    const uint64_t* tsw;
    uint64_t q_old, half, qd_neg, pinv, pinvp;
    uint32_t strided;
};

// Column access in fp64: the inverse transforms' row-pass output (lazy, below
// 4q) in, centred; the digits' column-pass output out, in [0, 2q)
struct ColInF64 {
    const GF_GLOBAL uint64_t* src;
    uint32_t logN2, b0;
    double q, qinv, c2;   // c2 = 2^52 + 2q
    GF_HD uint64_t raw(uint32_t s, uint32_t k, uint32_t) const { return gat(src, (k << logN2) + b0 + s); }
    GF_HD double conv(uint64_t w) const { return f64_center(f64_from_u52(w, c2), q, qinv); }
};
struct ColOutF64 {
    GF_GLOBAL uint64_t* dst;
    uint32_t logN2, b0;
    double q, qinv;
    GF_HD void operator()(uint32_t s, uint32_t k, uint32_t, double v) const {
        gat(dst, (k << logN2) + b0 + s) = u52_from_f64(f64_center(v, q, qinv) + q);
    }
};

// one thread's unit of every source: element u of source i (reduced, below
// 2^60) as its low and high 30 bits
template <uint32_t SP>
struct ModupUnit {
    uint32_t lo[MODUP_UNIT][SP], hi[MODUP_UNIT][SP];
};

// chunk 0 is one full radix-4 pass with one unit per thread
template <uint32_t S1>
GF_HD constexpr bool modup_chunk0_fits() {
    return chunk_size(S1, MODUP_RM, 0) == MODUP_RM &&
           ((1u << (MODUP_LOG_TILE - S1)) << (S1 - MODUP_RM)) == MODUP_THREADS;
}

// A target's column-pass twiddles: every column of a column pass takes the
// same ones, the first 2^S1 pairs (16 bytes each, fp64 or integer) of the
// target's forward table.  The kernel requests its share of them before the
// target's conversion and stores them into an LDS tile (tl) before the first
// pass; the passes behind the first block barrier read them from there.
GF_HD Words2 modup_tw_pair(const ModupJob& j, uint32_t t, uint32_t i) {
    const uint64_t twd = uniform_at(j.ttwd, t);
    const uint64_t* tab = twd ? (const uint64_t*)(uintptr_t)twd : (const uint64_t*)(uintptr_t)uniform_at(j.ttw, t);
    return gat((const GF_GLOBAL Words2*)gcast(tab), i);
}
GF_HD void modup_tw_put(uint64_t* tl, uint32_t i, const Words2& p) {
    tl[2 * i] = p.w[0];
    tl[2 * i + 1] = p.w[1];
}
// A column pass's first forward pass (chunk 0) takes twiddles 1, 2, 3 in
// every thread: read through the constant address space (scalar loads, no
// VGPRs, requested early)
struct TwUniform {
    static constexpr bool kRows = false;
    const uint64_t* t;
    GF_HD void get(uint32_t i, uint64_t& w, uint64_t& wp) const {
        w = uniform_at(t, 2 * i);
        wp = uniform_at(t, 2 * i + 1);
    }
};
struct TwUniformF64 {
    static constexpr bool kRows = false;
    const uint64_t* t;
    GF_HD void get(uint32_t i, double& w, double& wq) const {
        w = f64_of_bits(uniform_at(t, 2 * i));
        wq = f64_of_bits(uniform_at(t, 2 * i + 1));
    }
};
struct TwLds {
    static constexpr bool kRows = false;
    const uint64_t* t;
    GF_HD void get(uint32_t i, uint64_t& w, uint64_t& wp) const {
        w = t[2 * i];
        wp = t[2 * i + 1];
    }
};
struct TwLdsF64 {
    static constexpr bool kRows = false;
    const uint64_t* t;
    GF_HD void get(uint32_t i, double& w, double& wq) const {
        w = f64_of_bits(t[2 * i]);
        wq = f64_of_bits(t[2 * i + 1]);
    }
};

// Source i's column-pass twiddles: the first 2^S1 pairs of its inverse table
// (as modup_tw_pair for the targets)
GF_HD Words2 modup_src_tw_pair(const ModupJob& j, uint32_t i, uint32_t k) {
    const uint64_t* tab = j.stwd[i] ? j.stwd[i] : j.stw[i];
    return gat((const GF_GLOBAL Words2*)gcast(tab), k);
}

// source i's inverse column pass over the tile (all threads), its last pass
// (chunk 0) into the unit registers v (see UnitOut).  The first pass reads
// the table's own twiddles; stage_tw() then stores the source's twiddles
// into the LDS tile tl, which the passes behind the first barrier read.
template <uint32_t S1, bool ALL, class Sync, class StageTw>
GF_HD void modup_source_pass(const ModupJob& j, uint32_t i, const NttGeom& g, uint32_t tile, uint64_t* lds,
                             uint32_t tid, uint32_t nthr, Sync sync, uint64_t* v, const uint64_t* tl,
                             StageTw stage_tw) {
    static_assert(modup_chunk0_fits<S1>(), "chunk 0 must be a full radix-4 pass, one unit per thread");
    using P = PassOf<false, S1, MODUP_RM, 0>;
    static_assert(P::NC >= 2, "a column pass of two or more passes");
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC;
    const TileDesc d{LOGC, C, 1, 1, 0, true, false};
    if (j.stwd[i]) {
        const double q = j.sqd[i], qinv = j.sqinv[i];
        const ColInF64 gin{gcast(j.src[i]), g.logN2, tile * C, q, qinv, F64_2P52 + 2 * q};
        tile_pass_f64<false, S1, P::ST0, P::R, f64_center_after<false, S1, MODUP_RM>(0)>(
            d, TwF64{gcast((const double*)j.stwd[i])}, q, qinv, tid, nthr, gin, LdsOutF64{lds});
        stage_tw();
        sync();
        tile_chunks_f64<false, S1, MODUP_RM, 1>(lds, d, TwLdsF64{tl}, TwLdsF64{tl}, q, qinv, tid, nthr, sync, gin,
                                                UnitOutF64<ALL, MODUP_UNIT>{v});
        return;
    }
    const ColIn gin{gcast(j.src[i]), g.logN2, tile * C};
    tile_pass<false, false, S1, P::ST0, P::R>(d, TwPaired{gcast(j.stw[i])}, j.sq[i], tid, nthr, gin, LdsOut{lds});
    stage_tw();
    sync();
    tile_chunks<false, false, S1, MODUP_RM, 1>(lds, d, TwLds{tl}, TwLds{tl}, j.sq[i], tid, nthr, sync, gin,
                                               UnitOut<ALL, MODUP_UNIT>{v});
}
// source i's unit v (its inverse transform's last pass) into the
// conversion's registers
template <uint32_t SP, uint32_t S1>
GF_HD void modup_take_unit(const ModupJob& j, uint32_t i, const uint64_t* v, ModupUnit<SP>& x) {
    if (j.stwd[i]) {
        const double q = j.sqd[i], qinv = j.sqinv[i], w = j.spost[i], wq = j.spostq[i];
#pragma unroll
        for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
            // the last pass left |x| < 3q: times n^-1 QHatInv, into [0, q)
            const double y = f64_mulmod(f64_of_bits(v[u]), w, wq, q);
            const uint64_t v = u52_from_f64(f64_reduce(y, q, qinv));
            x.lo[u][i] = split30_lo(v);
            x.hi[u][i] = split30_hi(v);
        }
        return;
    }
    const uint64_t post = j.post[i], postp = j.postp[i], q = j.sq[i];
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
        // the inverse transform's final store: times n^-1 QHatInv, reduced
        const uint64_t r = shoup(v[u], post, postp, q);
        x.lo[u][i] = split30_lo(r);
        x.hi[u][i] = split30_hi(r);
    }
}
// target t's conversion of this thread's unit into cv: the exact column sums
// over the part's sources, REDC without its correction ([0, 2 q_t)), as the
// conversion kernel leaves them for the forward transforms (with c60 the
// R = 2^60 reduction of the columns, below 2.2 q_t); for an fp64 target as
// the double v - q_t (|.| <= 1.25 q_t)
template <uint32_t SP, uint32_t S1, bool C60>
GF_HD void modup_convert_unit(const ModupJob& j, uint32_t t, ModupUnit<SP>& x, uint64_t* cv) {
    const uint64_t q = uniform_at(j.tq, t), qn = uniform_at(j.tqneg, t);
    const bool f64 = uniform_at(j.ttwd, t) != 0;
    const double c2 = f64 ? F64_2P52 + f64_of_bits(uniform_at(j.tqf, 2 * t)) : 0.0;
    static_assert(!C60 || SP <= 6, "redc60_lazy takes at most 6 terms");
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
#pragma unroll
        for (uint32_t i = 0; i < SP; ++i) {
            // in place, so that each target's sums take the 32-bit halves as
            // they are (no 64-bit copies made once before the target loop)
            x.lo[u][i] = opaque32(x.lo[u][i]);
            x.hi[u][i] = opaque32(x.hi[u][i]);
        }
        auto cst = [&](uint32_t s) { return uniform_at(j.tab, s * j.sc + t); };
        uint64_t v;
        if constexpr (C60) v = redc60_lazy(dot30_cols<SP>(x.lo[u], x.hi[u], cst), q, qn);
        else v = redc_lazy(dot30<SP>(x.lo[u], x.hi[u], cst), q, qn);
        cv[u] = f64 ? bits_of_f64(f64_from_u52(v, c2)) : v;
    }
}
// Rescale mode: the dropped tower's value at this thread's 4 positions
// (column s, rows k0 + u TL of the tile), from the sources already in
// registers and dt, dr (coefficient form) read at those positions
template <uint32_t SP, uint32_t S1>
GF_HD void modup_drop_unit(const ModupJob& j, const NttGeom& g, uint32_t tile, uint32_t unit, ModupUnit<SP>& x,
                           uint64_t* dv) {
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC, TL = 1u << (S1 - MODUP_RM);
    const uint32_t s = unit & (C - 1), k0 = unit >> LOGC;
    const uint64_t q = j.q_old;
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
        const uint32_t gi = ((k0 + u * TL) << g.logN2) + tile * C + s;
        const U128 acc = dot30<SP>(x.lo[u], x.hi[u], [&](uint32_t k) { return uniform_at(j.tabd, k * j.strided); });
        const uint64_t cd = redc(acc, q, j.qd_neg);
        dv[u] = addmod(gat(gcast(j.dt), gi), shoup(submod(gat(gcast(j.dr), gi), cd, q), j.pinv, j.pinvp, q), q);
    }
}
// Rescale mode: target t's prologue value of this thread's unit into cv,
// congruent to rs_lazy_value's (below 1.44 q_t + 1; for an fp64 target as the
// centred double)
template <uint32_t SP, uint32_t S1>
GF_HD void modup_rs_unit(const ModupJob& j, uint32_t t, ModupUnit<SP>& x, const uint64_t* dv, uint64_t* cv) {
    static_assert(sizeof(RsTarget) == 7 * 8, "RsTarget layout");
    static_assert(SP + 2 <= 8, "the column sums take at most 8 terms");
    const uint64_t* w = (const uint64_t*)j.tg + 7 * t;
    const uint64_t q = uniform_at(w, 0), qneg = uniform_at(w, 1);
    const bool f64 = uniform_at(j.ttwd, t) != 0;
    const double qd = f64 ? f64_of_bits(uniform_at(j.tqf, 2 * t)) : 0.0;
    const double qinv = f64 ? f64_of_bits(uniform_at(j.tqf, 2 * t + 1)) : 0.0;
#pragma unroll
    for (uint32_t u = 0; u < MODUP_UNIT; ++u) {
#pragma unroll
        for (uint32_t i = 0; i < SP; ++i) {
            x.lo[u][i] = opaque32(x.lo[u][i]);
            x.hi[u][i] = opaque32(x.hi[u][i]);
        }
        // the sources, dv (below q_old < 2^60) and the centring flag as one
        // sum of SP + 2 terms: REDC leaves it below 1.44 q + 1
        uint32_t al[SP + 2], ah[SP + 2];
#pragma unroll
        for (uint32_t i = 0; i < SP; ++i) {
            al[i] = x.lo[u][i];
            ah[i] = x.hi[u][i];
        }
        al[SP] = split30_lo(dv[u]);
        ah[SP] = split30_hi(dv[u]);
        al[SP + 1] = dv[u] > j.half ? 1u : 0u;
        ah[SP + 1] = 0;
        const U128 acc = dot30<SP + 2>(al, ah, [&](uint32_t s) {
            return s < SP ? uniform_at(j.tab, s * j.sc + t) : uniform_at(j.tsw, 2 * t + (s - SP));
        });
        const uint64_t v = redc_lazy(acc, q, qneg);
        cv[u] = f64 ? bits_of_f64(f64_center(f64_from_u52(v, F64_2P52 + 2 * qd), qd, qinv)) : v;
    }
}

// target t's forward column pass from the unit registers cv (see UnitIn),
// stored to HBM unreduced (the digits' row pass continues from these
// values): the first pass with the table's own twiddles, the others with
// their copy in tl (visible after the first pass's barrier)
template <uint32_t S1, bool ALL, class Sync>
GF_HD void modup_target_pass(const ModupJob& j, uint32_t t, const NttGeom& g, uint32_t tile, uint64_t* lds,
                             uint32_t tid, uint32_t nthr, Sync sync, const uint64_t* cv, const uint64_t* tl) {
    static_assert(modup_chunk0_fits<S1>(), "chunk 0 must be a full radix-4 pass, one unit per thread");
    using P = PassOf<true, S1, MODUP_RM, 0>;
    static_assert(P::NC >= 2, "a column pass of two or more passes");
    constexpr uint32_t LOGC = MODUP_LOG_TILE - S1, C = 1u << LOGC;
    const TileDesc d{LOGC, C, 1, 1, 0, true, false};
    const uint64_t twd = uniform_at(j.ttwd, t);
    if (twd) {
        const double q = f64_of_bits(uniform_at(j.tqf, 2 * t)), qinv = f64_of_bits(uniform_at(j.tqf, 2 * t + 1));
        const ColOutF64 gout{gcast(j.dst + (size_t)t * g.n), g.logN2, tile * C, q, qinv};
        const UnitInF64<ALL, MODUP_UNIT> in{cv};
        tile_pass_f64<true, S1, P::ST0, P::R, f64_center_after<true, S1, MODUP_RM>(0)>(
            d, TwUniformF64{(const uint64_t*)(uintptr_t)twd}, q, qinv, tid, nthr, in, LdsOutF64{lds});
        sync();
        tile_chunks_f64<true, S1, MODUP_RM, 1>(lds, d, TwLdsF64{tl}, TwLdsF64{tl}, q, qinv, tid, nthr, sync, in, gout);
        return;
    }
    NttJob jt{};
    jt.q = uniform_at(j.tq, t);
    const ColOut<true> gout{gcast(j.dst + (size_t)t * g.n), g.logN2, tile * C, jt};
    const TwUniform tw{(const uint64_t*)(uintptr_t)uniform_at(j.ttw, t)};
    const UnitIn<ALL, MODUP_UNIT> in{cv};
    auto run = [&](auto LZc) {
        constexpr bool LZ = decltype(LZc)::value;
        tile_pass<true, LZ, S1, P::ST0, P::R>(d, tw, jt.q, tid, nthr, in, LdsOut{lds});
        sync();
        tile_chunks<true, LZ, S1, MODUP_RM, 1>(lds, d, TwLds{tl}, TwLds{tl}, jt.q, tid, nthr, sync, in, gout);
    };
    if (uniform_at(j.tred, t)) run(std::true_type{});
    else run(std::false_type{});
}

}  // namespace gf
