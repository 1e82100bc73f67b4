// Tile engine for the negacyclic NTT used by every transform in the library.
//
// Convention (the reference's kernels/ntt_ct.cu:19-112 and kernels/intt_ct.cu):
// forward = Cooley-Tukey, standard order in, bit-reversed order out, twiddle of
// stage m, group i = tw[m + i] with tw[k] = psi^brv(k); inverse = Gentleman-
// Sande with psi^-brv(k), followed by a multiply by n^-1 (here folded into the
// last store together with any caller constant).  Outputs are in [0, q).
//
// Decomposition: n = N1 * N2.  The first log2(N1) CT stages only couple
// elements with the same index mod N2 ("columns"); the last log2(N2) stages
// only couple elements inside one contiguous row of N2.  Phase A transforms a
// tile of C columns, phase B a tile of R rows; each phase is one global read
// and one global write of the tower.  Inside a phase the tile is processed in
// radix-2^r register passes (r <= 4): the first pass loads from HBM, the last
// stores to HBM, and LDS carries the exchange between passes.  For n <= 4096
// the whole tower is one tile (single phase).
//
// Butterflies are lazy (Harvey), with the 4q-bounded Shoup product of
// modarith.hpp: forward values stay in [0, 8q), inverse values in [0, 4q),
// reduced once at the final store (needs 8q < 2^64, i.e. q < 2^61).  Forward
// transforms over a modulus 2^32 < q < 2^57 (LZ) drop the forward butterfly's
// conditional subtraction altogether: each stage adds at most 4q to the bound,
// so inputs below 4q stay below (4 + 4 logn) q <= 72q < 2^64 and one Barrett-style reduction
// (reduce_red) at the final store replaces the per-stage ones.
//
// Twiddle index for CT stage st (local to the phase), sub-transform s, group i:
//     (G(s) << st) + i,   G(s) = g0 + s * gstep
// phase A / single phase: G = 1 (the table prefix, stage m = 2^st);
// phase B, row a:         G = N1 + a   (global stage m = N1 2^st).
#pragma once
#include <type_traits>

#include "modarith.hpp"

namespace gf {

// Device code reads towers and twiddle tables through global-address-space
// pointers: global loads take a scalar base plus a 32-bit offset and are
// counted apart from LDS traffic (generic pointers would compile to flat
// accesses that do neither).  The host build of this header (the CPU
// emulation) uses plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define GF_GLOBAL __attribute__((address_space(1)))
#else
#define GF_GLOBAL
#endif
template <class T>
GF_HD const GF_GLOBAL T* gcast(const T* p) { return (const GF_GLOBAL T*)p; }
template <class T>
GF_HD GF_GLOBAL T* gcast(T* p) { return (GF_GLOBAL T*)p; }
// Element i of a tower-sized array as a base plus a 32-bit byte offset (a
// tower is at most 1 MiB): with a workgroup-uniform base, loads and stores take
// the scalar-base form with one 32-bit VGPR offset, not a 64-bit VGPR address
// built per element.
template <class T>
GF_HD T& gat(GF_GLOBAL T* base, uint32_t i) {
    return *(GF_GLOBAL T*)((GF_GLOBAL char*)base + i * (uint32_t)sizeof(T));
}
template <class T>
GF_HD const T& gat(const GF_GLOBAL T* base, uint32_t i) {
    return *(const GF_GLOBAL T*)((const GF_GLOBAL char*)base + i * (uint32_t)sizeof(T));
}

struct NttJob {
    const uint64_t* in;   // tower read by the first phase
    uint64_t* out;        // tower written by every phase (may equal in)
    const uint64_t* tw;   // twiddles: paired table (w, Shoup precon) or w alone
    const uint64_t* twp;  // Shoup precons when tw is not paired, else unused
    uint64_t q;
    uint64_t post, postp; // inverse only: final multiplier (n^-1 * c) and precon
    // forward only, optional epilogue: out = NTT(in) + e0*c0 + e1*c1 (mod q)
    const uint64_t* e0;
    const uint64_t* e1;
    uint64_t c0, c0p, c1, c1p;
    // paired tables only: the twiddles of phase B's run-unit pass in lane
    // order (build_row_table); required by two-phase transforms of paired jobs
    const uint64_t* twr;
    // forward only: floor(2^64 / q) for a modulus 2^32 < q < 2^57 (lazy_red_of),
    // else 0.  Such a modulus leaves room in a 64-bit word for the whole
    // transform's growth, so the butterflies skip their conditional subtraction
    // (inputs below 4q stay below (4 + 4 logn) q <= 72 q) and the final store reduces once.
    uint64_t red;
    // with red: store those unreduced values, for a consumer that reduces them
    // itself (the key inner product); ignored with an epilogue
    uint32_t lazy_out;
    // inverse phase B of a modulus below 2^50 with inputs below 4q: its fp64
    // pairs and their row table (ntt_f64.hpp; null: the integer passes), q
    // and fl(1/q).  Its output is then in [0, 2q).
    const uint64_t* twd;
    const uint64_t* twrd;
    double fq, fqinv;
};

// floor(2^64 / q) for the moduli whose forward transforms run without
// per-butterfly reductions, else 0 (NttJob::red)
inline uint64_t lazy_red_of(uint64_t q) {
    if (q <= (1ull << 32) || q >= (1ull << 57)) return 0;
    return (uint64_t)(((unsigned __int128)1 << 64) / q);
}

struct NttGeom {
    uint32_t n, logn;
    uint32_t logN1, logN2;  // phase A covers logN1 stages, phase B logN2
    uint32_t C;             // columns per phase-A tile
    uint32_t R;             // rows per phase-B tile
    uint32_t single;        // 1: whole tower in one tile
    uint32_t rmax;          // largest register radix log2 (tuning knob, 2..4)
};

enum { NTT_LOG_TILE = 12, NTT_TILE = 1 << NTT_LOG_TILE, NTT_SINGLE_MAX = 4096, NTT_THREADS = 256 };

// row-tile padding: one spare word every 16, so strided radix units hit distinct LDS banks
GF_HD constexpr uint32_t padded(uint32_t k) { return k + (k >> 4); }
enum { NTT_TILE_PADDED = NTT_TILE + NTT_TILE / 16 };

inline NttGeom make_geom(uint32_t n, uint32_t rmax = 3) {
    NttGeom g{};
    g.n = n;
    g.rmax = rmax;
    uint32_t l = 0;
    while ((1u << l) < n) ++l;
    g.logn = l;
    if (n <= NTT_SINGLE_MAX) {
        g.single = 1;
        g.logN1 = l;
        g.logN2 = 0;
        g.C = 1;
        g.R = 1;
        return g;
    }
    g.single = 0;
    g.logN2 = (l + 1) / 2;
    if (g.logN2 > 8) g.logN2 = 8;
    g.logN1 = l - g.logN2;
    g.C = NTT_TILE >> g.logN1;
    g.R = NTT_TILE >> g.logN2;
    return g;
}

// The device scheduler may not move instructions across this point (a
// scheduling hint only).  Put between a batch of global loads and the work on
// them: left to itself, the scheduler interleaves each load with its use
// under register pressure and waits for the last loads one at a time, a
// memory round trip each.
#if defined(__HIP_DEVICE_COMPILE__)
#define GF_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define GF_SCHED_FENCE() ((void)0)
#endif

// two consecutive words as one 16-byte access
struct alignas(16) Words2 {
    uint64_t w[2];
};

// twiddle accessors: paired (w, wp) table -> one 16-byte load; or two arrays
struct TwPaired {
    static constexpr bool kRows = false;
    const GF_GLOBAL uint64_t* t;
    GF_HD void get(uint32_t i, uint64_t& w, uint64_t& wp) const {
        w = t[2 * i];
        wp = t[2 * i + 1];
    }
};
struct TwSplit {
    static constexpr bool kRows = false;
    const GF_GLOBAL uint64_t* w;
    const GF_GLOBAL uint64_t* p;
    GF_HD void get(uint32_t i, uint64_t& ww, uint64_t& wp) const {
        ww = w[i];
        wp = p[i];
    }
};
template <bool PAIRED>
GF_HD auto tw_of(const NttJob& j) {
    if constexpr (PAIRED) return TwPaired{gcast(j.tw)};
    else return TwSplit{gcast(j.tw), gcast(j.twp)};
}

// radix-2^r chunking of S stages: ceil(S/RM) chunks, the larger ones first
GF_HD constexpr uint32_t num_chunks(uint32_t S, uint32_t RM) { return (S + RM - 1) / RM; }
GF_HD constexpr uint32_t chunk_size(uint32_t S, uint32_t RM, uint32_t c) {
    return S / num_chunks(S, RM) + (c < S % num_chunks(S, RM) ? 1u : 0u);
}
GF_HD constexpr uint32_t chunk_start(uint32_t S, uint32_t RM, uint32_t c) {
    return c * (S / num_chunks(S, RM)) + (c < S % num_chunks(S, RM) ? c : S % num_chunks(S, RM));
}

// Phase B's last forward pass (its first inverse pass) is the last chunk,
// stages ST0 .. S2-1 (R = S2 - ST0 of them); its units are runs of 2^R
// consecutive elements, 2^(S2-R) units per row, and at local stage c the
// unit g of row a uses twiddle index ((N1 + a) << (ST0 + c)) + (g << c) + v,
// v < 2^c.  Every lane reads a different twiddle there, so the row table
// stores them in lane order,
//     entry(a, c, v, g) = a PR + ((2^c - 1) << LOGU) + (v << LOGU) + g,
//     LOGU = S2 - R,  PR = (2^R - 1) << LOGU,
// and the 2^LOGU lanes of one row read 2^LOGU consecutive 16-byte entries
// per load.
GF_HD constexpr uint32_t row_pass_r(uint32_t S2, uint32_t RM) { return chunk_size(S2, RM, num_chunks(S2, RM) - 1); }
inline size_t row_table_words(const NttGeom& g) {
    if (g.single) return 0;
    const uint32_t R = row_pass_r(g.logN2, g.rmax), LOGU = g.logN2 - R;
    return (size_t)2 * ((size_t)1 << g.logN1) * (((size_t)1 << R) - 1) * ((size_t)1 << LOGU);
}
// std_paired: (w, wp) pairs of the standard table tw[k]; out: row_table_words(g)
inline void build_row_table(const uint64_t* std_paired, const NttGeom& g, uint64_t* out) {
    if (g.single) return;
    const uint32_t S2 = g.logN2, R = row_pass_r(S2, g.rmax), ST0 = S2 - R, LOGU = S2 - R;
    const uint32_t N1 = 1u << g.logN1, PR = ((1u << R) - 1) << LOGU;
    for (uint32_t a = 0; a < N1; ++a)
        for (uint32_t c = 0; c < R; ++c)
            for (uint32_t v = 0; v < (1u << c); ++v)
                for (uint32_t u = 0; u < (1u << LOGU); ++u) {
                    const size_t e = (size_t)a * PR + (((1u << c) - 1) << LOGU) + ((size_t)v << LOGU) + u;
                    const size_t k = ((size_t)(N1 + a) << (ST0 + c)) + ((size_t)u << c) + v;
                    out[2 * e] = std_paired[2 * k];
                    out[2 * e + 1] = std_paired[2 * k + 1];
                }
}
// the row table of one phase-B tile (rows a0 ..): entry(s, c, v, g) relative to row a0
template <uint32_t R, uint32_t LOGU>
struct TwRows {
    static constexpr bool kRows = true;
    static constexpr uint32_t PR = ((1u << R) - 1) << LOGU;
    const GF_GLOBAL uint64_t* t;
    GF_HD void get_rp(uint32_t c, uint32_t s, uint32_t g, uint32_t v, uint64_t& w, uint64_t& wp) const {
        const uint32_t e = s * PR + (((1u << c) - 1) << LOGU) + (v << LOGU) + g;
        w = t[2 * e];
        wp = t[2 * e + 1];
    }
};

// Runtime description of how a tile is laid out and which twiddles it uses.
// Element k of sub-transform s lives at a[k*KS + s*SS]; sub-transform s uses
// twiddle group base G(s) = g0 + s*gstep.
struct TileDesc {
    uint32_t logT, KS, SS;
    uint32_t g0;   // twiddle indices stay below n <= 2^17
    uint32_t gstep;
    bool sub_fastest;   // unit order: consecutive threads on consecutive sub-transforms
    bool pad;           // element k of a sub-transform at padded(k) (KS must be 1)
};

// Element access of a register pass.  A pass names each element by its
// sub-transform s, its index k inside the sub-transform and its LDS slot `at`;
// the access policy decides where the value comes from or goes to.
struct LdsIn {
    const uint64_t* a;
    GF_HD uint64_t operator()(uint32_t, uint32_t, uint32_t at) const { return a[at]; }
};
struct LdsOut {
    uint64_t* a;
    GF_HD void operator()(uint32_t, uint32_t, uint32_t at, uint64_t v) const { a[at] = v; }
};
// the LDS accessors of the passes between a phase's first and last
struct LdsPlain {
    GF_HD LdsIn in(const uint64_t* a) const { return LdsIn{a}; }
    GF_HD LdsOut out(uint64_t* a) const { return LdsOut{a}; }
};

// Accessors that name an element by its unit instead (kUnit): element u of
// unit uid, for a pass whose one end is the values a thread holds in
// registers.  pass_in / pass_out pick the form the accessor takes.
template <class A, class = void>
struct IsUnitAcc : std::false_type {};
template <class A>
struct IsUnitAcc<A, std::void_t<decltype(A::kUnit)>> : std::true_type {};
template <class IN>
GF_HD auto pass_in(const IN& in, uint32_t uid, uint32_t u, uint32_t s, uint32_t k, uint32_t at) {
    if constexpr (IsUnitAcc<IN>::value) return in.unit(uid, u);
    else return in(s, k, at);
}
template <class OUT, class V>
GF_HD void pass_out(const OUT& out, uint32_t uid, uint32_t u, uint32_t s, uint32_t k, uint32_t at, V v) {
    if constexpr (IsUnitAcc<OUT>::value) out.unit(uid, u, v);
    else out(s, k, at, v);
}
// A pass's unit of U elements in registers: element u at v[u].  ALL = false:
// the thread's own unit (a pass with one unit per thread, so uid is the
// thread's and the index is the unrolled constant u); ALL = true: every unit
// of the tile at v + U uid, for a host emulation that runs the tile's units
// in one thread.
template <bool ALL, uint32_t U = 4>
struct UnitIn {
    static constexpr bool kUnit = true;
    const uint64_t* v;
    GF_HD uint64_t unit(uint32_t uid, uint32_t u) const { return v[(ALL ? uid * U : 0) + u]; }
};
template <bool ALL, uint32_t U = 4>
struct UnitOut {
    static constexpr bool kUnit = true;
    uint64_t* v;
    GF_HD void unit(uint32_t uid, uint32_t u, uint64_t x) const { v[(ALL ? uid * U : 0) + u] = x; }
};

// One radix-2^R pass (stages ST0 .. ST0+R-1 of an S-stage transform) over all
// units of the tile owned by this thread; fully unrolled, the R-stage unit
// stays in registers.  `in` supplies the unit's values, `out` takes them.
template <bool FWD, bool LZ, uint32_t S, uint32_t ST0, uint32_t R, class TW, class IN, class OUT>
GF_HD void tile_pass(const TileDesc& d, const TW& tw, uint64_t q, uint32_t tid, uint32_t nthr,
                     const IN& in, const OUT& out) {
    static_assert(FWD || !LZ, "the lazy-growth butterflies are forward only");
    constexpr uint32_t LOGTL = S - ST0 - R;   // log2(t_last)
    constexpr uint32_t TL = 1u << LOGTL;
    constexpr uint32_t LOGPS = S - R;         // log2(units per sub-transform)
    constexpr uint32_t R8 = 1u << R;
    const uint32_t T = 1u << d.logT;
    const uint32_t U = T << LOGPS;
    const uint64_t q4 = 4 * q, nq = 0 - q;
    for (uint32_t uid = tid; uid < U; uid += nthr) {
        uint32_t s, rest;
        if (d.sub_fastest) {
            s = uid & (T - 1);
            rest = uid >> d.logT;
        } else {
            rest = uid & ((1u << LOGPS) - 1);
            s = uid >> LOGPS;
        }
        const uint32_t l = rest & (TL - 1);
        const uint32_t g = rest >> LOGTL;
        const uint32_t k0 = (g << (S - ST0)) + l;
        const uint32_t sbase = s * d.SS;
        uint32_t at[R8];
#pragma unroll
        for (uint32_t u = 0; u < R8; ++u)
            at[u] = sbase + (d.pad ? padded(k0 + u * TL) : (k0 + u * TL) * d.KS);
        uint64_t x[R8];
#pragma unroll
        for (uint32_t u = 0; u < R8; ++u) x[u] = pass_in(in, uid, u, s, k0 + u * TL, at[u]);
        const uint32_t G = d.g0 + s * d.gstep;
        if (FWD) {
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) {
                const uint32_t h = 1u << (R - 1 - c);
                const uint32_t tb = (G << (ST0 + c)) + (g << c);
#pragma unroll
                for (uint32_t u = 0; u < R8; ++u) {
                    if (u & h) continue;
                    uint64_t w, wp;
                    if constexpr (TW::kRows) tw.get_rp(c, s, g, u >> (R - c), w, wp);
                    else tw.get(tb + (u >> (R - c)), w, wp);
                    // xv in [0, 4q) after the csub: outputs in [0, 8q).
                    // LZ: xv below the stage's bound B, outputs below B + 4q.
                    const uint64_t xv = LZ ? x[u] : csub(x[u], q4);
                    const uint64_t t = shoup_lazy4(x[u + h], w, wp, nq);  // [0, 4q)
                    x[u] = xv + t;
                    x[u + h] = xv - t + q4;
                }
            }
        } else {
#pragma unroll
            for (int c = (int)R - 1; c >= 0; --c) {
                const uint32_t h = 1u << (R - 1 - c);
                const uint32_t tb = (G << (ST0 + c)) + (g << c);
#pragma unroll
                for (uint32_t u = 0; u < R8; ++u) {
                    if (u & h) continue;
                    uint64_t w, wp;
                    if constexpr (TW::kRows) tw.get_rp((uint32_t)c, s, g, u >> (R - c), w, wp);
                    else tw.get(tb + (u >> (R - c)), w, wp);
                    const uint64_t xv = x[u], yv = x[u + h];           // [0, 4q)
                    x[u] = csub(xv + yv, q4);                          // [0, 4q)
                    x[u + h] = shoup_lazy4(xv - yv + q4, w, wp, nq);   // [0, 4q)
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < R8; ++u) pass_out(out, uid, u, s, k0 + u * TL, at[u], x[u]);
    }
}

// Pass I (in execution order) of an S-stage phase: its chunk, and whether its
// units are runs of consecutive elements with stride TL >= 8, so that a wave
// touching element k of 8 consecutive units reads 64 contiguous bytes.
template <bool FWD, uint32_t S, uint32_t RM, uint32_t I>
struct PassOf {
    static constexpr uint32_t NC = num_chunks(S, RM);
    static constexpr uint32_t C = FWD ? I : NC - 1 - I;
    static constexpr uint32_t ST0 = chunk_start(S, RM, C), R = chunk_size(S, RM, C);
    static constexpr bool wide_runs = S - ST0 - R >= 3;
};

// All passes of one S-stage phase, chunk order fixed at compile time.  The
// first pass reads through `gin` and the last writes through `gout` (global
// memory, or LDS when the caller stages that end); the passes in between go
// through LDS (accessors from lio) with a barrier after each.  A one-pass
// phase never touches LDS.
template <bool FWD, bool LZ, uint32_t S, uint32_t RM, uint32_t I, class TW, class TWL, class Sync, class GIN, class GOUT,
          class LIO = LdsPlain>
GF_HD void tile_chunks(uint64_t* a, const TileDesc& d, const TW& tw, const TWL& twl, uint64_t q, uint32_t tid,
                       uint32_t nthr, Sync sync, const GIN& gin, const GOUT& gout, LIO lio = LIO{}) {
    using P = PassOf<FWD, S, RM, I>;
    if constexpr (I < P::NC) {
        constexpr bool first = I == 0, last = I + 1 == P::NC;
        // the last chunk (stages up to S - 1) reads its twiddles through twl
        auto run = [&](const auto& in, const auto& out) {
            if constexpr (P::C + 1 == P::NC) tile_pass<FWD, LZ, S, P::ST0, P::R>(d, twl, q, tid, nthr, in, out);
            else tile_pass<FWD, LZ, S, P::ST0, P::R>(d, tw, q, tid, nthr, in, out);
        };
        if constexpr (first && last) {
            run(gin, gout);
        } else if constexpr (first) {
            run(gin, lio.out(a));
        } else if constexpr (last) {
            run(lio.in(a), gout);
        } else {
            run(lio.in(a), lio.out(a));
        }
        if constexpr (!last) {
            sync();
            tile_chunks<FWD, LZ, S, RM, I + 1>(a, d, tw, twl, q, tid, nthr, sync, gin, gout, lio);
        }
    }
}

// final store value: forward [0,8q) -> [0,q); inverse: times post (n^-1 c)
template <bool FWD>
GF_HD uint64_t ntt_final(uint64_t v, const NttJob& j) {
    if (FWD) return csub(csub(csub(v, 4 * j.q), 2 * j.q), j.q);
    return shoup(v, j.post, j.postp, j.q);
}
// forward final store with the optional linear epilogue at global index gi.
// LZ: v is below 72q; with the epilogue's two lazy products below 80q < 2^64,
// reduced once (or stored as is for a lazy_out job without epilogue).
// the epilogue's value for a forward transform's output v and the operands
// x0 = e0[gi], x1 = e1[gi] (j.e0 set)
template <bool LZ>
GF_HD uint64_t ntt_ep_value(uint64_t v, const NttJob& j, uint64_t x0, uint64_t x1) {
    if constexpr (LZ) {
        const uint64_t nq = 0 - j.q;
        v += shoup_lazy4(x0, j.c0, j.c0p, nq);
        v += shoup_lazy4(x1, j.c1, j.c1p, nq);
        return reduce_red(v, j.q, (uint32_t)j.red);
    } else {
        uint64_t r = csub(csub(csub(v, 4 * j.q), 2 * j.q), j.q);
        r = addmod(r, shoup(x0, j.c0, j.c0p, j.q), j.q);
        return addmod(r, shoup(x1, j.c1, j.c1p, j.q), j.q);
    }
}
template <bool LZ>
GF_HD uint64_t ntt_final_ep(uint64_t v, const NttJob& j, uint64_t gi) {
    if (j.e0) return ntt_ep_value<LZ>(v, j, gat(gcast(j.e0), gi), gat(gcast(j.e1), gi));
    if constexpr (LZ) {
        if (j.lazy_out) return v;
        return reduce_red(v, j.q, (uint32_t)j.red);
    } else {
        return csub(csub(csub(v, 4 * j.q), 2 * j.q), j.q);
    }
}

// ---- phase bodies (shared by the kernels and the host emulation) ----
// Every phase reads its tile from HBM in its first register pass and writes it
// in its last, straight from registers, wherever those accesses come in runs
// of consecutive words; LDS then only carries the exchange between passes.
// All reads of a tile precede the barrier before its first write, so in-place
// transforms (in == out) are safe.

// Phase A access: element k of column s of the tile at word (k << logN2) + b0 + s.
// Lanes run over consecutive columns, so both ends are always direct.
struct ColIn {
    const GF_GLOBAL uint64_t* src;
    uint32_t logN2, b0;
    GF_HD uint64_t operator()(uint32_t s, uint32_t k, uint32_t) const { return gat(src, (k << logN2) + b0 + s); }
};
template <bool FWD>
struct ColOut {
    GF_GLOBAL uint64_t* dst;
    uint32_t logN2, b0;
    const NttJob& j;
    // forward phase A hands lazy values in [0,8q) to phase B
    GF_HD void operator()(uint32_t s, uint32_t k, uint32_t, uint64_t v) const {
        gat(dst, (k << logN2) + b0 + s) = FWD ? v : ntt_final<false>(v, j);
    }
};

// Phase A: tile = columns [b0, b0+C) x all N1 = 2^S1 rows.
template <bool FWD, bool LZ, bool PAIRED, uint32_t S1, uint32_t RM, class Sync>
GF_HD void phaseA_body(const NttJob& j, const NttGeom& g, uint32_t tile, uint64_t* lds,
                       uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t logC = NTT_LOG_TILE - S1, C = 1u << logC;
    const uint32_t b0 = tile * C;
    const TileDesc d{logC, C, 1, 1, 0, true, false};
    // forward: phase A is the first phase (reads in); inverse: the second
    const ColIn gin{gcast(FWD ? j.in : j.out), g.logN2, b0};
    const ColOut<FWD> gout{gcast(j.out), g.logN2, b0, j};
    const auto tw = tw_of<PAIRED>(j);
    tile_chunks<FWD, LZ, S1, RM, 0>(lds, d, tw, tw, j.q, tid, nthr, sync, gin, gout);
}

// Row access (phase B, single phase): element k of row s at word base + s*N + k.
struct RowIn {
    const GF_GLOBAL uint64_t* src;   // tile start
    uint32_t N;
    GF_HD uint64_t operator()(uint32_t s, uint32_t k, uint32_t) const { return gat(src, s * N + k); }
};
// final store of a row tile: forward values through the epilogue (FIN = 2),
// inverse values through the n^-1 multiply (FIN = 1) or lazy (FIN = 0)
template <int FIN, bool LZ>
struct RowOut {
    GF_GLOBAL uint64_t* dst;         // tower start
    uint32_t off, N;
    const NttJob& j;
    GF_HD void operator()(uint32_t s, uint32_t k, uint32_t, uint64_t v) const {
        const uint32_t gi = off + s * N + k;
        gat(dst, gi) = FIN == 2 ? ntt_final_ep<LZ>(v, j, gi) : FIN == 1 ? ntt_final<false>(v, j) : v;
    }
};

// Word indices of the units a wave owns in a row-tile pass of radix 2^R whose
// units are short runs (TL < 8).  The loop of tile_pass gives a wave 64
// consecutive units per round, and those hold 64 << R consecutive words of the
// tile (units with the same group are interleaved by TL and adjacent in
// number), so f(idx) visits exactly the wave's own words with lanes on
// consecutive idx: a coalesced copy that needs only a wave-level fence.
// A tile of fewer than 64 units leaves the wave's last round partial: the
// words past the tile then belong to no unit and are skipped.
template <uint32_t R, uint32_t WORDS, class F>
GF_HD void wave_units(uint32_t tid, uint32_t nthr, F f) {
    const uint32_t W = nthr < 64 ? nthr : 64;
    const uint32_t lane = tid & (W - 1);
    constexpr uint32_t U = WORDS >> R;
    for (uint32_t u0 = tid - lane; u0 < U; u0 += nthr) {
#pragma unroll
        for (uint32_t m = 0; m < (1u << R); ++m) {
            const uint32_t idx = (u0 << R) + m * W + lane;
            if (U % 64 == 0 || idx < WORDS) f(idx);
        }
    }
}

// The stores of a forward row tile with an epilogue (j.e0 set), staged in LDS
// as wave_units visits them: the epilogue operands of 2 of a lane's
// words are requested before any is used (per word, the loads would
// otherwise wait in turn behind the previous word's arithmetic, a memory
// round trip each; 2 at a time keep the kernel's register count).
template <uint32_t R, uint32_t WORDS, bool LZ>
GF_HD void row_out_ep(const NttJob& j, GF_GLOBAL uint64_t* dst, uint32_t off, const uint64_t* lds, uint32_t tid,
                      uint32_t nthr) {
    const uint32_t W = nthr < 64 ? nthr : 64;
    const uint32_t lane = tid & (W - 1);
    constexpr uint32_t U = WORDS >> R, M = 1u << R, MB = M < 2 ? M : 2;
    const GF_GLOBAL uint64_t* e0 = gcast(j.e0);
    const GF_GLOBAL uint64_t* e1 = gcast(j.e1);
    for (uint32_t u0 = tid - lane; u0 < U; u0 += nthr) {
#pragma unroll
        for (uint32_t mb = 0; mb < M; mb += MB) {
            uint64_t x0[MB], x1[MB];
            GF_SCHED_FENCE();   // this batch's loads stay behind the previous batch's stores
#pragma unroll
            for (uint32_t m = 0; m < MB; ++m) {
                const uint32_t idx = (u0 << R) + (mb + m) * W + lane;
                if (U % 64 == 0 || idx < WORDS) {
                    x0[m] = gat(e0, off + idx);
                    x1[m] = gat(e1, off + idx);
                }
            }
            GF_SCHED_FENCE();
#pragma unroll
            for (uint32_t m = 0; m < MB; ++m) {
                const uint32_t idx = (u0 << R) + (mb + m) * W + lane;
                if (U % 64 == 0 || idx < WORDS)
                    gat(dst, off + idx) = ntt_ep_value<LZ>(lds[padded(idx)], j, x0[m], x1[m]);
            }
        }
    }
}

// A row tile whose first or last pass holds short runs (TL < 8) stages that
// end through LDS, wave by wave (no block barrier).
// ROWTAB: the job's row table (j.twr) serves the last chunk; off = a0 << S for
// the tile's first row a0.  KEEP: the tile's values stay in LDS after the last
// pass (unreduced, element idx of the tile at lds[padded(idx)]) for a consumer
// in the same workgroup, and dst is unused.  IN_LDS: the tile's input is
// already in LDS in that layout (written by the workgroup before a block
// barrier), and src is unused.
template <bool FWD, bool LZ, bool PAIRED, uint32_t S, uint32_t RM, uint32_t ROWS, int FIN, bool ROWTAB,
          class Sync, bool KEEP = false, bool IN_LDS = false>
GF_HD void row_tile(const NttJob& j, const TileDesc& d, const GF_GLOBAL uint64_t* src, GF_GLOBAL uint64_t* dst,
                    uint32_t off, uint64_t* lds, uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t N = 1u << S, NC = num_chunks(S, RM), WORDS = ROWS * N;
    using P0 = PassOf<FWD, S, RM, 0>;
    using PL = PassOf<FWD, S, RM, NC - 1>;
    constexpr bool in_direct = P0::wide_runs;
    constexpr bool out_direct = PL::wide_runs;
    static_assert(!KEEP || !out_direct, "a kept tile ends in LDS");
    static_assert(!IN_LDS || !in_direct, "a tile given in LDS starts with a staged pass");
    if constexpr (!in_direct && !IN_LDS) {
        wave_units<P0::R, WORDS>(tid, nthr, [&](uint32_t idx) { lds[padded(idx)] = gat(src, off + idx); });
        sync.wave();
    }
    const RowIn gin{src + off, N};
    const RowOut<FIN, LZ> gout{dst, off, N, j};
    const auto tw = tw_of<PAIRED>(j);
    auto chunks = [&](const auto& twl) {
        if constexpr (in_direct && out_direct)
            tile_chunks<FWD, LZ, S, RM, 0>(lds, d, tw, twl, j.q, tid, nthr, sync, gin, gout);
        else if constexpr (in_direct)
            tile_chunks<FWD, LZ, S, RM, 0>(lds, d, tw, twl, j.q, tid, nthr, sync, gin, LdsOut{lds});
        else if constexpr (out_direct)
            tile_chunks<FWD, LZ, S, RM, 0>(lds, d, tw, twl, j.q, tid, nthr, sync, LdsIn{lds}, gout);
        else
            tile_chunks<FWD, LZ, S, RM, 0>(lds, d, tw, twl, j.q, tid, nthr, sync, LdsIn{lds}, LdsOut{lds});
    };
    if constexpr (ROWTAB) {
        constexpr uint32_t RL = row_pass_r(S, RM), LOGU = S - RL;
        using TR = TwRows<RL, LOGU>;
        chunks(TR{gcast(j.twr) + (size_t)2 * (off >> S) * TR::PR});
    } else {
        chunks(tw);
    }
    if constexpr (!out_direct && !KEEP) {
        sync.wave();
        if constexpr (FIN == 2) {
            if (j.e0) {
                row_out_ep<PL::R, WORDS, LZ>(j, dst, off, lds, tid, nthr);
                return;
            }
        }
        wave_units<PL::R, WORDS>(tid, nthr, [&](uint32_t idx) { gout(0, idx, 0, lds[padded(idx)]); });
    }
}

// Phase B: tile = rows [a0, a0+R), each a contiguous run of N2 = 2^S2.
// IN_LDS (inverse): the tile's input is already in LDS (row_tile).
template <bool FWD, bool LZ, bool PAIRED, uint32_t S2, uint32_t RM, class Sync, bool IN_LDS = false>
GF_HD void phaseB_body(const NttJob& j, const NttGeom& g, uint32_t tile, uint64_t* lds,
                       uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t N2 = 1u << S2, logR = NTT_LOG_TILE - S2, R = 1u << logR;
    const uint32_t N1 = 1u << g.logN1;
    const uint32_t a0 = tile * R;
    const TileDesc d{logR, 1, padded(N2), N1 + a0, 1, false, true};
    // forward: phase B is the second phase (reads out); inverse: the first.
    // Inverse phase B hands lazy values in [0,4q) to phase A.
    row_tile<FWD, LZ, PAIRED, S2, RM, R, FWD ? 2 : 0, PAIRED, Sync, false, IN_LDS>(
        j, d, gcast(FWD ? j.out : j.in), gcast(j.out), a0 << S2, lds, tid, nthr, sync);
}

// Single phase: the whole tower (n = 2^S <= NTT_SINGLE_MAX) is one tile.
template <bool FWD, bool LZ, bool PAIRED, uint32_t S, uint32_t RM, class Sync>
GF_HD void single_body(const NttJob& j, uint64_t* lds, uint32_t tid, uint32_t nthr, Sync sync) {
    const TileDesc d{0, 1, padded(1u << S), 1, 0, false, true};
    row_tile<FWD, LZ, PAIRED, S, RM, 1, FWD ? 2 : 1, false>(j, d, gcast(j.in), gcast(j.out), 0, lds, tid, nthr, sync);
}

// Calls f(integral_constant S1, integral_constant S2) for the phase split of
// logn (two-phase sizes) -- the set of instantiated geometries.
template <class F>
inline bool with_split(uint32_t logN1, uint32_t logN2, F&& f) {
#define GF_SPLIT(a, b) if (logN1 == a && logN2 == b) { f(std::integral_constant<uint32_t, a>{}, std::integral_constant<uint32_t, b>{}); return true; }
    GF_SPLIT(6, 7) GF_SPLIT(7, 7) GF_SPLIT(7, 8) GF_SPLIT(8, 8) GF_SPLIT(9, 8)
#undef GF_SPLIT
    return false;
}
template <class F>
inline bool with_radix(uint32_t r, F&& f) {
    if (r == 2) { f(std::integral_constant<uint32_t, 2>{}); return true; }
    if (r == 3) { f(std::integral_constant<uint32_t, 3>{}); return true; }
    if (r == 4) { f(std::integral_constant<uint32_t, 4>{}); return true; }
    return false;
}
template <class F>
inline bool with_single(uint32_t logn, F&& f) {
#define GF_ONE(a) if (logn == a) { f(std::integral_constant<uint32_t, a>{}); return true; }
    GF_ONE(1) GF_ONE(2) GF_ONE(3) GF_ONE(4) GF_ONE(5) GF_ONE(6) GF_ONE(7) GF_ONE(8) GF_ONE(9)
    GF_ONE(10) GF_ONE(11) GF_ONE(12)
#undef GF_ONE
    return false;
}

}  // namespace gf
