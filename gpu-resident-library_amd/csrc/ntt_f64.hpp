// Transform passes in double precision for moduli below 2^50.
//
// A residue is held as a double whose value is an integer representative
// (signed, |x| < 2^53, so exact).  On gfx950 a double multiply, FMA, add or
// round-to-integer issues at the rate of a 32-bit integer multiply, and the
// modular product below takes 6 of them against the 9 multiplies and ~13
// other instructions of the 64-bit integer Shoup product; the butterfly adds
// are single instructions.
//
// Product y w mod q (f64_mulmod), for an integer |y| < 2^53 and a twiddle w
// held centred (|w| <= q/2) together with wq = fl(w / q):
//     k = rint(fl(y wq)),  p = fl(y w),  e = y w - p  (exact: fma(y, w, -p)),
//     r = p - k q          (fma(-k, q, p): an integer below 2^53, so exact),
//     return r + e         (= y w - k q exactly)
// fl(y wq) lies within |y| 2^-53 of y w / q (w/q and the product each round
// by at most half an ulp, |w/q| <= 1/2), so |y w / q - k| < 1/2 + |y| 2^-53
// and the result has |y w - k q| < q (1/2 + |y| 2^-53), below 1.5 q.  The
// value p - k q is that result minus e, with |e| <= 2^48 (p < 2^101): below
// 2^53, so the FMA rounds nothing.
//
// Bounds, in multiples of q (q < 2^50, so 8 q < 2^53):
//   forward (Cooley-Tukey) butterfly x +- w y: each stage adds at most 1.5;
//   inverse (Gentleman-Sande) x + y, w (x - y): sums double, products < 1.5.
// A pass centres its outputs (x - rint(x / q) q, then |x| <= q/2 + 2) where
// the next pass could otherwise exceed 7.5 q; every value and every product
// input then stays below 2^53.
#pragma once
#include "ntt_engine.hpp"

namespace gf {

#if defined(__clang__)
#define GF_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define GF_NO_CONTRACT
#endif

// An input accessor with a raw form (HasRaw): raw(s, k, at) fetches the word,
// conv(w) turns it into the pass's double
template <class A, class = void>
struct HasRaw : std::false_type {};
template <class A>
struct HasRaw<A, std::void_t<decltype(&A::raw)>> : std::true_type {};

// moduli whose transforms may run in double precision
GF_HD constexpr bool f64_modulus(uint64_t q) { return q < (1ull << 50); }

GF_HD double f64_of_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
GF_HD uint64_t bits_of_f64(double x) { return __builtin_bit_cast(uint64_t, x); }

constexpr double F64_2P52 = 4503599627370496.0;   // 2^52
// v < 2^52 as a double, minus c: the word with the exponent of 2^52 or-ed in
// is the double 2^52 + v; subtracting c2 = 2^52 + c (exact for c < 2^52)
// leaves v - c
GF_HD double f64_from_u52(uint64_t v, double c2) {
    GF_NO_CONTRACT
    return f64_of_bits(v | 0x4330000000000000ull) - c2;
}
// an integer 0 <= x < 2^52 as a 64-bit word: 2^52 + x has x as its mantissa
GF_HD uint64_t u52_from_f64(double x) {
    GF_NO_CONTRACT
    return bits_of_f64(x + F64_2P52) & 0x000FFFFFFFFFFFFFull;
}
GF_HD double f64_mulmod(double y, double w, double wq, double q) {
    GF_NO_CONTRACT
    const double k = __builtin_rint(y * wq);
    const double p = y * w;
    const double e = __builtin_fma(y, w, -p);
    const double r = __builtin_fma(-k, q, p);
    return r + e;
}
// centred representative: |result| <= q/2 + 2 for any integer |x| < 2^53
// (fl(x / q) is within 8 2^-52 of x / q, so rint misses by one only next to
// a half-integer)
GF_HD double f64_center(double x, double q, double qinv) {
    GF_NO_CONTRACT
    return __builtin_fma(-__builtin_rint(x * qinv), q, x);
}
// the representative in [0, q)
GF_HD double f64_reduce(double x, double q, double qinv) {
    GF_NO_CONTRACT
    const double c = f64_center(x, q, qinv);
    return c < 0 ? c + q : c;
}

// (w, fl(w/q)) pairs, w centred, in the index layout of the paired integer tables
// (kAhead: a table in global memory, whose forward pass twiddles of the first
// stages are requested together with the pass's data)
struct TwF64 {
    static constexpr bool kRows = false;
    static constexpr bool kAhead = true;
    const GF_GLOBAL double* t;
    GF_HD void get(uint32_t i, double& w, double& wq) const {
        w = t[2 * i];
        wq = t[2 * i + 1];
    }
};
// the row table of one phase-B tile (TwRows' lane-order layout) in fp64 pairs
template <uint32_t R, uint32_t LOGU>
struct TwRowsF64 {
    static constexpr bool kRows = true;
    static constexpr bool kAhead = true;
    static constexpr uint32_t PR = ((1u << R) - 1) << LOGU;
    const GF_GLOBAL double* t;
    GF_HD void get_rp(uint32_t c, uint32_t s, uint32_t g, uint32_t v, double& w, double& wq) const {
        const uint32_t e = s * PR + (((1u << c) - 1) << LOGU) + (v << LOGU) + g;
        w = t[2 * e];
        wq = t[2 * e + 1];
    }
};
inline void f64_twiddle(uint64_t w, uint64_t q, double& wd, double& wq) {
    const int64_t c = w > q / 2 ? (int64_t)w - (int64_t)q : (int64_t)w;
    wd = (double)c;
    wq = wd / (double)q;
}

// LDS holds the doubles' bit patterns
struct LdsInF64 {
    const uint64_t* a;
    GF_HD double operator()(uint32_t, uint32_t, uint32_t at) const { return f64_of_bits(a[at]); }
};
struct LdsOutF64 {
    uint64_t* a;
    GF_HD void operator()(uint32_t, uint32_t, uint32_t at, double v) const { a[at] = bits_of_f64(v); }
};
struct LdsPlainF64 {
    GF_HD LdsInF64 in(const uint64_t* a) const { return LdsInF64{a}; }
    GF_HD LdsOutF64 out(uint64_t* a) const { return LdsOutF64{a}; }
};
// a unit in registers (UnitIn, UnitOut) holding the doubles' bit patterns
template <bool ALL, uint32_t U = 4>
struct UnitInF64 {
    static constexpr bool kUnit = true;
    const uint64_t* v;
    GF_HD double unit(uint32_t uid, uint32_t u) const { return f64_of_bits(v[(ALL ? uid * U : 0) + u]); }
};
template <bool ALL, uint32_t U = 4>
struct UnitOutF64 {
    static constexpr bool kUnit = true;
    uint64_t* v;
    GF_HD void unit(uint32_t uid, uint32_t u, double x) const { v[(ALL ? uid * U : 0) + u] = bits_of_f64(x); }
};

template <class TW, class = void>
struct TwAhead : std::false_type {};
template <class TW>
struct TwAhead<TW, std::void_t<decltype(TW::kAhead)>> : std::integral_constant<bool, TW::kAhead> {};

// One radix-2^R pass over the tile's units, as tile_pass (same unit and
// twiddle indexing) on doubles.  CENTER: outputs centred before `out`.  An
// inverse pass of more than two stages centres its unit after every second
// stage as well.
template <bool FWD, uint32_t S, uint32_t ST0, uint32_t R, bool CENTER, class TW, class IN, class OUT>
GF_HD void tile_pass_f64(const TileDesc& d, const TW& tw, double q, double qinv, uint32_t tid, uint32_t nthr,
                         const IN& in, const OUT& out) {
    constexpr uint32_t LOGTL = S - ST0 - R;
    constexpr uint32_t TL = 1u << LOGTL;
    constexpr uint32_t LOGPS = S - R;
    constexpr uint32_t R8 = 1u << R;
    const uint32_t T = 1u << d.logT;
    const uint32_t U = T << LOGPS;
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
        const uint32_t G = d.g0 + s * d.gstep;
        // a forward pass's first stage takes one twiddle: requested with the
        // data; from a global table, so are the 2 and 4 of stages 1 and 2
        // (one memory round trip instead of three before stage 3)
        constexpr uint32_t AH = FWD && TwAhead<TW>::value ? (R < 3 ? R : 3) : 1;
        double w0 = 0, wq0 = 0;
        double wa[6], wqa[6];   // stage c < AH: its pair v at (1 << c) - 2 + v
        if constexpr (FWD) {
            if constexpr (TW::kRows) tw.get_rp(0, s, g, 0, w0, wq0);
            else tw.get((G << ST0) + g, w0, wq0);
#pragma unroll
            for (uint32_t c = 1; c < AH; ++c) {
#pragma unroll
                for (uint32_t v = 0; v < (1u << c); ++v) {
                    double& w = wa[(1u << c) - 2 + v];
                    double& wq = wqa[(1u << c) - 2 + v];
                    if constexpr (TW::kRows) tw.get_rp(c, s, g, v, w, wq);
                    else tw.get((G << (ST0 + c)) + (g << c) + v, w, wq);
                }
            }
        }
        double x[R8];
        if constexpr (HasRaw<IN>::value) {
            uint64_t r[R8];
#pragma unroll
            for (uint32_t u = 0; u < R8; ++u) r[u] = in.raw(s, k0 + u * TL, at[u]);
            GF_SCHED_FENCE();
#pragma unroll
            for (uint32_t u = 0; u < R8; ++u) x[u] = in.conv(r[u]);
        } else {
#pragma unroll
            for (uint32_t u = 0; u < R8; ++u) x[u] = pass_in(in, uid, u, s, k0 + u * TL, at[u]);
        }
        if (FWD) {
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) {
                const uint32_t h = 1u << (R - 1 - c);
                const uint32_t tb = (G << (ST0 + c)) + (g << c);
#pragma unroll
                for (uint32_t u = 0; u < R8; ++u) {
                    if (u & h) continue;
                    double w = w0, wq = wq0;
                    if (c > 0 && c < AH) {
                        w = wa[(1u << c) - 2 + (u >> (R - c))];
                        wq = wqa[(1u << c) - 2 + (u >> (R - c))];
                    } else if (c > 0) {
                        if constexpr (TW::kRows) tw.get_rp(c, s, g, u >> (R - c), w, wq);
                        else tw.get(tb + (u >> (R - c)), w, wq);
                    }
                    const double m = f64_mulmod(x[u + h], w, wq, q);
                    const double a = x[u];
                    x[u] = a + m;
                    x[u + h] = a - m;
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
                    double w, wq;
                    if constexpr (TW::kRows) tw.get_rp((uint32_t)c, s, g, u >> (R - c), w, wq);
                    else tw.get(tb + (u >> (R - c)), w, wq);
                    const double a = x[u], b = x[u + h];
                    x[u] = a + b;
                    x[u + h] = f64_mulmod(a - b, w, wq, q);
                }
                if (c > 0 && (R - c) % 2 == 0) {
#pragma unroll
                    for (uint32_t u = 0; u < R8; ++u) x[u] = f64_center(x[u], q, qinv);
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < R8; ++u) {
            if constexpr (CENTER) x[u] = f64_center(x[u], q, qinv);
            pass_out(out, uid, u, s, k0 + u * TL, at[u], x[u]);
        }
    }
}

// Whether fp64 pass I (execution order) of an S-stage phase in radix-2^RM
// chunks centres its outputs.  The last pass never does (its consumer
// reduces).  Inverse, inputs centred: every pass (sums double per stage, and
// with the centring after every second stage inside a pass its values stay
// below 3 q).  Forward, inputs |x| <= 1.25 q: a pass of R stages adds
// 1.5 R q, and a pass centres when the next one would take the bound past
// 7.5 q.
template <bool FWD, uint32_t S, uint32_t RM>
GF_HD constexpr bool f64_center_after(uint32_t I) {
    const uint32_t NC = num_chunks(S, RM);
    if (I + 1 >= NC) return false;
    if (!FWD) return true;
    double b = 1.25;
    for (uint32_t c = 0; c + 1 < NC; ++c) {
        b += 1.5 * chunk_size(S, RM, c);
        const bool ctr = b + 1.5 * chunk_size(S, RM, c + 1) > 7.5;
        if (c == I) return ctr;
        if (ctr) b = 0.51;
    }
    return false;
}
// the bound (multiples of q) of the last forward pass's outputs
template <uint32_t S, uint32_t RM>
GF_HD constexpr double f64_fwd_out_bound() {
    const uint32_t NC = num_chunks(S, RM);
    double b = 1.25;
    for (uint32_t c = 0; c < NC; ++c) {
        b += 1.5 * chunk_size(S, RM, c);
        if (c + 1 < NC && b + 1.5 * chunk_size(S, RM, c + 1) > 7.5) b = 0.51;
    }
    return b;
}

// All passes of one phase in fp64, as tile_chunks: the first reads through
// gin, the last writes through gout, LDS (bit patterns) in between; the last
// chunk (stages up to S - 1) reads its twiddles through twl.
template <bool FWD, uint32_t S, uint32_t RM, uint32_t I, class TW, class TWL, class Sync, class GIN, class GOUT,
          class LIO = LdsPlainF64>
GF_HD void tile_chunks_f64(uint64_t* a, const TileDesc& d, const TW& tw, const TWL& twl, double q, double qinv,
                           uint32_t tid, uint32_t nthr, Sync sync, const GIN& gin, const GOUT& gout, LIO lio = LIO{}) {
    using P = PassOf<FWD, S, RM, I>;
    if constexpr (I < P::NC) {
        static_assert(!FWD || f64_fwd_out_bound<S, RM>() <= 7.5, "forward bound");
        constexpr bool first = I == 0, last = I + 1 == P::NC;
        constexpr bool ctr = f64_center_after<FWD, S, RM>(I);
        auto run = [&](const auto& in, const auto& out) {
            if constexpr (P::C + 1 == P::NC) tile_pass_f64<FWD, S, P::ST0, P::R, ctr>(d, twl, q, qinv, tid, nthr, in, out);
            else tile_pass_f64<FWD, S, P::ST0, P::R, ctr>(d, tw, q, qinv, tid, nthr, in, out);
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
            tile_chunks_f64<FWD, S, RM, I + 1>(a, d, tw, twl, q, qinv, tid, nthr, sync, gin, gout, lio);
        }
    }
}

// Row access in fp64.  In: a tower's words (below 4q; CTR: centred after
// the conversion, else below 2q and shifted by q), from HBM or staged raw in
// LDS.  Out: the centred value plus q as a word in [0, 2q), to HBM or LDS.
template <bool CTR>
struct RowInF64 {
    const GF_GLOBAL uint64_t* src;   // tile start
    uint32_t N;
    double c2, q, qinv;
    GF_HD uint64_t raw(uint32_t s, uint32_t k, uint32_t) const { return gat(src, s * N + k); }
    GF_HD double conv(uint64_t w) const {
        const double x = f64_from_u52(w, c2);
        return CTR ? f64_center(x, q, qinv) : x;
    }
};
template <bool CTR>
struct LdsConvIn {
    const uint64_t* a;
    double c2, q, qinv;
    GF_HD double operator()(uint32_t, uint32_t, uint32_t at) const {
        const double x = f64_from_u52(a[at], c2);
        return CTR ? f64_center(x, q, qinv) : x;
    }
};
struct RowOutU52 {
    GF_GLOBAL uint64_t* dst;   // tower start
    uint32_t off, N;
    double q, qinv;
    GF_HD void operator()(uint32_t s, uint32_t k, uint32_t, double v) const {
        gat(dst, off + s * N + k) = u52_from_f64(f64_center(v, q, qinv) + q);
    }
};
struct LdsOutU52 {
    uint64_t* a;
    double q, qinv;
    GF_HD void operator()(uint32_t, uint32_t, uint32_t at, double v) const {
        a[at] = u52_from_f64(f64_center(v, q, qinv) + q);
    }
};

// A row tile in fp64, with row_tile's layout, staging and twiddle tables
// (twd: fp64 pairs, twrd: their row table), for the two uses that take it:
//   forward, KEEP: the keyswitch digits' row pass, inputs below 2q, the tile
//     left in LDS for the inner product (element idx at lds[padded(idx)], in
//     [0, 2q));
//   inverse: a phase B, inputs below 4q, its output in [0, 2q) to dst for
//     phase A.
template <bool FWD, uint32_t S, uint32_t RM, uint32_t ROWS, bool KEEP, class Sync, bool IN_LDS = false>
GF_HD void row_tile_f64(const uint64_t* twd, const uint64_t* twrd, double q, double qinv, const TileDesc& d,
                        const GF_GLOBAL uint64_t* src, GF_GLOBAL uint64_t* dst, uint32_t off, uint64_t* lds,
                        uint32_t tid, uint32_t nthr, Sync sync) {
    constexpr uint32_t N = 1u << S, NC = num_chunks(S, RM), WORDS = ROWS * N;
    using P0 = PassOf<FWD, S, RM, 0>;
    using PL = PassOf<FWD, S, RM, NC - 1>;
    constexpr bool in_direct = P0::wide_runs, out_direct = PL::wide_runs;
    static_assert(FWD == KEEP, "fp64 row tiles: forward kept in LDS, or inverse to HBM");
    static_assert(KEEP ? !out_direct : out_direct, "fp64 row tile shape");
    const double c2 = F64_2P52 + (FWD ? q : 2 * q);
    static_assert(!IN_LDS || !in_direct, "a tile given in LDS starts with a staged pass");
    if constexpr (!in_direct && !IN_LDS) {
        wave_units<P0::R, WORDS>(tid, nthr, [&](uint32_t idx) { lds[padded(idx)] = gat(src, off + idx); });
        sync.wave();
    }
    constexpr uint32_t RL = row_pass_r(S, RM), LOGU = S - RL;
    using TR = TwRowsF64<RL, LOGU>;
    const TR twl{gcast((const double*)twrd) + (size_t)2 * (off >> S) * TR::PR};
    const TwF64 tw{gcast((const double*)twd)};
    auto outs = [&](const auto& gin) {
        if constexpr (KEEP) tile_chunks_f64<FWD, S, RM, 0>(lds, d, tw, twl, q, qinv, tid, nthr, sync, gin,
                                                           LdsOutU52{lds, q, qinv});
        else tile_chunks_f64<FWD, S, RM, 0>(lds, d, tw, twl, q, qinv, tid, nthr, sync, gin,
                                            RowOutU52{dst, off, N, q, qinv});
    };
    if constexpr (in_direct) outs(RowInF64<!FWD>{src + off, N, c2, q, qinv});
    else outs(LdsConvIn<!FWD>{lds, c2, q, qinv});
}

// Inverse phase B in fp64 (tile = rows [a0, a0 + R), as phaseB_body) for a
// job with fp64 tables: output in [0, 2q) for phase A.  IN_LDS: the tile's
// input is already in LDS (row_tile).
template <uint32_t S2, uint32_t RM, class Sync, bool IN_LDS = false>
GF_HD void phaseB_inv_f64(const NttJob& j, const NttGeom& g, uint32_t tile, uint64_t* lds, uint32_t tid,
                          uint32_t nthr, Sync sync) {
    constexpr uint32_t N2 = 1u << S2, logR = NTT_LOG_TILE - S2, R = 1u << logR;
    const uint32_t N1 = 1u << g.logN1;
    const uint32_t a0 = tile * R;
    const TileDesc d{logR, 1, padded(N2), N1 + a0, 1, false, true};
    row_tile_f64<false, S2, RM, R, false, Sync, IN_LDS>(j.twd, j.twrd, j.fq, j.fqinv, d, gcast(j.in), gcast(j.out),
                                                         a0 << S2, lds, tid, nthr, sync);
}

}  // namespace gf
