// Host emulation of the library's device tile logic (test infrastructure).
// Compiles the same headers the HIP kernels use with a plain C++ compiler and
// runs each workgroup sequentially, so index and arithmetic mistakes show up
// on the CPU against the oracle before any GPU run.
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ksi.hpp"
#include "modup.hpp"
#include "ntt_engine.hpp"
#include "tensor.hpp"

using namespace gf;

namespace {
struct NoSync {
    void operator()() const {}
    void wave() const {}
};

template <bool PAIRED, uint32_t RM, bool LZ>
void emu_ntt_t(const NttJob& j0, uint32_t n, int fwd) {
    NttGeom g = make_geom(n, RM);
    NttJob j = j0;
    std::vector<uint64_t> rows;   // paired jobs carry phase B's row table, as the plans do
    if (PAIRED && !g.single) {
        rows.resize(row_table_words(g));
        build_row_table(j.tw, g, rows.data());
        j.twr = rows.data();
    }
    std::vector<uint64_t> lds(NTT_TILE_PADDED > padded(n) ? NTT_TILE_PADDED : padded(n));
    if (g.single) {
        with_single(g.logn, [&](auto Sc) {
            constexpr uint32_t S = decltype(Sc)::value;
            if (fwd) single_body<true, LZ, PAIRED, S, RM>(j, lds.data(), 0, 1, NoSync{});
            else single_body<false, false, PAIRED, S, RM>(j, lds.data(), 0, 1, NoSync{});
        });
        return;
    }
    with_split(g.logN1, g.logN2, [&](auto A, auto B) {
        constexpr uint32_t S1 = decltype(A)::value, S2 = decltype(B)::value;
        const uint32_t tilesA = (1u << S2) / g.C, tilesB = (1u << S1) / g.R;
        if (fwd) {
            for (uint32_t t = 0; t < tilesA; ++t) phaseA_body<true, LZ, PAIRED, S1, RM>(j, g, t, lds.data(), 0, 1, NoSync{});
            for (uint32_t t = 0; t < tilesB; ++t) phaseB_body<true, LZ, PAIRED, S2, RM>(j, g, t, lds.data(), 0, 1, NoSync{});
        } else {
            for (uint32_t t = 0; t < tilesB; ++t) phaseB_body<false, false, PAIRED, S2, RM>(j, g, t, lds.data(), 0, 1, NoSync{});
            for (uint32_t t = 0; t < tilesA; ++t) phaseA_body<false, false, PAIRED, S1, RM>(j, g, t, lds.data(), 0, 1, NoSync{});
        }
    });
}

template <bool PAIRED>
void emu_ntt_r(const NttJob& j, uint32_t n, int fwd, uint32_t rmax) {
    with_radix(rmax, [&](auto Rc) {
        // the kernels' per-job choice: forward jobs with red skip the butterfly reductions
        if (fwd && j.red) emu_ntt_t<PAIRED, decltype(Rc)::value, true>(j, n, fwd);
        else emu_ntt_t<PAIRED, decltype(Rc)::value, false>(j, n, fwd);
    });
}
}  // namespace

// In-place transform of x (n residues mod q).  paired != 0 builds an
// interleaved (w, precon) table from tw/twp; rmax = register radix log2.
// lazy (forward only): 1 = red = lazy_red_of(q) (0 for moduli outside its
// range), 2 = also lazy_out.  e0/e1 (may be null): the forward epilogue
// out = NTT(x) + e0 c0 + e1 c1 with Shoup precons c0p/c1p.
extern "C" void emu_ntt_ex(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* tw, const uint64_t* twp,
                           int fwd, uint64_t post, uint64_t postp, int paired, uint32_t rmax, int lazy,
                           const uint64_t* e0, const uint64_t* e1, uint64_t c0, uint64_t c1) {
    std::vector<uint64_t> t2;
    NttJob j{x, x, tw, twp, q, post, postp};
    if (paired) {
        t2.resize(2 * (std::size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            t2[2 * i] = tw[i];
            t2[2 * i + 1] = twp[i];
        }
        j.tw = t2.data();
        j.twp = nullptr;
    }
    if (fwd && lazy) {
        j.red = lazy_red_of(q);
        j.lazy_out = lazy == 2;
    }
    if (fwd && e0) {
        j.e0 = e0;
        j.e1 = e1;
        j.c0 = c0;
        j.c0p = host_precon(c0, q);
        j.c1 = c1;
        j.c1p = host_precon(c1, q);
    }
    if (paired) emu_ntt_r<true>(j, n, fwd, rmax);
    else emu_ntt_r<false>(j, n, fwd, rmax);
}
extern "C" void emu_ntt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* tw, const uint64_t* twp,
                        int fwd, uint64_t post, uint64_t postp, int paired, uint32_t rmax) {
    emu_ntt_ex(x, n, q, tw, twp, fwd, post, postp, paired, rmax, 0, nullptr, nullptr, 0, 0);
}
extern "C" uint64_t emu_lazy_red_of(uint64_t q) { return lazy_red_of(q); }

namespace {
// fp64 pairs (ntt_f64.hpp) of a paired integer table's twiddles
std::vector<uint64_t> f64_pairs_of(const uint64_t* paired, uint32_t n, uint64_t q) {
    std::vector<uint64_t> out(2 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) {
        double wd, wq;
        f64_twiddle(paired[2 * k], q, wd, wq);
        std::memcpy(&out[2 * k], &wd, 8);
        std::memcpy(&out[2 * k + 1], &wq, 8);
    }
    return out;
}

template <uint32_t RM, bool LZ>
void emu_ksi_t(uint32_t n, uint64_t q, const uint64_t* t2, uint32_t np, uint32_t xf, uint64_t* src,
               const uint64_t* kb, const uint64_t* ka, uint64_t* o0, uint64_t* o1, uint32_t pairs,
               uint32_t lazy_out, int f64, const uint64_t* it2, uint64_t post, uint64_t postp, uint64_t* x0,
               uint64_t* x1) {
    const NttGeom g = make_geom(n, RM);
    std::vector<uint64_t> rows(row_table_words(g));
    build_row_table(t2, g, rows.data());
    std::vector<uint64_t> t2d, rowsd;
    KsiJob J{};
    if (f64) {
        // the row passes in fp64, with digits below 2q as the fp64 ModUp leaves them
        t2d = f64_pairs_of(t2, n, q);
        rowsd.resize(row_table_words(g));
        build_row_table(t2d.data(), g, rowsd.data());
        J.twd = t2d.data();
        J.twrd = rowsd.data();
        J.fq = (double)q;
        J.fqinv = 1.0 / J.fq;
    }
    J.np = np;
    J.xf = xf;
    J.tw = t2;
    J.twr = rows.data();
    J.q = q;
    J.qneg = host_qneg(q);
    J.red = lazy_red_of(q);
    J.d30 = ksi_d30(q, J.red);
    J.o0 = o0;
    J.o1 = o1;
    J.pairs = pairs;
    J.lazy_out = lazy_out;
    std::vector<uint64_t> irows;
    if (it2) {   // the inverse row pass of both outputs inside the fused step, into x0 / x1
        irows.resize(row_table_words(g));
        build_row_table(it2, g, irows.data());
        uint64_t* xs[2] = {x0, x1};
        for (int c = 0; c < 2; ++c) {
            J.rp[c] = NttJob{c ? o1 : o0, xs[c], it2, nullptr, q, post, postp};
            J.rp[c].twr = irows.data();
        }
        J.rp_on = 1;
    }
    std::vector<uint64_t> lds(NTT_TILE_PADDED > (size_t)np * KSI_PART_WORDS ? NTT_TILE_PADDED : (size_t)np * KSI_PART_WORDS);
    with_split(g.logN1, g.logN2, [&](auto A, auto B) {
        constexpr uint32_t S1 = decltype(A)::value, S2 = decltype(B)::value;
        for (uint32_t p = 0; p < np; ++p) {
            J.src[p] = src + (size_t)p * n;
            J.kb[p] = kb + (size_t)p * n;
            J.ka[p] = ka + (size_t)p * n;
            if ((xf >> p) & 1) {   // the column pass, in place, as the plans' NTT_FA step
                NttJob t{src + (size_t)p * n, src + (size_t)p * n, t2, nullptr, q, 0, 0};
                t.red = J.red;
                for (uint32_t tile = 0; tile < (1u << S2) / g.C; ++tile)
                    phaseA_body<true, LZ, true, S1, RM>(t, g, tile, lds.data(), 0, 1, NoSync{});
                if (f64)
                    for (uint32_t k = 0; k < n; ++k) src[(size_t)p * n + k] %= q;
            }
        }
        for (uint32_t tile = 0; tile < n / KSI_TILE; ++tile)
            ksi_body<S2, RM, LZ>(J, 1u << S1, tile, lds.data(), 0, 1, 0, 1, 0, 1, NoSync{}, NoSync{});
        if (it2)   // the inverse transforms' column pass after it, as the fused ModDown runs it
            for (int c = 0; c < 2; ++c)
                for (uint32_t tile = 0; tile < (1u << S2) / g.C; ++tile)
                    phaseA_body<false, false, true, S1, RM>(J.rp[c], g, tile, lds.data(), 0, 1, NoSync{});
    });
}
}  // namespace

// Fused digit row pass + key inner product over np parts: part p of src holds
// a coefficient-form tower (bit p of xf: transformed here, column pass then the
// fused row pass) or an evaluation-form digit; o = sum_p NTT?(src_p) k_p R^-1.
// f64 != 0 (q below 2^50): the column passes' output reduced below q, then
// the row passes in fp64.  iw != null: the step also runs the inverse
// transforms' row pass of both outputs (tables iw/iwp, final factor
// post/postp) into x0 / x1 instead of storing o0 / o1, and the column pass
// follows here: x_c = INTT(o_c) times post.
extern "C" void emu_ksi(uint32_t n, uint64_t q, const uint64_t* tw, const uint64_t* twp, uint32_t rmax,
                        uint32_t np, uint32_t xf, uint64_t* src, const uint64_t* kb, const uint64_t* ka,
                        uint64_t* o0, uint64_t* o1, uint32_t pairs, uint32_t lazy_out, int f64, const uint64_t* iw,
                        const uint64_t* iwp, uint64_t post, uint64_t postp, uint64_t* x0, uint64_t* x1) {
    std::vector<uint64_t> t2(2 * (std::size_t)n), it2;
    for (uint32_t i = 0; i < n; ++i) {
        t2[2 * i] = tw[i];
        t2[2 * i + 1] = twp[i];
    }
    if (iw) {
        it2.resize(2 * (std::size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            it2[2 * i] = iw[i];
            it2[2 * i + 1] = iwp[i];
        }
    }
    const uint64_t* it = iw ? it2.data() : nullptr;
    with_radix(rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        if (lazy_red_of(q))
            emu_ksi_t<RM, true>(n, q, t2.data(), np, xf, src, kb, ka, o0, o1, pairs, lazy_out, f64, it, post, postp,
                                x0, x1);
        else
            emu_ksi_t<RM, false>(n, q, t2.data(), np, xf, src, kb, ka, o0, o1, pairs, lazy_out, f64, it, post, postp,
                                 x0, x1);
    });
}

// The tensor product fused with the inverse row pass of its t2, as the
// fused kernel runs it per row tile (tensor_tile_lds, then phase B from LDS,
// fp64 when f64 != 0), then the integer phase A over x: t0, t1, t2 and
// x = INTT(t2) times the final factor post (post, postp).
extern "C" void emu_tensor_ib(uint32_t n, uint64_t q, const uint64_t* iw, const uint64_t* iwp, uint64_t post,
                              uint64_t postp, uint32_t rmax, int f64, const uint64_t* a0, const uint64_t* a1,
                              const uint64_t* b0, const uint64_t* b1, uint64_t* t0, uint64_t* t1, uint64_t* t2,
                              uint64_t* x) {
    std::vector<uint64_t> tw2(2 * (std::size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        tw2[2 * i] = iw[i];
        tw2[2 * i + 1] = iwp[i];
    }
    with_radix(rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        const NttGeom g = make_geom(n, RM);
        const std::vector<uint64_t> t2d = f64_pairs_of(tw2.data(), n, q);
        std::vector<uint64_t> rows(row_table_words(g)), rowsd(row_table_words(g)), lds(NTT_TILE_PADDED);
        build_row_table(tw2.data(), g, rows.data());
        build_row_table(t2d.data(), g, rowsd.data());
        const TensorJob tj{a0, a1, b0, b1, t0, t1, t2, q, host_qneg(q), host_r2(q), 1};
        NttJob j{t2, x, tw2.data(), nullptr, q, post, postp};
        j.twr = rows.data();
        if (f64) {
            j.twd = t2d.data();
            j.twrd = rowsd.data();
            j.fq = (double)q;
            j.fqinv = 1.0 / j.fq;
        }
        with_split(g.logN1, g.logN2, [&](auto A, auto B) {
            constexpr uint32_t S1 = decltype(A)::value, S2 = decltype(B)::value;
            if constexpr ((NTT_TILE >> S2) << S2 == NTT_TILE) {
                for (uint32_t tile = 0; tile < (1u << S1) / g.R; ++tile) {
                    std::fill(lds.begin(), lds.end(), ~0ull);   // the row pass sees only what the tensor left
                    tensor_tile_lds(tj, tile * NTT_TILE, lds.data(), 0, 1);
                    if (f64) phaseB_inv_f64<S2, RM, NoSync, true>(j, g, tile, lds.data(), 0, 1, NoSync{});
                    else phaseB_body<false, false, true, S2, RM, NoSync, true>(j, g, tile, lds.data(), 0, 1, NoSync{});
                }
                for (uint32_t tile = 0; tile < (1u << S2) / g.C; ++tile)
                    phaseA_body<false, false, true, S1, RM>(j, g, tile, lds.data(), 0, 1, NoSync{});
            }
        });
    });
}

// Inverse transform of a two-phase ring with phase B in fp64 (q below 2^50,
// inputs below 4q) and the integer phase A after it, in place; post, postp:
// the final factor of phase A (n^-1 c).  Output in [0, q).
extern "C" void emu_intt_f64(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* iw, const uint64_t* iwp,
                             uint64_t post, uint64_t postp, uint32_t rmax) {
    std::vector<uint64_t> t2(2 * (std::size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        t2[2 * i] = iw[i];
        t2[2 * i + 1] = iwp[i];
    }
    with_radix(rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        const NttGeom g = make_geom(n, RM);
        const std::vector<uint64_t> t2d = f64_pairs_of(t2.data(), n, q);
        std::vector<uint64_t> rowsd(row_table_words(g)), lds(NTT_TILE_PADDED);
        build_row_table(t2d.data(), g, rowsd.data());
        NttJob j{x, x, t2.data(), nullptr, q, post, postp};
        j.twd = t2d.data();
        j.twrd = rowsd.data();
        j.fq = (double)q;
        j.fqinv = 1.0 / j.fq;
        with_split(g.logN1, g.logN2, [&](auto A, auto B) {
            constexpr uint32_t S1 = decltype(A)::value, S2 = decltype(B)::value;
            for (uint32_t tile = 0; tile < (1u << S1) / g.R; ++tile)
                phaseB_inv_f64<S2, RM>(j, g, tile, lds.data(), 0, 1, NoSync{});
            for (uint32_t k = 0; k < n; ++k)
                if (x[k] >= 2 * q) std::abort();   // phase B hands phase A values in [0, 2q)
            for (uint32_t tile = 0; tile < (1u << S2) / g.C; ++tile)
                phaseA_body<false, false, true, S1, RM>(j, g, tile, lds.data(), 0, 1, NoSync{});
        });
    });
}
// the key inner product's output with unreduced digits: sum_s a[s] b[s]
// accumulated as 128 bits (mul128/add128), then redc_red
extern "C" uint64_t emu_inner_red(const uint64_t* a, const uint64_t* b, uint32_t ns, uint64_t q) {
    U128 T{0, 0};
    for (uint32_t s = 0; s < ns; ++s) T = add128(T, mul128(a[s], b[s]));
    return redc_red(T, q, host_qneg(q), (uint32_t)lazy_red_of(q));
}
extern "C" uint64_t emu_reduce_red(uint64_t x, uint64_t q) { return reduce_red(x, q, (uint32_t)lazy_red_of(q)); }

extern "C" uint64_t emu_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return mulmod(a, b, q, host_qneg(q), host_r2(q));
}
extern "C" uint64_t emu_shoup_lazy4(uint64_t x, uint64_t w, uint64_t q) {
    return shoup_lazy4(x, w, host_precon(w, q), 0 - q);
}
extern "C" uint64_t emu_csub(uint64_t x, uint64_t m) { return csub(x, m); }
// sum_{s<ns} a[s] b[s] through the column accumulator (ns in 1..8) -> out[0] lo, out[1] hi
extern "C" void emu_dot30(const uint64_t* a, const uint64_t* b, uint32_t ns, uint64_t* out) {
    uint32_t lo[8], hi[8];
    for (uint32_t s = 0; s < ns; ++s) {
        lo[s] = split30_lo(a[s]);
        hi[s] = split30_hi(a[s]);
    }
    U128 T{0, 0};
    auto f = [&](uint32_t s) { return b[s]; };
    switch (ns) {
        case 1: T = dot30<1>(lo, hi, f); break;
        case 2: T = dot30<2>(lo, hi, f); break;
        case 3: T = dot30<3>(lo, hi, f); break;
        case 4: T = dot30<4>(lo, hi, f); break;
        case 5: T = dot30<5>(lo, hi, f); break;
        case 6: T = dot30<6>(lo, hi, f); break;
        case 7: T = dot30<7>(lo, hi, f); break;
        case 8: T = dot30<8>(lo, hi, f); break;
        default: break;
    }
    out[0] = T.lo;
    out[1] = T.hi;
}
extern "C" uint64_t emu_shoup(uint64_t x, uint64_t w, uint64_t q) {
    return shoup(x, w, host_precon(w, q), q);
}
extern "C" uint64_t emu_mulhi(uint64_t a, uint64_t b) { return mulhi64(a, b); }
extern "C" uint64_t emu_redc(uint64_t lo, uint64_t hi, uint64_t q) {
    return redc(U128{lo, hi}, q, host_qneg(q));
}

namespace {
std::vector<uint64_t> paired(const uint64_t* tw, const uint64_t* twp, uint32_t n) {
    std::vector<uint64_t> t2(2 * (std::size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        t2[2 * i] = tw[i];
        t2[2 * i + 1] = twp[i];
    }
    return t2;
}

template <uint32_t SP, uint32_t S1, bool C60>
void emu_modup_t(const ModupJob& j, const NttGeom& g) {
    // one thread runs the tile's 256 units: their registers side by side
    std::vector<uint64_t> lds(MODUP_TILE), v((size_t)MODUP_THREADS * MODUP_UNIT), tl(2u << S1);
    std::vector<ModupUnit<SP>> units(MODUP_THREADS);
    for (uint32_t tile = 0; tile < g.n / MODUP_TILE; ++tile) {
        for (uint32_t i = 0; i < SP; ++i) {
            modup_source_pass<S1, true>(j, i, g, tile, lds.data(), 0, 1, NoSync{}, v.data(), tl.data(), [&] {
                for (uint32_t k = 0; k < (1u << S1); ++k) modup_tw_put(tl.data(), k, modup_src_tw_pair(j, i, k));
            });
            for (uint32_t u = 0; u < MODUP_THREADS; ++u)
                modup_take_unit<SP, S1>(j, i, v.data() + (size_t)u * MODUP_UNIT, units[u]);
        }
        for (uint32_t t = 0; t < j.sc; ++t) {
            for (uint32_t u = 0; u < MODUP_THREADS; ++u)
                modup_convert_unit<SP, S1, C60>(j, t, units[u], v.data() + (size_t)u * MODUP_UNIT);
            for (uint32_t i = 0; i < (1u << S1); ++i) modup_tw_put(tl.data(), i, modup_tw_pair(j, t, i));
            modup_target_pass<S1, true>(j, t, g, tile, lds.data(), 0, 1, NoSync{}, v.data(), tl.data());
        }
    }
}
template <uint32_t SP, uint32_t S1>
void emu_moddown_t(const ModupJob& j, const NttGeom& g) {
    std::vector<uint64_t> lds(MODUP_TILE), v((size_t)MODUP_THREADS * MODUP_UNIT), tl(2u << S1);
    std::vector<ModupUnit<SP>> units(MODUP_THREADS);
    std::vector<uint64_t> dv((size_t)MODUP_THREADS * MODUP_UNIT);
    for (uint32_t tile = 0; tile < g.n / MODUP_TILE; ++tile) {
        for (uint32_t i = 0; i < SP; ++i) {
            modup_source_pass<S1, true>(j, i, g, tile, lds.data(), 0, 1, NoSync{}, v.data(), tl.data(), [&] {
                for (uint32_t k = 0; k < (1u << S1); ++k) modup_tw_put(tl.data(), k, modup_src_tw_pair(j, i, k));
            });
            for (uint32_t u = 0; u < MODUP_THREADS; ++u)
                modup_take_unit<SP, S1>(j, i, v.data() + (size_t)u * MODUP_UNIT, units[u]);
        }
        for (uint32_t u = 0; u < MODUP_THREADS; ++u)
            modup_drop_unit<SP, S1>(j, g, tile, u, units[u], dv.data() + (size_t)u * MODUP_UNIT);
        for (uint32_t t = 0; t < j.sc; ++t) {
            for (uint32_t u = 0; u < MODUP_THREADS; ++u)
                modup_rs_unit<SP, S1>(j, t, units[u], dv.data() + (size_t)u * MODUP_UNIT,
                                      v.data() + (size_t)u * MODUP_UNIT);
            for (uint32_t i = 0; i < (1u << S1); ++i) modup_tw_put(tl.data(), i, modup_tw_pair(j, t, i));
            modup_target_pass<S1, true>(j, t, g, tile, lds.data(), 0, 1, NoSync{}, v.data(), tl.data());
        }
    }
}
}  // namespace

// The fused ModDown + rescale prologue of one component (the fused ModUp
// kernel's rescale mode): src[k] (k < K) the inverse row-pass output of P
// tower k (modulus pm[k], inverse tables sw/swp, final factor post/postp);
// the dropped tower q_old with its coefficient-form operands dt, dr, the
// conversion column tabd[K] and P^-1 mod q_old pinv; surviving towers t < sc
// with moduli tq[t], forward tables ttw/ttwp, constants tab[K][sc] and s2[t];
// dst[t] receives target t's forward column-pass output.  ref != 0: the
// unfused order (each P tower's inverse column pass, the prologue kernel's
// arithmetic per position, each target's forward column pass).  f64 != 0
// (fused only): targets below 2^50 take the fp64 passes.
extern "C" void emu_moddown(uint32_t n, uint32_t K, uint32_t sc, const uint64_t* const* src, const uint64_t* pm,
                            const uint64_t* const* sw, const uint64_t* const* swp, const uint64_t* post,
                            const uint64_t* postp, const uint64_t* tq, const uint64_t* const* ttw,
                            const uint64_t* const* ttwp, const uint64_t* tab, const uint64_t* tabd,
                            const uint64_t* dt, const uint64_t* dr, uint64_t q_old, uint64_t pinv,
                            const uint64_t* s2, uint64_t* dst, int ref, int f64) {
    const NttGeom g = make_geom(n, 3);
    std::vector<std::vector<uint64_t>> stab(K), ttab(sc), td(sc);
    for (uint32_t k = 0; k < K; ++k) stab[k] = paired(sw[k], swp[k], n);
    for (uint32_t t = 0; t < sc; ++t) ttab[t] = paired(ttw[t], ttwp[t], n);
    std::vector<RsTarget> tg(sc);
    std::vector<uint64_t> tqn(sc), tred(sc), ttw_addr(sc), ttwd(sc, 0), tqf(2 * (size_t)sc), tsw(2 * (size_t)sc);
    for (uint32_t t = 0; t < sc; ++t) {
        const uint64_t q = tq[t];
        tg[t] = RsTarget{q, host_qneg(q), q > q_old ? q - q_old : (q_old - q) % q, q > q_old ? 1u : 0u, s2[t],
                         host_precon(s2[t], q), host_precon(1, q)};
        rs_sw_consts(q, q_old, s2[t], tsw[2 * t], tsw[2 * t + 1]);
        tqn[t] = host_qneg(q);
        tred[t] = lazy_red_of(q);
        ttw_addr[t] = (uint64_t)(uintptr_t)ttab[t].data();
        const double qd = (double)q, qi = 1.0 / qd;
        std::memcpy(&tqf[2 * t], &qd, 8);
        std::memcpy(&tqf[2 * t + 1], &qi, 8);
        if (f64 && f64_modulus(q)) {
            td[t] = f64_pairs_of(ttab[t].data(), n, q);
            ttwd[t] = (uint64_t)(uintptr_t)td[t].data();
        }
    }
    const uint64_t pinvp = host_precon(pinv, q_old), qd_neg = host_qneg(q_old), half = q_old >> 1;
    if (!ref) {
        ModupJob J{};
        for (uint32_t k = 0; k < K; ++k) {
            J.src[k] = src[k];
            J.stw[k] = stab[k].data();
            J.sq[k] = pm[k];
            J.post[k] = post[k];
            J.postp[k] = postp[k];
        }
        J.dst = dst;
        J.tab = tab;
        J.tq = tq;
        J.tqneg = tqn.data();
        J.ttw = ttw_addr.data();
        J.tred = tred.data();
        J.ttwd = ttwd.data();
        J.tqf = tqf.data();
        J.sp = K;
        J.sc = sc;
        J.dt = dt;
        J.dr = dr;
        J.tabd = tabd;
        J.strided = 1;
        J.tg = tg.data();
        J.tsw = tsw.data();
        J.q_old = q_old;
        J.half = half;
        J.qd_neg = qd_neg;
        J.pinv = pinv;
        J.pinvp = pinvp;
        with_split(g.logN1, g.logN2, [&](auto A, auto) {
            constexpr uint32_t S1 = decltype(A)::value;
            switch (K) {
                case 1: emu_moddown_t<1, S1>(J, g); break;
                case 2: emu_moddown_t<2, S1>(J, g); break;
                case 3: emu_moddown_t<3, S1>(J, g); break;
                case 4: emu_moddown_t<4, S1>(J, g); break;
                case 5: emu_moddown_t<5, S1>(J, g); break;
                case 6: emu_moddown_t<6, S1>(J, g); break;
            }
        });
        return;
    }
    std::vector<uint64_t> x((size_t)K * n), lds(NTT_TILE_PADDED);
    with_split(g.logN1, g.logN2, [&](auto A, auto) {
        constexpr uint32_t S1 = decltype(A)::value;
        const uint32_t tilesA = g.n / NTT_TILE;
        for (uint32_t k = 0; k < K; ++k) {
            std::copy(src[k], src[k] + n, x.begin() + (size_t)k * n);
            NttJob jn{nullptr, x.data() + (size_t)k * n, stab[k].data(), nullptr, pm[k], post[k], postp[k]};
            for (uint32_t tile = 0; tile < tilesA; ++tile)
                phaseA_body<false, false, true, S1, 3>(jn, g, tile, lds.data(), 0, 1, NoSync{});
        }
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t lo[8], hi[8];
            for (uint32_t k = 0; k < K; ++k) {
                lo[k] = split30_lo(x[(size_t)k * n + i]);
                hi[k] = split30_hi(x[(size_t)k * n + i]);
            }
            auto dot = [&](auto b) {
                switch (K) {
                    case 1: return dot30<1>(lo, hi, b);
                    case 2: return dot30<2>(lo, hi, b);
                    case 3: return dot30<3>(lo, hi, b);
                    case 4: return dot30<4>(lo, hi, b);
                    case 5: return dot30<5>(lo, hi, b);
                    default: return dot30<6>(lo, hi, b);
                }
            };
            const uint64_t cd = redc(dot([&](uint32_t k) { return tabd[k]; }), q_old, qd_neg);
            const uint64_t dv = addmod(dt[i], shoup(submod(dr[i], cd, q_old), pinv, pinvp, q_old), q_old);
            for (uint32_t t = 0; t < sc; ++t)
                dst[(size_t)t * n + i] = rs_lazy_value(dv, dv > half, tg[t], dot([&](uint32_t k) { return tab[k * sc + t]; }));
        }
        for (uint32_t t = 0; t < sc; ++t) {
            uint64_t* y = dst + (size_t)t * n;
            NttJob jt{y, y, ttab[t].data(), nullptr, tq[t], 0, 0};
            jt.red = tred[t];
            for (uint32_t tile = 0; tile < tilesA; ++tile) {
                if (jt.red) phaseA_body<true, true, true, S1, 3>(jt, g, tile, lds.data(), 0, 1, NoSync{});
                else phaseA_body<true, false, true, S1, 3>(jt, g, tile, lds.data(), 0, 1, NoSync{});
            }
        }
    });
}

// The fused ModUp over one part: src[i] (i < sp) holds the inverse row-pass
// output of source i (modulus sq[i], tables tw/twp of sw[i]), post/postp its
// final factors; targets j < sc with moduli tq[j] (tables ttw/ttwp[j]), tab
// [sp][sc] in Montgomery form; dst[j] receives target j's column-pass output.
// ref != 0 runs the unfused kernels' order instead: each source's whole
// inverse column pass (standard tiles, radix rmax), the conversion with the
// conversion kernel's arithmetic, each target's forward column pass.
// f64 != 0 (fused only): sources and targets below 2^50 take the fp64 passes.
// c60 != 0 (fused only): the conversion's R = 2^60 column reduction, with the
// constants rescaled from tab (times 2^-4).
extern "C" void emu_modup(uint32_t n, uint32_t sp, uint32_t sc, const uint64_t* const* src, const uint64_t* sq,
                          const uint64_t* const* sw, const uint64_t* const* swp, const uint64_t* post,
                          const uint64_t* postp, const uint64_t* tq, const uint64_t* const* ttw,
                          const uint64_t* const* ttwp, const uint64_t* tab, uint64_t* dst, int ref, uint32_t rmax,
                          int f64, int c60) {
    const NttGeom g = make_geom(n, rmax);
    std::vector<std::vector<uint64_t>> stab(sp), ttab(sc);
    for (uint32_t i = 0; i < sp; ++i) stab[i] = paired(sw[i], swp[i], n);
    for (uint32_t t = 0; t < sc; ++t) ttab[t] = paired(ttw[t], ttwp[t], n);
    std::vector<uint64_t> tqn(sc), tred(sc), ttw_addr(sc);
    for (uint32_t t = 0; t < sc; ++t) {
        tqn[t] = host_qneg(tq[t]);
        tred[t] = lazy_red_of(tq[t]);
        ttw_addr[t] = (uint64_t)(uintptr_t)ttab[t].data();
    }
    if (!ref) {
        // fp64 pairs of the same twiddles (the engine's root tables), as words
        auto f64_pairs = [&](const uint64_t* w, uint64_t q) {
            std::vector<uint64_t> out(2 * (size_t)n);
            for (uint32_t k = 0; k < n; ++k) {
                double wd, wq;
                f64_twiddle(w[k], q, wd, wq);
                std::memcpy(&out[2 * k], &wd, 8);
                std::memcpy(&out[2 * k + 1], &wq, 8);
            }
            return out;
        };
        std::vector<std::vector<uint64_t>> sd(sp), td(sc);
        std::vector<uint64_t> ttwd(sc, 0), tqf(2 * (size_t)sc);
        ModupJob J{};
        for (uint32_t i = 0; i < sp; ++i) {
            J.src[i] = src[i];
            J.stw[i] = stab[i].data();
            J.sq[i] = sq[i];
            J.post[i] = post[i];
            J.postp[i] = postp[i];
            if (f64 && f64_modulus(sq[i])) {
                sd[i] = f64_pairs(sw[i], sq[i]);
                J.stwd[i] = sd[i].data();
                J.sqd[i] = (double)sq[i];
                J.sqinv[i] = 1.0 / J.sqd[i];
                f64_twiddle(post[i], sq[i], J.spost[i], J.spostq[i]);
            }
        }
        for (uint32_t t = 0; t < sc; ++t) {
            const double qd = (double)tq[t], qi = 1.0 / qd;
            std::memcpy(&tqf[2 * t], &qd, 8);
            std::memcpy(&tqf[2 * t + 1], &qi, 8);
            if (f64 && f64_modulus(tq[t])) {
                td[t] = f64_pairs(ttw[t], tq[t]);
                ttwd[t] = (uint64_t)(uintptr_t)td[t].data();
            }
        }
        std::vector<uint64_t> tab60((size_t)sp * sc);
        for (uint32_t i = 0; i < sp; ++i)
            for (uint32_t t = 0; t < sc; ++t)
                tab60[(size_t)i * sc + t] = host_mulmod(tab[(size_t)i * sc + t], host_invmod(16, tq[t]), tq[t]);
        J.dst = dst;
        J.tab = c60 ? tab60.data() : tab;
        J.c60 = c60 ? 1u : 0u;
        J.tq = tq;
        J.tqneg = tqn.data();
        J.ttw = ttw_addr.data();
        J.tred = tred.data();
        J.ttwd = ttwd.data();
        J.tqf = tqf.data();
        J.sp = sp;
        J.sc = sc;
        auto go = [&](auto C60) {
            constexpr bool CC = decltype(C60)::value;
            with_split(g.logN1, g.logN2, [&](auto A, auto) {
                constexpr uint32_t S1 = decltype(A)::value;
                switch (sp) {
                    case 1: emu_modup_t<1, S1, CC>(J, g); break;
                    case 2: emu_modup_t<2, S1, CC>(J, g); break;
                    case 3: emu_modup_t<3, S1, CC>(J, g); break;
                    case 4: emu_modup_t<4, S1, CC>(J, g); break;
                    case 5: emu_modup_t<5, S1, CC>(J, g); break;
                    case 6: emu_modup_t<6, S1, CC>(J, g); break;
                }
            });
        };
        if (c60) go(std::true_type{});
        else go(std::false_type{});
        return;
    }
    std::vector<uint64_t> x((size_t)sp * n), lds(NTT_TILE_PADDED);
    with_radix(rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        with_split(g.logN1, g.logN2, [&](auto A, auto) {
            constexpr uint32_t S1 = decltype(A)::value;
            const uint32_t tilesA = g.n / NTT_TILE;
            for (uint32_t i = 0; i < sp; ++i) {
                std::copy(src[i], src[i] + n, x.begin() + (size_t)i * n);
                NttJob jn{nullptr, x.data() + (size_t)i * n, stab[i].data(), nullptr, sq[i], post[i], postp[i]};
                for (uint32_t tile = 0; tile < tilesA; ++tile)
                    phaseA_body<false, false, true, S1, RM>(jn, g, tile, lds.data(), 0, 1, NoSync{});
            }
            for (uint32_t t = 0; t < sc; ++t) {
                uint64_t* y = dst + (size_t)t * n;
                for (uint32_t k = 0; k < n; ++k) {
                    if (sp <= 8) {
                        uint32_t lo[8], hi[8];
                        for (uint32_t i = 0; i < sp; ++i) {
                            lo[i] = split30_lo(x[(size_t)i * n + k]);
                            hi[i] = split30_hi(x[(size_t)i * n + k]);
                        }
                        U128 acc{0, 0};
                        auto b = [&](uint32_t i) { return tab[(size_t)i * sc + t]; };
                        switch (sp) {
                            case 1: acc = dot30<1>(lo, hi, b); break;
                            case 2: acc = dot30<2>(lo, hi, b); break;
                            case 3: acc = dot30<3>(lo, hi, b); break;
                            case 4: acc = dot30<4>(lo, hi, b); break;
                            case 5: acc = dot30<5>(lo, hi, b); break;
                            case 6: acc = dot30<6>(lo, hi, b); break;
                        }
                        y[k] = redc_lazy(acc, tq[t], tqn[t]);
                    }
                }
                NttJob jt{y, y, ttab[t].data(), nullptr, tq[t], 0, 0};
                jt.red = tred[t];
                for (uint32_t tile = 0; tile < tilesA; ++tile) {
                    if (jt.red) phaseA_body<true, true, true, S1, RM>(jt, g, tile, lds.data(), 0, 1, NoSync{});
                    else phaseA_body<true, false, true, S1, RM>(jt, g, tile, lds.data(), 0, 1, NoSync{});
                }
            }
        });
    });
}
