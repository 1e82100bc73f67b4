// Tile engine for the negacyclic NTT used by every transform in the library.
//
// Convention (the reference's kernels/ntt_ct.cu:19-112 and kernels/intt_ct.cu):
// forward = Cooley-Tukey, standard order in, bit-reversed order out, twiddle of
// stage m, group i = tw[m + i] with tw[k] = psi^brv(k); inverse = Gentleman-
// Sande with psi^-brv(k), followed by a multiply by n^-1 (here folded into the
// last store together with any caller constant).
//
// Decomposition: n = N1 * N2.  The first log2(N1) CT stages only couple
// elements with the same index mod N2 ("columns"); the last log2(N2) stages
// only couple elements inside one contiguous row of N2.  Phase A transforms a
// tile of C columns, phase B a tile of R rows; each phase is one global read
// and one global write of the tower.  Inside a phase the tile lives in LDS and
// is processed in radix-2^r register passes (r <= 3), one LDS round trip per
// pass.  For n <= 4096 the whole tower is one tile (single phase).
//
// Twiddle index for CT stage st (local to the phase), sub-transform s, group i:
//     (G(s) << st) + i,   G(s) = g0 + s * gstep
// phase A / single phase: G = 1 (the table prefix, stage m = 2^st);
// phase B, row a:         G = N1 + a   (global stage m = N1 2^st).
#pragma once
#include "modarith.hpp"

namespace gf {

struct NttJob {
    const uint64_t* in;   // tower read by the first phase
    uint64_t* out;        // tower written by every phase (may equal in)
    const uint64_t* tw;   // psi^brv (forward) or psi^-brv (inverse), n entries
    const uint64_t* twp;  // Shoup precons of tw
    uint64_t q;
    uint64_t post, postp; // inverse only: final multiplier (n^-1 * c) and precon
};

struct NttGeom {
    uint32_t n, logn;
    uint32_t logN1, logN2;  // phase A covers logN1 stages, phase B logN2
    uint32_t C;             // columns per phase-A tile
    uint32_t R;             // rows per phase-B tile
    uint32_t single;        // 1: whole tower in one tile
};

enum { NTT_LOG_TILE = 12, NTT_TILE = 1 << NTT_LOG_TILE, NTT_SINGLE_MAX = 4096, NTT_THREADS = 256 };

inline NttGeom make_geom(uint32_t n) {
    NttGeom g{};
    g.n = n;
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

// radix-2^r chunking of S stages: ceil(S/3) chunks, the larger ones first
GF_HD uint32_t num_chunks(uint32_t S) { return (S + 2) / 3; }
GF_HD uint32_t chunk_size(uint32_t S, uint32_t c) {
    uint32_t nc = num_chunks(S), base = S / nc, extra = S % nc;
    return base + (c < extra ? 1u : 0u);
}
GF_HD uint32_t chunk_start(uint32_t S, uint32_t c) {
    uint32_t st = 0;
    for (uint32_t i = 0; i < c; ++i) st += chunk_size(S, i);
    return st;
}

// One radix-2^r pass over every unit of the tile owned by this thread.
// Element k of sub-transform s lives at a[k*KS + s*SS].
template <bool FWD>
GF_HD void tile_pass(uint64_t* a, uint32_t S, uint32_t st0, uint32_t r, uint32_t logT,
                     uint32_t KS, uint32_t SS, const uint64_t* tw, const uint64_t* twp,
                     uint64_t g0, uint32_t gstep, uint64_t q, bool sub_fastest, uint32_t tid,
                     uint32_t nthr) {
    const uint32_t T = 1u << logT;
    const uint32_t logtl = S - st0 - r;          // log2(t_last)
    const uint32_t tl = 1u << logtl;
    const uint32_t per_sub = 1u << (S - r);     // units per sub-transform
    const uint32_t U = T * per_sub;
    const uint32_t R8 = 1u << r;
    for (uint32_t uid = tid; uid < U; uid += nthr) {
        uint32_t s, rest;
        if (sub_fastest) {
            s = uid & (T - 1);
            rest = uid >> logT;
        } else {
            rest = uid & (per_sub - 1);
            s = uid >> (S - r);
        }
        const uint32_t l = rest & (tl - 1);
        const uint32_t g = rest >> logtl;
        const uint32_t k0 = (g << (S - st0)) + l;
        uint64_t* base = a + (uint64_t)k0 * KS + (uint64_t)s * SS;
        const uint64_t step = (uint64_t)tl * KS;
        uint64_t x[8];
        for (uint32_t u = 0; u < R8; ++u) x[u] = base[u * step];
        const uint64_t G = g0 + (uint64_t)s * gstep;
        if (FWD) {
            for (uint32_t c = 0; c < r; ++c) {
                const uint32_t h = 1u << (r - 1 - c);
                for (uint32_t u = 0; u < R8; ++u) {
                    if (u & h) continue;
                    const uint64_t ti = (G << (st0 + c)) + ((uint64_t)g << c) + (u >> (r - c));
                    const uint64_t y = shoup(x[u + h], tw[ti], twp[ti], q);
                    const uint64_t xv = x[u];
                    x[u] = addmod(xv, y, q);
                    x[u + h] = submod(xv, y, q);
                }
            }
        } else {
            for (int c = (int)r - 1; c >= 0; --c) {
                const uint32_t h = 1u << (r - 1 - c);
                for (uint32_t u = 0; u < R8; ++u) {
                    if (u & h) continue;
                    const uint64_t ti = (G << (st0 + c)) + ((uint64_t)g << c) + (u >> (r - c));
                    const uint64_t xv = x[u], yv = x[u + h];
                    x[u] = addmod(xv, yv, q);
                    x[u + h] = shoup(submod(xv, yv, q), tw[ti], twp[ti], q);
                }
            }
        }
        for (uint32_t u = 0; u < R8; ++u) base[u * step] = x[u];
    }
}

// All passes of one phase.  Sync is a functor run between passes (the
// workgroup barrier on the device, nothing in the sequential host emulation).
template <bool FWD, class Sync>
GF_HD void tile_transform(uint64_t* a, uint32_t S, uint32_t logT, uint32_t KS, uint32_t SS,
                          const uint64_t* tw, const uint64_t* twp, uint64_t g0, uint32_t gstep,
                          uint64_t q, bool sub_fastest, uint32_t tid, uint32_t nthr, Sync sync) {
    const uint32_t nc = num_chunks(S);
    for (uint32_t i = 0; i < nc; ++i) {
        const uint32_t c = FWD ? i : nc - 1 - i;
        tile_pass<FWD>(a, S, chunk_start(S, c), chunk_size(S, c), logT, KS, SS, tw, twp, g0, gstep, q,
                       sub_fastest, tid, nthr);
        sync();
    }
}

// ---- phase bodies (shared by the kernels and the host emulation) ----
// Phase A: tile = columns [b0, b0+C) x all N1 rows.
template <bool FWD, class Sync>
GF_HD void phaseA_body(const NttJob& j, const NttGeom& g, uint32_t tile, uint64_t* lds,
                       uint32_t tid, uint32_t nthr, Sync sync) {
    const uint32_t C = g.C, N1 = 1u << g.logN1, logN2 = g.logN2, logC = NTT_LOG_TILE - g.logN1;
    const uint32_t b0 = tile * C;
    const uint64_t* src = FWD ? j.in : j.out;   // forward: first phase; inverse: second
    for (uint32_t idx = tid; idx < C * N1; idx += nthr) {
        const uint32_t row = idx >> logC, c = idx & (C - 1);
        lds[idx] = src[((uint64_t)row << logN2) + b0 + c];
    }
    sync();
    tile_transform<FWD>(lds, g.logN1, logC, C, 1, j.tw, j.twp, 1, 0, j.q, true, tid, nthr, sync);
    for (uint32_t idx = tid; idx < C * N1; idx += nthr) {
        const uint32_t row = idx >> logC, c = idx & (C - 1);
        uint64_t v = lds[idx];
        if (!FWD) v = shoup(v, j.post, j.postp, j.q);
        j.out[((uint64_t)row << logN2) + b0 + c] = v;
    }
}

// Phase B: tile = rows [a0, a0+R), each a contiguous run of N2.
template <bool FWD, class Sync>
GF_HD void phaseB_body(const NttJob& j, const NttGeom& g, uint32_t tile, uint64_t* lds,
                       uint32_t tid, uint32_t nthr, Sync sync) {
    const uint32_t R = g.R, N2 = 1u << g.logN2, N1 = 1u << g.logN1;
    const uint32_t a0 = tile * R;
    const uint64_t* src = FWD ? j.out : j.in;   // forward: second phase; inverse: first
    const uint64_t off = (uint64_t)a0 << g.logN2;
    for (uint32_t idx = tid; idx < R * N2; idx += nthr) lds[idx] = src[off + idx];
    sync();
    tile_transform<FWD>(lds, g.logN2, NTT_LOG_TILE - g.logN2, 1, N2, j.tw, j.twp, (uint64_t)N1 + a0,
                        1, j.q, false, tid, nthr, sync);
    for (uint32_t idx = tid; idx < R * N2; idx += nthr) j.out[off + idx] = lds[idx];
}

// Single phase: the whole tower (n <= NTT_SINGLE_MAX) is one tile.
template <bool FWD, class Sync>
GF_HD void single_body(const NttJob& j, const NttGeom& g, uint64_t* lds, uint32_t tid,
                       uint32_t nthr, Sync sync) {
    const uint32_t n = g.n;
    for (uint32_t idx = tid; idx < n; idx += nthr) lds[idx] = j.in[idx];
    sync();
    tile_transform<FWD>(lds, g.logn, 0, 1, n, j.tw, j.twp, 1, 0, j.q, false, tid, nthr, sync);
    for (uint32_t idx = tid; idx < n; idx += nthr) {
        uint64_t v = lds[idx];
        if (!FWD) v = shoup(v, j.post, j.postp, j.q);
        j.out[idx] = v;
    }
}

}  // namespace gf
