// The tensor product of two ciphertexts (kernels/tensor.cu:28-51), per
// position and per row tile of an inverse transform.
#pragma once
#include "ntt_engine.hpp"

namespace gf {

// t0 = a0 b0, t1 = a0 b1 + a1 b0, t2 = a1 b1 on tower q (Montgomery
// constants qneg, r2 = R^2 mod q), all arrays tower-contiguous
struct TensorJob {
    const uint64_t *c0a, *c1a, *c0b, *c1b;
    uint64_t *t0, *t1, *t2;
    uint64_t q, qneg, r2;
    uint32_t pairs;   // 1: every pointer 16-byte aligned (two residues per lane and access)
};

GF_HD void tensor_one(const TensorJob& j, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t& t0,
                      uint64_t& t1, uint64_t& t2) {
    t0 = mulmod(a0, b0, j.q, j.qneg, j.r2);
    const U128 m = add128(mul128(a0, b1), mul128(a1, b0));
    t1 = redc(mul128(redc(m, j.q, j.qneg), j.r2), j.q, j.qneg);
    t2 = mulmod(a1, b1, j.q, j.qneg, j.r2);
}

// The tensor product over one row tile (NTT_TILE positions from base), with
// t0, t1, t2 stored as the tensor kernel stores them and t2 also left in LDS
// at lds[padded(k)] for the tile's position k: the layout in which an inverse
// transform's row pass takes its tile (row_tile with IN_LDS).  Two positions
// per 16-byte access (j.pairs), the operands of 4 pairs requested at once.
GF_HD void tensor_tile_lds(const TensorJob& j, uint32_t base, uint64_t* lds, uint32_t tid, uint32_t nthr) {
    constexpr uint32_t NP = NTT_TILE / 2, NB = 4;
    const GF_GLOBAL Words2* a0p = (const GF_GLOBAL Words2*)gcast(j.c0a + base);
    const GF_GLOBAL Words2* a1p = (const GF_GLOBAL Words2*)gcast(j.c1a + base);
    const GF_GLOBAL Words2* b0p = (const GF_GLOBAL Words2*)gcast(j.c0b + base);
    const GF_GLOBAL Words2* b1p = (const GF_GLOBAL Words2*)gcast(j.c1b + base);
    GF_GLOBAL Words2* t0p = (GF_GLOBAL Words2*)gcast(j.t0 + base);
    GF_GLOBAL Words2* t1p = (GF_GLOBAL Words2*)gcast(j.t1 + base);
    GF_GLOBAL Words2* t2p = (GF_GLOBAL Words2*)gcast(j.t2 + base);
    for (uint32_t pb = tid; pb < NP; pb += NB * nthr) {
        Words2 a0[NB], a1[NB], b0[NB], b1[NB];
#pragma unroll
        for (uint32_t e = 0; e < NB; ++e) {
            const uint32_t p = pb + e * nthr;
            if (p < NP) {
                a0[e] = a0p[p];
                a1[e] = a1p[p];
                b0[e] = b0p[p];
                b1[e] = b1p[p];
            }
        }
#pragma unroll
        for (uint32_t e = 0; e < NB; ++e) {
            const uint32_t p = pb + e * nthr;
            if (p >= NP) continue;
            Words2 r0, r1, r2;
#pragma unroll
            for (uint32_t k = 0; k < 2; ++k)
                tensor_one(j, a0[e].w[k], a1[e].w[k], b0[e].w[k], b1[e].w[k], r0.w[k], r1.w[k], r2.w[k]);
            t0p[p] = r0;
            t1p[p] = r1;
            t2p[p] = r2;
            lds[padded(2 * p)] = r2.w[0];
            lds[padded(2 * p + 1)] = r2.w[1];
        }
    }
}

// one row tile's fused job: the tensor product of tower i and the inverse
// transform's row pass of its t2 (r.in == t.t2)
struct TensorIbJob {
    TensorJob t;
    NttJob r;
};

}  // namespace gf
