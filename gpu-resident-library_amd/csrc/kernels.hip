// HIP kernels for gfx950 (CDNA4): the pointwise RNS kernels of the multiply /
// keyswitch / rescale / rotate path, and the reference's single-tower
// extern "C" launchers built on the same code (the transforms are in
// kernels_ntt_*.hip, the fused row pass + inner product in kernels_ksi.hip).
//
// All arithmetic is exact integer work on the VALU (64-bit products assembled
// from 32-bit multiplies, see modarith.hpp).  Wave64, 256-thread workgroups.
#include <hip/hip_runtime.h>

#include "../../include/gpufhe_kernels.h"
#include "ctr_sampler.hpp"
#include "ntt_kernels.hpp"

namespace gf {

// ------------------------------------------------------ pointwise kernels --
#define PW_FOR(n)                                                                      \
    for (uint32_t e_ = 0; e_ < PW_EPT; ++e_)                                           \
        for (uint32_t i = (blockIdx.x * PW_EPT + e_) * PW_THREADS + threadIdx.x; i < (n); \
             i = 0xFFFFFFFFu)

// The same four elements per lane as PW_FOR, taken as two adjacent pairs
// (i, i+1) with i even: one 16-byte access per pair halves the memory
// instructions of the address and data path (TA/TD), which bounds the
// streaming kernels with several operand arrays.  Needs an even n and
// 16-byte-aligned arrays (the jobs' `pairs` flag).
#define PW2_FOR(n)                                                                        \
    for (uint32_t e_ = 0; e_ < PW_EPT / 2; ++e_)                                          \
        for (uint32_t i = ((blockIdx.x * (PW_EPT / 2) + e_) * PW_THREADS + threadIdx.x) * 2; \
             i < (n); i = 0xFFFFFFFFu)

// Word i of a job's array as a global access with the job's (workgroup-
// uniform) base and a 32-bit byte offset (gat): no flat access, no 64-bit
// address per element.  Every array a job names spans fewer than 2^29 words.
template <class T>
__device__ __forceinline__ const T& ga(const T* p, uint32_t i) { return gat(gcast(p), i); }
template <class T>
__device__ __forceinline__ T& ga(T* p, uint32_t i) { return gat(gcast(p), i); }

struct alignas(16) Pair {
    uint64_t v[2];
};
__device__ __forceinline__ Pair ld2(const uint64_t* p, uint32_t i) { return ga((const Pair*)p, i / 2); }
__device__ __forceinline__ void st2(uint64_t* p, uint32_t i, const Pair& x) { ga((Pair*)p, i / 2) = x; }

__device__ __forceinline__ void tensor_body(const TensorJob& j, uint32_t n) {
    if (j.pairs) {
        PW2_FOR(n) {
            const Pair a0 = ld2(j.c0a, i), a1 = ld2(j.c1a, i), b0 = ld2(j.c0b, i), b1 = ld2(j.c1b, i);
            Pair t0, t1, t2;
#pragma unroll
            for (int k = 0; k < 2; ++k) tensor_one(j, a0.v[k], a1.v[k], b0.v[k], b1.v[k], t0.v[k], t1.v[k], t2.v[k]);
            st2(j.t0, i, t0);
            st2(j.t1, i, t1);
            st2(j.t2, i, t2);
        }
    } else {
        PW_FOR(n) {
            uint64_t t0, t1, t2;
            tensor_one(j, ga(j.c0a, i), ga(j.c1a, i), ga(j.c0b, i), ga(j.c1b, i), t0, t1, t2);
            ga(j.t0, i) = t0;
            ga(j.t1, i) = t1;
            ga(j.t2, i) = t2;
        }
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_tensor(const TensorJob* __restrict__ jobs,
                                                       uint32_t off, uint32_t n) {
    tensor_body(jobs[off + blockIdx.y], n);
}

// Tables that every lane of a block reads at the same index (conversion
// matrices, target moduli, job pointers) are read through the constant
// address space, so they become scalar loads shared by the wave.
typedef const __attribute__((address_space(4))) uint64_t* cu64p;
__device__ __forceinline__ cu64p uni(const uint64_t* p) { return (cu64p)p; }

// Calls f(integral_constant<NS>) for a source count 1..CONV_UNROLLED known
// only at run time (uniform per block), so the source column lives in
// registers and every multiply-accumulate is unrolled; larger counts take
// f(integral_constant<0>), a loop that re-reads its sources.  (The register
// budget of a kernel is that of its largest case, hence the cap.)
enum { CONV_UNROLLED = 8 };
template <class F>
__device__ __forceinline__ void with_nsrc(uint32_t ns, F&& f) {
    switch (ns) {
#define GF_NS(k) case k: f(std::integral_constant<uint32_t, k>{}); break;
        GF_NS(1) GF_NS(2) GF_NS(3) GF_NS(4) GF_NS(5) GF_NS(6) GF_NS(7) GF_NS(8)
#undef GF_NS
        default: f(std::integral_constant<uint32_t, 0>{}); break;
    }
}

// The sources of one element for a templated count NS, split at bit 30 for
// dot30 (contexts whose moduli are all below 2^60); NS = 0 holds nothing.
template <uint32_t NS>
struct SrcCols {
    uint32_t lo[NS > 0 ? NS : 1], hi[NS > 0 ? NS : 1];
    __device__ __forceinline__ void load(const uint64_t* src, uint32_t n, uint32_t i) {
#pragma unroll
        for (uint32_t s = 0; s < NS; ++s) {
            const uint64_t v = ga(src, s * n + i);
            lo[s] = split30_lo(v);
            hi[s] = split30_hi(v);
        }
    }
};

// sum_s x_s tab[s*stride + t] as 128 bits (Montgomery constants in tab, one
// REDC by the caller).  NS > 0: column sums over the pre-split sources;
// NS = 0 (more than 8 sources, or a modulus of 2^60 or more): runtime count,
// full 128-bit products, sources read from memory.
template <uint32_t NS>
__device__ __forceinline__ U128 dot_src(const uint64_t* src, uint32_t n, uint32_t i, uint32_t ns, cu64p tab,
                                        uint32_t stride, uint32_t t, const SrcCols<NS>& x) {
    if constexpr (NS > 0) {
        return dot30<NS>(x.lo, x.hi, [&](uint32_t s) { return tab[s * stride + t]; });
    } else {
        U128 acc{0, 0};
        for (uint32_t s = 0; s < ns; ++s) acc = add128(acc, mul128(ga(src, s * n + i), tab[s * stride + t]));
        return acc;
    }
}

// Targets are left in [0, 2q) (REDC without its correction) where the sums
// come from dot30 (at most 8 sources below 2^60: the sum is below 2^63 q, so
// the uncorrected REDC is below 1.5 q + 1); every consumer of a conversion is
// a forward transform, whose butterflies take inputs below 4q.  The generic
// path (up to 16 sources, moduli up to 2^61) reduces fully.
template <uint32_t NS>
__device__ __forceinline__ void conv_one(const ConvJob& j, uint32_t n, uint32_t i) {
    SrcCols<NS> x;
    x.load(j.src, n, i);
    const cu64p tab = uni(j.tab), tq = uni(j.tq), tqn = uni(j.tqneg);
    for (uint32_t t = j.j0; t < j.j1; ++t) {
        const U128 acc = dot_src<NS>(j.src, n, i, j.nsrc, tab, j.ntgt, t, x);
        ga(j.dst, t * n + i) = NS > 0 ? redc_lazy(acc, tq[t], tqn[t]) : redc(acc, tq[t], tqn[t]);
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_conv(const ConvJob* __restrict__ jobs,
                                                     uint32_t off, uint32_t n) {
    const ConvJob j = jobs[off + blockIdx.y];
    with_nsrc(j.wide ? 0 : j.nsrc, [&](auto NSc) {
        constexpr uint32_t NS = decltype(NSc)::value;
        PW_FOR(n) { conv_one<NS>(j, n, i); }
    });
}

// An accumulated inner-product sum to [0, q).  With j.red the digits may be
// the forward transforms' unreduced values (below 72q, NttJob::lazy_out):
// the sum stays below 16 * 72 q^2 < 2^125 for q < 2^57, REDC without its
// correction leaves less than 2^62, and reduce_red finishes.
__device__ __forceinline__ uint64_t inner_out(const U128& T, const InnerJob& j) {
    if (j.red) return redc_red(T, j.q, j.qneg, (uint32_t)j.red);
    return redc(T, j.q, j.qneg);
}

__device__ __forceinline__ uint32_t brv_bits(uint32_t v, uint32_t logn) {
    return logn ? (__brev(v) >> (32 - logn)) : 0;
}
// evaluation-domain automorphism of index k: out[i] = in[perm_index(i)]
// (slots in bit-reversed order; k odd, below 2n)
__device__ __forceinline__ uint32_t perm_index(uint32_t i, uint64_t k, uint32_t n, uint32_t logn) {
    const uint64_t e = ((2ull * brv_bits(i, logn) + 1) * k) & (2ull * n - 1);
    return brv_bits((uint32_t)((e - 1) >> 1), logn);
}

__global__ __launch_bounds__(PW_THREADS) void k_inner(const InnerJob* __restrict__ jobs,
                                                      uint32_t off, uint32_t n) {
    const InnerJob& j = jobs[off + blockIdx.y];
    const uint32_t np = j.np;
    uint64_t* o0 = j.o0;
    uint64_t* o1 = j.o1;
    if (j.gk) {
        const uint32_t logn = 31 - __builtin_clz(n);
        PW_FOR(n) {
            const uint32_t src = perm_index(i, j.gk, n, logn);
            U128 a0{0, 0}, a1{0, 0};
            for (uint32_t p = 0; p < np; ++p) {
                const uint64_t d = ga(j.dig[p], src);
                a0 = add128(a0, mul128(d, ga(j.kb[p], i)));
                a1 = add128(a1, mul128(d, ga(j.ka[p], i)));
            }
            ga(o0, i) = inner_out(a0, j);
            ga(o1, i) = inner_out(a1, j);
        }
    } else if (j.pairs) {
        PW2_FOR(n) {
            U128 a0[2] = {{0, 0}, {0, 0}}, a1[2] = {{0, 0}, {0, 0}};
            for (uint32_t p = 0; p < np; ++p) {
                const Pair d = ld2(j.dig[p], i), kb = ld2(j.kb[p], i), ka = ld2(j.ka[p], i);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    a0[k] = add128(a0[k], mul128(d.v[k], kb.v[k]));
                    a1[k] = add128(a1[k], mul128(d.v[k], ka.v[k]));
                }
            }
            st2(o0, i, Pair{{inner_out(a0[0], j), inner_out(a0[1], j)}});
            st2(o1, i, Pair{{inner_out(a1[0], j), inner_out(a1[1], j)}});
        }
    } else {
        PW_FOR(n) {
            U128 a0{0, 0}, a1{0, 0};
            for (uint32_t p = 0; p < np; ++p) {
                const uint64_t d = ga(j.dig[p], i);
                a0 = add128(a0, mul128(d, ga(j.kb[p], i)));
                a1 = add128(a1, mul128(d, ga(j.ka[p], i)));
            }
            ga(o0, i) = inner_out(a0, j);
            ga(o1, i) = inner_out(a1, j);
        }
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_combine(const CombineJob* __restrict__ jobs,
                                                        uint32_t off, uint32_t n) {
    const CombineJob j = jobs[off + blockIdx.y];
    PW_FOR(n) {
        uint64_t v = shoup(submod(ga(j.res, i), ga(j.conv, i), j.q), j.pinv, j.pinvp, j.q);
        if (j.t) v = addmod(v, ga(j.t, i), j.q);
        ga(j.out, i) = v;
    }
}

__device__ __forceinline__ uint64_t switch_one(uint64_t x, const SwitchJob& j) {
    const uint64_t d = x > j.half ? j.diff : 0;
    if (j.new_larger) return x + d;
    const uint64_t xm = shoup(x, 1, j.red, j.q_new);   // x mod q_new, exact
    return xm >= d ? xm - d : xm + j.q_new - d;
}

__global__ __launch_bounds__(PW_THREADS) void k_switch(const SwitchJob* __restrict__ jobs,
                                                       uint32_t off, uint32_t n) {
    const SwitchJob j = jobs[off + blockIdx.y];
    PW_FOR(n) { ga(j.out, i) = switch_one(ga(j.in, i), j); }
}

__global__ __launch_bounds__(PW_THREADS) void k_fuse(const FuseJob* __restrict__ jobs,
                                                     uint32_t off, uint32_t n) {
    const FuseJob j = jobs[off + blockIdx.y];
    PW_FOR(n) {
        ga(j.out, i) = addmod(shoup(ga(j.a, i), j.s1, j.s1p, j.q), shoup(ga(j.b, i), j.s2, j.s2p, j.q), j.q);
    }
}

template <uint32_t K>
__device__ __forceinline__ void rs_one(const RsJob& j, uint32_t n, uint32_t i) {
    SrcCols<K> x;
    x.load(j.xp, n, i);
    // the dropped tower, from the sources already in registers
    const uint64_t cd = redc(dot_src<K>(j.xp, n, i, j.K, uni(j.tabd), j.strided, 0, x), j.q_old, j.qd_neg);
    const uint64_t dv =
        addmod(ga(j.dt, i), shoup(submod(ga(j.dr, i), cd, j.q_old), j.pinv, j.pinvp, j.q_old), j.q_old);
    const bool hi = dv > j.half;
    const cu64p tab = uni(j.tab);
    static_assert(sizeof(RsTarget) == 7 * 8, "RsTarget layout");
    for (uint32_t t = j.j0; t < j.j1; ++t) {
        const cu64p w = uni((const uint64_t*)j.tg) + 7 * t;
        RsTarget tg;
        tg.q = w[0];
        tg.qneg = w[1];
        tg.diff = w[2];
        tg.new_larger = w[3];
        tg.s2 = w[4];
        tg.s2p = w[5];
        tg.red = w[6];
        const U128 acc = dot_src<K>(j.xp, n, i, j.K, tab, j.ntgt, t, x);
        if constexpr (K > 0) {
            // the output only feeds the epilogue's forward transform, which
            // takes inputs below 4q (the fused ModDown computes the same value)
            ga(j.out, t * n + i) = rs_lazy_value(dv, hi, tg, acc);
            continue;
        }
        const uint64_t conv = redc(acc, tg.q, tg.qneg);
        // centred switch q_old -> q_t (kernels/switch_modulus.cu:11-27 semantics);
        // dv mod q_t by a Shoup reduction with w = 1 (exact for any dv < 2^64)
        uint64_t sw;
        if (tg.new_larger) {
            sw = dv + (hi ? tg.diff : 0);
        } else {
            const uint64_t d = hi ? tg.diff : 0;
            const uint64_t xm = shoup(dv, 1, tg.red, tg.q);
            sw = xm >= d ? xm - d : xm + tg.q - d;
        }
        ga(j.out, t * n + i) = addmod(shoup(sw, tg.s2, tg.s2p, tg.q), conv, tg.q);
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_rs(const RsJob* __restrict__ jobs, uint32_t off,
                                                   uint32_t n) {
    const RsJob j = jobs[off + blockIdx.y];
    with_nsrc(j.wide ? 0 : j.K, [&](auto Kc) {
        constexpr uint32_t K = decltype(Kc)::value;
        PW_FOR(n) { rs_one<K>(j, n, i); }
    });
}

__global__ __launch_bounds__(PW_THREADS) void k_embed(const EmbedJob* __restrict__ jobs,
                                                      uint32_t off, uint32_t n) {
    const EmbedJob j = jobs[off + blockIdx.y];
    PW_FOR(n) {
        const int64_t v = j.src[i];
        const uint64_t m = (uint64_t)(v < 0 ? -v : v);
        const uint64_t r = m < j.q ? m : m % j.q;   // |v| is small: the remainder is not taken in practice
        j.dst[i] = (v < 0 && r) ? j.q - r : r;
    }
}

// The counter sampler (ctr_sampler.hpp) over one polynomial per job: every
// coefficient drawn on its own; the Gaussian table is read through the
// constant address space (all lanes read the same threshold).
__global__ __launch_bounds__(PW_THREADS) void k_sample(const SampleJob* __restrict__ jobs, uint32_t off, uint32_t n,
                                                       uint64_t seed, const uint64_t* __restrict__ T, uint32_t K) {
    const SampleJob j = jobs[off + blockIdx.y];
    const cu64p tab = (cu64p)T;
    PW_FOR(n) {
        const uint64_t r = ctr_draw(seed, j.domain, i);
        if (j.kind == 0) j.coef[i] = ctr_ternary(r);
        else if (j.kind == 1) j.coef[i] = ctr_gauss(r, K, [&](uint32_t x) { return tab[x]; });
        else j.res[i] = reduce_red(r, j.q, (uint32_t)j.red);
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_pkgen(const PkJob* __restrict__ jobs, uint32_t off, uint32_t n) {
    const PkJob j = jobs[off + blockIdx.y];
    PW_FOR(n) {
        j.b[i] = submod(shoup(j.e[i], j.ns, j.nsp, j.q), mulmod(j.a[i], j.s[i], j.q, j.qneg, j.r2), j.q);
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_ekgen(const EkJob* __restrict__ jobs,
                                                      uint32_t off, uint32_t n) {
    const EkJob j = jobs[off + blockIdx.y];
    PW_FOR(n) {
        const uint64_t u = j.u[i];
        uint64_t a = addmod(mulmod(j.pka[i], u, j.q, j.qneg, j.r2), shoup(j.e1[i], j.ns, j.nsp, j.q), j.q);
        uint64_t b = addmod(mulmod(j.pkb[i], u, j.q, j.qneg, j.r2), shoup(j.e0[i], j.ns, j.nsp, j.q), j.q);
        if (j.sold) b = addmod(b, shoup(j.sold[i], j.pq, j.pqp, j.q), j.q);
        if (j.plain) {
            j.a_out[i] = a;
            j.b_out[i] = b;
        } else {
            // Montgomery form: REDC(x R^2) = x R
            j.a_out[i] = redc(mul128(a, j.r2), j.q, j.qneg);
            j.b_out[i] = redc(mul128(b, j.r2), j.q, j.qneg);
        }
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_dec(const DecJob* __restrict__ jobs, uint32_t off, uint32_t n) {
    const DecJob j = jobs[off + blockIdx.y];
    PW_FOR(n) { ga(j.out, i) = addmod(ga(j.c0, i), mulmod(ga(j.c1, i), ga(j.s, i), j.q, j.qneg, j.r2), j.q); }
}

__global__ __launch_bounds__(PW_THREADS) void k_ptacc(const PtAccJob* __restrict__ jobs,
                                                      uint32_t off, uint32_t n) {
    const PtAccJob& j = jobs[off + blockIdx.y];
    const uint32_t m = j.m;
    const uint64_t q = j.q, qneg = j.qneg, r2 = j.r2;
    PW_FOR(n) {
        U128 a0{0, 0}, a1{0, 0};
        for (uint32_t t = 0; t < m; ++t) {
            const uint64_t p = ga(j.pt[t], i);
            a0 = add128(a0, mul128(p, ga(j.x0[t], i)));
            a1 = add128(a1, mul128(p, ga(j.x1[t], i)));
        }
        // REDC(REDC(sum) R^2) = sum mod q (needs m q < 2^64)
        ga(j.o0, i) = redc(mul128(redc(a0, q, qneg), r2), q, qneg);
        ga(j.o1, i) = redc(mul128(redc(a1, q, qneg), r2), q, qneg);
    }
}

__global__ __launch_bounds__(PW_THREADS) void k_add(const AddJob* __restrict__ jobs,
                                                    uint32_t off, uint32_t n) {
    const AddJob j = jobs[off + blockIdx.y];
    PW_FOR(n) { ga(j.out, i) = addmod(ga(j.a, i), ga(j.b, i), j.q); }
}

struct PermList {
    const PermJob* p;
    uint32_t off;
    __device__ PermJob get(uint32_t i) const { return p[off + i]; }
};
struct PermOne {
    PermJob j;
    __device__ PermJob get(uint32_t) const { return j; }
};
template <class J>
__global__ __launch_bounds__(PW_THREADS) void k_perm(J jobs, uint32_t n, uint32_t logn) {
    const PermJob j = jobs.get(blockIdx.y);
    PW_FOR(n) { ga(j.out, i) = ga(j.in, perm_index(i, j.k, n, logn)); }
}

__global__ __launch_bounds__(PW_THREADS) void k_copy(const CopyJob* __restrict__ jobs,
                                                     uint32_t off, uint32_t n) {
    const CopyJob j = jobs[off + blockIdx.y];
    PW_FOR(n) { ga(j.out, i) = ga(j.in, i); }
}

// synthetic uniform residues: splitmix64 of (seed, tower, index) reduced mod q
__global__ __launch_bounds__(PW_THREADS) void k_fill(uint64_t* d, const uint64_t* __restrict__ mods,
                                                     uint32_t n, uint64_t seed) {
    const uint32_t t = blockIdx.y;
    const uint64_t q = mods[t];
    PW_FOR(n) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)t * n + i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        d[(uint64_t)t * n + i] = z % q;
    }
}

static dim3 pw_grid(uint32_t n, uint32_t nj) {
    return dim3((n + PW_THREADS * PW_EPT - 1) / (PW_THREADS * PW_EPT), nj);
}

#define GF_PW_LAUNCH(fn, kern, T)                                               \
    void fn(const T* d, uint32_t nj, uint32_t n, hipStream_t s) {               \
        for (uint32_t off = 0; off < nj; off += MAXY) {                         \
            uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;                   \
            kern<<<pw_grid(n, cnt), PW_THREADS, 0, s>>>(d, off, n);             \
        }                                                                       \
    }
GF_PW_LAUNCH(launch_tensor, k_tensor, TensorJob)
GF_PW_LAUNCH(launch_conv, k_conv, ConvJob)
GF_PW_LAUNCH(launch_inner, k_inner, InnerJob)
GF_PW_LAUNCH(launch_combine, k_combine, CombineJob)
GF_PW_LAUNCH(launch_switch, k_switch, SwitchJob)
GF_PW_LAUNCH(launch_fuse, k_fuse, FuseJob)
GF_PW_LAUNCH(launch_add, k_add, AddJob)
GF_PW_LAUNCH(launch_copy, k_copy, CopyJob)
GF_PW_LAUNCH(launch_rs, k_rs, RsJob)
GF_PW_LAUNCH(launch_embed, k_embed, EmbedJob)
GF_PW_LAUNCH(launch_ekgen, k_ekgen, EkJob)
GF_PW_LAUNCH(launch_pkgen, k_pkgen, PkJob)

void launch_sample(const SampleJob* d, uint32_t nj, uint32_t n, uint64_t seed, const uint64_t* T, uint32_t K,
                   hipStream_t s) {
    for (uint32_t off = 0; off < nj; off += MAXY) {
        const uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;
        k_sample<<<pw_grid(n, cnt), PW_THREADS, 0, s>>>(d, off, n, seed, T, K);
    }
}
GF_PW_LAUNCH(launch_dec, k_dec, DecJob)
GF_PW_LAUNCH(launch_ptacc, k_ptacc, PtAccJob)

void launch_fill(uint64_t* d, const uint64_t* d_mods, uint32_t towers, uint32_t n, uint64_t seed,
                 hipStream_t s) {
    if (towers) k_fill<<<pw_grid(n, towers), PW_THREADS, 0, s>>>(d, d_mods, n, seed);
}

void launch_perm(const PermJob* d, uint32_t nj, uint32_t n, uint32_t logn, hipStream_t s) {
    for (uint32_t off = 0; off < nj; off += MAXY) {
        uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;
        k_perm<<<pw_grid(n, cnt), PW_THREADS, 0, s>>>(PermList{d, off}, n, logn);
    }
}
void launch_perm_one(const PermJob& j, uint32_t n, uint32_t logn, hipStream_t s) {
    k_perm<<<pw_grid(n, 1), PW_THREADS, 0, s>>>(PermOne{j}, n, logn);
}

// ------------------------------------------- reference single-tower ABI ---
// Montgomery constants of an odd modulus q < 2^63, derived on the device (the
// reference's launcher signatures pass the moduli in device memory only):
// qneg = -q^-1 mod 2^64 by Newton's iteration (x = q is q^-1 to 3 bits, as
// q^2 = 1 mod 8, and each step doubles the correct bits: 6 steps give 64+);
// r2 = 2^128 mod q as REDC(r * r) with r = 2^96 mod q, itself 2^b mod q
// (2^(b-1) <= q < 2^b, so 2^b - q) doubled 96 - b times.
__device__ __forceinline__ void mont_consts(uint64_t q, uint64_t& qneg, uint64_t& r2) {
    uint64_t x = q;
    for (int k = 0; k < 6; ++k) x *= 2 - q * x;
    qneg = (uint64_t)0 - x;
    const uint32_t b = 64 - (uint32_t)__builtin_clzll(q);
    uint64_t r = (1ull << b) - q;
    for (uint32_t k = b; k < 96; ++k) r = csub(r + r, q);
    r2 = redc(mul128(r, r), q, qneg);
}

// The modulus of this workgroup's tower and its Montgomery constants, derived
// once per workgroup by one lane and shared through LDS.
struct WgMont {
    uint64_t q, qneg, r2;
};
__device__ __forceinline__ WgMont wg_mont(const uint64_t* mods, uint32_t t) {
    __shared__ uint64_t sh[3];
    if (threadIdx.x == 0) {
        const uint64_t q = mods[t];
        uint64_t qneg, r2;
        mont_consts(q, qneg, r2);
        sh[0] = q;
        sh[1] = qneg;
        sh[2] = r2;
    }
    __syncthreads();
    return WgMont{sh[0], sh[1], sh[2]};
}
__device__ __forceinline__ bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// exact 128-bit remainder with mu = floor((2^128-1)/p): Barrett quotient from
// the top limbs of a*mu, error at most a few p, removed by the final loop
__device__ __forceinline__ uint64_t barrett128(U128 a, uint64_t p, uint64_t mu_lo, uint64_t mu_hi) {
    const U128 ll = mul128(a.lo, mu_lo);
    const U128 lh = mul128(a.lo, mu_hi);
    const U128 hl = mul128(a.hi, mu_lo);
    U128 mid{ll.hi, 0};
    mid = add128(mid, U128{lh.lo, 0});
    mid = add128(mid, U128{hl.lo, 0});
    const uint64_t qhat = a.hi * mu_hi + lh.hi + hl.hi + mid.hi;
    uint64_t r = a.lo - qhat * p;
    while (r >= p) r -= p;
    return r;
}

__device__ uint64_t abi_conv_barrett(const uint64_t* src, const uint64_t* qhi, const uint64_t* qhip,
                                     const uint64_t* q, const uint64_t* qhmp, uint64_t p, uint64_t mlo,
                                     uint64_t mhi, uint32_t sizeQ, uint32_t sizeP, uint32_t t, uint32_t n,
                                     uint32_t r) {
    U128 acc{0, 0};
    for (uint32_t i = 0; i < sizeQ; ++i) {
        const uint64_t xi = shoup(src[(uint64_t)i * n + r], qhi[i], qhip[i], q[i]);
        acc = add128(acc, mul128(xi, qhmp[(uint64_t)i * sizeP + t]));
    }
    return barrett128(acc, p, mlo, mhi);
}

__global__ void k_ref_mult(const uint64_t* a, const uint64_t* b, uint64_t* r, uint64_t q,
                           uint64_t qneg, uint64_t r2, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = mulmod(a[i], b[i], q, qneg, r2);
}
__global__ void k_ref_switch(uint64_t* v, SwitchJob j, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = switch_one(v[i], j);
}
__global__ void k_ref_fuse(uint64_t* a, const uint64_t* b, uint64_t s1, uint64_t s1p, uint64_t s2,
                           uint64_t s2p, uint64_t q, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = addmod(shoup(a[i], s1, s1p, q), shoup(b[i], s2, s2p, q), q);
}
// LaunchApproxSwitchCRTBasis (kernels/basis_convert.cu): each workgroup first
// converts the QHatModp table to Montgomery form in LDS (one REDC of x * r2 per
// entry, with each target's constants derived on the device), then every lane
// converts PW_EPT coefficients: sources x_i = src_i * QHatInv_i mod q_i (Shoup
// with the caller's precons), then per target one exact sum and one REDC.  The
// sum of sizeQ products x_i * tab < sizeQ * q_max * p_t stays below p_t 2^64
// when sizeQ * q_max < 2^64, so the REDC is exact after one subtraction; other
// shapes (and more than ABI_CONV_MAXP targets) keep the Barrett form with the
// caller's mu, the reference's arithmetic.
enum { ABI_CONV_MAXP = 64 };
__device__ uint64_t abi_conv_barrett(const uint64_t* src, const uint64_t* qhi, const uint64_t* qhip,
                                     const uint64_t* q, const uint64_t* qhmp, uint64_t p, uint64_t mlo,
                                     uint64_t mhi, uint32_t sizeQ, uint32_t sizeP, uint32_t t, uint32_t n,
                                     uint32_t r);
template <uint32_t NS>
__device__ __forceinline__ void abi_conv_elem(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                              const uint64_t* qhi, const uint64_t* qhip, const uint64_t* q,
                                              const uint64_t* tab, const uint64_t* pq, const uint64_t* pn,
                                              uint32_t sizeQ, uint32_t sizeP, uint32_t n, uint32_t r) {
    if constexpr (NS > 0) {
        uint32_t lo[NS], hi[NS];
#pragma unroll
        for (uint32_t i = 0; i < NS; ++i) {
            const uint64_t x = shoup(ga(src, i * n + r), qhi[i], qhip[i], q[i]);
            lo[i] = split30_lo(x);
            hi[i] = split30_hi(x);
        }
        for (uint32_t t = 0; t < sizeP; ++t) {
            const U128 acc = dot30<NS>(lo, hi, [&](uint32_t i) { return tab[i * sizeP + t]; });
            ga(dst, t * n + r) = redc(acc, pq[t], pn[t]);
        }
    } else {
        for (uint32_t t = 0; t < sizeP; ++t) {
            U128 acc{0, 0};
            for (uint32_t i = 0; i < sizeQ; ++i)
                acc = add128(acc, mul128(shoup(ga(src, i * n + r), qhi[i], qhip[i], q[i]), tab[i * sizeP + t]));
            ga(dst, t * n + r) = redc(acc, pq[t], pn[t]);
        }
    }
}
__global__ __launch_bounds__(PW_THREADS) void k_ref_approx_switch(
    const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, const uint64_t* __restrict__ qhi,
    const uint64_t* __restrict__ qhip, const uint64_t* __restrict__ q, const uint64_t* __restrict__ qhmp,
    const uint64_t* __restrict__ p, const uint64_t* __restrict__ mlo, const uint64_t* __restrict__ mhi,
    uint32_t sizeQ, uint32_t sizeP, uint32_t n) {
    __shared__ uint64_t s_tab[CONV_MAX_SRC * ABI_CONV_MAXP], s_pq[ABI_CONV_MAXP], s_pn[ABI_CONV_MAXP],
        s_r2[ABI_CONV_MAXP], s_src[3 * CONV_MAX_SRC];
    __shared__ uint32_t s_mode;   // 0: Barrett fallback, 1: generic REDC sums, 2: dot30 column sums
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        uint64_t qmax = 0, all = 0;
        for (uint32_t i = 0; i < sizeQ; ++i) {
            qmax = q[i] > qmax ? q[i] : qmax;
            all |= q[i];
        }
        for (uint32_t t = 0; t < sizeP && t < ABI_CONV_MAXP; ++t) all |= p[t];
        const bool fits = sizeQ >= 1 && sizeQ <= CONV_MAX_SRC && sizeP <= ABI_CONV_MAXP &&
                          qmax < ~(uint64_t)0 / sizeQ;
        s_mode = !fits ? 0u : (sizeQ <= CONV_UNROLLED && (all >> 60) == 0) ? 2u : 1u;
    }
    __syncthreads();
    const uint32_t mode = s_mode;
    if (mode == 0) {
        PW_FOR(n) {
            for (uint32_t t = 0; t < sizeP; ++t)
                ga(dst, t * n + i) = abi_conv_barrett(src, qhi, qhip, q, qhmp, p[t], mlo[t], mhi[t], sizeQ, sizeP, t, n, i);
        }
        return;
    }
    for (uint32_t t = tid; t < sizeP; t += PW_THREADS) {
        uint64_t qn, r2;
        mont_consts(p[t], qn, r2);
        s_pq[t] = p[t];
        s_pn[t] = qn;
        s_r2[t] = r2;
    }
    for (uint32_t i = tid; i < sizeQ; i += PW_THREADS) {
        s_src[i] = q[i];
        s_src[CONV_MAX_SRC + i] = qhi[i];
        s_src[2 * CONV_MAX_SRC + i] = qhip[i];
    }
    __syncthreads();
    for (uint32_t e = tid; e < sizeQ * sizeP; e += PW_THREADS) {
        const uint32_t t = e % sizeP;
        s_tab[e] = redc(mul128(qhmp[e], s_r2[t]), s_pq[t], s_pn[t]);   // x R mod p_t
    }
    __syncthreads();
    const uint64_t *sq = s_src, *sh = s_src + CONV_MAX_SRC, *shp = s_src + 2 * CONV_MAX_SRC;
    if (mode == 2) {
        with_nsrc(sizeQ, [&](auto NSc) {
            constexpr uint32_t NS = decltype(NSc)::value;
            PW_FOR(n) { abi_conv_elem<NS>(src, dst, sh, shp, sq, s_tab, s_pq, s_pn, sizeQ, sizeP, n, i); }
        });
    } else {
        PW_FOR(n) { abi_conv_elem<0>(src, dst, sh, shp, sq, s_tab, s_pq, s_pn, sizeQ, sizeP, n, i); }
    }
}
__global__ void k_ref_ksmacc(uint64_t* o0, uint64_t* o1, const uint64_t* d, const uint64_t* b,
                             const uint64_t* a, uint64_t q, uint64_t qneg, uint64_t r2, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    o0[k] = addmod(o0[k], mulmod(d[k], b[k], q, qneg, r2), q);
    o1[k] = addmod(o1[k], mulmod(d[k], a[k], q, qneg, r2), q);
}
// LaunchKSMultAccBatched (kernels/ks_fastcore.cu:71-81): tower blockIdx.y,
// PW_EPT positions per lane, the tower's Montgomery constants derived once per
// workgroup; two positions per 16-byte access when every array is aligned.
__global__ __launch_bounds__(PW_THREADS) void k_ref_ksmacc_batched(uint64_t* o0, uint64_t* o1,
                                                                   const uint64_t* const* digits,
                                                                   const uint64_t* bv, const uint64_t* av,
                                                                   const uint32_t* keyRow,
                                                                   const uint64_t* mods, uint32_t n) {
    const uint32_t t = blockIdx.y;
    const WgMont m = wg_mont(mods, t);
    const uint64_t* d = digits[t];
    const size_t ko = (size_t)keyRow[t] * n, to = (size_t)t * n;
    uint64_t *r0 = o0 + to, *r1 = o1 + to;
    const uint64_t *b = bv + ko, *a = av + ko;
    auto one = [&](uint64_t x0, uint64_t x1, uint64_t dv, uint64_t kb, uint64_t ka, uint64_t& y0, uint64_t& y1) {
        y0 = addmod(x0, mulmod(dv, kb, m.q, m.qneg, m.r2), m.q);
        y1 = addmod(x1, mulmod(dv, ka, m.q, m.qneg, m.r2), m.q);
    };
    if ((n & 1) == 0 && al16(d) && al16(b) && al16(a) && al16(r0) && al16(r1)) {
        PW2_FOR(n) {
            const Pair x0 = ld2(r0, i), x1 = ld2(r1, i), dv = ld2(d, i), kb = ld2(b, i), ka = ld2(a, i);
            Pair y0, y1;
#pragma unroll
            for (int k = 0; k < 2; ++k) one(x0.v[k], x1.v[k], dv.v[k], kb.v[k], ka.v[k], y0.v[k], y1.v[k]);
            st2(r0, i, y0);
            st2(r1, i, y1);
        }
    } else {
        PW_FOR(n) { one(ga(r0, i), ga(r1, i), ga(d, i), ga(b, i), ga(a, i), ga(r0, i), ga(r1, i)); }
    }
}
__global__ void k_ref_moddown(uint64_t* ans, const uint64_t* a, const uint64_t* b, uint64_t s,
                              uint64_t sp, uint64_t q, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) ans[k] = shoup(submod(a[k], b[k], q), s, sp, q);
}
__global__ void k_ref_automorph(uint64_t* out, const uint64_t* in, uint32_t n, uint64_t kk, uint64_t q) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t M = 2ull * n;
    const uint64_t idx = ((uint64_t)j * (kk % M)) % M;
    const uint64_t v = in[j];
    if (idx < n) out[idx] = v;
    else out[idx - n] = v ? q - v : 0;
}
__global__ void k_ref_add(uint64_t* a, const uint64_t* b, uint64_t q, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = addmod(a[k], b[k], q);
}
__global__ void k_ref_ptmulacc(uint64_t* acc0, uint64_t* acc1, const uint64_t* b0, const uint64_t* b1,
                               const uint64_t* d, uint64_t q, uint64_t qneg, uint64_t r2, uint32_t n) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    acc0[k] = addmod(acc0[k], mulmod(b0[k], d[k], q, qneg, r2), q);
    acc1[k] = addmod(acc1[k], mulmod(b1[k], d[k], q, qneg, r2), q);
}
// LaunchTensor (kernels/tensor.cu:28-51): the plan's tensor kernel body on
// tower blockIdx.y, with that tower's Montgomery constants derived once per
// workgroup.
__global__ __launch_bounds__(PW_THREADS) void k_ref_tensor(const uint64_t* c0a, const uint64_t* c1a,
                                                           const uint64_t* c0b, const uint64_t* c1b, uint64_t* t0,
                                                           uint64_t* t1, uint64_t* t2, const uint64_t* mods,
                                                           uint32_t n) {
    const WgMont m = wg_mont(mods, blockIdx.y);
    const size_t o = (size_t)blockIdx.y * n;
    TensorJob j{c0a + o, c1a + o, c0b + o, c1b + o, t0 + o, t1 + o, t2 + o, m.q, m.qneg, m.r2, 0};
    j.pairs = (n & 1) == 0 && al16(j.c0a) && al16(j.c1a) && al16(j.c0b) && al16(j.c1b) && al16(j.t0) &&
              al16(j.t1) && al16(j.t2);
    tensor_body(j, n);
}
// LaunchCombine (kernels/tensor.cu:53-61)
__global__ __launch_bounds__(PW_THREADS) void k_ref_combine(uint64_t* r0, uint64_t* r1, const uint64_t* t0,
                                                            const uint64_t* t1, const uint64_t* ba0,
                                                            const uint64_t* ba1, const uint64_t* mods,
                                                            uint32_t n) {
    const uint64_t q = mods[blockIdx.y];
    const size_t o = (size_t)blockIdx.y * n;
    uint64_t *x0 = r0 + o, *x1 = r1 + o;
    const uint64_t *u0 = t0 + o, *u1 = t1 + o, *b0 = ba0 + o, *b1 = ba1 + o;
    if ((n & 1) == 0 && al16(x0) && al16(x1) && al16(u0) && al16(u1) && al16(b0) && al16(b1)) {
        PW2_FOR(n) {
            const Pair a = ld2(u0, i), b = ld2(b0, i), c = ld2(u1, i), d = ld2(b1, i);
            st2(x0, i, Pair{{addmod(a.v[0], b.v[0], q), addmod(a.v[1], b.v[1], q)}});
            st2(x1, i, Pair{{addmod(c.v[0], d.v[0], q), addmod(c.v[1], d.v[1], q)}});
        }
    } else {
        PW_FOR(n) {
            ga(x0, i) = addmod(ga(u0, i), ga(b0, i), q);
            ga(x1, i) = addmod(ga(u1, i), ga(b1, i), q);
        }
    }
}

}  // namespace gf

using namespace gf;

static inline uint32_t blocks256(uint32_t n) { return (n + 255) / 256; }

extern "C" {

void LaunchNTT_CT(uint64_t* d_x, const uint64_t* d_roots, const uint64_t* d_precon, uint32_t n,
                  uint64_t q, hipStream_t s) {
    NttJob j{d_x, d_x, d_roots, d_precon, q, 0, 0};
    launch_ntt_one(j, make_geom(n), true, s);
}

void LaunchINTT_GS(uint64_t* d_x, const uint64_t* d_invroots, const uint64_t* d_invprecon, uint32_t n,
                   uint64_t q, uint64_t n_inv, uint64_t n_inv_precon, hipStream_t s) {
    NttJob j{d_x, d_x, d_invroots, d_invprecon, q, n_inv, n_inv_precon};
    launch_ntt_one(j, make_geom(n), false, s);
}

void LaunchRNSMultTower(const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_r, uint64_t q, uint32_t n,
                        hipStream_t stream) {
    k_ref_mult<<<blocks256(n), 256, 0, stream>>>(d_a, d_b, d_r, q, host_qneg(q), host_r2(q), n);
}

void LaunchSwitchModulus(uint64_t* d_v, uint64_t q_old, uint64_t q_new, uint32_t n, hipStream_t s) {
    k_ref_switch<<<blocks256(n), 256, 0, s>>>(d_v, make_switch(nullptr, nullptr, q_old, q_new), n);
}

void LaunchRescaleFuse(uint64_t* d_a, const uint64_t* d_b, uint64_t s1, uint64_t s2, uint64_t q,
                       uint32_t n, hipStream_t s) {
    k_ref_fuse<<<blocks256(n), 256, 0, s>>>(d_a, d_b, s1 % q, host_precon(s1 % q, q), s2 % q,
                                            host_precon(s2 % q, q), q, n);
}

void LaunchApproxSwitchCRTBasis(const uint64_t* d_src, uint64_t* d_dst, const uint64_t* d_QHatInvModq,
                                const uint64_t* d_QHatInvModqPrecon, const uint64_t* d_q,
                                const uint64_t* d_QHatModp, const uint64_t* d_p, const uint64_t* d_mu_lo,
                                const uint64_t* d_mu_hi, uint32_t sizeQ, uint32_t sizeP, uint32_t n,
                                hipStream_t s) {
    k_ref_approx_switch<<<pw_grid(n, 1), PW_THREADS, 0, s>>>(d_src, d_dst, d_QHatInvModq, d_QHatInvModqPrecon,
                                                     d_q, d_QHatModp, d_p, d_mu_lo, d_mu_hi, sizeQ,
                                                     sizeP, n);
}

void LaunchKSMultAcc(uint64_t* d_out0, uint64_t* d_out1, const uint64_t* d_digit, const uint64_t* d_b,
                     const uint64_t* d_a, uint64_t q, uint32_t n, hipStream_t s) {
    k_ref_ksmacc<<<blocks256(n), 256, 0, s>>>(d_out0, d_out1, d_digit, d_b, d_a, q, host_qneg(q),
                                              host_r2(q), n);
}

void LaunchKSMultAccBatched(uint64_t* d_out0, uint64_t* d_out1, const uint64_t* const* d_digits,
                            const uint64_t* d_bv, const uint64_t* d_av, const uint32_t* d_keyRow,
                            const uint64_t* d_mods, uint32_t towers, uint32_t n, hipStream_t s) {
    if (towers) k_ref_ksmacc_batched<<<pw_grid(n, towers), PW_THREADS, 0, s>>>(d_out0, d_out1, d_digits, d_bv,
                                                                    d_av, d_keyRow, d_mods, n);
}

void LaunchModDownCombine(uint64_t* d_ans, const uint64_t* d_a, const uint64_t* d_b, uint64_t sc,
                          uint64_t q, uint32_t n, hipStream_t st) {
    k_ref_moddown<<<blocks256(n), 256, 0, st>>>(d_ans, d_a, d_b, sc % q, host_precon(sc % q, q), q, n);
}

void LaunchAutomorphPermute(uint64_t* d_out, const uint64_t* d_in, uint32_t n, uint64_t k, uint64_t q,
                            hipStream_t s) {
    k_ref_automorph<<<blocks256(n), 256, 0, s>>>(d_out, d_in, n, k, q);
}

void LaunchAddInto(uint64_t* d_a, const uint64_t* d_b, uint64_t q, uint32_t n, hipStream_t s) {
    k_ref_add<<<blocks256(n), 256, 0, s>>>(d_a, d_b, q, n);
}

void LaunchPtMulAcc(uint64_t* d_acc0, uint64_t* d_acc1, const uint64_t* d_b0, const uint64_t* d_b1,
                    const uint64_t* d_d, uint64_t q, uint32_t n, hipStream_t s) {
    k_ref_ptmulacc<<<blocks256(n), 256, 0, s>>>(d_acc0, d_acc1, d_b0, d_b1, d_d, q, host_qneg(q),
                                                host_r2(q), n);
}

void LaunchTensor(const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                  uint64_t* t0, uint64_t* t1, uint64_t* t2, const uint64_t* d_mods, uint32_t sizeQ,
                  uint32_t n, hipStream_t s) {
    if (sizeQ) k_ref_tensor<<<pw_grid(n, sizeQ), PW_THREADS, 0, s>>>(c0a, c1a, c0b, c1b, t0, t1, t2, d_mods, n);
}

void LaunchCombine(uint64_t* r0, uint64_t* r1, const uint64_t* t0, const uint64_t* t1, const uint64_t* ba0,
                   const uint64_t* ba1, const uint64_t* d_mods, uint32_t sizeQ, uint32_t n, hipStream_t s) {
    if (sizeQ) k_ref_combine<<<pw_grid(n, sizeQ), PW_THREADS, 0, s>>>(r0, r1, t0, t1, ba0, ba1, d_mods, n);
}

}  // extern "C"
