// Internal job descriptors and launch helpers shared by the kernels and the
// host orchestration.  Every batched kernel walks a device array of jobs, one
// job per (operation, tower); the grid's y dimension indexes the job, so one
// launch covers any mix of ciphertexts, components and levels at one ring size.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <stdexcept>
#include <string>

#include "ksi.hpp"
#include "modarith.hpp"
#include "modup.hpp"
#include "ntt_engine.hpp"
#include "tensor.hpp"

namespace gf {

enum { MAXP = 16, PW_THREADS = 256, PW_EPT = 4, CONV_MAX_SRC = 16 };

// per-modulus constants (row = index in the Q|P modulus list of a context)
struct RowConst {
    uint64_t q, qneg, r2, r1, ninv, ninvp;
};

// dst[j] = (sum_i src[i] * tab[i*ntgt + j]) * R^-1 mod p_j for targets j in
// [j0, j1); tab holds the conversion constants in Montgomery form, so one REDC
// of the 128-bit sum gives the remainder (left in [0, 2 p_j) when wide is 0:
// the outputs only feed forward transforms).  src/dst are tower-contiguous.
struct ConvJob {
    const uint64_t* src;
    uint64_t* dst;
    const uint64_t* tab;
    const uint64_t* tq;     // target moduli [ntgt]
    const uint64_t* tqneg;  // target -q^-1 mod 2^64 [ntgt]
    uint32_t nsrc, ntgt, j0, j1;
    uint32_t wide;          // 1: some modulus is 2^60 or more (full 128-bit products)
};

// o0 = sum_p dig[p] * kb[p], o1 = sum_p dig[p] * ka[p] (keys in Montgomery form);
// with gk != 0 the digits are read through the evaluation-domain automorphism
// of index gk (digit p at position i is dig[p][perm_gk(i)]), which is the
// inner product of a hoisted rotation without a permuted copy of the digits
struct InnerJob {
    const uint64_t* dig[MAXP];
    const uint64_t* kb[MAXP];
    const uint64_t* ka[MAXP];
    uint64_t* o0;
    uint64_t* o1;
    uint64_t q, qneg;
    uint64_t red;     // lazy_red_of(q): digits of this row may be unreduced (NttJob::lazy_out)
    uint64_t gk;      // 0, or the automorphism index the digits are read through
    uint32_t np;
    uint32_t pairs;   // as TensorJob::pairs (unused with gk)
};

// out = t + (res - conv) * pinv   (ApproxModDown combine fused with r = t + ba)
struct CombineJob {
    const uint64_t* res;
    const uint64_t* conv;
    const uint64_t* t;      // may be null: out = (res - conv) * pinv
    uint64_t* out;
    uint64_t q, pinv, pinvp;
};

// centred modulus switch q_old -> q_new
struct SwitchJob {
    const uint64_t* in;
    uint64_t* out;
    uint64_t q_old, q_new, half, diff;
    uint64_t red;   // floor(2^64 / q_new), for x mod q_new without a divide
    uint32_t new_larger;
};

// out = a s1 + b s2
struct FuseJob {
    const uint64_t* a;
    const uint64_t* b;
    uint64_t* out;
    uint64_t q, s1, s1p, s2, s2p;
};

// out = a + b
struct AddJob {
    const uint64_t* a;
    const uint64_t* b;
    uint64_t* out;
    uint64_t q;
};

// evaluation-domain automorphism: out[i] = in[perm_k(i)]
struct PermJob {
    const uint64_t* in;
    uint64_t* out;
    uint64_t k;
};

// Fused ModDown + rescale prologue of one component.  First the dropped tower
// of r = t + ba in coefficient form (q_old = q_last):
//   drop = dt + (dr - REDC(sum_k xp[k] * tabd[k*strided])) * pinv
// with dt = INTT(t_last), dr = INTT(res_last), xp = INTT'd P towers scaled by
// n^-1 PHatInv; then for the surviving towers i in [j0, j1)
//   v_i = s2_i * switch(drop -> q_i) + REDC(sum_k xp[k] * tab[k*ntgt + i]),
// left in [0, 4 q_i) when wide is 0 (the forward transform takes it so);
// tab folds -P^-1 s1 into PHatModq, so NTT(v_i) + t_i s1 + res_i P^-1 s1 is
// the rescaled output tower.
struct RsJob {
    const uint64_t* xp;
    const uint64_t* dt;
    const uint64_t* dr;
    uint64_t* out;
    const uint64_t* tab;
    const RsTarget* tg;
    const uint64_t* tabd;   // ModDown conversion column of the dropped tower
    uint64_t q_old, half, qd_neg, pinv, pinvp;
    uint32_t K, ntgt, j0, j1, strided;
    uint32_t wide;          // as ConvJob::wide
};

// one tower of a freshly sampled small polynomial: dst = src mod q (signed src)
struct EmbedJob {
    const int64_t* src;
    uint64_t* dst;
    uint64_t q;
};
// switching-key rows of one part and one tower (src/evalkeygen.cpp:42-96):
// a = pkA u + ns e1, b = pkB u + ns e0 (+ (P mod q) sOld on the part's own Q
// towers, sold null elsewhere), written in Montgomery form (x R mod q)
struct EkJob {
    const uint64_t *pka, *pkb, *u, *e0, *e1, *sold;
    uint64_t *a_out, *b_out;
    uint64_t q, qneg, r2, ns, nsp, pq, pqp;
    // 1: outputs in plain form -- an encryption (src/encdec.cpp:41-64): u = v,
    // sold = the message, pq = 1, so b_out = c0 and a_out = c1
    uint32_t plain;
};

// one polynomial drawn by the counter sampler (ctr_sampler.hpp): kind 0
// ternary and 1 Gaussian into signed coefficients `coef`; kind 2 uniform mod q
// into the residues `res` (red = floor(2^64 / q), q > 2^32)
struct SampleJob {
    int64_t* coef;
    uint64_t* res;
    uint64_t q, red;
    uint32_t domain, kind;
};
// a public key's b = ns e - a s (mod q), every operand in evaluation form
struct PkJob {
    const uint64_t *a, *s, *e;
    uint64_t* b;
    uint64_t q, qneg, r2, ns, nsp;
};

// out = c0 + c1 s (mod q): the phase of a decryption before its inverse
// transform (src/encdec.cpp:68-81)
struct DecJob {
    const uint64_t *c0, *c1, *s;
    uint64_t* out;
    uint64_t q, qneg, r2;
};

// plaintext-weighted sum of m ciphertext towers (linear transforms):
// out0 = sum_j pt_j x0_j, out1 = sum_j pt_j x1_j mod q (the accumulate of
// kernels/stage_acc.cu:12-28 over all terms at once)
struct PtAccJob {
    const uint64_t* pt[MAXP];
    const uint64_t* x0[MAXP];
    const uint64_t* x1[MAXP];
    uint64_t* o0;
    uint64_t* o1;
    uint64_t q, qneg, r2;
    uint32_t m;
};

// copy one tower
struct CopyJob {
    const uint64_t* in;
    uint64_t* out;
};

// every pointer 16-byte aligned (pairs of residues load as one dwordx4)
inline uint32_t aligned16(std::initializer_list<const void*> ps) {
    for (const void* p : ps)
        if ((uintptr_t)p & 15) return 0;
    return 1;
}

inline TensorJob tensor_job(const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                            uint64_t* t0, uint64_t* t1, uint64_t* t2, const RowConst& rc, uint32_t n) {
    const uint32_t pairs = n % 2 == 0 && aligned16({c0a, c1a, c0b, c1b, t0, t1, t2});
    return TensorJob{c0a, c1a, c0b, c1b, t0, t1, t2, rc.q, rc.qneg, rc.r2, pairs};
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("gpufhe: ") + what + ": " + hipGetErrorString(e));
}

// batched launches (jobs in device memory)
// which: 3 = the whole transform; 1 / 2 = only its first / second kernel
// (two-phase rings; a single-phase ring has one kernel, run when bit 0 is set)
void launch_ntt(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, bool fwd, hipStream_t s,
                uint32_t which = 3);
void launch_ntt_one(const NttJob& job, const NttGeom& g, bool fwd, hipStream_t s);
void launch_tensor(const TensorJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_conv(const ConvJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_inner(const InnerJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_combine(const CombineJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_switch(const SwitchJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_fuse(const FuseJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_add(const AddJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_perm(const PermJob* d, uint32_t nj, uint32_t n, uint32_t logn, hipStream_t s);
void launch_perm_one(const PermJob& j, uint32_t n, uint32_t logn, hipStream_t s);
void launch_copy(const CopyJob* d, uint32_t nj, uint32_t n, hipStream_t s);
// T: the Gaussian table's K thresholds (device memory)
void launch_sample(const SampleJob* d, uint32_t nj, uint32_t n, uint64_t seed, const uint64_t* T, uint32_t K,
                   hipStream_t s);
void launch_pkgen(const PkJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_fill(uint64_t* d, const uint64_t* d_mods, uint32_t towers, uint32_t n, uint64_t seed,
                 hipStream_t s);
void launch_rs(const RsJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_embed(const EmbedJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_ekgen(const EkJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_dec(const DecJob* d, uint32_t nj, uint32_t n, hipStream_t s);
void launch_ptacc(const PtAccJob* d, uint32_t nj, uint32_t n, hipStream_t s);
// fused digit row pass + key inner product: regions = LDS regions (at least every
// job's transformed-part count), waves = workgroup size in waves (1..KSI_WAVES)
void launch_ksi(const KsiJob* d, uint32_t nj, const NttGeom& g, uint32_t regions, uint32_t waves, hipStream_t s);
void launch_tensor_ib(const TensorIbJob* d, uint32_t nj, const NttGeom& g, hipStream_t s);
// the fused ModUp over jobs of sp sources each
void launch_modup(const ModupJob* d, uint32_t nj, const NttGeom& g, uint32_t arg, hipStream_t s);
void launch_moddown(const ModupJob* d, uint32_t nj, const NttGeom& g, uint32_t sp, hipStream_t s);

inline SwitchJob make_switch(const uint64_t* in, uint64_t* out, uint64_t q_old, uint64_t q_new) {
    SwitchJob j;
    j.in = in;
    j.out = out;
    j.q_old = q_old;
    j.q_new = q_new;
    j.half = q_old >> 1;
    j.new_larger = q_new > q_old;
    j.diff = q_new > q_old ? q_new - q_old : (q_old - q_new) % q_new;
    j.red = host_precon(1, q_new);
    return j;
}

}  // namespace gf
