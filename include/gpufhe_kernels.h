/*
 * Device-layer C ABI: the reference's extern "C" kernel launchers, re-built
 * for gfx950.  Same names, argument order and meaning; cudaStream_t becomes
 * hipStream_t (the only signature change).  All pointers are caller-owned
 * device pointers, tower-major [tower][n] uint64 residues in [0, q); every
 * launcher is asynchronous on the given stream and returns nothing (as in the
 * reference, errors surface at the caller's next synchronisation).
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 */
#ifndef GPUFHE_KERNELS_H
#define GPUFHE_KERNELS_H
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/ntt.h:8 -- forward CT NTT, standard in, bit-reversed out, [0,q). */
void LaunchNTT_CT(uint64_t* d_x, const uint64_t* d_roots, const uint64_t* d_precon, uint32_t n,
                  uint64_t q, hipStream_t s);

/* include/intt.h:7 -- inverse GS NTT, bit-reversed in, standard out, x n^-1. */
void LaunchINTT_GS(uint64_t* d_x, const uint64_t* d_invroots, const uint64_t* d_invprecon,
                   uint32_t n, uint64_t q, uint64_t n_inv, uint64_t n_inv_precon, hipStream_t s);

/* include/rns_arith.h:7 -- r = a*b mod q, one tower. */
void LaunchRNSMultTower(const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_r, uint64_t q,
                        uint32_t n, hipStream_t stream);

/* include/switch_modulus.h:7 -- centred switch q_old -> q_new in place. */
void LaunchSwitchModulus(uint64_t* d_v, uint64_t q_old, uint64_t q_new, uint32_t n, hipStream_t s);

/* include/rescale.h:9 -- a = a*s1 + b*s2 mod q, one tower. */
void LaunchRescaleFuse(uint64_t* d_a, const uint64_t* d_b, uint64_t s1, uint64_t s2, uint64_t q,
                       uint32_t n, hipStream_t s);

/* include/basis_convert.h:9 -- ApproxSwitchCRTBasis sizeQ -> sizeP towers,
 * QHatModp flattened [sizeQ][sizeP], Barrett mu = floor((2^128-1)/p) as lo/hi. */
void LaunchApproxSwitchCRTBasis(const uint64_t* d_src, uint64_t* d_dst,
                                const uint64_t* d_QHatInvModq, const uint64_t* d_QHatInvModqPrecon,
                                const uint64_t* d_q, const uint64_t* d_QHatModp, const uint64_t* d_p,
                                const uint64_t* d_mu_lo, const uint64_t* d_mu_hi, uint32_t sizeQ,
                                uint32_t sizeP, uint32_t n, hipStream_t s);

/* include/ks_fastcore.h:9 -- out0 += digit*b, out1 += digit*a mod q, one tower. */
void LaunchKSMultAcc(uint64_t* d_out0, uint64_t* d_out1, const uint64_t* d_digit,
                     const uint64_t* d_b, const uint64_t* d_a, uint64_t q, uint32_t n,
                     hipStream_t s);

/* include/ks_fastcore.h:13 -- the same over `towers` towers; digits is a
 * device array of per-tower digit pointers, keyRow the eval-key row of each. */
void LaunchKSMultAccBatched(uint64_t* d_out0, uint64_t* d_out1, const uint64_t* const* d_digits,
                            const uint64_t* d_bv, const uint64_t* d_av, const uint32_t* d_keyRow,
                            const uint64_t* d_mods, uint32_t towers, uint32_t n, hipStream_t s);

/* include/moddown.h:7 -- ans = (a - b) * s mod q, one tower. */
void LaunchModDownCombine(uint64_t* d_ans, const uint64_t* d_a, const uint64_t* d_b, uint64_t s,
                          uint64_t q, uint32_t n, hipStream_t st);

/* include/automorph.h:4 -- coefficient-domain sigma_k: out[(j k) mod 2n] = +-in[j]. */
void LaunchAutomorphPermute(uint64_t* d_out, const uint64_t* d_in, uint32_t n, uint64_t k,
                            uint64_t q, hipStream_t s);

/* include/stage_acc.h:8 -- a = a + b mod q, one tower. */
void LaunchAddInto(uint64_t* d_a, const uint64_t* d_b, uint64_t q, uint32_t n, hipStream_t s);

/* include/stage_acc.h:4 -- acc0 += b0*d, acc1 += b1*d mod q, one tower. */
void LaunchPtMulAcc(uint64_t* d_acc0, uint64_t* d_acc1, const uint64_t* d_b0, const uint64_t* d_b1,
                    const uint64_t* d_d, uint64_t q, uint32_t n, hipStream_t s);

/* kernels/tensor.cu:28 (header-less in the reference, declared by
 * tests/test_evalmult_rescale.cpp:13-14) -- t0=a0b0, t1=a0b1+a1b0, t2=a1b1
 * over sizeQ towers, d_mods = device array of the sizeQ moduli. */
void LaunchTensor(const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
                  const uint64_t* c1b, uint64_t* t0, uint64_t* t1, uint64_t* t2,
                  const uint64_t* d_mods, uint32_t sizeQ, uint32_t n, hipStream_t s);

/* kernels/tensor.cu:53 (declared by tests/test_evalmult_rescale.cpp:15-16) --
 * r0 = t0 + ba0, r1 = t1 + ba1 over sizeQ towers. */
void LaunchCombine(uint64_t* r0, uint64_t* r1, const uint64_t* t0, const uint64_t* t1,
                   const uint64_t* ba0, const uint64_t* ba1, const uint64_t* d_mods,
                   uint32_t sizeQ, uint32_t n, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
