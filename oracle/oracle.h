/*
 * CPU oracle for the device-resident CKKS multiply / rescale / rotate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker (or as the timed CPU baseline).
 *
 * Plain C99 + unsigned __int128.  Every function is a restatement of the
 * reference algorithm named in its comment (paths are relative to the
 * reference repository root).  All arithmetic is exact modular integer work;
 * the results are therefore uniquely defined by the inputs (moduli, roots,
 * keys), which is what makes bit-exact parity checks meaningful.
 *
 * Pinning (how each function is checked independently of its own formula):
 *   or_mt64_*            C++ standard known answer (10000th draw of mt19937_64)
 *                        and libstdc++'s std::mt19937_64.
 *   or_is_prime, or_native_primes, or_native_root, or_rescale_consts,
 *   or_ks_constants      compared with the reference's own src/params.cpp and
 *                        src/ks_constants.cpp compiled into oracle/_ref, and with
 *                        Python big-integer restatements.
 *   or_ntt               brute-force negacyclic DFT at n <= 2^10 (closed form
 *                        X[k] = a(psi^(2 brv(k)+1))), committed golden vectors.
 *   or_intt              INTT(NTT(x)) == x at every n tested.
 *   or_switch_modulus    closed form [centred(x)]_{q_new}, both directions,
 *                        edge x = floor(q_old/2) (strict >).
 *   or_rescale_*         CRT big-integer identity y_i = ((X - x~_l)/q_l) mod q_i.
 *   or_approx_switch     big-integer sum (sum_i [x_i QHatInv_i]_{q_i} QHat_i) mod p_j.
 *   or_keyswitch         decryption identity ba0 + ba1*s == d*s_old + small, and
 *                        digit-sum identity sum_p digit_p * gadget_p == d*P.
 *   or_evalmult_rescale  exact small-integer product recovery (maxerr 0).
 *   or_automorph_*       eval-domain permutation identity and slot rotation.
 *   or_rotate_hoisted    equals or_rotate exactly for k = 1 (identity
 *                        automorphism) at every level; rotated slots decode
 *                        to the rotated message for k = 5^r.
 *   or_linear_transform  one term with pt = 1 equals the (hoisted) rotation;
 *                        the unrotated term equals pt (.) c; linear in the
 *                        plaintexts; decodes to sum_j diag_j (.) rot_j(z).
 *   or_keygen, or_evalkeygen, or_encrypt, or_decrypt
 *                        RLWE decryption identities (noise bounds), enc/dec
 *                        round trip with maxerr 0.
 *   or_ctr_draw          published SplitMix64 known answers (sequential
 *                        outputs from seed 1234567), domain separation.
 *   or_ctr_gauss         distribution: moments against sigma and the
 *                        histogram of or_gauss_sample (same table).
 *   or_keygen_ctr, or_evalkeygen_ctr
 *                        ternary / uniform sample statistics, the RLWE
 *                        identity b + a s = ns e with e small, and a switching
 *                        key's decryption identity under the generated pk.
 * No function in this file is "parity unpinned".
 */
#ifndef GPUFHE_ORACLE_H
#define GPUFHE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar arithmetic ---- */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_powmod(uint64_t b, uint64_t e, uint64_t q);
uint64_t or_invmod(uint64_t a, uint64_t q);
uint64_t or_shoup_precon(uint64_t w, uint64_t q);
uint64_t or_shoup_mulmod(uint64_t x, uint64_t w, uint64_t wp, uint64_t q);

/* ---- mt19937_64 (standard engine) ---- */
typedef struct { uint64_t mt[312]; int idx; } or_mt64;
void or_mt64_seed(or_mt64* g, uint64_t seed);
uint64_t or_mt64_next(or_mt64* g);
/* discrete Gaussian (cumulative-table sampler), see or_gauss_sample */
int64_t or_gauss_sample(or_mt64* g, double sigma);

/* ---- parameters ---- */
int or_is_prime(uint64_t x);
int or_native_primes(uint64_t* out, uint32_t count, uint32_t bits, uint32_t n,
                     const uint64_t* avoid, uint32_t navoid);
uint64_t or_native_root(uint32_t n, uint64_t q);
void or_rescale_consts(uint64_t* s1, uint64_t* s2, const uint64_t* mod, uint32_t last);
void or_root_tables(uint32_t n, uint64_t q, uint64_t root, uint64_t* fr, uint64_t* fp,
                    uint64_t* ir, uint64_t* ip, uint64_t* ninv_pair);

/* ---- per-tower primitives ---- */
void or_ntt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* fr, const uint64_t* fp);
void or_intt(uint64_t* x, uint32_t n, uint64_t q, const uint64_t* ir, const uint64_t* ip,
             uint64_t ninv, uint64_t ninv_p);
void or_rns_mult(const uint64_t* a, const uint64_t* b, uint64_t* r, uint64_t q, uint32_t n);
void or_switch_modulus(uint64_t* v, uint32_t n, uint64_t q_old, uint64_t q_new);
void or_rescale_fuse(uint64_t* a, const uint64_t* b, uint64_t s1, uint64_t s2, uint64_t q,
                     uint32_t n);
void or_approx_switch(const uint64_t* src, uint64_t* dst, const uint64_t* QHatInv,
                      const uint64_t* QHatInvPre, const uint64_t* q, const uint64_t* QHatModp,
                      const uint64_t* p, uint32_t sizeQ, uint32_t sizeP, uint32_t n);
void or_ks_macc(uint64_t* out0, uint64_t* out1, const uint64_t* digit, const uint64_t* b,
                const uint64_t* a, uint64_t q, uint32_t n);
void or_moddown_combine(uint64_t* ans, const uint64_t* a, const uint64_t* b, uint64_t s,
                        uint64_t q, uint32_t n);
void or_tensor(const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
               const uint64_t* c1b, uint64_t* t0, uint64_t* t1, uint64_t* t2,
               const uint64_t* mods, uint32_t sizeQ, uint32_t n);
void or_combine(uint64_t* r0, uint64_t* r1, const uint64_t* t0, const uint64_t* t1,
                const uint64_t* ba0, const uint64_t* ba1, const uint64_t* mods,
                uint32_t sizeQ, uint32_t n);
void or_add_into(uint64_t* a, const uint64_t* b, uint64_t q, uint32_t n);
void or_automorph_coeff(uint64_t* out, const uint64_t* in, uint32_t n, uint64_t k, uint64_t q);

/* ---- keyswitch constants for one level (reference layout) ----
 * Flat layout, allocated by or_ks_constants, freed by or_ks_constants_free:
 *   per part p (numPart parts): sizePart[p], sizeCompl[p], startIdx[p]
 *   partQHatInv/Prec/SrcMod   : [numPart][alpha]          (row stride alpha)
 *   partQHatModp              : [numPart][alpha*(sizeQl+sizeP)], part block holds
 *                               [sizePart x sizeCompl] packed (row stride sizeCompl)
 *   partComplMod/MuLo/MuHi    : [numPart][sizeQl+sizeP]
 *   pHatInv/Prec              : [sizeP]
 *   pHatModq                  : [sizeP][sizeQl]
 *   mdMuLo/Hi, pInvModq       : [sizeQl]
 */
typedef struct {
    uint32_t sizeQl, sizeP, numPart, alpha;
    uint32_t *sizePart, *sizeCompl, *startIdx;
    uint64_t *partQHatInv, *partQHatInvPrec, *partSrcMod, *partQHatModp;
    uint64_t *partComplMod, *partBMuLo, *partBMuHi;
    uint64_t *pHatInv, *pHatInvPrec, *pHatModq, *mdBMuLo, *mdBMuHi, *pInvModq;
} or_ksc;
or_ksc* or_ks_constants(const uint64_t* moduliQ, uint32_t sizeQl, const uint64_t* moduliP,
                        uint32_t sizeP, uint32_t numPart, uint32_t alpha);
void or_ks_constants_free(or_ksc* K);

/* ---- a parameter set with root tables for every Q and P modulus ---- */
typedef struct {
    uint32_t n, L, K, alpha, dnum;
    uint64_t *mod, *root;          /* L + K entries: Q moduli then P moduli */
    uint64_t *fr, *fp, *ir, *ip;   /* (L+K)*n */
    uint64_t *ninv, *ninvp;        /* L+K */
} or_params;
or_params* or_params_create(uint32_t n, const uint64_t* qmods, uint32_t L,
                            const uint64_t* pmods, uint32_t K, const uint64_t* roots,
                            uint32_t dnum);
void or_params_free(or_params* P);
/* parts used at a level of sizeQl towers (fixed part width alpha) */
uint32_t or_num_parts(const or_params* P, uint32_t sizeQl);

/* ---- composed operations (level = sizeQl towers of Q) ---- */
int or_keyswitch(const or_params* P, uint32_t sizeQl, const uint64_t* d, const uint64_t* av,
                 const uint64_t* bv, uint64_t* ba0, uint64_t* ba1);
int or_ks_digits(const or_params* P, uint32_t sizeQl, const uint64_t* d, uint64_t* digits);
int or_rescale_raw(const or_params* P, uint64_t* c, uint32_t towers);
int or_evalmult_rescale(const or_params* P, uint32_t sizeQl, const uint64_t* c0a,
                        const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                        const uint64_t* av, const uint64_t* bv, uint64_t* out0,
                        uint64_t* out1);
int or_evalmult_norescale(const or_params* P, uint32_t sizeQl, const uint64_t* c0a,
                          const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                          const uint64_t* av, const uint64_t* bv, uint64_t* r0,
                          uint64_t* r1);
int or_automorph_eval(const or_params* P, uint64_t* v, uint32_t towers, uint32_t k,
                      int p_towers);
int or_rotate(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
              uint32_t k, const uint64_t* av, const uint64_t* bv, uint64_t* out0,
              uint64_t* out1);
/* sum_j pts[j] * rot_kj(c0, c1); av[j] NULL (with ks[j] = 1) for the unrotated term */
int or_linear_transform(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
                        uint32_t m, const uint32_t* ks, const uint64_t* const* av,
                        const uint64_t* const* bv, const uint64_t* const* pts, uint64_t* out0,
                        uint64_t* out1);
/* m rotations of one ciphertext sharing one digit decomposition of c1 */
int or_rotate_hoisted(const or_params* P, uint32_t sizeQl, const uint64_t* c0, const uint64_t* c1,
                      uint32_t m, const uint32_t* ks, const uint64_t* const* av,
                      const uint64_t* const* bv, uint64_t* const* out0, uint64_t* const* out1);

/* ---- keys, encryption (sampler described in oracle.c) ---- */
void or_keygen(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
               uint64_t ns, double sigma, uint64_t seed, uint32_t hamming, uint64_t* s,
               uint64_t* pkA, uint64_t* pkB);
void or_evalkeygen_sold(uint32_t n, uint32_t sizeQ, uint32_t sizeP, uint32_t alpha,
                        uint32_t numPart, const uint64_t* modQP, const uint64_t* rootQP,
                        const uint64_t* sOld, const uint64_t* pkA, const uint64_t* pkB,
                        uint64_t ns, double sigma, uint64_t seed, uint64_t* av,
                        uint64_t* bv);
/* counter sampler of the device key generation (see oracle.c) and the key
 * generators that use it */
uint64_t or_ctr_draw(uint64_t seed, uint32_t domain, uint32_t index);
int64_t or_ctr_gauss(uint64_t seed, uint32_t domain, uint32_t index, double sigma);
void or_keygen_ctr(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
                   uint64_t ns, double sigma, uint64_t seed, uint64_t* s, uint64_t* pkA,
                   uint64_t* pkB);
void or_evalkeygen_ctr(uint32_t n, uint32_t sizeQ, uint32_t sizeP, uint32_t alpha,
                       uint32_t numPart, const uint64_t* modQP, const uint64_t* rootQP,
                       const uint64_t* sOld, const uint64_t* pkA, const uint64_t* pkB,
                       uint64_t ns, double sigma, uint64_t seed, uint64_t* av,
                       uint64_t* bv);
void or_encrypt(uint32_t n, const uint64_t* mods, const uint64_t* roots, uint32_t count,
                const int64_t* m, const uint64_t* pkA, const uint64_t* pkB, uint64_t ns,
                double sigma, uint64_t seed, uint64_t* c0, uint64_t* c1);
void or_decrypt(uint32_t n, uint64_t q0, uint64_t root0, const uint64_t* c0,
                const uint64_t* c1, const uint64_t* s, int64_t* mout);

#ifdef __cplusplus
}
#endif
#endif
