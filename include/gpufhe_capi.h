/*
 * Handle-based C ABI of the device-resident CKKS engine (gfx950).
 *
 * The reference has no library-level EvalMult / Rescale / Rotate / KeySwitch:
 * its callers compose them from kernel launchers and the C++ resident API
 * (tests/test_evalmult_rescale.cpp:69-73, bench/bench_evalmult_resident.cpp:96-100,
 * src/rotate.cpp:118-132, src/keyswitch_resident.cpp:139-244).  These entry
 * points expose those compositions as first-class calls, over the same data
 * layout (tower-major [tower][n] uint64, evaluation form), plus batch plans
 * that run one operation over many ciphertexts in one launch per stage.
 *
 * Conventions: pointers named d_* / c* / o* / ws are device pointers; host
 * arrays are named h_*.  Device arrays need 8-byte alignment; when every array
 * of a stage is 16-byte aligned (as gf_malloc returns) the streaming kernels
 * move two residues per access, which is faster, with identical results.
 * Functions returning int give 0 on success and -1 on error (message in
 * gf_last_error()); constructors return NULL on error.
 * Operations are asynchronous on the given stream; the caller synchronises.
 */
#ifndef GPUFHE_CAPI_H
#define GPUFHE_CAPI_H
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gf_ctx gf_ctx;
typedef struct gf_key gf_key;
typedef struct gf_plan gf_plan;

const char* gf_last_error(void);

/* ---- device memory and streams (plumbing for callers without a runtime) ---- */
int gf_malloc(void** d_ptr, size_t bytes);
int gf_free(void* d_ptr);
int gf_h2d(void* d_dst, const void* h_src, size_t bytes);
int gf_d2h(void* h_dst, const void* d_src, size_t bytes);
int gf_d2d(void* d_dst, const void* d_src, size_t bytes);
int gf_device_sync(void);
int gf_stream_create(hipStream_t* s);
int gf_stream_destroy(hipStream_t s);
int gf_stream_sync(hipStream_t s);

/* ---- parameters ----
 * Native parameter generation of src/params.cpp:29-75 (OpenFHE-free path). */
int gf_native_primes(uint64_t* h_out, uint32_t count, uint32_t bits, uint32_t n,
                     const uint64_t* h_avoid, uint32_t navoid);
uint64_t gf_native_root(uint32_t n, uint64_t q);
void gf_native_rescale_consts(uint64_t* h_s1, uint64_t* h_s2, const uint64_t* h_mod,
                              uint32_t last);

/* Resident context for Q = q[0..L), P = p[0..K), one 2n-th root per modulus
 * (Q then P), dnum digits (part width alpha = ceil(L/dnum)).  Uploads root
 * tables and every level's ModUp / ModDown / rescale constants once.
 * Replaces ks_context_create's table and constant setup
 * (src/keyswitch_resident.cpp:32-101) for all levels at once. */
gf_ctx* gf_ctx_create(uint32_t n, const uint64_t* h_q, uint32_t L, const uint64_t* h_p,
                      uint32_t K, const uint64_t* h_roots, uint32_t dnum);
void gf_ctx_destroy(gf_ctx* c);
/* bytes of device workspace one in-flight operation needs (the DeviceKSWork
 * of src/keyswitch_resident.cpp:103-114, extended to EvalMult + Rescale) */
size_t gf_ws_bytes(const gf_ctx* c);

/* Switching key (relinearisation or rotation): host arrays av, bv laid out
 * [nparts][L+K][n] in evaluation form, as produced by evalkeygen_host_sold
 * (src/evalkeygen.cpp:42-96).  Uploaded once, kept resident. */
gf_key* gf_key_create(gf_ctx* c, const uint64_t* h_av, const uint64_t* h_bv, uint32_t nparts,
                      hipStream_t s);
/* Frees the key and evicts the context's cached single-op plans that read it
 * (they hold its device pointers). */
void gf_key_destroy(gf_key* k);
/* number of single-op plans the context holds in its cache (one per
 * operation and argument set called so far) */
uint32_t gf_ctx_cached_plans(gf_ctx* c);

/* Context with an explicit part width alpha (KeySwitchConstants::alpha of
 * include/keyswitch.h:17-19; ks_context_create builds one per level). */
gf_ctx* gf_ctx_create_alpha(uint32_t n, const uint64_t* h_q, uint32_t L, const uint64_t* h_p,
                            uint32_t K, const uint64_t* h_roots, uint32_t alpha);
uint32_t gf_ctx_towers(const gf_ctx* c);
/* Key rows in the reference's eval-key layout: h_av/h_bv [nparts][rows][n]
 * with rows = fullQ + K; Q tower i reads row i, P tower k reads row fullQ+k
 * (src/keyswitch_resident.cpp:84-93, KeySwitchConstants::evalKeyTowers). */
gf_key* gf_key_create_rows(gf_ctx* c, const uint64_t* h_av, const uint64_t* h_bv, uint32_t nparts,
                           uint32_t rows, uint32_t fullQ, hipStream_t s);
/* Synthetic switching key of uniform residues, generated on the device from
 * seed (benchmark key material: a switching key is computationally uniform). */
gf_key* gf_key_create_uniform(gf_ctx* c, uint32_t nparts, uint64_t seed, hipStream_t s);
/* Switching key generated on the device (evalkeygen_host_sold of
 * src/evalkeygen.cpp:42-96 without the host round trips): per part
 * av = pkA u + ns e1, bv = pkB u + ns e0 + (P mod q) sOld on the part's Q
 * towers, over all L+K towers of the context.  d_sold, d_pka, d_pkb: device
 * arrays [L+K][n] in evaluation form.  u, e0, e1 are drawn on the host from
 * mt19937_64(seed) in the reference's order with the library's Gaussian
 * (sigma), so the key equals evalkeygen_host_sold's for the same inputs. */
gf_key* gf_key_generate(gf_ctx* c, const uint64_t* d_sold, const uint64_t* d_pka, const uint64_t* d_pkb,
                        uint64_t ns, double sigma, uint64_t seed, hipStream_t s);
/* The same switching key with u, e0, e1 drawn on the device by the counter
 * sampler: draw(seed, domain, i) = SplitMix64's output function of
 * seed + (domain 2^32 + i + 1) 0x9E3779B97F4A7C15; ternary = draw mod 3 - 1,
 * Gaussian = the cumulative table's index of draw >> 1, negated when draw & 1;
 * part p uses domains 0x20000 + 3p (u), + 1 (e0), + 2 (e1).  Nothing is
 * sampled or computed on the host. */
gf_key* gf_key_generate_ctr(gf_ctx* c, const uint64_t* d_sold, const uint64_t* d_pka, const uint64_t* d_pkb,
                            uint64_t ns, double sigma, uint64_t seed, hipStream_t s);
/* Public key generation on the device (src/keygen.cpp:55-112, the reference's
 * keygen_host) with the counter sampler: s ternary (domain 1), e Gaussian
 * (domain 2), a of tower t uniform mod q_t (domain 0x10000 + t, taken as
 * evaluation form), b = ns e - a s, over all L + K towers of the context.
 * d_s, d_pka, d_pkb: device arrays [L+K][n], written in evaluation form;
 * returns after the key is complete.  Moduli must exceed 2^32. */
int gf_keygen(gf_ctx* c, uint64_t seed, double sigma, uint64_t ns, uint64_t* d_s, uint64_t* d_pka,
              uint64_t* d_pkb, hipStream_t s);
/* rows of a resident key in plain form, h_av/h_bv [nparts][rows][n] */
int gf_key_download(const gf_key* k, uint64_t* h_av, uint64_t* h_bv, hipStream_t s);
/* Encryption into resident buffers (encrypt_host of src/encdec.cpp:41-64 with
 * the algebra on the device): c0 = pkB v + ns e0 + m, c1 = pkA v + ns e1 over
 * the first `level` Q towers.  h_m: n signed message coefficients on the host;
 * d_pka, d_pkb: device [level][n], evaluation form; d_c0, d_c1: device
 * [level][n].  v, e0, e1 are drawn from mt19937_64(seed) in the reference's
 * order with the library's Gaussian (sigma), so the result equals
 * encrypt_host's for the same inputs.  Returns after the device work. */
int gf_encrypt(gf_ctx* c, uint32_t level, const uint64_t* d_pka, const uint64_t* d_pkb, const int64_t* h_m,
               uint64_t ns, double sigma, uint64_t seed, uint64_t* d_c0, uint64_t* d_c1, hipStream_t s);
/* Decryption from resident buffers (decrypt_host of src/encdec.cpp:68-81):
 * tower 0 of c0 + c1 s, inverse-transformed and centred into h_m (n values on
 * the host).  d_s: the secret in evaluation form (at least tower 0). */
int gf_decrypt(gf_ctx* c, const uint64_t* d_c0, const uint64_t* d_c1, const uint64_t* d_s, int64_t* h_m,
               hipStream_t s);
void gf_key_shape(const gf_key* k, uint32_t* nparts, uint32_t* rows, uint32_t* fullQ);

/* ---- tower utilities ----
 * Device root tables of (n, q, root) in the reference layout
 * (src/keyswitch.cpp:22-35): fr = psi^brv, ir = psi^-brv, Shoup precons.
 * Shared process-wide and never freed (src/keyswitch_resident.cpp:47-68). */
int gf_root_tables(uint32_t n, uint64_t q, uint64_t root, const uint64_t** fr, const uint64_t** fp,
                   const uint64_t** ir, const uint64_t** ip, uint64_t* ninv, uint64_t* ninv_precon);
/* In-place forward (inverse = 0) or inverse NTT of `towers` consecutive
 * towers, tower t under h_mods[t] / h_roots[t] (one launch per tower; no
 * allocation, so it may be captured). */
int gf_ntt_towers(uint32_t n, uint64_t* d_x, uint32_t towers, const uint64_t* h_mods,
                  const uint64_t* h_roots, int inverse, hipStream_t s);
/* Galois automorphism sigma_k on evaluation-form towers, out of place
 * (the INTT -> LaunchAutomorphPermute -> NTT of src/rotate.cpp:70-84 as
 * one permutation of the bit-reversed evaluation points). */
int gf_automorph(uint32_t n, const uint64_t* d_in, uint64_t* d_out, uint32_t towers, uint64_t k,
                 hipStream_t s);

/* ---- single operations (level = number of Q towers of the inputs) ---- */
/* tensor -> keyswitch -> combine -> rescale; outputs have level-1 towers */
int gf_evalmult_rescale(gf_ctx* c, gf_key* rlk, uint32_t level, const uint64_t* c0a,
                        const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                        uint64_t* o0, uint64_t* o1, uint64_t* ws, hipStream_t s);
/* gf_evalmult_rescale that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace and the plan's job arrays
 * live in the library's static device array (gf_noalloc_scratch_bytes: 160 MiB
 * per GPU, held until the process ends), sized for one operation up to
 * n = 2^16, L = 24, K = 6, dnum = 4; larger contexts are rejected.  Every call
 * on a GPU shares that array, so calls that overlap on one GPU (other streams,
 * other threads) race; calls on one stream are ordered.  The context must have
 * been created on the current device.  Nothing in the array needs to start at
 * zero: every word an operation reads it wrote first. */
int gf_evalmult_rescale_noalloc(gf_ctx* c, gf_key* rlk, uint32_t level, const uint64_t* c0a,
                                const uint64_t* c1a, const uint64_t* c0b, const uint64_t* c1b,
                                uint64_t* o0, uint64_t* o1, hipStream_t s);
size_t gf_noalloc_scratch_bytes(void);
/* tensor -> keyswitch -> combine (no rescale) */
int gf_evalmult(gf_ctx* c, gf_key* rlk, uint32_t level, const uint64_t* c0a, const uint64_t* c1a,
                const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0, uint64_t* o1, uint64_t* ws,
                hipStream_t s);
/* hybrid keyswitch of d: (ba0, ba1) over Q_level */
int gf_keyswitch(gf_ctx* c, gf_key* k, uint32_t level, const uint64_t* d, uint64_t* ba0,
                 uint64_t* ba1, uint64_t* ws, hipStream_t s);
/* drop the last tower of both components: outputs have level-1 towers */
int gf_rescale(gf_ctx* c, uint32_t level, const uint64_t* c0, const uint64_t* c1, uint64_t* o0,
               uint64_t* o1, uint64_t* ws, hipStream_t s);
/* one component: d_o (level-1 towers) may equal d_c (in place, as
 * rescale_resident_raw, src/keyswitch_resident.cpp:227-244).  ws: at least
 * gf_rescale_ws_bytes (a gf_ws_bytes workspace also serves) */
int gf_rescale_comp(gf_ctx* c, uint32_t level, const uint64_t* d_c, uint64_t* d_o, uint64_t* ws,
                    hipStream_t s);
/* scratch of gf_rescale_comp: L towers of n words */
size_t gf_rescale_ws_bytes(const gf_ctx* c);
/* (sigma_k(c0) + ba0, ba1), (ba0, ba1) = KS(sigma_k(c1)); k odd, < 2n */
int gf_rotate(gf_ctx* c, gf_key* rk, uint32_t level, uint64_t k, const uint64_t* c0,
              const uint64_t* c1, uint64_t* o0, uint64_t* o1, uint64_t* ws, hipStream_t s);
/* componentwise sum (EvalAdd) */
int gf_add(gf_ctx* c, uint32_t level, const uint64_t* a0, const uint64_t* a1, const uint64_t* b0,
           const uint64_t* b1, uint64_t* o0, uint64_t* o1, hipStream_t s);
/* plaintext ops; pt: `level` towers in evaluation form (the reference's pt_eval,
 * src/ptops.cpp:1-5).  mul: (c0 pt, c1 pt) as ct_mul_pt_host
 * (src/ptops.cpp:20-26) on the device; add: (c0 + pt, c1).  Outputs may
 * equal the inputs. */
int gf_mul_plain(gf_ctx* c, uint32_t level, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt,
                 uint64_t* o0, uint64_t* o1, hipStream_t s);
int gf_add_plain(gf_ctx* c, uint32_t level, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt,
                 uint64_t* o0, uint64_t* o1, hipStream_t s);
/* single-ciphertext forms of gf_plan_rotate_hoisted / gf_plan_linear_transform
 * below (same arguments with count = 1; o0/o1/ws_rot hold m entries for the
 * hoisted rotation) */
int gf_rotate_hoisted(gf_ctx* c, uint32_t level, const uint64_t* c0, const uint64_t* c1, uint32_t m,
                      gf_key* const* rks, const uint64_t* h_k, uint64_t* const* o0, uint64_t* const* o1,
                      uint64_t* ws_base, uint64_t* const* ws_rot, hipStream_t s);
int gf_linear_transform(gf_ctx* c, uint32_t level, const uint64_t* c0, const uint64_t* c1, uint32_t m,
                        gf_key* const* rks, const uint64_t* h_k, const uint64_t* const* pts, uint64_t* o0,
                        uint64_t* o1, uint64_t* ws_base, uint64_t* const* ws_rot, hipStream_t s);

/* ---- batch plans: one launch per stage for `count` operations ----
 * Host arrays of device pointers, one entry per operation; levels may differ.
 * Each operation needs its own workspace of gf_ws_bytes(). */
gf_plan* gf_plan_evalmult_rescale(gf_ctx* c, gf_key* rlk, uint32_t count, const uint32_t* h_levels,
                                  const uint64_t* const* c0a, const uint64_t* const* c1a,
                                  const uint64_t* const* c0b, const uint64_t* const* c1b,
                                  uint64_t* const* o0, uint64_t* const* o1, uint64_t* const* ws);
/* per-operation key handles (e.g. one rotation key per distinct index) */
gf_plan* gf_plan_rotate(gf_ctx* c, gf_key* const* rks, uint32_t count, const uint32_t* h_levels,
                        const uint64_t* h_k, const uint64_t* const* c0, const uint64_t* const* c1,
                        uint64_t* const* o0, uint64_t* const* o1, uint64_t* const* ws);
/* hoisted rotations: each of `count` ciphertexts rotated by all m indices
 * h_k[0..m) (rotation keys rks[0..m)); the digit decomposition of c1 is done
 * once per ciphertext in ws_base[i].  Outputs and per-rotation workspaces
 * ws_rot are indexed [i*m + j].  Not in the reference (it decomposes
 * sigma_k(c1) for every rotation, src/rotate.cpp:118-132): results decrypt to
 * the same rotation with different (equally small) noise. */
gf_plan* gf_plan_rotate_hoisted(gf_ctx* c, uint32_t count, const uint32_t* h_levels,
                                const uint64_t* const* c0, const uint64_t* const* c1, uint32_t m,
                                gf_key* const* rks, const uint64_t* h_k, uint64_t* const* o0,
                                uint64_t* const* o1, uint64_t* const* ws_base, uint64_t* const* ws_rot);
/* plaintext-weighted sums of rotations (the diagonal method of linear
 * transforms): o_i = sum_j pts[j] * rot_{h_k[j]}(c_i) for each of `count`
 * ciphertexts, the m rotations of one ciphertext hoisted over one
 * decomposition; rks[j] may be NULL with h_k[j] = 1 for the unrotated term.
 * pts[j]: device plaintexts in evaluation form (at least the ciphertext's
 * towers), shared by all ciphertexts.  The accumulate generalises
 * LaunchPtMulAcc (kernels/stage_acc.cu:12-28) to all terms in one pass.
 * ws_rot indexed [i*m + j]; m <= 16. */
gf_plan* gf_plan_linear_transform(gf_ctx* c, uint32_t count, const uint32_t* h_levels,
                                  const uint64_t* const* c0, const uint64_t* const* c1, uint32_t m,
                                  gf_key* const* rks, const uint64_t* h_k, const uint64_t* const* pts,
                                  uint64_t* const* o0, uint64_t* const* o1, uint64_t* const* ws_base,
                                  uint64_t* const* ws_rot);
gf_plan* gf_plan_add(gf_ctx* c, uint32_t count, const uint32_t* h_levels,
                     const uint64_t* const* a0, const uint64_t* const* a1,
                     const uint64_t* const* b0, const uint64_t* const* b1, uint64_t* const* o0,
                     uint64_t* const* o1);
/* batch forms of gf_mul_plain / gf_add_plain: one plaintext per operation */
gf_plan* gf_plan_mul_plain(gf_ctx* c, uint32_t count, const uint32_t* h_levels, const uint64_t* const* c0,
                           const uint64_t* const* c1, const uint64_t* const* pt, uint64_t* const* o0,
                           uint64_t* const* o1);
gf_plan* gf_plan_add_plain(gf_ctx* c, uint32_t count, const uint32_t* h_levels, const uint64_t* const* c0,
                           const uint64_t* const* c1, const uint64_t* const* pt, uint64_t* const* o0,
                           uint64_t* const* o1);
int gf_plan_run(gf_plan* p, hipStream_t s);
/* the same work with every job (one tower of one stage) as its own launch:
 * the launch structure of the reference's resident path, for comparison */
int gf_plan_run_per_job(gf_plan* p, hipStream_t s);
/* capture the plan into a hipGraph on first call (s must not be the null
 * stream), then replay it */
int gf_plan_run_graph(gf_plan* p, hipStream_t s);
uint32_t gf_plan_launches(const gf_plan* p);
void gf_plan_destroy(gf_plan* p);

/* ---- plan introspection, for timing one stage on its own stream ----
 * kind: 0 fwd NTT, 1 inverse NTT, 2 tensor, 3 basis conversion, 4 key inner
 * product, 5 ModDown combine, 6 switch modulus, 7 rescale fuse, 8 add,
 * 9 automorphism, 10 copy, 11 dropped-tower build, 12 fused ModDown/rescale
 * prologue.  jobs = towers (or conversion chunks) per launch. */
uint32_t gf_plan_num_steps(const gf_plan* p);
int gf_plan_step_info(const gf_plan* p, uint32_t i, uint32_t* kind, uint32_t* jobs);
int gf_plan_run_step(gf_plan* p, uint32_t i, hipStream_t s);
/* Kernels in step i (2 for a transform of a ring split into two passes, else
 * 1) and one of them alone (which = 1 or 2): for per-kernel timing. */
uint32_t gf_plan_step_kernels(const gf_plan* p, uint32_t i);
int gf_plan_run_step_kernel(gf_plan* p, uint32_t i, uint32_t which, hipStream_t s);

/* ---- tuning ----
 * NTT register radix log2 (2, 3 or 4) used by plans built afterwards; the
 * autotuner times each on a synthetic 256 MiB tower batch (forward and
 * inverse, in place) and keeps the fastest; ms_out is per 96 towers. */
int gf_ctx_set_ntt_radix(gf_ctx* c, uint32_t rmax);
uint32_t gf_ctx_get_ntt_radix(const gf_ctx* c);
int gf_ctx_autotune(gf_ctx* c, hipStream_t s, uint32_t* rmax_out, float* ms_out);
/* Plan knobs, searched on a batch of `batch` (1..256) EvalMult+Rescale at the
 * context's top level, the batch plans will run (each candidate builds the
 * plan and times the stages the knob shapes): "ksi_extra" (one
 * extra wave in the fused row pass + inner product workgroups, 0/1),
 * "conv_chunk" and "rs_chunk" (targets per ModUp-conversion / rescale-prologue
 * job), "modup_fused" (ModUp as one kernel per part and column tile, 0/1),
 * "modup_f64" (that kernel's column passes and the inverse row pass before
 * it in double precision for moduli below 2^50, 0/1), "ksi_f64" (with
 * modup_f64, the fused inner product's row passes as well, 0/1),
 * "moddown_fused" (EvalMult+Rescale's P-tower inverse column pass, rescale
 * prologue and surviving towers' forward column pass as one kernel, 0/1),
 * "tensor_ib" (with modup_fused, the tensor product and the inverse row pass
 * of its t2 as one kernel per row tile, 0/1), "ksi_rowpass" (with
 * moddown_fused, the P part's inverse row pass inside the P rows' inner
 * product launch, from its LDS, 0/1).
 * The report is JSON with every candidate's time, the radix search of
 * gf_ctx_autotune included; valid until the next tuning call.  Knobs (and
 * "ntt_radix", "rechunk_wgs", "modup_min_wgs") can be set by name; plans built
 * afterwards use them, and results are identical for every setting. */
int gf_ctx_autotune_plan(gf_ctx* c, uint32_t batch, hipStream_t s);
const char* gf_ctx_tuning_report(const gf_ctx* c);
int gf_ctx_set_knob(gf_ctx* c, const char* name, uint32_t value);
int gf_ctx_get_knob(const gf_ctx* c, const char* name, uint32_t* value);
/* 1 (default): EvalMult+Rescale merges ModDown and the rescale through NTT
 * linearity (one forward transform per surviving tower); 0: the reference's
 * stage sequence (ModDown NTTs, combine, INTT/switch/NTT/fuse).  Results are
 * identical; only plans built afterwards are affected. */
int gf_ctx_set_fusion(gf_ctx* c, int on);

/* ---- synthetic data: fill `level` towers at d with residues mod q_0..q_{level-1}
 * derived from (seed, index) by a 64-bit mixing hash (benchmark inputs) ---- */
int gf_fill_uniform(gf_ctx* c, uint64_t* d, uint32_t level, uint64_t seed, hipStream_t s);

/* ---- events (plumbing for timing on a caller stream) ---- */
int gf_event_create(hipEvent_t* e);
int gf_event_destroy(hipEvent_t e);
int gf_event_record(hipEvent_t e, hipStream_t s);
int gf_event_elapsed_ms(float* ms, hipEvent_t start, hipEvent_t stop);

/* ---- stream capture: record what a caller launches on s (plans, the
 * launchers of gpufhe_kernels.h, single-op calls whose plans are already
 * cached) into a replayable graph, as the reference's benchmark does with
 * cudaStreamBeginCapture (bench/bench_evalmult_resident.cpp:93-102) ---- */
int gf_capture_begin(hipStream_t s);
int gf_capture_end(hipStream_t s, hipGraphExec_t* out);
int gf_graph_launch(hipGraphExec_t g, hipStream_t s);
int gf_graph_destroy(hipGraphExec_t g);

/* ---- wire format: ciphertexts and switching keys with a parameter header ----
 * The reference keeps raw tower-major arrays (src/device_ciphertext.cpp:43-61);
 * these frame the same arrays with the ring size, tower count, moduli, a
 * scale and an FNV-1a checksum, and refuse data that does not match.
 * Ciphertext functions work on host arrays (towers x n each). */
size_t gf_ct_packed_bytes(uint32_t n, uint32_t towers);
int gf_ct_pack(const uint64_t* h_c0, const uint64_t* h_c1, uint32_t n, uint32_t towers,
               const uint64_t* h_moduli, double scale, void* out, size_t cap);
int gf_ct_unpack_header(const void* buf, size_t len, uint32_t* n, uint32_t* towers, double* scale);
int gf_ct_unpack(const void* buf, size_t len, uint64_t* h_moduli, uint64_t* h_c0, uint64_t* h_c1);
/* a resident key to bytes (plain form) and back onto a context with the same moduli */
size_t gf_key_packed_bytes(const gf_key* k);
int gf_key_pack(const gf_key* k, void* out, size_t cap, hipStream_t s);
gf_key* gf_key_unpack(gf_ctx* c, const void* buf, size_t len, hipStream_t s);

/* ---- serving: continuous batching over resident ciphertexts ----
 * No counterpart in the reference (it serves nothing).  A server owns a few
 * lanes (stream + job arena) and a pool of per-operation workspace slots
 * capped by ws_budget_bytes (admission by HBM footprint).  Requests wait in
 * arrival order; each step retires finished batches and, for every idle lane,
 * admits the waiting requests of the oldest request's ring that fit, as one
 * plan.  Batches on different lanes overlap.  Times are milliseconds on the
 * server clock (completion = device event).  Single-threaded: call a server
 * from one thread at a time. */
enum { GF_OP_EVALMULT_RESCALE = 0, GF_OP_ROTATE = 1, GF_OP_ADD = 2 };
typedef struct gf_server gf_server;
typedef struct {
    uint32_t ring;        /* from gf_server_add_ring */
    uint32_t op;          /* GF_OP_* */
    uint32_t level;       /* Q towers of the inputs */
    uint32_t rot;         /* rotation-key index of the ring (GF_OP_ROTATE) */
    const uint64_t *a0, *a1, *b0, *b1;   /* device inputs (b unused by rotate) */
    uint64_t *o0, *o1;    /* device outputs, or NULL: result in the server's slot,
                             valid until the slot serves another request */
    double arrival_ms;    /* submit: < 0 means now; trace: offset from the start */
} gf_request;
typedef struct {
    uint64_t id;
    double arrival_ms, launch_ms, done_ms;
    uint32_t batch_size, ring, op, level;
} gf_completion;

gf_server* gf_server_create(size_t ws_budget_bytes, uint32_t lanes, uint32_t max_batch);
void gf_server_destroy(gf_server* srv);
/* a ring: context, relinearisation key (may be NULL), nrot rotation keys with
 * their automorphism indices */
int gf_server_add_ring(gf_server* srv, gf_ctx* c, gf_key* rlk, gf_key* const* rot_keys,
                       const uint64_t* h_rot_k, uint32_t nrot, uint32_t* ring_out);
size_t gf_server_slot_bytes(const gf_server* srv, uint32_t ring);
double gf_server_now_ms(const gf_server* srv);
int gf_server_submit(gf_server* srv, const gf_request* req, uint64_t* id_out);
/* retire finished batches and launch new ones; never blocks on the device */
int gf_server_step(gf_server* srv, uint32_t* launched);
uint32_t gf_server_poll(gf_server* srv, gf_completion* out, uint32_t max);
uint32_t gf_server_pending(const gf_server* srv);
int gf_server_drain(gf_server* srv);
/* replay `count` requests in real time (arrival_ms ascending, offsets from
 * the call); out[i] receives request i's completion */
int gf_server_run_trace(gf_server* srv, uint32_t count, const gf_request* reqs,
                        gf_completion* out);
int gf_server_stats(const gf_server* srv, uint64_t* batches, uint64_t* admitted,
                    size_t* slot_bytes_used);

#ifdef __cplusplus
}
#endif
#endif
