// C++ host API of the device-resident CKKS engine (namespace gpufhe).
//
// Same types, function names, argument order and meaning as the reference's
// C++ API, so its callers recompile against this header unchanged except for
// cudaStream_t -> hipStream_t.  Everything here is built on the extern "C"
// layer (gpufhe_capi.h, gpufhe_kernels.h); the compute is the library's HIP
// kernels.  Reference headers replaced (paths relative to the reference repo):
//   include/device_ciphertext.h, include/ckks_ops.h, include/keyswitch.h,
//   include/keyswitch_resident.h, include/keygen.h, plus the header-less
//   functions the reference's tests declare themselves (encode_host,
//   decode_host, native_*, pt_to_eval_host, ct_add_ct_host, rotate_ct_host,
//   automorphism_eval_host; e.g. tests/test_native_e2e.cpp:13-20).
//
// Errors: std::runtime_error (allocation, copy, unknown modulus, invalid
// arguments), as the reference.  Device work is asynchronous on the given
// stream; the caller synchronises.  A context is read-only and may be shared
// by streams; one DeviceKSWork per in-flight operation.
#pragma once
#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <vector>

struct gf_ctx;
struct gf_key;

namespace gpufhe {

enum class Scheme : uint8_t { CKKS, BFV, BGV };
enum class Format : uint8_t { COEFFICIENT, EVALUATION };

// include/device_ciphertext.h:15-43 -- owns c0/c1 (towers x n u64 each) in
// HBM.  Move-only.  upload/to_host are the only host transfers.
class DeviceCiphertext {
public:
    DeviceCiphertext(Scheme scheme, uint32_t n, uint32_t num_towers);
    ~DeviceCiphertext();
    DeviceCiphertext(const DeviceCiphertext&) = delete;
    DeviceCiphertext& operator=(const DeviceCiphertext&) = delete;
    DeviceCiphertext(DeviceCiphertext&&) noexcept;
    DeviceCiphertext& operator=(DeviceCiphertext&&) noexcept;

    void upload(const std::vector<uint64_t>& c0, const std::vector<uint64_t>& c1);
    void to_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1) const;

    uint64_t* c0() { return d_c0_; }
    uint64_t* c1() { return d_c1_; }
    const uint64_t* c0() const { return d_c0_; }
    const uint64_t* c1() const { return d_c1_; }
    uint32_t n() const { return n_; }
    uint32_t num_towers() const { return towers_; }
    uint32_t capacity_towers() const { return cap_; }
    Format format() const { return format_; }
    void set_format(Format f) { format_ = f; }
    // only the count changes; the buffers keep their size
    void drop_tower() {
        if (towers_) --towers_;
    }
    Scheme scheme() const { return scheme_; }

private:
    Scheme scheme_;
    Format format_ = Format::EVALUATION;
    uint32_t n_ = 0, towers_ = 0, cap_ = 0;
    uint64_t* d_c0_ = nullptr;
    uint64_t* d_c1_ = nullptr;
};

// include/ckks_ops.h:11-36 -- per-tower root tables for a relin-free chain.
// The pointers refer to the library's process-wide table cache (shared with
// every context, never freed), so the destructor frees nothing.
struct DeviceTables {
    std::vector<const uint64_t*> d_froots, d_fprecon, d_iroots, d_iprecon;
    std::vector<uint64_t> q, root;
    std::vector<uint64_t> ninv, ninv_precon;
    uint32_t n = 0;
    ~DeviceTables();
};
// roots: the first 2n-th root g^((q-1)/2n), g = 2, 3, ... (src/ckks_ops.cpp:32-38)
void build_device_tables(DeviceTables& T, uint32_t n, const std::vector<uint64_t>& moduli);
// a *= b per tower, both components (src/ckks_ops.cpp:88-98)
void mul_resident(DeviceCiphertext& a, const DeviceCiphertext& b, const DeviceTables& T,
                  hipStream_t stream);
// drop the last tower: y_i = a_i s1_i + NTT_i(switch(INTT(a_last))) s2_i, both
// components; scratch and dropCoeff are n-word device buffers (src/ckks_ops.cpp:101-133)
void rescale_resident(DeviceCiphertext& ct, const DeviceTables& T, const std::vector<uint64_t>& s1,
                      const std::vector<uint64_t>& s2, hipStream_t stream, uint64_t* scratch,
                      uint64_t* dropCoeff);

// include/keyswitch.h:11-55 -- hybrid keyswitch constants and eval key of one
// level, in the reference's layout.
struct KeySwitchConstants {
    uint32_t n = 0;
    uint32_t sizeQl = 0;
    uint32_t sizeP = 0;
    uint32_t numPart = 0;
    uint32_t alpha = 0;
    uint32_t fullQ = 0;
    std::vector<uint64_t> qMod;
    std::vector<uint64_t> pMod;
    std::vector<uint32_t> sizePart, sizeCompl, startIdx;
    std::vector<std::vector<uint64_t>> partQHatInv, partQHatInvPrec, partSrcMod, partQHatModp,
        partComplMod, partBMuLo, partBMuHi;
    // eval key per part: [evalKeyTowers][n], Q tower i at row i, P tower k at fullQ + k
    std::vector<std::vector<uint64_t>> av, bv;
    uint32_t evalKeyTowers = 0;
    std::vector<uint64_t> pHatInv, pHatInvPrec, pHatModq, mdBMuLo, mdBMuHi, pInvModq;
    // 2n-th root per modulus; a modulus absent here uses native_root
    std::vector<uint64_t> rootModList, rootValList;
};
struct KeySwitchResult {
    std::vector<uint64_t> ba0, ba1;
};
// src/ks_constants.cpp:31-102
void compute_keyswitch_constants(KeySwitchConstants& K, const std::vector<uint64_t>& moduliQ,
                                 const std::vector<uint64_t>& moduliP, uint32_t numPart);
// host vectors in and out (src/keyswitch.cpp:58-198): uploads, switches on the
// device, downloads.  Same result as keyswitch_resident.
KeySwitchResult keyswitch_core_resident(const std::vector<uint64_t>& aTowers,
                                        const KeySwitchConstants& K);

namespace detail {
struct KSImpl;
}
// include/keyswitch_resident.h:11-35 -- shared read-only device state.
struct DeviceKSContext {
    uint32_t n = 0, sizeQl = 0, sizeP = 0, sizeQlP = 0, numPart = 0, fullQ = 0, alpha = 0;
    std::vector<uint64_t> modList;                   // Q then P moduli of this context
    std::vector<const uint64_t*> d_fr, d_fp, d_ir, d_ip;   // root tables per modList entry
    std::vector<uint64_t> ninv, ninv_p;
    std::vector<uint32_t> sizePart, sizeCompl, startIdx;
    std::vector<uint64_t> pInvModqHost, qModHost, pModHost, rootList;
    std::shared_ptr<detail::KSImpl> impl;            // engine context + resident key
    gf_ctx* engine() const;
    gf_key* key() const;
};
// one per in-flight operation (include/keyswitch_resident.h:29-37)
struct DeviceKSWork {
    uint64_t* d_ws = nullptr;
    size_t bytes = 0;
};
DeviceKSContext ks_context_create(const KeySwitchConstants& K);
DeviceKSWork ks_work_create(const DeviceKSContext& C);
void ks_context_destroy(DeviceKSContext& C);
void ks_work_destroy(DeviceKSWork& W);

// src/keyswitch_resident.cpp:139-225 -- d_a (sizeQl towers, eval) -> (d_ba0, d_ba1)
void keyswitch_resident(const uint64_t* d_a, uint64_t* d_ba0, uint64_t* d_ba1, const DeviceKSContext& C,
                        DeviceKSWork& W, hipStream_t s);
// src/keyswitch_resident.cpp:227-244 -- one component in place, towers -> towers-1.
// s1/s2 must be the rescale constants of these moduli (native_rescale_consts);
// scratch/drop are accepted for signature compatibility: the batched rescale
// needs towers - 1 towers of scratch, so ks_context_create and every
// ks_work_create each allocate one L-tower scratch into a pool of the context,
// and a stream takes one from the pool on its first call and keeps it (a stream
// beyond the pool's count allocates one on its first call).
void rescale_resident_raw(uint64_t* d_c, uint32_t towers, const DeviceKSContext& C,
                          const std::vector<uint64_t>& s1, const std::vector<uint64_t>& s2,
                          uint64_t* d_scratch, uint64_t* d_drop, hipStream_t s);

// ---- first-class operations on resident ciphertexts ----
// Levels follow the inputs' tower counts; out must have at least the result's
// tower count, which it is lowered to.  `C` holds the relinearisation key
// (EvalMult) or the rotation key for the Galois index k (Rotate: k odd, < 2n,
// k = 5^r mod 2n rotates the slots by r).  Rescale and Rotate work in place.
void EvalMult(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out,
              const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s);
void Rescale(DeviceCiphertext& ct, const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s);
void EvalMultRescale(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out,
                     const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s);
void Rotate(DeviceCiphertext& ct, uint32_t k, const DeviceKSContext& rotC, DeviceKSWork& W,
            hipStream_t s);
void KeySwitch(const uint64_t* d_a, uint32_t towers, uint64_t* d_ba0, uint64_t* d_ba1,
               const DeviceKSContext& C, DeviceKSWork& W, hipStream_t s);
void EvalAdd(const DeviceCiphertext& a, const DeviceCiphertext& b, DeviceCiphertext& out,
             const DeviceKSContext& C, hipStream_t s);
// plaintext ops (src/ptops.cpp:1-26 on the device): d_pt holds a.num_towers()
// towers in evaluation form (pt_to_eval_host, uploaded); out may be a
void EvalMultPlain(const DeviceCiphertext& a, const uint64_t* d_pt, DeviceCiphertext& out,
                   const DeviceKSContext& C, hipStream_t s);
void EvalAddPlain(const DeviceCiphertext& a, const uint64_t* d_pt, DeviceCiphertext& out,
                  const DeviceKSContext& C, hipStream_t s);
// Hoisted rotations: ct.c1 is decomposed once (in W) and rotated by every
// Galois index ks[j] with the key of rotC[j] (rotation contexts of one
// parameter set); Wrot: one workspace per index.  outs[j] receives the
// rotation by ks[j].  Decrypts to the same rotations as Rotate, with other
// (equally small) noise, since the decomposition is of c1, not of sigma(c1).
void RotateHoisted(const DeviceCiphertext& ct, const std::vector<uint32_t>& ks,
                   const std::vector<const DeviceKSContext*>& rotC, std::vector<DeviceCiphertext>& outs,
                   DeviceKSWork& W, std::vector<DeviceKSWork>& Wrot, hipStream_t s);
// out = sum_j pt_j * rot_ks[j](ct) (diagonal method; kernels/stage_acc.cu:12-28
// for all terms at once), rotations hoisted as above; rotC[j] may be null
// with ks[j] = 1 for the unrotated term (at least one term must have a key).
// d_pts[j]: ct.num_towers() towers in evaluation form; at most 16 terms.
// Wrot[j] serves term j (it may be empty for the unrotated term); out may be ct.
void LinearTransform(const DeviceCiphertext& ct, const std::vector<uint32_t>& ks,
                     const std::vector<const DeviceKSContext*>& rotC, const std::vector<const uint64_t*>& d_pts,
                     DeviceCiphertext& out, DeviceKSWork& W, std::vector<DeviceKSWork>& Wrot, hipStream_t s);

// ---- keys, encryption, rotation (include/keygen.h:6-45) ----
struct KeyPairHost {
    std::vector<uint64_t> s;         // sizeQ x n, eval
    std::vector<uint64_t> pkA, pkB;  // sizeQ x n, eval
};
// 0 = uniform ternary secret, else h nonzero coefficients (src/keygen.cpp:18,53)
void set_secret_hamming_weight(uint32_t h);
KeyPairHost keygen_host(uint32_t n, const std::vector<uint64_t>& moduli, const std::vector<uint64_t>& roots,
                        uint64_t ns, double sigma, uint64_t seed);
void evalkeygen_host_sold(KeySwitchConstants& K, const std::vector<uint64_t>& sQP,
                          const std::vector<uint64_t>& sOld, const std::vector<uint64_t>& pkA_QP,
                          const std::vector<uint64_t>& pkB_QP, const std::vector<uint64_t>& PModq_QP,
                          const std::vector<uint64_t>& modQP, const std::vector<uint64_t>& rootQP,
                          uint64_t ns, double sigma, uint64_t seed);
void evalkeygen_host(KeySwitchConstants& K, const std::vector<uint64_t>& sQP,
                     const std::vector<uint64_t>& pkA_QP, const std::vector<uint64_t>& pkB_QP,
                     const std::vector<uint64_t>& PModq_QP, const std::vector<uint64_t>& modQP,
                     const std::vector<uint64_t>& rootQP, uint64_t ns, double sigma, uint64_t seed);
void encrypt_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, const std::vector<int64_t>& m,
                  const std::vector<uint64_t>& pkA, const std::vector<uint64_t>& pkB, uint32_t n,
                  const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root, uint64_t ns,
                  double sigma, uint64_t seed);
void decrypt_host(std::vector<int64_t>& mout, const std::vector<uint64_t>& c0, const std::vector<uint64_t>& c1,
                  const std::vector<uint64_t>& s, uint32_t n, const std::vector<uint64_t>& mod,
                  const std::vector<uint64_t>& root);

void automorphism_eval_device(uint64_t* d_v, uint32_t towers, uint32_t n, uint32_t k,
                              const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root,
                              uint64_t* d_scratch, hipStream_t s);
void rotate_ct_device(uint64_t* d_c0, uint64_t* d_c1, uint32_t k, const DeviceKSContext& C, DeviceKSWork& W,
                      uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod,
                      const std::vector<uint64_t>& root, uint64_t* d_scratch, uint64_t* d_b0, uint64_t* d_b1,
                      hipStream_t s);
void rotate_ct_resident(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, uint32_t k,
                        const KeySwitchConstants& Krot, uint32_t towers, uint32_t n,
                        const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root);
void rotate_ct_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, uint32_t k,
                    const KeySwitchConstants& Krot, uint32_t towers, uint32_t n,
                    const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root);
void automorphism_eval_host(std::vector<uint64_t>& v, uint32_t towers, uint32_t n, uint32_t k,
                            const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root);
void pt_to_eval_host(std::vector<uint64_t>& out, const std::vector<int64_t>& m, uint32_t towers, uint32_t n,
                     const std::vector<uint64_t>& mod, const std::vector<uint64_t>& root);
// src/ptops.cpp:13-19
void ct_add_ct_host(std::vector<uint64_t>& a0, std::vector<uint64_t>& a1, const std::vector<uint64_t>& b0,
                    const std::vector<uint64_t>& b1, uint32_t towers, uint32_t n,
                    const std::vector<uint64_t>& mod);
// src/ptops.cpp:20-26 -- (c0 pt, c1 pt), pt in evaluation form
void ct_mul_pt_host(std::vector<uint64_t>& c0, std::vector<uint64_t>& c1, const std::vector<uint64_t>& ptEval,
                    uint32_t towers, uint32_t n, const std::vector<uint64_t>& mod);

// ---- parameters (src/params.cpp:29-75) ----
void native_primes(std::vector<uint64_t>& out, uint32_t count, uint32_t bits, uint32_t n,
                   const std::vector<uint64_t>& avoid);
uint64_t native_root(uint32_t n, uint64_t q);
void native_rescale_consts(std::vector<uint64_t>& s1, std::vector<uint64_t>& s2,
                           const std::vector<uint64_t>& mod, uint32_t last);

// ---- CKKS canonical encoding (src/encode.cpp:13-44), O(n log n) ----
void encode_host(std::vector<int64_t>& m, const std::vector<std::complex<double>>& z, uint32_t N,
                 double Delta);
void decode_host(std::vector<std::complex<double>>& z, const std::vector<int64_t>& m, uint32_t N,
                 double Delta);

}  // namespace gpufhe
