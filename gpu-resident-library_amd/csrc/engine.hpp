// Host-side engine: resident parameter context, resident keys, per-op
// workspaces and "plans" (precomputed job arrays + the launch sequence of one
// operation over one batch of ciphertexts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

namespace gf {

// RAII device allocation
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    void alloc(size_t b) {
        release();
        if (b) hip_check(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

// device root tables of one modulus (shared process-wide, never freed)
struct RootTable {
    const uint64_t *fr, *fp, *ir, *ip;   // reference layout (separate arrays)
    const uint64_t *fw2, *iv2;           // interleaved (w, precon) pairs used by the plans
    // moduli below 2^50: the same twiddles as fp64 pairs (w centred, fl(w/q);
    // ntt_f64.hpp), else null
    const uint64_t *fw2d = nullptr, *iv2d = nullptr;
    uint64_t ninv, ninvp;
};
RootTable root_table(uint32_t n, uint64_t q, uint64_t root);
// phase-B row tables (twiddles of the run-unit pass in lane order) for radix rmax
struct RowTables {
    const uint64_t* fwd = nullptr;
    const uint64_t* inv = nullptr;
    const uint64_t* fwd_d = nullptr;   // of the fp64 pairs (moduli below 2^50), else null
    const uint64_t* inv_d = nullptr;
};
RowTables row_tables(uint32_t n, uint64_t q, uint64_t root, uint32_t rmax);

// constants of one level (sizeQl = l towers of Q)
struct PartConst {
    uint32_t st, sp, sc;
    std::vector<uint32_t> compl_rows;          // context rows of the complement towers
    std::vector<uint64_t> post, postp;         // n^-1 * QHatInv_i (INTT folding), per source
    const uint64_t* d_tab = nullptr;           // QHatModp * R mod target, [sp][sc]
    const uint64_t* d_tab60 = nullptr;         // the same times 2^60 instead of R = 2^64
    uint32_t c60 = 0;                          // 1: the sources sum to at most 1.2 2^60 (ModupJob::c60)
    const uint64_t* d_tq = nullptr;            // target moduli [sc]
    const uint64_t* d_tqneg = nullptr;
    const uint64_t* d_ttw = nullptr;           // [sc] addresses of the targets' paired forward tables
    const uint64_t* d_tred = nullptr;          // [sc] lazy_red_of(target modulus)
    const uint64_t* d_ttwd = nullptr;          // [sc] addresses of the targets' fp64 forward pairs, or 0
    const uint64_t* d_tqf = nullptr;           // [2 sc] target q and fl(1/q) as doubles
};
struct LevelConst {
    uint32_t l = 0, np = 0;
    std::vector<PartConst> parts;
    std::vector<uint64_t> pd_post, pd_postp;   // n^-1 * PHatInv_k, per P tower
    const uint64_t* d_tabMD = nullptr;         // PHatModq * R mod q_i, [K][l]
    const uint64_t* d_tqMD = nullptr;          // Q moduli [l]
    const uint64_t* d_tqnegMD = nullptr;
    std::vector<uint64_t> pinv, pinvp;         // P^-1 mod q_i
    std::vector<uint64_t> s1, s1p, s2, s2p;    // rescale constants dropping tower l-1
    std::vector<uint64_t> ps1, ps1p;           // P^-1 s1 mod q_i (fused ModDown + rescale)
    const uint64_t* d_tabRS = nullptr;         // PHatModq * (-P^-1 s1) * R mod q_i, [K][l-1]
    const RsTarget* d_rstg = nullptr;          // [l-1]
    const uint64_t* d_rs_sw = nullptr;         // [2 (l-1)] rs_sw_consts of each surviving tower
    // the surviving towers as targets of the fused ModDown (ModupJob tables, [l-1])
    const uint64_t* d_rs_tq = nullptr;
    const uint64_t* d_rs_tqneg = nullptr;
    const uint64_t* d_rs_ttw = nullptr;
    const uint64_t* d_rs_tred = nullptr;
    const uint64_t* d_rs_ttwd = nullptr;
    const uint64_t* d_rs_tqf = nullptr;        // [2 (l-1)]
};

// Resident parameter set: Q moduli (L), P moduli (K), part width alpha.
class Ctx {
public:
    Ctx(uint32_t n, const std::vector<uint64_t>& q, const std::vector<uint64_t>& p,
        const std::vector<uint64_t>& roots, uint32_t alpha);
    uint32_t n, logn, L, K, alpha, dnum;
    NttGeom geom;
    std::vector<uint64_t> mods, roots;        // L + K rows: Q then P
    std::vector<RowConst> rc;
    std::vector<RootTable> rt;
    std::vector<LevelConst> lv;               // index l = 1..L (0 unused)
    DevBuf consts;                            // all per-level device constants
    const uint64_t* d_zeros = nullptr;        // L + K zero words (a target table with no fp64 target)
    // EvalMult+Rescale with merged ModDown/rescale, and keyswitches whose
    // digit row pass feeds the key inner product directly (KSI)
    bool fused = true;
    uint32_t wide = 0;                        // 1: a modulus is 2^60 or more (generic conversion sums)
    std::vector<RowTables> rtr;               // per row: phase-B row tables for geom.rmax
    // register radix of transforms in plans built from now on (2..4)
    void set_radix(uint32_t r);
    // the row table matching a paired table of this context (fw2, iv2, or
    // their fp64 pairs fw2d, iv2d), else null
    const uint64_t* row_table_of(const uint64_t* paired) const;
    // the fp64 pairs of the same twiddles as a paired table (fw2 -> fw2d,
    // iv2 -> iv2d), null for moduli of 2^50 and above
    const uint64_t* f64_table_of(const uint64_t* paired) const;
    uint32_t row_q(uint32_t i) const { return i; }
    uint32_t row_p(uint32_t k) const { return L + k; }
    uint32_t num_parts(uint32_t l) const { return (l + alpha - 1) / alpha; }
    // workspace (uint64 words) needed by one operation at the top level
    size_t ws_words() const;
    size_t rescale_ws_words() const;   // scratch of a one-component rescale (L towers)
    // times the NTT variants on a synthetic tower batch and keeps the fastest
    // register radix in geom (plans built afterwards use it); returns it
    uint32_t autotune(hipStream_t s, float* best_ms = nullptr);
    // Plan knobs, read when a plan is built: one extra wave per fused row pass
    // + inner product workgroup (it only takes part in the inner product), and
    // the targets per job of the ModUp conversion and of the rescale prologue.
    uint32_t ksi_extra = 0, conv_chunk = 32, rs_chunk = 32;
    // 1: ModUp's inverse column pass, conversion and digit column pass run as
    // one kernel (modup.hpp) where the shape allows it
    uint32_t modup_fused = 1;
    // 1: plans with the fused ModUp run the transforms of moduli below 2^50
    // around it in double precision (ntt_f64.hpp): the inverse row pass
    // before it and its column passes
    uint32_t modup_f64 = 1;
    // 1 (with modup_f64): the row passes of the fused inner product after it
    // as well (fp64 butterflies issue in about half the cycles of the integer
    // Shoup butterflies)
    uint32_t ksi_f64 = 1;
    // 1: the tensor product and the inverse row pass of its t2 (in front of
    // the fused ModUp) run as one kernel per row tile, t2 handed over in LDS
    uint32_t tensor_ib = 1;
    // 1: with the fused ModDown, the P rows' inner product runs the inverse
    // row pass of its two output rows itself (the rows handed over in LDS)
    uint32_t ksi_rowpass = 1;
    // 1: the fused EvalMult+Rescale runs the P towers' inverse column pass,
    // the rescale prologue and the surviving towers' forward column pass as
    // one kernel (the fused ModUp kernel's rescale mode), where the fused
    // ModUp runs; stages below modup_min_wgs workgroups run them apart
    uint32_t moddown_fused = 1;
    // a plan whose fused ModUp stage would launch fewer workgroups than this
    // runs it unfused (a single operation's stage cannot fill the chip)
    uint32_t modup_min_wgs = 2048;
    // conversion and prologue stages that would launch fewer workgroups than
    // this are re-split into 8-target jobs, whatever the chunk knobs say
    uint32_t rechunk_wgs = 8192;
    // times each plan knob's candidates on a batch of `batch` EvalMult+Rescale
    // at the top level (the stage shapes plans of that batch run) and keeps
    // the fastest; the search (every candidate's time) is left in tune_report
    // as JSON
    void autotune_plan(hipStream_t s, uint32_t batch);
    std::string tune_report;
    // sets or reads a knob by name ("ntt_radix", "ksi_extra", "conv_chunk",
    // "rs_chunk", "modup_fused", "modup_f64", "ksi_f64", "moddown_fused", "tensor_ib", "ksi_rowpass",
    // "modup_min_wgs", "rechunk_wgs")
    void set_knob(const std::string& name, uint32_t v);
    uint32_t knob(const std::string& name) const;
    std::string radix_report;   // the radix search of autotune(), as JSON
    size_t tower_bytes() const { return (size_t)n * 8; }
};

// Resident switching key: rows_key = fullQ + K rows per part, Montgomery form.
struct DevKey {
    uint32_t nparts = 0, rows = 0, fullQ = 0;
    DevBuf a, b;  // [nparts][rows][n]
};
std::unique_ptr<DevKey> upload_key(const Ctx& C, const uint64_t* av, const uint64_t* bv,
                                   uint32_t nparts, uint32_t rows, uint32_t fullQ,
                                   hipStream_t s);
// ctr: u, e0, e1 drawn on the device by the counter sampler (ctr_sampler.hpp),
// else on the host in the reference's order (sampler.hpp)
std::unique_ptr<DevKey> generate_key(const Ctx& C, const uint64_t* sold, const uint64_t* pka, const uint64_t* pkb,
                                     uint64_t ns, double sigma, uint64_t seed, hipStream_t s, bool ctr = false);
// public key on the device with the counter sampler: s (evaluation form), a,
// b = ns e - a s over all L + K towers of the context ([L + K][n] each)
void keygen_device(const Ctx& C, uint64_t seed, double sigma, uint64_t ns, uint64_t* s_out, uint64_t* a_out,
                   uint64_t* b_out, hipStream_t st);
void download_key(const Ctx& C, const DevKey& k, uint64_t* av, uint64_t* bv, hipStream_t s);
// encryption of the coefficient vector h_m (n signed values) under the public
// key (pka, pkb: device [level][n], evaluation form) into device c0, c1
void encrypt_device(const Ctx& C, uint32_t level, const uint64_t* pka, const uint64_t* pkb, const int64_t* h_m,
                    uint64_t ns, double sigma, uint64_t seed, uint64_t* c0, uint64_t* c1, hipStream_t s);
// decryption of (c0, c1) under the secret s (device towers, evaluation form;
// tower 0 is used) into n centred coefficients on the host
void decrypt_device(const Ctx& C, const uint64_t* c0, const uint64_t* c1, const uint64_t* sk, int64_t* h_m,
                    hipStream_t s);

// One launch step of a plan
struct Step {
    // NTT_FA: a forward transform's first (column) pass only; KSI: the fused
    // row pass of the ModUp digits and the key inner product
    // NTT_IB: an inverse transform's first (row) pass only; MODUP: the fused
    // ModUp (modup.hpp), arg = the jobs' source count
    // TENSOR_IB: the tensor product fused with the inverse row pass of its t2
    enum Kind { NTT_F, NTT_I, TENSOR, CONV, INNER, COMBINE, SWITCH, FUSE, ADD, PERM, COPY, RSV, PTACC,
                NTT_FA, KSI, NTT_IB, MODUP, MODDOWN, NTT_FB, TENSOR_IB } kind;
    size_t off;     // byte offset of the first job in the plan's device buffer
    uint32_t count;
    // KSI: LDS regions (bits 0-7) and waves per workgroup (bits 8-15); the
    // jobs of one KSI step share their transformed-part count
    uint32_t arg = 0;
};

class Plan {
public:
    std::vector<Step> steps;
    DevBuf dev;                  // job arrays, when the plan owns them
    const void* base = nullptr;  // where the job arrays live (dev.p or a caller arena)
    uint32_t n = 0, logn = 0;
    NttGeom geom{};
    void run(hipStream_t s) const;
    // which: 3 = every kernel of step i; 1 / 2 = only its first / second kernel
    void run_step(size_t i, hipStream_t s, uint32_t which = 3) const;
    uint32_t step_kernels(size_t i) const;   // 2 for a two-phase transform, else 1
    void run_jobs(size_t i, hipStream_t s, uint32_t which, uint32_t first, uint32_t cnt) const;
    // every job of every step as its own launch: the one-launch-per-tower
    // structure of the reference's resident path, for comparison
    void run_per_job(hipStream_t s) const;
    // hipGraph of the whole sequence (captured on first use)
    hipGraphExec_t graph = nullptr;
    void run_graph(hipStream_t s);
    ~Plan();
};

// Collects jobs stage by stage for any number of operations, then uploads.
class PlanBuilder {
public:
    explicit PlanBuilder(const Ctx& C) : C_(C) {}
    // per-op recipes (all buffers are device pointers, tower-major)
    void keyswitch(uint32_t l, const uint64_t* d, uint64_t* ba0, uint64_t* ba1, const DevKey& key,
                   uint64_t* ws, const uint64_t* t0 = nullptr, const uint64_t* t1 = nullptr);
    void evalmult(uint32_t l, const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
                  const uint64_t* c1b, uint64_t* r0, uint64_t* r1, const DevKey& key, uint64_t* ws);
    void rescale(uint32_t l, const uint64_t* c0, const uint64_t* c1, uint64_t* o0, uint64_t* o1,
                 uint64_t* ws);
    // one component (o may equal c: every surviving tower is read before it is written)
    void rescale_one(uint32_t l, const uint64_t* c, uint64_t* o, uint64_t* ws);
    void evalmult_rescale(uint32_t l, const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
                          const uint64_t* c1b, uint64_t* o0, uint64_t* o1, const DevKey& key,
                          uint64_t* ws);
    void rotate(uint32_t l, uint64_t k, const uint64_t* c0, const uint64_t* c1, uint64_t* o0,
                uint64_t* o1, const DevKey& key, uint64_t* ws);
    // m rotations of (c0, c1) sharing one digit decomposition (ws_base), one
    // workspace per rotation
    void rotate_hoisted(uint32_t l, uint32_t m, const uint64_t* ks, const uint64_t* c0, const uint64_t* c1,
                        uint64_t* const* o0, uint64_t* const* o1, const DevKey* const* keys, uint64_t* ws_base,
                        uint64_t* const* ws_rot);
    void add(uint32_t l, const uint64_t* a0, const uint64_t* a1, const uint64_t* b0,
             const uint64_t* b1, uint64_t* o0, uint64_t* o1);
    // sum_j pts[j] * rot_ks[j](c0, c1) with hoisted rotations; keys[j] null and
    // ks[j] = 1 for the unrotated term
    void linear_transform(uint32_t l, uint32_t m, const uint64_t* ks, const uint64_t* c0, const uint64_t* c1,
                          const uint64_t* const* pts, uint64_t* o0, uint64_t* o1, const DevKey* const* keys,
                          uint64_t* ws_base, uint64_t* const* ws_rot);
    // plaintext (evaluation form, at least l towers) times / plus a ciphertext
    void mul_plain(uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                   uint64_t* o1);
    void add_plain(uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                   uint64_t* o1);
    std::unique_ptr<Plan> finish();
    // the same, with the job arrays staged in pinned host memory h_stage and
    // copied on stream s into the caller's device arena (both `cap` bytes): no
    // allocation and no host synchronisation.  The arena and the staging area
    // must stay untouched until the plan's work on s has completed.
    std::unique_ptr<Plan> finish_into(void* d_arena, void* h_stage, size_t cap, hipStream_t s);
    // the plan with its job arrays laid out in `blob` (host memory) only: the
    // caller copies them to device memory and sets Plan::base
    std::unique_ptr<Plan> finish_host(std::vector<char>& blob) { return layout(blob); }

private:
    const Ctx& C_;
    // stage lists in execution order
    std::vector<TensorJob> tensor_;
    std::vector<TensorIbJob> tensor_ib_;
    std::vector<PermJob> perm_;
    std::vector<NttJob> intt_up_;
    std::vector<NttJob> intt_upb_;   // inverse transforms whose column pass runs in the fused ModUp
    std::vector<ModupJob> modup_;
    struct ModupSrc {
        uint32_t l, pi;
        const uint64_t* x;
        uint64_t* dp;
    };
    std::vector<ModupSrc> modup_src_;   // what each fused ModUp job converts, to run it unfused
    std::vector<ConvJob> conv_up_;
    std::vector<NttJob> ntt_dig_;
    std::vector<NttJob> ntt_diga_;   // digits whose row pass runs inside the KSI step
    std::vector<InnerJob> inner_;
    std::vector<KsiJob> ksi_;
    std::vector<NttJob> intt_p_;
    std::vector<ConvJob> conv_dn_;
    std::vector<NttJob> ntt_conv_;
    std::vector<CombineJob> combine_;
    std::vector<AddJob> add_rot_;
    std::vector<CopyJob> copy_rot_;
    std::vector<NttJob> intt_drop_;
    std::vector<SwitchJob> switch_;
    std::vector<NttJob> ntt_sw_;
    std::vector<FuseJob> fuse_;
    std::vector<RsJob> rs_;
    std::vector<NttJob> ntt_v_;
    std::vector<NttJob> intt_pb_;    // P towers whose inverse column pass runs in the fused ModDown
    std::vector<ModupJob> moddown_;
    std::vector<NttJob> ntt_vb_;     // epilogue transforms whose column pass runs in the fused ModDown
    struct ModdownSrc {
        uint32_t l, c;
        const uint64_t *xp, *dt, *dr, *t, *res;
        uint64_t *vc, *out;
    };
    std::vector<ModdownSrc> moddown_src_;   // what each fused ModDown job computes, to run it apart
    // the rescale tail of one component of the fused EvalMult+Rescale: the
    // prologue and the epilogue transforms (apart), or the fused ModDown and
    // the epilogue's row pass
    void rescale_tail(const ModdownSrc& m, bool fused);
    bool use_moddown(uint32_t l) const;
    void evalmult_rescale_fused(uint32_t l, const uint64_t* c0a, const uint64_t* c1a,
                                const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0, uint64_t* o1,
                                const DevKey& key, uint64_t* ws);
    // lazy_rows: accumulator rows below it only feed the fused rescale
    // epilogue, so the fused inner product may leave them unreduced
    // p_rowpass: the P towers' inverse transforms stop after their row pass
    // (their column pass runs in the fused ModDown)
    void keyswitch_core(uint32_t l, const uint64_t* d, const DevKey& key, uint64_t* ws,
                        const uint64_t* extra0, const uint64_t* extra1, uint64_t* dx0, uint64_t* dx1,
                        uint32_t lazy_rows = 0, bool p_rowpass = false);
    void modup(uint32_t l, const uint64_t* d, uint64_t* ws, const uint64_t* extra0, const uint64_t* extra1,
               uint64_t* dx0, uint64_t* dx1, bool rowpass_in_ksi = false);
    // the digits' row pass fused with the key inner product (two-phase rings,
    // up to KSI_MAXP parts, fused schedule on)
    bool use_ksi(uint32_t l) const;
    bool use_modup(uint32_t l) const;
    void modup_unfused(uint32_t l, uint32_t pi, const uint64_t* x, uint64_t* dp, bool colpass_only);
    void ksi_moddown_src(uint32_t l, const uint64_t* d, const DevKey& key, uint64_t* ws, uint32_t lazy_rows,
                         bool p_rowpass = false);
    const uint64_t* digit_tower(uint32_t l, const uint64_t* d, const uint64_t* dig, uint32_t pi,
                                uint32_t i) const;
    // gk != 0: the digits are read through the automorphism of index gk
    template <class F>
    void inner_moddown_src(uint32_t l, const DevKey& key, uint64_t* ws, F digit, uint64_t gk = 0);
    void moddown_tail(uint32_t l, uint64_t* ws, uint64_t* ba0, uint64_t* ba1, const uint64_t* t0,
                      const uint64_t* t1);
    std::vector<PtAccJob> ptacc_;
    std::vector<AddJob> add_;
    std::vector<CopyJob> copy_;
    std::unique_ptr<Plan> layout(std::vector<char>& blob);
    void rescale_comp(uint32_t l, const uint64_t* c, uint64_t* o, uint64_t* drop, uint64_t* sw);
};

// workspace offsets inside one op's workspace (in towers)
struct WsLayout {
    size_t t0, t1, t2, x, dig, res0, res1, xp, drop, dt, dr, total;
    static WsLayout make(const Ctx& C);
};

}  // namespace gf
