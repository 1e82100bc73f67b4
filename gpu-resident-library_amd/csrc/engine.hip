// Host-side engine implementation (see engine.hpp).
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <algorithm>
#include <sstream>
#include <tuple>

#include "ctr_sampler.hpp"
#include "engine.hpp"
#include "sampler.hpp"

namespace gf {

// ------------------------------------------------------------ root tables --
// fr[i] = psi^brv(i), ir[i] = psi^-brv(i) with Shoup precons, n^-1: the table
// layout of the reference (src/keyswitch.cpp:22-35).  Tables depend only on
// (n, q, psi); they are built once per process and shared by every context
// (the reference caches them the same way, src/keyswitch_resident.cpp:47-68).
RootTable root_table(uint32_t n, uint64_t q, uint64_t root) {
    static std::mutex mu;
    static std::map<std::tuple<uint32_t, uint64_t, uint64_t>, RootTable> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(n, q, root);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    uint32_t logn = 0;
    while ((1u << logn) < n) ++logn;
    // moduli below 2^50 also get the fp64 pairs of both tables (ntt_f64.hpp)
    const bool f64 = f64_modulus(q);
    std::vector<uint64_t> h((size_t)(f64 ? 12 : 8) * n), pw(n);
    const uint64_t bases[2] = {root, host_invmod(root, q)};
    for (int s = 0; s < 2; ++s) {
        pw[0] = 1;
        for (uint32_t i = 1; i < n; ++i) pw[i] = host_mulmod(pw[i - 1], bases[s], q);
        uint64_t* tr = h.data() + (size_t)(2 * s) * n;
        uint64_t* tp = tr + n;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t r = 0, v = i;
            for (uint32_t k = 0; k < logn; ++k) { r = (r << 1) | (v & 1); v >>= 1; }
            tr[i] = pw[r];
            tp[i] = host_precon(pw[r], q);
            uint64_t* pr = h.data() + (size_t)(4 + 2 * s) * n;   // paired copy
            pr[2 * i] = tr[i];
            pr[2 * i + 1] = tp[i];
            if (f64) {
                double wd, wq;
                f64_twiddle(tr[i], q, wd, wq);
                uint64_t* pd = h.data() + (size_t)(8 + 2 * s) * n;
                std::memcpy(pd + 2 * i, &wd, 8);
                std::memcpy(pd + 2 * i + 1, &wq, 8);
            }
        }
    }
    void* d = nullptr;
    hip_check(hipMalloc(&d, h.size() * 8), "hipMalloc root table");
    hip_check(hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice), "upload root table");
    const uint64_t* b = (const uint64_t*)d;
    RootTable t;
    t.fr = b;
    t.fp = b + n;
    t.ir = b + 2 * (size_t)n;
    t.ip = b + 3 * (size_t)n;
    t.fw2 = b + 4 * (size_t)n;
    t.iv2 = b + 6 * (size_t)n;
    t.fw2d = f64 ? b + 8 * (size_t)n : nullptr;
    t.iv2d = f64 ? b + 10 * (size_t)n : nullptr;
    t.ninv = host_invmod(n, q);
    t.ninvp = host_precon(t.ninv, q);
    cache.emplace(key, t);
    return t;
}

// Phase-B row tables (build_row_table) of one modulus for the geometry of
// radix rmax, forward and inverse, built from the paired tables on first use
// and kept for the process like the root tables.  Null for single-phase n.
RowTables row_tables(uint32_t n, uint64_t q, uint64_t root, uint32_t rmax) {
    static std::mutex mu;
    static std::map<std::tuple<uint32_t, uint64_t, uint64_t, uint32_t>, RowTables> cache;
    std::lock_guard<std::mutex> lk(mu);
    const NttGeom g = make_geom(n, rmax);
    if (g.single) return RowTables{};
    auto key = std::make_tuple(n, q, root, rmax);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const RootTable t = root_table(n, q, root);
    const size_t words = row_table_words(g);
    // forward, inverse, and for moduli below 2^50 their fp64 pairs (entries
    // of two words either way, so the same permutation)
    const uint64_t* src[4] = {t.fw2, t.iv2, t.fw2d, t.iv2d};
    const int nt = t.fw2d ? 4 : 2;
    std::vector<uint64_t> pair(2 * (size_t)n), out(nt * words);
    void* d = nullptr;
    hip_check(hipMalloc(&d, out.size() * 8), "hipMalloc row table");
    for (int k = 0; k < nt; ++k) {
        hip_check(hipMemcpy(pair.data(), src[k], pair.size() * 8, hipMemcpyDeviceToHost), "read table");
        build_row_table(pair.data(), g, out.data() + k * words);
    }
    hip_check(hipMemcpy(d, out.data(), out.size() * 8, hipMemcpyHostToDevice), "upload row table");
    const uint64_t* b = (const uint64_t*)d;
    RowTables r{b, b + words, nt == 4 ? b + 2 * words : nullptr, nt == 4 ? b + 3 * words : nullptr};
    cache.emplace(key, r);
    return r;
}

static uint64_t prod_mod(const std::vector<uint64_t>& mods, size_t b, size_t e, uint64_t q,
                         size_t skip) {
    uint64_t r = 1 % q;
    for (size_t k = b; k < e; ++k)
        if (k != skip) r = host_mulmod(r, mods[k] % q, q);
    return r;
}

// ------------------------------------------------------------------- Ctx --
Ctx::Ctx(uint32_t n_, const std::vector<uint64_t>& q, const std::vector<uint64_t>& p,
         const std::vector<uint64_t>& roots_, uint32_t alpha_)
    : n(n_), L((uint32_t)q.size()), K((uint32_t)p.size()), alpha(alpha_) {
    if (n < 2 || (n & (n - 1)) || n > (1u << 17)) throw std::runtime_error("gpufhe: n must be a power of two in [2, 2^17]");
    if (L == 0 || K == 0) throw std::runtime_error("gpufhe: need at least one Q and one P modulus");
    if (roots_.size() != L + K) throw std::runtime_error("gpufhe: need one root per modulus");
    if (alpha == 0 || alpha > L) throw std::runtime_error("gpufhe: bad part width");
    if (alpha > CONV_MAX_SRC || K > CONV_MAX_SRC)
        throw std::runtime_error("gpufhe: part width and P tower count must be <= 16");
    logn = 0;
    while ((1u << logn) < n) ++logn;
    geom = make_geom(n);
    dnum = (L + alpha - 1) / alpha;
    if (dnum > MAXP) throw std::runtime_error("gpufhe: too many digits");
    mods = q;
    mods.insert(mods.end(), p.begin(), p.end());
    roots = roots_;
    uint64_t qmax = 0;
    for (uint64_t m : mods) {
        if (m < 3 || (m >> 61)) throw std::runtime_error("gpufhe: moduli must be odd and < 2^61");
        if ((m - 1) % (2ull * n)) throw std::runtime_error("gpufhe: modulus is not 1 mod 2n");
        qmax = m > qmax ? m : qmax;
    }
    wide = (qmax >> 60) ? 1u : 0u;
    // one REDC of a sum of up to max(alpha, K, dnum) products must stay below q 2^64
    const uint32_t terms = alpha > K ? (alpha > dnum ? alpha : dnum) : (K > dnum ? K : dnum);
    if ((unsigned __int128)terms * qmax >= ((unsigned __int128)1 << 64))
        throw std::runtime_error("gpufhe: moduli too large for the single-REDC accumulators");
    rc.resize(L + K);
    rt.resize(L + K);
    for (uint32_t r = 0; r < L + K; ++r) {
        const uint64_t m = mods[r];
        rt[r] = root_table(n, m, roots[r]);
        rc[r] = RowConst{m, host_qneg(m), host_r2(m), host_r1(m), rt[r].ninv, rt[r].ninvp};
    }
    set_radix(geom.rmax);

    // per-level constants; device tables go into one buffer
    std::vector<uint64_t> blob;
    struct Fix { const uint64_t** dst; size_t off; };
    std::vector<Fix> fix;
    auto put = [&](const std::vector<uint64_t>& v, const uint64_t** dst) {
        fix.push_back({dst, blob.size()});
        blob.insert(blob.end(), v.begin(), v.end());
    };
    lv.resize(L + 1);
    for (uint32_t l = 1; l <= L; ++l) {
        LevelConst& c = lv[l];
        c.l = l;
        c.np = num_parts(l);
        c.parts.resize(c.np);
        for (uint32_t pi = 0; pi < c.np; ++pi) {
            PartConst& P = c.parts[pi];
            P.st = alpha * pi;
            P.sp = (l > P.st + alpha) ? alpha : l - P.st;
            for (uint32_t i = 0; i < l; ++i)
                if (i < P.st || i >= P.st + P.sp) P.compl_rows.push_back(row_q(i));
            for (uint32_t k = 0; k < K; ++k) P.compl_rows.push_back(row_p(k));
            P.sc = (uint32_t)P.compl_rows.size();
            std::vector<uint64_t> tab((size_t)P.sp * P.sc), tab60((size_t)P.sp * P.sc), tq(P.sc), tqn(P.sc);
            unsigned __int128 ssum = 0;
            for (uint32_t i = 0; i < P.sp; ++i) ssum += mods[P.st + i];
            // the R = 2^60 column reduction keeps the targets below 2.2 q when
            // the sources (reduced residues) sum to at most 1.2 2^60
            P.c60 = (5 * ssum <= 6 * ((unsigned __int128)1 << 60) && P.sp <= 6 && qmax < (1ull << 60)) ? 1u : 0u;
            for (uint32_t i = 0; i < P.sp; ++i) {
                const uint64_t qi = mods[P.st + i];
                const uint64_t qhat = prod_mod(mods, P.st, P.st + P.sp, qi, P.st + i);
                const uint64_t v = host_mulmod(host_invmod(qhat, qi), rt[P.st + i].ninv, qi);
                P.post.push_back(v);
                P.postp.push_back(host_precon(v, qi));
                for (uint32_t j = 0; j < P.sc; ++j) {
                    const uint64_t t = mods[P.compl_rows[j]];
                    const uint64_t qh = prod_mod(mods, P.st, P.st + P.sp, t, P.st + i);
                    tab[(size_t)i * P.sc + j] = host_mulmod(qh, host_r1(t), t);
                    tab60[(size_t)i * P.sc + j] = host_mulmod(qh, (1ull << 60) % t, t);
                }
            }
            for (uint32_t j = 0; j < P.sc; ++j) {
                tq[j] = mods[P.compl_rows[j]];
                tqn[j] = host_qneg(tq[j]);
            }
            std::vector<uint64_t> ttw(P.sc), tred(P.sc), ttwd(P.sc), tqf(2 * (size_t)P.sc);
            for (uint32_t j = 0; j < P.sc; ++j) {
                ttw[j] = (uint64_t)(uintptr_t)rt[P.compl_rows[j]].fw2;
                tred[j] = lazy_red_of(tq[j]);
                ttwd[j] = (uint64_t)(uintptr_t)rt[P.compl_rows[j]].fw2d;
                const double qd = (double)tq[j], qi = 1.0 / qd;
                std::memcpy(&tqf[2 * j], &qd, 8);
                std::memcpy(&tqf[2 * j + 1], &qi, 8);
            }
            put(tab, &P.d_tab);
            put(tab60, &P.d_tab60);
            put(tq, &P.d_tq);
            put(tqn, &P.d_tqneg);
            put(ttw, &P.d_ttw);
            put(tred, &P.d_tred);
            put(ttwd, &P.d_ttwd);
            put(tqf, &P.d_tqf);
        }
        // ModDown P -> Q_l
        std::vector<uint64_t> tab((size_t)K * l), tq(l), tqn(l);
        for (uint32_t k = 0; k < K; ++k) {
            const uint64_t pk = mods[L + k];
            const uint64_t v = host_mulmod(host_invmod(prod_mod(mods, L, L + K, pk, L + k), pk), rt[L + k].ninv, pk);
            c.pd_post.push_back(v);
            c.pd_postp.push_back(host_precon(v, pk));
            for (uint32_t i = 0; i < l; ++i)
                tab[(size_t)k * l + i] = host_mulmod(prod_mod(mods, L, L + K, mods[i], L + k), host_r1(mods[i]), mods[i]);
        }
        for (uint32_t i = 0; i < l; ++i) {
            tq[i] = mods[i];
            tqn[i] = host_qneg(mods[i]);
            const uint64_t pinv = host_invmod(prod_mod(mods, L, L + K, mods[i], (size_t)-1), mods[i]);
            c.pinv.push_back(pinv);
            c.pinvp.push_back(host_precon(pinv, mods[i]));
        }
        put(tab, &c.d_tabMD);
        put(tq, &c.d_tqMD);
        put(tqn, &c.d_tqnegMD);
        // rescale dropping tower l-1: s1 = q_{l-1}^-1 mod q_i, s2 = -s1 mod q_i
        // (s2 = floor(Q' [Q'^-1]_{q_l} / q_l) mod q_i is congruent to -s1, src/params.cpp:50-75)
        if (l >= 2) {
            const uint64_t ql = mods[l - 1];
            std::vector<uint64_t> tabRS((size_t)K * (l - 1)), tg, rsw(2 * (size_t)(l - 1));
            for (uint32_t i = 0; i + 1 < l; ++i) {
                const uint64_t qi = mods[i];
                const uint64_t s1 = host_invmod(ql % qi, qi);
                const uint64_t s2 = s1 ? qi - s1 : 0;
                c.s1.push_back(s1);
                c.s1p.push_back(host_precon(s1, qi));
                c.s2.push_back(s2);
                c.s2p.push_back(host_precon(s2, qi));
                // fused ModDown + rescale: P^-1 s1 (epilogue) and -P^-1 s1 folded into PHatModq
                const uint64_t ps1 = host_mulmod(c.pinv[i], s1, qi);
                c.ps1.push_back(ps1);
                c.ps1p.push_back(host_precon(ps1, qi));
                const uint64_t neg = ps1 ? qi - ps1 : 0;
                for (uint32_t k = 0; k < K; ++k)
                    tabRS[(size_t)k * (l - 1) + i] = host_mulmod(tab[(size_t)k * l + i], neg, qi);
                const SwitchJob sj = make_switch(nullptr, nullptr, ql, qi);
                const RsTarget t{qi, host_qneg(qi), sj.diff, sj.new_larger, s2, host_precon(s2, qi), sj.red};
                const uint64_t* tw = (const uint64_t*)&t;
                tg.insert(tg.end(), tw, tw + sizeof(RsTarget) / 8);
                rs_sw_consts(qi, ql, s2, rsw[2 * i], rsw[2 * i + 1]);
            }
            put(tabRS, &c.d_tabRS);
            put(tg, (const uint64_t**)&c.d_rstg);
            put(rsw, &c.d_rs_sw);
            // the surviving towers as the fused ModDown's targets
            std::vector<uint64_t> rq(l - 1), rqn(l - 1), rtw(l - 1), rred(l - 1), rtwd(l - 1), rqf(2 * (size_t)(l - 1));
            for (uint32_t i = 0; i + 1 < l; ++i) {
                rq[i] = mods[i];
                rqn[i] = host_qneg(mods[i]);
                rtw[i] = (uint64_t)(uintptr_t)rt[i].fw2;
                rred[i] = lazy_red_of(mods[i]);
                rtwd[i] = (uint64_t)(uintptr_t)rt[i].fw2d;
                const double qd = (double)mods[i], qi = 1.0 / qd;
                std::memcpy(&rqf[2 * i], &qd, 8);
                std::memcpy(&rqf[2 * i + 1], &qi, 8);
            }
            put(rq, &c.d_rs_tq);
            put(rqn, &c.d_rs_tqneg);
            put(rtw, &c.d_rs_ttw);
            put(rred, &c.d_rs_tred);
            put(rtwd, &c.d_rs_ttwd);
            put(rqf, &c.d_rs_tqf);
        }
    }
    put(std::vector<uint64_t>(L + K, 0), &d_zeros);
    consts.alloc(blob.size() * 8);
    hip_check(hipMemcpy(consts.p, blob.data(), blob.size() * 8, hipMemcpyHostToDevice), "upload constants");
    for (auto& f : fix) *f.dst = consts.as<uint64_t>() + f.off;
}

WsLayout WsLayout::make(const Ctx& C) {
    WsLayout w{};
    const size_t L = C.L, K = C.K, W = L + K;
    w.t0 = 0;
    w.t1 = L;
    w.t2 = 2 * L;
    w.x = 3 * L;                         // 2L: INTT'd part towers, later the ModDown conversions
    w.dig = 5 * L;
    size_t digsz = (size_t)C.dnum * W;
    if (digsz < 2 * (L - 1)) digsz = 2 * (L - 1);   // also holds the rescale's switched towers
    w.res0 = w.dig + digsz;
    w.res1 = w.res0 + W;
    w.xp = w.res1 + W;
    w.drop = w.xp + 2 * K;
    w.dt = w.drop + 2;     // INTT of t0/t1's last tower (fused rescale)
    w.dr = w.dt + 2;       // INTT of res0/res1's last tower
    w.total = w.dr + 2;
    return w;
}
size_t Ctx::ws_words() const { return WsLayout::make(*this).total * (size_t)n; }
size_t Ctx::rescale_ws_words() const { return (size_t)L * n; }

void Ctx::set_radix(uint32_t r) {
    if (r < 2 || r > 4) throw std::runtime_error("gpufhe: NTT radix log2 must be 2, 3 or 4");
    std::vector<RowTables> t(L + K);
    for (uint32_t i = 0; i < L + K; ++i) t[i] = row_tables(n, mods[i], roots[i], r);
    geom = make_geom(n, r);
    rtr = std::move(t);
}

const uint64_t* Ctx::row_table_of(const uint64_t* paired) const {
    if (!paired) return nullptr;
    for (uint32_t i = 0; i < L + K; ++i) {
        if (paired == rt[i].fw2) return rtr[i].fwd;
        if (paired == rt[i].iv2) return rtr[i].inv;
        if (paired == rt[i].fw2d) return rtr[i].fwd_d;
        if (paired == rt[i].iv2d) return rtr[i].inv_d;
    }
    return nullptr;
}
const uint64_t* Ctx::f64_table_of(const uint64_t* paired) const {
    for (uint32_t i = 0; i < L + K; ++i) {
        if (paired == rt[i].fw2) return rt[i].fw2d;
        if (paired == rt[i].iv2) return rt[i].iv2d;
    }
    return nullptr;
}

uint32_t Ctx::autotune(hipStream_t s, float* best_ms) {
    // 256 MiB of towers (enough workgroups to fill the chip several times
    // over, as the batched stages do) cycling through every modulus, forward
    // then inverse, in place; each radix is timed twice, interleaved, and
    // keeps its faster round
    const uint32_t towers = std::max<uint32_t>(96, (1u << 25) / n);
    DevBuf buf((size_t)towers * n * 8);
    // job arrays per radix (the row tables differ): [r - 2][fwd towers, inv towers]
    std::vector<NttJob> jobs;
    for (uint32_t r = 2; r <= 4; ++r)
        for (int inv = 0; inv < 2; ++inv)
            for (uint32_t t = 0; t < towers; ++t) {
                const uint32_t m = t % (L + K);
                const RowTables rw = row_tables(n, mods[m], roots[m], r);
                uint64_t* x = buf.as<uint64_t>() + (size_t)t * n;
                NttJob j = inv ? NttJob{x, x, rt[m].iv2, nullptr, mods[m], rt[m].ninv, rt[m].ninvp}
                               : NttJob{x, x, rt[m].fw2, nullptr, mods[m], 0, 0};
                j.twr = inv ? rw.inv : rw.fwd;
                if (!inv) j.red = lazy_red_of(mods[m]);   // as the plans' forward transforms
                jobs.push_back(j);
            }
    DevBuf dj(jobs.size() * sizeof(NttJob));
    hip_check(hipMemcpy(dj.p, jobs.data(), dj.bytes, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemsetAsync(buf.p, 0, buf.bytes, s), "clear");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    uint32_t best = geom.rmax;
    float bms = 1e30f;
    float per[5] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    for (int round = 0; round < 2; ++round)
        for (uint32_t r = 2; r <= 4; ++r) {
            NttGeom g = make_geom(n, r);
            const NttJob* fw = dj.as<NttJob>() + (size_t)(r - 2) * 2 * towers;
            const NttJob* iv = fw + towers;
            launch_ntt(fw, towers, g, true, s);   // warm
            launch_ntt(iv, towers, g, false, s);
            hip_check(hipEventRecord(e0, s), "event");
            for (int rep = 0; rep < 3; ++rep) {
                launch_ntt(fw, towers, g, true, s);
                launch_ntt(iv, towers, g, false, s);
            }
            hip_check(hipEventRecord(e1, s), "event");
            hip_check(hipEventSynchronize(e1), "event sync");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            if (ms < per[r]) per[r] = ms;
            if (ms < bms) { bms = ms; best = r; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_radix(best);
    {
        std::ostringstream o;
        o << "{\"workload_towers\": " << towers << ", \"ms_per_fwd_inv_pass\": {";
        for (uint32_t r = 2; r <= 4; ++r) o << (r > 2 ? ", " : "") << "\"" << r << "\": " << per[r] / 3;
        o << "}, \"chosen\": " << best << "}";
        radix_report = o.str();
    }
    // reported per 96 towers, forward and inverse
    if (best_ms) *best_ms = bms / 3 * 96.0f / (float)towers;
    return best;
}

void Ctx::set_knob(const std::string& name, uint32_t v) {
    if (name == "ntt_radix") {
        set_radix(v);
    } else if (name == "ksi_extra") {
        if (v > 1) throw std::runtime_error("gpufhe: ksi_extra must be 0 or 1");
        ksi_extra = v;
    } else if (name == "conv_chunk" || name == "rs_chunk") {
        if (v < 1 || v > 64) throw std::runtime_error("gpufhe: chunk knobs must be in [1, 64]");
        (name == "conv_chunk" ? conv_chunk : rs_chunk) = v;
    } else if (name == "modup_fused") {
        if (v > 1) throw std::runtime_error("gpufhe: modup_fused must be 0 or 1");
        modup_fused = v;
    } else if (name == "modup_f64" || name == "ksi_f64" || name == "moddown_fused" || name == "tensor_ib" ||
               name == "ksi_rowpass") {
        if (v > 1) throw std::runtime_error("gpufhe: " + name + " must be 0 or 1");
        (name == "modup_f64"     ? modup_f64
         : name == "ksi_f64"     ? ksi_f64
         : name == "tensor_ib"   ? tensor_ib
         : name == "ksi_rowpass" ? ksi_rowpass
                                 : moddown_fused) = v;
    } else if (name == "modup_min_wgs") {
        modup_min_wgs = v;
    } else if (name == "rechunk_wgs") {
        rechunk_wgs = v;
    } else {
        throw std::runtime_error("gpufhe: unknown tuning knob " + name);
    }
}
uint32_t Ctx::knob(const std::string& name) const {
    if (name == "ntt_radix") return geom.rmax;
    if (name == "ksi_extra") return ksi_extra;
    if (name == "conv_chunk") return conv_chunk;
    if (name == "rs_chunk") return rs_chunk;
    if (name == "modup_fused") return modup_fused;
    if (name == "modup_f64") return modup_f64;
    if (name == "ksi_f64") return ksi_f64;
    if (name == "tensor_ib") return tensor_ib;
    if (name == "ksi_rowpass") return ksi_rowpass;
    if (name == "moddown_fused") return moddown_fused;
    if (name == "modup_min_wgs") return modup_min_wgs;
    if (name == "rechunk_wgs") return rechunk_wgs;
    throw std::runtime_error("gpufhe: unknown tuning knob " + name);
}

// Coordinate search over the plan knobs on one batch of B EvalMult+Rescale at
// the top level (uniform key and inputs): for each knob in turn, every
// candidate builds the whole plan, runs it once, and times the steps the knob
// shapes (the fused row pass + inner product steps, the conversion steps, the
// rescale prologue) over 5 runs, twice, interleaved; the knob keeps its
// fastest candidate before the next knob is searched.  Stage timing keeps
// the other stages' noise out of the comparison; B should be the batch the
// plans will run (a chunk width that wins at one batch can lose at another).
void Ctx::autotune_plan(hipStream_t s, uint32_t B) {
    const uint32_t l = L;
    if (B < 1 || B > 256) throw std::runtime_error("gpufhe: tuning batch must be in [1, 256]");
    std::ostringstream rep;
    rep << "{\"ntt_radix\": " << (radix_report.empty() ? "null" : radix_report);
    if (l < 2) {
        rep << "}";
        tune_report = rep.str();
        return;
    }
    DevBuf dm(mods.size() * 8);
    hip_check(hipMemcpy(dm.p, mods.data(), dm.bytes, hipMemcpyHostToDevice), "upload moduli");
    DevKey key;
    key.nparts = dnum;
    key.rows = L + K;
    key.fullQ = L;
    const size_t per = (size_t)key.rows * n;
    key.a.alloc((size_t)dnum * per * 8);
    key.b.alloc((size_t)dnum * per * 8);
    for (uint32_t p = 0; p < dnum; ++p) {
        launch_fill(key.a.as<uint64_t>() + p * per, dm.as<uint64_t>(), key.rows, n, 9001 + 2 * p, s);
        launch_fill(key.b.as<uint64_t>() + p * per, dm.as<uint64_t>(), key.rows, n, 9002 + 2 * p, s);
    }
    const size_t T = (size_t)l * n, To = (size_t)(l - 1) * n, W = ws_words();
    DevBuf in((size_t)B * 4 * T * 8), out((size_t)B * 2 * To * 8), ws((size_t)B * W * 8);
    for (uint32_t i = 0; i < 4 * B; ++i) launch_fill(in.as<uint64_t>() + i * T, dm.as<uint64_t>(), l, n, 77 + i, s);
    hip_check(hipStreamSynchronize(s), "tuning inputs");
    auto build = [&]() {
        PlanBuilder b(*this);
        for (uint32_t i = 0; i < B; ++i) {
            const uint64_t* x = in.as<uint64_t>() + (size_t)4 * i * T;
            uint64_t* o = out.as<uint64_t>() + (size_t)2 * i * To;
            b.evalmult_rescale(l, x, x + T, x + 2 * T, x + 3 * T, o, o + To, key, ws.as<uint64_t>() + i * W);
        }
        return b.finish();
    };
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    // the steps of the given kinds, timed together
    auto time_steps = [&](const Plan& P, std::initializer_list<Step::Kind> kinds) {
        P.run(s);
        hip_check(hipEventRecord(e0, s), "event");
        for (int r = 0; r < 5; ++r)
            for (size_t i = 0; i < P.steps.size(); ++i)
                for (Step::Kind k : kinds)
                    if (P.steps[i].kind == k) P.run_step(i, s);
        hip_check(hipEventRecord(e1, s), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        return ms / 5;
    };
    auto tune = [&](const char* name, uint32_t& knob, const std::vector<uint32_t>& cands,
                    std::initializer_list<Step::Kind> kinds) {
        std::vector<float> best(cands.size(), 1e30f);
        for (int round = 0; round < 2; ++round)
            for (size_t c = 0; c < cands.size(); ++c) {
                knob = cands[c];
                auto P = build();
                best[c] = std::min(best[c], time_steps(*P, kinds));
            }
        size_t bi = 0;
        for (size_t c = 1; c < cands.size(); ++c)
            if (best[c] < best[bi]) bi = c;
        knob = cands[bi];
        rep << ", \"" << name << "\": {\"stage_ms\": {";
        for (size_t c = 0; c < cands.size(); ++c) rep << (c ? ", " : "") << "\"" << cands[c] << "\": " << best[c];
        rep << "}, \"chosen\": " << knob << "}";
    };
    rep << ", \"batch\": " << B << ", \"level\": " << l;
    if (fused && !geom.single && n % KSI_TILE == 0) tune("ksi_extra", ksi_extra, {0, 1}, {Step::KSI});
    // the conversion chunk width matters where the ModUp runs unfused: tuned
    // with it unfused, then the fused ModUp is weighed against that
    const uint32_t mf = modup_fused;
    modup_fused = 0;
    tune("conv_chunk", conv_chunk, {8, 16, 32}, {Step::CONV});
    modup_fused = mf;
    if (fused && !geom.single && n % KSI_TILE == 0 && n % MODUP_TILE == 0 && !wide && alpha <= MODUP_MAXSRC) {
        // fp64 passes (the fused ModUp's column passes and the inverse row
        // pass before it; then the fused inner product's row passes), timed
        // with the ModUp fused
        modup_fused = 1;
        tune("modup_f64", modup_f64, {0, 1}, {Step::NTT_IB, Step::MODUP});
        if (modup_f64) tune("ksi_f64", ksi_f64, {0, 1}, {Step::KSI});
        modup_fused = mf;
        tune("modup_fused", modup_fused, {0, 1}, {Step::NTT_IB, Step::NTT_I, Step::MODUP, Step::CONV, Step::NTT_FA});
        // the tensor product fused with the row pass in front of the fused ModUp
        if (modup_fused) tune("tensor_ib", tensor_ib, {0, 1}, {Step::TENSOR, Step::NTT_IB, Step::TENSOR_IB});
    }
    if (fused) {
        // the prologue's chunk width matters where the ModDown runs apart:
        // tuned that way, then the fused ModDown is weighed against it
        const uint32_t md = moddown_fused;
        moddown_fused = 0;
        tune("rs_chunk", rs_chunk, {8, 16, 32}, {Step::RSV});
        moddown_fused = md;
        if (modup_fused && !geom.single && n % KSI_TILE == 0 && n % MODUP_TILE == 0 && !wide &&
            alpha <= MODUP_MAXSRC && K <= MODUP_MAXSRC)
            tune("moddown_fused", moddown_fused, {0, 1},
                 {Step::NTT_IB, Step::NTT_I, Step::MODDOWN, Step::RSV, Step::NTT_F, Step::NTT_FB});
        // the P part's inverse row pass inside the P rows' inner product
        // (listed first: on a tie the tuner keeps the first candidate, and
        // this one moves fewer bytes in one launch fewer)
        if (moddown_fused && !geom.single && n % KSI_TILE == 0)
            tune("ksi_rowpass", ksi_rowpass, {1, 0}, {Step::KSI, Step::NTT_IB});
    }
    rep << "}";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_report = rep.str();
}

// ------------------------------------------------------------------ keys --
std::unique_ptr<DevKey> upload_key(const Ctx& C, const uint64_t* av, const uint64_t* bv,
                                   uint32_t nparts, uint32_t rows, uint32_t fullQ, hipStream_t s) {
    if (fullQ < C.L || rows != fullQ + C.K) throw std::runtime_error("gpufhe: key row layout does not match the context");
    if (nparts < C.dnum) throw std::runtime_error("gpufhe: key has fewer parts than the context needs");
    auto k = std::make_unique<DevKey>();
    k->nparts = nparts;
    k->rows = rows;
    k->fullQ = fullQ;
    const size_t words = (size_t)nparts * rows * C.n;
    k->a.alloc(words * 8);
    k->b.alloc(words * 8);
    hip_check(hipMemcpyAsync(k->a.p, av, words * 8, hipMemcpyHostToDevice, s), "upload key");
    hip_check(hipMemcpyAsync(k->b.p, bv, words * 8, hipMemcpyHostToDevice, s), "upload key");
    // to Montgomery form (x R mod q) on the device, rows this context can address
    std::vector<FuseJob> jobs;
    for (uint32_t p = 0; p < nparts; ++p)
        for (uint32_t r = 0; r < rows; ++r) {
            uint32_t row;
            if (r < C.L) row = r;
            else if (r >= fullQ) row = C.L + (r - fullQ);
            else continue;
            const uint64_t q = C.mods[row], R = C.rc[row].r1;
            for (int ab = 0; ab < 2; ++ab) {
                uint64_t* t = (ab ? k->b.as<uint64_t>() : k->a.as<uint64_t>()) + ((size_t)p * rows + r) * C.n;
                jobs.push_back(FuseJob{t, t, t, q, R, host_precon(R, q), 0, 0});
            }
        }
    DevBuf dj(jobs.size() * sizeof(FuseJob));
    hip_check(hipMemcpyAsync(dj.p, jobs.data(), dj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    launch_fuse(dj.as<FuseJob>(), (uint32_t)jobs.size(), C.n, s);
    hip_check(hipStreamSynchronize(s), "key conversion");
    return k;
}

// Switching key generated on the device (src/evalkeygen.cpp:42-96): the
// small polynomials u, e0, e1 of every part are drawn on the host in the
// reference's order, embedded into every Q and P tower, transformed and
// combined with the public key and sOld on the device, and stored in
// Montgomery form.  sold, pka, pkb: device arrays [L+K][n], evaluation form.
// The counter sampler's Gaussian table for sigma in device memory
static DevBuf gauss_table_dev(double sigma, uint32_t& K, hipStream_t s) {
    const GaussSampler gs(sigma);
    K = (uint32_t)gs.K();
    DevBuf T(gs.table().size() * 8);
    hip_check(hipMemcpyAsync(T.p, gs.table().data(), T.bytes, hipMemcpyHostToDevice, s), "upload Gaussian table");
    hip_check(hipStreamSynchronize(s), "upload Gaussian table");
    return T;
}
static uint64_t floor_2_64_over(uint64_t q) {
    if (q <= (1ull << 32)) throw std::runtime_error("gpufhe: the counter sampler needs moduli above 2^32");
    return (uint64_t)(((unsigned __int128)1 << 64) / q);
}

// Public key generation on the device (src/keygen.cpp:55-112) with the
// counter sampler: s ternary (domain CTR_SECRET), e Gaussian (CTR_ERROR), a of
// tower t uniform (CTR_UNIFORM + t, taken as evaluation form), over all L + K
// towers; s and e embedded and transformed, b = ns e - a s.
void keygen_device(const Ctx& C, uint64_t seed, double sigma, uint64_t ns, uint64_t* s_out, uint64_t* a_out,
                   uint64_t* b_out, hipStream_t st) {
    const uint32_t n = C.n, W = C.L + C.K;
    uint32_t K = 0;
    DevBuf T = gauss_table_dev(sigma, K, st);
    DevBuf coef((size_t)2 * n * 8), et((size_t)W * n * 8);
    int64_t* sc = coef.as<int64_t>();
    int64_t* ec = sc + n;
    std::vector<SampleJob> sj{SampleJob{sc, nullptr, 0, 0, CTR_SECRET, 0}, SampleJob{ec, nullptr, 0, 0, CTR_ERROR, 1}};
    std::vector<EmbedJob> ej;
    std::vector<NttJob> nj;
    std::vector<PkJob> pj;
    for (uint32_t t = 0; t < W; ++t) {
        const size_t o = (size_t)t * n;
        const uint64_t q = C.mods[t];
        sj.push_back(SampleJob{nullptr, a_out + o, q, floor_2_64_over(q), CTR_UNIFORM + t, 2});
        ej.push_back(EmbedJob{sc, s_out + o, q});
        ej.push_back(EmbedJob{ec, et.as<uint64_t>() + o, q});
        for (uint64_t* x : {s_out + o, et.as<uint64_t>() + o}) {
            NttJob J{x, x, C.rt[t].fw2, nullptr, q, 0, 0};
            J.twr = C.rtr[t].fwd;
            nj.push_back(J);
        }
        const uint64_t nsq = ns % q;
        pj.push_back(PkJob{a_out + o, s_out + o, et.as<uint64_t>() + o, b_out + o, q, C.rc[t].qneg, C.rc[t].r2, nsq,
                           host_precon(nsq, q)});
    }
    DevBuf dsj(sj.size() * sizeof(SampleJob)), dej(ej.size() * sizeof(EmbedJob)), dnj(nj.size() * sizeof(NttJob)),
        dpj(pj.size() * sizeof(PkJob));
    hip_check(hipMemcpyAsync(dsj.p, sj.data(), dsj.bytes, hipMemcpyHostToDevice, st), "upload jobs");
    hip_check(hipMemcpyAsync(dej.p, ej.data(), dej.bytes, hipMemcpyHostToDevice, st), "upload jobs");
    hip_check(hipMemcpyAsync(dnj.p, nj.data(), dnj.bytes, hipMemcpyHostToDevice, st), "upload jobs");
    hip_check(hipMemcpyAsync(dpj.p, pj.data(), dpj.bytes, hipMemcpyHostToDevice, st), "upload jobs");
    launch_sample(dsj.as<SampleJob>(), (uint32_t)sj.size(), n, seed, T.as<uint64_t>(), K, st);
    launch_embed(dej.as<EmbedJob>(), (uint32_t)ej.size(), n, st);
    launch_ntt(dnj.as<NttJob>(), (uint32_t)nj.size(), C.geom, true, st);
    launch_pkgen(dpj.as<PkJob>(), (uint32_t)pj.size(), n, st);
    hip_check(hipGetLastError(), "key generation launch");
    hip_check(hipStreamSynchronize(st), "key generation");
}

std::unique_ptr<DevKey> generate_key(const Ctx& C, const uint64_t* sold, const uint64_t* pka, const uint64_t* pkb,
                                     uint64_t ns, double sigma, uint64_t seed, hipStream_t s, bool ctr) {
    const uint32_t n = C.n, W = C.L + C.K, parts = C.dnum;
    auto k = std::make_unique<DevKey>();
    k->nparts = parts;
    k->rows = W;
    k->fullQ = C.L;
    const size_t per = (size_t)W * n;
    k->a.alloc((size_t)parts * per * 8);
    k->b.alloc((size_t)parts * per * 8);
    // small polynomials (3 per part), then their towers [part][3][W][n]
    DevBuf coef((size_t)parts * 3 * n * 8), tow((size_t)parts * 3 * per * 8);
    DevBuf T, dsj;
    if (ctr) {
        // drawn on the device: part p's u (ternary), e0, e1 (Gaussian) in
        // domains CTR_KEY + 3p + {0, 1, 2}
        uint32_t K = 0;
        T = gauss_table_dev(sigma, K, s);
        std::vector<SampleJob> sj;
        for (uint32_t p = 0; p < parts; ++p)
            for (uint32_t v = 0; v < 3; ++v)
                sj.push_back(SampleJob{coef.as<int64_t>() + (size_t)(3 * p + v) * n, nullptr, 0, 0, CTR_KEY + 3 * p + v,
                                       v == 0 ? 0u : 1u});
        dsj.alloc(sj.size() * sizeof(SampleJob));
        hip_check(hipMemcpyAsync(dsj.p, sj.data(), dsj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
        launch_sample(dsj.as<SampleJob>(), (uint32_t)sj.size(), n, seed, T.as<uint64_t>(), K, s);
    } else {
        // drawn on the host in the reference's order (sampler.hpp)
        std::vector<int64_t> u, e0, e1;
        sample_switch_key(n, parts, sigma, seed, u, e0, e1);
        std::vector<int64_t> hc((size_t)parts * 3 * n);
        for (uint32_t p = 0; p < parts; ++p) {
            std::copy(u.begin() + (size_t)p * n, u.begin() + (size_t)(p + 1) * n, hc.begin() + (size_t)(3 * p) * n);
            std::copy(e0.begin() + (size_t)p * n, e0.begin() + (size_t)(p + 1) * n, hc.begin() + (size_t)(3 * p + 1) * n);
            std::copy(e1.begin() + (size_t)p * n, e1.begin() + (size_t)(p + 1) * n, hc.begin() + (size_t)(3 * p + 2) * n);
        }
        hip_check(hipMemcpyAsync(coef.p, hc.data(), coef.bytes, hipMemcpyHostToDevice, s), "upload samples");
        hip_check(hipStreamSynchronize(s), "upload samples");   // hc is pageable and goes out of scope
    }
    std::vector<EmbedJob> ej;
    std::vector<NttJob> nj;
    std::vector<EkJob> kj;
    const int64_t* cb = coef.as<int64_t>();
    uint64_t* tb = tow.as<uint64_t>();
    for (uint32_t p = 0; p < parts; ++p) {
        const uint32_t st = C.alpha * p, en = std::min(C.L, st + C.alpha);
        for (uint32_t v = 0; v < 3; ++v)
            for (uint32_t t = 0; t < W; ++t) {
                uint64_t* x = tb + ((size_t)(3 * p + v) * W + t) * n;
                ej.push_back(EmbedJob{cb + (size_t)(3 * p + v) * n, x, C.mods[t]});
                NttJob J{x, x, C.rt[t].fw2, nullptr, C.mods[t], 0, 0};
                J.twr = C.rtr[t].fwd;
                nj.push_back(J);
            }
        for (uint32_t t = 0; t < W; ++t) {
            const uint64_t q = C.mods[t];
            uint64_t pq = 1 % q;
            for (uint32_t j = 0; j < C.K; ++j) pq = host_mulmod(pq, C.mods[C.L + j] % q, q);
            const uint64_t nsq = ns % q;
            const size_t o = (size_t)t * n;
            EkJob J{};
            J.pka = pka + o;
            J.pkb = pkb + o;
            J.u = tb + ((size_t)(3 * p) * W + t) * n;
            J.e0 = tb + ((size_t)(3 * p + 1) * W + t) * n;
            J.e1 = tb + ((size_t)(3 * p + 2) * W + t) * n;
            J.sold = (t >= st && t < en) ? sold + o : nullptr;
            J.a_out = k->a.as<uint64_t>() + (size_t)p * per + o;
            J.b_out = k->b.as<uint64_t>() + (size_t)p * per + o;
            J.q = q;
            J.qneg = C.rc[t].qneg;
            J.r2 = C.rc[t].r2;
            J.ns = nsq;
            J.nsp = host_precon(nsq, q);
            J.pq = pq;
            J.pqp = host_precon(pq, q);
            kj.push_back(J);
        }
    }
    DevBuf dej(ej.size() * sizeof(EmbedJob)), dnj(nj.size() * sizeof(NttJob)), dkj(kj.size() * sizeof(EkJob));
    hip_check(hipMemcpyAsync(dej.p, ej.data(), dej.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    hip_check(hipMemcpyAsync(dnj.p, nj.data(), dnj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    hip_check(hipMemcpyAsync(dkj.p, kj.data(), dkj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    launch_embed(dej.as<EmbedJob>(), (uint32_t)ej.size(), n, s);
    launch_ntt(dnj.as<NttJob>(), (uint32_t)nj.size(), C.geom, true, s);
    launch_ekgen(dkj.as<EkJob>(), (uint32_t)kj.size(), n, s);
    hip_check(hipGetLastError(), "key generation launch");
    hip_check(hipStreamSynchronize(s), "key generation");
    return k;
}

// Encryption on the device (src/encdec.cpp:41-64): v (ternary), e0, e1
// (Gaussian) drawn on the host in the reference's order; the message and the
// three small polynomials embedded into the `level` Q towers and transformed on
// the device; then c0 = pkB v + ns e0 + m, c1 = pkA v + ns e1 by the
// switching-key kernel in plain form (sold = m, P mod q replaced by 1).
void encrypt_device(const Ctx& C, uint32_t level, const uint64_t* pka, const uint64_t* pkb, const int64_t* h_m,
                    uint64_t ns, double sigma, uint64_t seed, uint64_t* c0, uint64_t* c1, hipStream_t s) {
    if (level < 1 || level > C.L) throw std::runtime_error("gpufhe: bad level");
    const uint32_t n = C.n;
    std::vector<int64_t> v, e0, e1;
    sample_encrypt(n, sigma, seed, v, e0, e1);
    // coefficients [m, v, e0, e1][n]; their towers [4][level][n]
    std::vector<int64_t> hc((size_t)4 * n);
    std::copy(h_m, h_m + n, hc.begin());
    std::copy(v.begin(), v.end(), hc.begin() + n);
    std::copy(e0.begin(), e0.end(), hc.begin() + 2 * (size_t)n);
    std::copy(e1.begin(), e1.end(), hc.begin() + 3 * (size_t)n);
    DevBuf coef(hc.size() * 8), tow((size_t)4 * level * n * 8);
    hip_check(hipMemcpyAsync(coef.p, hc.data(), coef.bytes, hipMemcpyHostToDevice, s), "upload samples");
    const int64_t* cb = coef.as<int64_t>();
    uint64_t* tb = tow.as<uint64_t>();
    auto tower = [&](uint32_t k, uint32_t t) { return tb + ((size_t)k * level + t) * n; };
    std::vector<EmbedJob> ej;
    std::vector<NttJob> nj;
    std::vector<EkJob> kj;
    for (uint32_t k = 0; k < 4; ++k)
        for (uint32_t t = 0; t < level; ++t) {
            ej.push_back(EmbedJob{cb + (size_t)k * n, tower(k, t), C.mods[t]});
            NttJob J{tower(k, t), tower(k, t), C.rt[t].fw2, nullptr, C.mods[t], 0, 0};
            J.twr = C.rtr[t].fwd;
            J.red = lazy_red_of(C.mods[t]);
            nj.push_back(J);
        }
    for (uint32_t t = 0; t < level; ++t) {
        const uint64_t q = C.mods[t], nsq = ns % q;
        const size_t o = (size_t)t * n;
        EkJob J{};
        J.pka = pka + o;
        J.pkb = pkb + o;
        J.u = tower(1, t);
        J.e0 = tower(2, t);
        J.e1 = tower(3, t);
        J.sold = tower(0, t);
        J.a_out = c1 + o;
        J.b_out = c0 + o;
        J.q = q;
        J.qneg = C.rc[t].qneg;
        J.r2 = C.rc[t].r2;
        J.ns = nsq;
        J.nsp = host_precon(nsq, q);
        J.pq = 1 % q;
        J.pqp = host_precon(J.pq, q);
        J.plain = 1;
        kj.push_back(J);
    }
    DevBuf dej(ej.size() * sizeof(EmbedJob)), dnj(nj.size() * sizeof(NttJob)), dkj(kj.size() * sizeof(EkJob));
    hip_check(hipMemcpyAsync(dej.p, ej.data(), dej.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    hip_check(hipMemcpyAsync(dnj.p, nj.data(), dnj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    hip_check(hipMemcpyAsync(dkj.p, kj.data(), dkj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    launch_embed(dej.as<EmbedJob>(), (uint32_t)ej.size(), n, s);
    launch_ntt(dnj.as<NttJob>(), (uint32_t)nj.size(), C.geom, true, s);
    launch_ekgen(dkj.as<EkJob>(), (uint32_t)kj.size(), n, s);
    hip_check(hipGetLastError(), "encryption launch");
    hip_check(hipStreamSynchronize(s), "encryption");   // before the staging buffers are freed
}

// Decryption (src/encdec.cpp:68-81): tower 0 only, INTT(c0 + c1 s) with n^-1,
// centred on the host
void decrypt_device(const Ctx& C, const uint64_t* c0, const uint64_t* c1, const uint64_t* sk, int64_t* h_m,
                    hipStream_t s) {
    const uint32_t n = C.n;
    DevBuf ph((size_t)n * 8), dj(sizeof(DecJob)), dn(sizeof(NttJob));
    const DecJob J{c0, c1, sk, ph.as<uint64_t>(), C.mods[0], C.rc[0].qneg, C.rc[0].r2};
    NttJob T{ph.as<uint64_t>(), ph.as<uint64_t>(), C.rt[0].iv2, nullptr, C.mods[0], C.rt[0].ninv, C.rt[0].ninvp};
    T.twr = C.rtr[0].inv;
    hip_check(hipMemcpyAsync(dj.p, &J, sizeof J, hipMemcpyHostToDevice, s), "upload jobs");
    hip_check(hipMemcpyAsync(dn.p, &T, sizeof T, hipMemcpyHostToDevice, s), "upload jobs");
    launch_dec(dj.as<DecJob>(), 1, n, s);
    launch_ntt(dn.as<NttJob>(), 1, C.geom, false, s);
    std::vector<uint64_t> h(n);
    hip_check(hipMemcpyAsync(h.data(), ph.p, (size_t)n * 8, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "decryption");
    const uint64_t q = C.mods[0];
    for (uint32_t k = 0; k < n; ++k) h_m[k] = h[k] > q / 2 ? (int64_t)h[k] - (int64_t)q : (int64_t)h[k];
}

// the key's rows back in plain form (x R^-1 mod q), [nparts][rows][n] on the host
void download_key(const Ctx& C, const DevKey& k, uint64_t* av, uint64_t* bv, hipStream_t s) {
    const uint32_t n = C.n;
    const size_t words = (size_t)k.nparts * k.rows * n;
    DevBuf ta(words * 8), tbuf(words * 8);
    std::vector<FuseJob> jobs;
    for (uint32_t p = 0; p < k.nparts; ++p)
        for (uint32_t r = 0; r < k.rows; ++r) {
            uint32_t row;
            if (r < C.L) row = r;
            else if (r >= k.fullQ) row = C.L + (r - k.fullQ);
            else continue;   // rows this context never reads stay as uploaded
            const uint64_t q = C.mods[row], rinv = host_invmod(C.rc[row].r1, q);
            const size_t o = ((size_t)p * k.rows + r) * n;
            jobs.push_back(FuseJob{k.a.as<uint64_t>() + o, k.a.as<uint64_t>() + o, ta.as<uint64_t>() + o, q, rinv,
                                   host_precon(rinv, q), 0, 0});
            jobs.push_back(FuseJob{k.b.as<uint64_t>() + o, k.b.as<uint64_t>() + o, tbuf.as<uint64_t>() + o, q, rinv,
                                   host_precon(rinv, q), 0, 0});
        }
    hip_check(hipMemcpyAsync(ta.p, k.a.p, ta.bytes, hipMemcpyDeviceToDevice, s), "copy key");
    hip_check(hipMemcpyAsync(tbuf.p, k.b.p, tbuf.bytes, hipMemcpyDeviceToDevice, s), "copy key");
    DevBuf dj(jobs.size() * sizeof(FuseJob));
    hip_check(hipMemcpyAsync(dj.p, jobs.data(), dj.bytes, hipMemcpyHostToDevice, s), "upload jobs");
    launch_fuse(dj.as<FuseJob>(), (uint32_t)jobs.size(), n, s);
    hip_check(hipMemcpyAsync(av, ta.p, ta.bytes, hipMemcpyDeviceToHost, s), "download key");
    hip_check(hipMemcpyAsync(bv, tbuf.p, tbuf.bytes, hipMemcpyDeviceToHost, s), "download key");
    hip_check(hipStreamSynchronize(s), "download key");
}

// ------------------------------------------------------------------ Plan --
Plan::~Plan() {
    if (graph) (void)hipGraphExecDestroy(graph);
}

void Plan::run(hipStream_t s) const {
    for (size_t i = 0; i < steps.size(); ++i) run_step(i, s);
}

uint32_t Plan::step_kernels(size_t i) const {
    const Step::Kind k = steps[i].kind;
    return (k == Step::NTT_F || k == Step::NTT_I) && !geom.single ? 2u : 1u;   // NTT_IB, NTT_FA: one pass
}

void Plan::run_step(size_t i, hipStream_t s, uint32_t which) const { run_jobs(i, s, which, 0, steps[i].count); }

void Plan::run_jobs(size_t i, hipStream_t s, uint32_t which, uint32_t first, uint32_t cnt) const {
    const char* base = (const char*)(this->base ? this->base : dev.p);
    const Step& st = steps[i];
    const char* p = base + st.off;
    if (st.kind != Step::NTT_F && st.kind != Step::NTT_I && !(which & 1)) return;
#define GF_J(T) ((const T*)p + first)
    switch (st.kind) {
        case Step::NTT_F: launch_ntt(GF_J(NttJob), cnt, geom, true, s, which); break;
        case Step::NTT_I: launch_ntt(GF_J(NttJob), cnt, geom, false, s, which); break;
        case Step::TENSOR: launch_tensor(GF_J(TensorJob), cnt, n, s); break;
        case Step::TENSOR_IB: launch_tensor_ib(GF_J(TensorIbJob), cnt, geom, s); break;
        case Step::CONV: launch_conv(GF_J(ConvJob), cnt, n, s); break;
        case Step::INNER: launch_inner(GF_J(InnerJob), cnt, n, s); break;
        case Step::COMBINE: launch_combine(GF_J(CombineJob), cnt, n, s); break;
        case Step::SWITCH: launch_switch(GF_J(SwitchJob), cnt, n, s); break;
        case Step::FUSE: launch_fuse(GF_J(FuseJob), cnt, n, s); break;
        case Step::ADD: launch_add(GF_J(AddJob), cnt, n, s); break;
        case Step::PERM: launch_perm(GF_J(PermJob), cnt, n, logn, s); break;
        case Step::COPY: launch_copy(GF_J(CopyJob), cnt, n, s); break;
        case Step::RSV: launch_rs(GF_J(RsJob), cnt, n, s); break;
        case Step::PTACC: launch_ptacc(GF_J(PtAccJob), cnt, n, s); break;
        case Step::NTT_FA: launch_ntt(GF_J(NttJob), cnt, geom, true, s, 1); break;
        case Step::KSI: launch_ksi(GF_J(KsiJob), cnt, geom, st.arg & 0xFF, st.arg >> 8, s); break;
        case Step::NTT_IB: launch_ntt(GF_J(NttJob), cnt, geom, false, s, 1); break;
        case Step::MODUP: launch_modup(GF_J(ModupJob), cnt, geom, st.arg, s); break;
        case Step::MODDOWN: launch_moddown(GF_J(ModupJob), cnt, geom, st.arg, s); break;
        case Step::NTT_FB: launch_ntt(GF_J(NttJob), cnt, geom, true, s, 2); break;
    }
#undef GF_J
}

void Plan::run_per_job(hipStream_t s) const {
    for (size_t i = 0; i < steps.size(); ++i)
        for (uint32_t j = 0; j < steps[i].count; ++j) run_jobs(i, s, 3, j, 1);
}

void Plan::run_graph(hipStream_t s) {
    if (!graph) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        run(s);
        hip_check(hipStreamEndCapture(s, &g), "end capture");
        hip_check(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
    }
    hip_check(hipGraphLaunch(graph, s), "graph launch");
}

// --------------------------------------------------------------- builder --
// Targets per conversion job: a job loads and splits its source column once
// for all of its targets, so large chunks save work; a stage with few jobs
// is split back into small chunks at layout (split_chunks) to fill the chip.
static const uint32_t CONV_CHUNK_SMALL = 8;

// Re-chunk conversion jobs ([j0, j1) target ranges) into CONV_CHUNK_SMALL
// pieces when the stage would launch fewer than min_wgs workgroups
// (Ctx::rechunk_wgs, by default 8192: 32 per CU).
template <class J>
static void split_chunks(std::vector<J>& v, uint32_t n, uint32_t min_wgs) {
    const size_t blocks = (n + PW_THREADS * PW_EPT - 1) / (PW_THREADS * PW_EPT);
    if (v.empty() || v.size() * blocks >= min_wgs) return;
    std::vector<J> out;
    for (const J& j : v)
        for (uint32_t a = j.j0; a < j.j1; a += CONV_CHUNK_SMALL) {
            J k = j;
            k.j0 = a;
            k.j1 = std::min(j.j1, a + CONV_CHUNK_SMALL);
            out.push_back(k);
        }
    v.swap(out);
}

// Hybrid keyswitch up to the ModDown sources (src/keyswitch_resident.cpp:139-201):
// ModUp of every digit, the key inner product over Q_l P, and the INTT of both
// accumulators' P towers scaled by n^-1 PHatInv.  Optional extra inverse
// transforms (plain n^-1) ride in the first INTT launch.
void PlanBuilder::keyswitch_core(uint32_t l, const uint64_t* d, const DevKey& key, uint64_t* ws,
                                 const uint64_t* extra0, const uint64_t* extra1, uint64_t* dx0,
                                 uint64_t* dx1, uint32_t lazy_rows, bool p_rowpass) {
    const Ctx& C = C_;
    if (l < 1 || l > C.L) throw std::runtime_error("gpufhe: bad level");
    if (key.nparts < C.lv[l].np) throw std::runtime_error("gpufhe: key has too few parts for this level");
    if (use_ksi(l)) {
        modup(l, d, ws, extra0, extra1, dx0, dx1, true);
        ksi_moddown_src(l, d, key, ws, lazy_rows, p_rowpass);
        return;
    }
    if (p_rowpass) throw std::runtime_error("gpufhe: the fused ModDown needs the fused inner product");
    modup(l, d, ws, extra0, extra1, dx0, dx1);
    const uint64_t* dig = ws + WsLayout::make(C).dig * C.n;
    inner_moddown_src(l, key, ws, [&](uint32_t pi, uint32_t i) { return digit_tower(l, d, dig, pi, i); });
}

bool PlanBuilder::use_ksi(uint32_t l) const {
    return C_.fused && !C_.geom.single && C_.n % KSI_TILE == 0 && C_.lv[l].np <= KSI_MAXP;
}
// the fused ModUp: column sums over at most MODUP_MAXSRC sources below 2^60
bool PlanBuilder::use_modup(uint32_t l) const {
    return C_.modup_fused && use_ksi(l) && !C_.wide && C_.alpha <= MODUP_MAXSRC && C_.n % MODUP_TILE == 0;
}
// the fused ModDown + rescale prologue: the fused ModUp's kernel over the K
// P towers
bool PlanBuilder::use_moddown(uint32_t l) const {
    return C_.moddown_fused && use_modup(l) && C_.K <= MODUP_MAXSRC && l >= 2;
}

// 3b-5. The digits' row pass fused with the inner product over Q_l P
// (src/keyswitch_resident.cpp:171-200), into the workspace accumulators;
// then their P towers to coefficient form (ModDown's sources), as
// inner_moddown_src.  Digit towers of part pi other than the input's own hold
// the column-pass output of their forward transform (modup with
// rowpass_in_ksi).
void PlanBuilder::ksi_moddown_src(uint32_t l, const uint64_t* d, const DevKey& key, uint64_t* ws,
                                  uint32_t lazy_rows, bool p_rowpass) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K, W = l + K;
    const LevelConst& lv = C.lv[l];
    const WsLayout w = WsLayout::make(C);
    const uint64_t* dig = ws + w.dig * n;
    uint64_t* res[2] = {ws + w.res0 * n, ws + w.res1 * n};
    uint64_t* xp = ws + w.xp * n;
    for (uint32_t i = 0; i < W; ++i) {
        const uint32_t row = i < l ? C.row_q(i) : C.row_p(i - l);
        const uint32_t krow = i < l ? i : key.fullQ + (i - l);
        KsiJob J{};
        J.np = lv.np;
        for (uint32_t pi = 0; pi < lv.np; ++pi) {
            const PartConst& P = lv.parts[pi];
            const bool own = i >= P.st && i < P.st + P.sp;
            J.src[pi] = digit_tower(l, d, dig, pi, i);
            if (!own) J.xf |= 1u << pi;
            const size_t ko = ((size_t)pi * key.rows + krow) * n;
            J.kb[pi] = key.b.as<uint64_t>() + ko;
            J.ka[pi] = key.a.as<uint64_t>() + ko;
        }
        J.o0 = res[0] + (size_t)i * n;
        J.o1 = res[1] + (size_t)i * n;
        J.tw = C.rt[row].fw2;
        J.q = C.rc[row].q;
        J.qneg = C.rc[row].qneg;
        J.red = lazy_red_of(J.q);
        J.d30 = ksi_d30(J.q, J.red);
        J.lazy_out = i < lazy_rows;
        J.pairs = aligned16({J.o0, J.o1});
        for (uint32_t pi = 0; pi < lv.np; ++pi) J.pairs &= aligned16({J.src[pi], J.kb[pi], J.ka[pi]});
        ksi_.push_back(J);
    }
    for (int c = 0; c < 2; ++c)
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t r = C.row_p(k);
            (p_rowpass ? intt_pb_ : intt_p_)
                .push_back(NttJob{res[c] + (size_t)(l + k) * n, xp + (size_t)(c * K + k) * n, C.rt[r].iv2, nullptr,
                                  C.mods[r], lv.pd_post[k], lv.pd_postp[k]});
        }
}

// 1-3. ModUp of d into the workspace digits (eval form); extra towers (the
// last tower of t0/t1 in the fused EvalMult+Rescale) ride along the INTT.
void PlanBuilder::modup(uint32_t l, const uint64_t* d, uint64_t* ws, const uint64_t* extra0,
                        const uint64_t* extra1, uint64_t* dx0, uint64_t* dx1, bool rowpass_in_ksi) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K, W = l + K;
    const LevelConst& lv = C.lv[l];
    const WsLayout w = WsLayout::make(C);
    uint64_t* x = ws + w.x * n;
    uint64_t* dig = ws + w.dig * n;
    // fused: the inverse transforms' row pass here; their column pass, the
    // conversion and the digits' column pass in one ModUp kernel per part
    const bool fuse = rowpass_in_ksi && use_modup(l);
    // 1. part towers to coefficient form, scaled by n^-1 QHatInv (ModUp's first factor)
    for (uint32_t i = 0; i < l; ++i) {
        const uint32_t pi = i / C.alpha, ii = i - pi * C.alpha;
        (fuse ? intt_upb_ : intt_up_)
            .push_back(NttJob{d + (size_t)i * n, x + (size_t)i * n, C.rt[i].iv2, nullptr, C.mods[i],
                              lv.parts[pi].post[ii], lv.parts[pi].postp[ii]});
    }
    const uint64_t* ex[2] = {extra0, extra1};
    uint64_t* dx[2] = {dx0, dx1};
    for (int c = 0; c < 2; ++c)
        if (ex[c])
            intt_up_.push_back(NttJob{ex[c], dx[c], C.rt[l - 1].iv2, nullptr, C.mods[l - 1], C.rt[l - 1].ninv,
                                      C.rt[l - 1].ninvp});
    // 2-3. ModUp: convert every part to its complement, back to evaluation form
    for (uint32_t pi = 0; pi < lv.np; ++pi) {
        const PartConst& P = lv.parts[pi];
        uint64_t* dp = dig + (size_t)pi * W * n;
        if (fuse) {
            modup_src_.push_back(ModupSrc{l, pi, x, dp});
            ModupJob J{};
            for (uint32_t i = 0; i < P.sp; ++i) {
                const uint32_t r = P.st + i;
                J.src[i] = x + (size_t)r * n;
                J.stw[i] = C.rt[r].iv2;
                J.sq[i] = C.mods[r];
                J.post[i] = P.post[i];
                J.postp[i] = P.postp[i];
                if (C.modup_f64 && C.rt[r].iv2d) {
                    J.stwd[i] = C.rt[r].iv2d;
                    J.sqd[i] = (double)C.mods[r];
                    J.sqinv[i] = 1.0 / J.sqd[i];
                    f64_twiddle(P.post[i], C.mods[r], J.spost[i], J.spostq[i]);
                }
            }
            J.dst = dp;
            J.ttwd = C.modup_f64 ? P.d_ttwd : C.d_zeros;
            J.tqf = P.d_tqf;
            J.c60 = P.c60;
            J.tab = P.c60 ? P.d_tab60 : P.d_tab;
            J.tq = P.d_tq;
            J.tqneg = P.d_tqneg;
            J.ttw = P.d_ttw;
            J.tred = P.d_tred;
            J.sp = P.sp;
            J.sc = P.sc;
            modup_.push_back(J);
            continue;
        }
        modup_unfused(l, pi, x, dp, rowpass_in_ksi);
    }
}

// The conversion of part pi (sources x, coefficient form) into the digit
// towers dp, and the digits' forward transforms (column pass only when their
// row pass runs in the KSI step)
void PlanBuilder::modup_unfused(uint32_t l, uint32_t pi, const uint64_t* x, uint64_t* dp, bool colpass_only) {
    const Ctx& C = C_;
    const uint32_t n = C.n;
    const PartConst& P = C.lv[l].parts[pi];
    for (uint32_t j0 = 0; j0 < P.sc; j0 += C.conv_chunk)
        conv_up_.push_back(ConvJob{x + (size_t)P.st * n, dp, P.d_tab, P.d_tq, P.d_tqneg, P.sp, P.sc, j0,
                                   j0 + C.conv_chunk < P.sc ? j0 + C.conv_chunk : P.sc, C.wide});
    for (uint32_t j = 0; j < P.sc; ++j) {
        const uint32_t r = P.compl_rows[j];
        (colpass_only ? ntt_diga_ : ntt_dig_)
            .push_back(NttJob{dp + (size_t)j * n, dp + (size_t)j * n, C.rt[r].fw2, nullptr, C.mods[r], 0, 0});
    }
}

// Tower i (of Q_l P) of digit pi: the input's own tower inside the part,
// else the part's converted complement (src/keyswitch_resident.cpp:191-194).
const uint64_t* PlanBuilder::digit_tower(uint32_t l, const uint64_t* d, const uint64_t* dig, uint32_t pi,
                                         uint32_t i) const {
    const Ctx& C = C_;
    const uint32_t n = C.n, W = l + C.K;
    const PartConst& P = C.lv[l].parts[pi];
    if (i >= P.st && i < P.st + P.sp) return d + (size_t)i * n;
    return dig + ((size_t)pi * W + (i < P.st ? i : i - P.sp)) * n;
}

// 4. inner product with the key over Q_l P (src/keyswitch_resident.cpp:171-200)
// into the workspace accumulators, 5. their P towers to coefficient form
// (ModDown's sources).  digit(pi, i) names the digit towers.
template <class F>
void PlanBuilder::inner_moddown_src(uint32_t l, const DevKey& key, uint64_t* ws, F digit, uint64_t gk) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K, W = l + K;
    const LevelConst& lv = C.lv[l];
    const WsLayout w = WsLayout::make(C);
    uint64_t* res[2] = {ws + w.res0 * n, ws + w.res1 * n};
    uint64_t* xp = ws + w.xp * n;
    for (uint32_t i = 0; i < W; ++i) {
        const uint32_t row = i < l ? C.row_q(i) : C.row_p(i - l);
        const uint32_t krow = i < l ? i : key.fullQ + (i - l);
        InnerJob J{};
        J.np = lv.np;
        for (uint32_t pi = 0; pi < lv.np; ++pi) {
            J.dig[pi] = digit(pi, i);
            const size_t ko = ((size_t)pi * key.rows + krow) * n;
            J.kb[pi] = key.b.as<uint64_t>() + ko;
            J.ka[pi] = key.a.as<uint64_t>() + ko;
        }
        J.o0 = res[0] + (size_t)i * n;
        J.o1 = res[1] + (size_t)i * n;
        J.q = C.rc[row].q;
        J.qneg = C.rc[row].qneg;
        J.red = lazy_red_of(J.q);
        J.gk = gk;
        J.pairs = n % 2 == 0 && aligned16({J.o0, J.o1});
        for (uint32_t pi = 0; pi < lv.np; ++pi) J.pairs &= aligned16({J.dig[pi], J.kb[pi], J.ka[pi]});
        inner_.push_back(J);
    }
    for (int c = 0; c < 2; ++c)
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t r = C.row_p(k);
            intt_p_.push_back(NttJob{res[c] + (size_t)(l + k) * n, xp + (size_t)(c * K + k) * n, C.rt[r].iv2,
                                     nullptr, C.mods[r], lv.pd_post[k], lv.pd_postp[k]});
        }
}

void PlanBuilder::keyswitch(uint32_t l, const uint64_t* d, uint64_t* ba0, uint64_t* ba1,
                            const DevKey& key, uint64_t* ws, const uint64_t* t0,
                            const uint64_t* t1) {
    keyswitch_core(l, d, key, ws, nullptr, nullptr, nullptr, nullptr);
    moddown_tail(l, ws, ba0, ba1, t0, t1);
}

// 6-8. ApproxModDown of both workspace accumulators, fused with
// out_c = t_c + ba_c (src/keyswitch_resident.cpp:203-224)
void PlanBuilder::moddown_tail(uint32_t l, uint64_t* ws, uint64_t* ba0, uint64_t* ba1, const uint64_t* t0,
                               const uint64_t* t1) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K;
    const LevelConst& lv = C.lv[l];
    const WsLayout w = WsLayout::make(C);
    uint64_t* res[2] = {ws + w.res0 * n, ws + w.res1 * n};
    uint64_t* xp = ws + w.xp * n;
    uint64_t* conv = ws + w.x * n;  // the part towers are dead once ModUp has converted them
    // 6-8. ApproxModDown of both accumulators, fused with r = t + ba (:203-224)
    uint64_t* outs[2] = {ba0, ba1};
    const uint64_t* ts[2] = {t0, t1};
    for (int c = 0; c < 2; ++c) {
        uint64_t* cv = conv + (size_t)c * l * n;
        for (uint32_t j0 = 0; j0 < l; j0 += C.conv_chunk)
            conv_dn_.push_back(ConvJob{xp + (size_t)c * K * n, cv, lv.d_tabMD, lv.d_tqMD, lv.d_tqnegMD, K, l,
                                       j0, j0 + C.conv_chunk < l ? j0 + C.conv_chunk : l, C.wide});
        for (uint32_t i = 0; i < l; ++i) {
            ntt_conv_.push_back(NttJob{cv + (size_t)i * n, cv + (size_t)i * n, C.rt[i].fw2, nullptr,
                                       C.mods[i], 0, 0});
            combine_.push_back(CombineJob{res[c] + (size_t)i * n, cv + (size_t)i * n,
                                          ts[c] ? ts[c] + (size_t)i * n : nullptr, outs[c] + (size_t)i * n,
                                          C.mods[i], lv.pinv[i], lv.pinvp[i]});
        }
    }
}

// EvalMult + Rescale with ModDown and the rescale merged through NTT linearity:
//   drop_c = INTT(t_c,last) + (INTT(res_c,last) - conv_c,last) P^-1      (coefficient form)
//   y_c,i  = t_c,i s1_i + res_c,i P^-1 s1_i + NTT_i(s2_i switch(drop_c) - P^-1 s1_i conv_c,i)
// which equals rescale(t + ModDown(res)) of the reference
// (tests/test_evalmult_rescale.cpp:69-73) residue for residue, with one
// forward transform per surviving tower instead of two.
void PlanBuilder::evalmult_rescale_fused(uint32_t l, const uint64_t* c0a, const uint64_t* c1a,
                                         const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0,
                                         uint64_t* o1, const DevKey& key, uint64_t* ws) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K, last = l - 1;
    const WsLayout w = WsLayout::make(C);
    uint64_t* t[3] = {ws + w.t0 * n, ws + w.t1 * n, ws + w.t2 * n};
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        tensor_.push_back(tensor_job(c0a + o, c1a + o, c0b + o, c1b + o, t[0] + o, t[1] + o, t[2] + o,
                                     C.rc[i], n));
    }
    uint64_t* dt = ws + w.dt * n;
    uint64_t* dr = ws + w.dr * n;
    const bool fmd = use_moddown(l);
    // accumulator rows below `last` only feed the epilogue's lazy Shoup products
    keyswitch_core(l, t[2], key, ws, t[0] + (size_t)last * n, t[1] + (size_t)last * n, dt, dt + n, last, fmd);
    uint64_t* res[2] = {ws + w.res0 * n, ws + w.res1 * n};
    uint64_t* xp = ws + w.xp * n;
    uint64_t* v = ws + w.dig * n;     // digits are dead after the inner product
    uint64_t* outs[2] = {o0, o1};
    for (int c = 0; c < 2; ++c) {
        intt_p_.push_back(NttJob{res[c] + (size_t)last * n, dr + (size_t)c * n, C.rt[last].iv2, nullptr,
                                 C.mods[last], C.rt[last].ninv, C.rt[last].ninvp});
        const ModdownSrc m{l, (uint32_t)c, xp + (size_t)c * K * n, dt + (size_t)c * n, dr + (size_t)c * n,
                           t[c], res[c], v + (size_t)c * last * n, outs[c]};
        rescale_tail(m, fmd);
    }
}

// The rescale tail of one component: from the K P towers (coefficient form,
// or after their inverse row pass when fused), the two dropped-tower
// operands and the accumulators, the surviving output towers.
void PlanBuilder::rescale_tail(const ModdownSrc& m, bool fused) {
    const Ctx& C = C_;
    const uint32_t n = C.n, K = C.K, l = m.l, last = l - 1;
    const LevelConst& lv = C.lv[l];
    if (fused) {
        ModupJob J{};
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t r = C.row_p(k);
            J.src[k] = m.xp + (size_t)k * n;
            J.stw[k] = C.rt[r].iv2;
            J.sq[k] = C.mods[r];
            J.post[k] = lv.pd_post[k];
            J.postp[k] = lv.pd_postp[k];
            if (C.modup_f64 && C.rt[r].iv2d) {
                J.stwd[k] = C.rt[r].iv2d;
                J.sqd[k] = (double)C.mods[r];
                J.sqinv[k] = 1.0 / J.sqd[k];
                f64_twiddle(lv.pd_post[k], C.mods[r], J.spost[k], J.spostq[k]);
            }
        }
        J.dst = m.out;   // surviving tower i's column-pass output at out + i n
        J.tab = lv.d_tabRS;
        J.tq = lv.d_rs_tq;
        J.tqneg = lv.d_rs_tqneg;
        J.ttw = lv.d_rs_ttw;
        J.tred = lv.d_rs_tred;
        J.ttwd = C.modup_f64 ? lv.d_rs_ttwd : C.d_zeros;
        J.tqf = lv.d_rs_tqf;
        J.sp = K;
        J.sc = last;
        J.dt = m.dt;
        J.dr = m.dr;
        J.tabd = lv.d_tabMD + last;   // PHatModq column of q_last, row stride l
        J.tg = lv.d_rstg;
        J.tsw = lv.d_rs_sw;
        J.q_old = C.mods[last];
        J.half = C.mods[last] >> 1;
        J.qd_neg = C.rc[last].qneg;
        J.pinv = lv.pinv[last];
        J.pinvp = lv.pinvp[last];
        J.strided = l;
        moddown_.push_back(J);
        moddown_src_.push_back(m);
    } else {
        for (uint32_t j0 = 0; j0 < last; j0 += C.rs_chunk) {
            RsJob R{};
            R.xp = m.xp;
            R.dt = m.dt;
            R.dr = m.dr;
            R.out = m.vc;
            R.tab = lv.d_tabRS;
            R.tg = lv.d_rstg;
            R.tabd = lv.d_tabMD + last;   // PHatModq column of q_last, row stride l
            R.q_old = C.mods[last];
            R.half = C.mods[last] >> 1;
            R.qd_neg = C.rc[last].qneg;
            R.pinv = lv.pinv[last];
            R.pinvp = lv.pinvp[last];
            R.K = K;
            R.ntgt = last;
            R.j0 = j0;
            R.j1 = j0 + C.rs_chunk < last ? j0 + C.rs_chunk : last;
            R.strided = l;
            R.wide = C.wide;
            rs_.push_back(R);
        }
    }
    for (uint32_t i = 0; i < last; ++i) {
        // fused: the column pass is done, this row pass reads and writes out
        NttJob J{fused ? m.out + (size_t)i * n : m.vc + (size_t)i * n, m.out + (size_t)i * n, C.rt[i].fw2, nullptr,
                 C.mods[i], 0, 0};
        J.e0 = m.t + (size_t)i * n;
        J.c0 = lv.s1[i];
        J.c0p = lv.s1p[i];
        J.e1 = m.res + (size_t)i * n;
        J.c1 = lv.ps1[i];
        J.c1p = lv.ps1p[i];
        (fused ? ntt_vb_ : ntt_v_).push_back(J);
    }
}

void PlanBuilder::evalmult(uint32_t l, const uint64_t* c0a, const uint64_t* c1a, const uint64_t* c0b,
                           const uint64_t* c1b, uint64_t* r0, uint64_t* r1, const DevKey& key,
                           uint64_t* ws) {
    const Ctx& C = C_;
    const uint32_t n = C.n;
    const WsLayout w = WsLayout::make(C);
    uint64_t* t0 = ws + w.t0 * n;
    uint64_t* t1 = ws + w.t1 * n;
    uint64_t* t2 = ws + w.t2 * n;
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        tensor_.push_back(tensor_job(c0a + o, c1a + o, c0b + o, c1b + o, t0 + o, t1 + o, t2 + o, C.rc[i], n));
    }
    keyswitch(l, t2, r0, r1, key, ws, t0, t1);
}

void PlanBuilder::rescale_comp(uint32_t l, const uint64_t* c, uint64_t* o, uint64_t* drop, uint64_t* sw) {
    const Ctx& C = C_;
    const uint32_t n = C.n, last = l - 1;
    const LevelConst& lv = C.lv[l];
    intt_drop_.push_back(NttJob{c + (size_t)last * n, drop, C.rt[last].iv2, nullptr, C.mods[last],
                                C.rt[last].ninv, C.rt[last].ninvp});
    for (uint32_t i = 0; i < last; ++i) {
        uint64_t* si = sw + (size_t)i * n;
        switch_.push_back(make_switch(drop, si, C.mods[last], C.mods[i]));
        ntt_sw_.push_back(NttJob{si, si, C.rt[i].fw2, nullptr, C.mods[i], 0, 0});
        fuse_.push_back(FuseJob{c + (size_t)i * n, si, o + (size_t)i * n, C.mods[i], lv.s1[i], lv.s1p[i],
                                lv.s2[i], lv.s2p[i]});
    }
}

void PlanBuilder::rescale(uint32_t l, const uint64_t* c0, const uint64_t* c1, uint64_t* o0, uint64_t* o1,
                          uint64_t* ws) {
    const Ctx& C = C_;
    if (l < 2 || l > C.L) throw std::runtime_error("gpufhe: rescale needs at least two towers");
    const uint32_t n = C.n;
    const WsLayout w = WsLayout::make(C);
    uint64_t* drop = ws + w.drop * n;
    uint64_t* sw = ws + w.dig * n;   // ModUp digits are dead by now
    rescale_comp(l, c0, o0, drop, sw);
    rescale_comp(l, c1, o1, drop + n, sw + (size_t)(l - 1) * n);
}

// ws: L towers at least (Ctx::rescale_ws_words): the dropped tower's
// coefficient form, then the l - 1 switched towers
void PlanBuilder::rescale_one(uint32_t l, const uint64_t* c, uint64_t* o, uint64_t* ws) {
    const Ctx& C = C_;
    if (l < 2 || l > C.L) throw std::runtime_error("gpufhe: rescale needs at least two towers");
    rescale_comp(l, c, o, ws, ws + C.n);
}

void PlanBuilder::evalmult_rescale(uint32_t l, const uint64_t* c0a, const uint64_t* c1a,
                                   const uint64_t* c0b, const uint64_t* c1b, uint64_t* o0, uint64_t* o1,
                                   const DevKey& key, uint64_t* ws) {
    const Ctx& C = C_;
    if (l < 2) throw std::runtime_error("gpufhe: EvalMult+Rescale needs at least two towers");
    if (C.fused) {
        evalmult_rescale_fused(l, c0a, c1a, c0b, c1b, o0, o1, key, ws);
        return;
    }
    const WsLayout w = WsLayout::make(C);
    uint64_t* r0 = ws + w.t0 * C.n;
    uint64_t* r1 = ws + w.t1 * C.n;
    evalmult(l, c0a, c1a, c0b, c1b, r0, r1, key, ws);
    rescale(l, r0, r1, o0, o1, ws);
}

void PlanBuilder::rotate(uint32_t l, uint64_t k, const uint64_t* c0, const uint64_t* c1, uint64_t* o0,
                         uint64_t* o1, const DevKey& key, uint64_t* ws) {
    const Ctx& C = C_;
    const uint32_t n = C.n;
    if (!(k & 1) || k >= 2ull * n) throw std::runtime_error("gpufhe: automorphism index must be odd and < 2n");
    const WsLayout w = WsLayout::make(C);
    uint64_t* a0 = ws + w.t0 * n;
    uint64_t* a1 = ws + w.t2 * n;
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        perm_.push_back(PermJob{c0 + o, a0 + o, k});
        perm_.push_back(PermJob{c1 + o, a1 + o, k});
    }
    // (sigma(c0) + ba0, ba1) with (ba0, ba1) = KS(sigma(c1)) (src/rotate.cpp:118-132)
    keyswitch(l, a1, o0, o1, key, ws, a0, nullptr);
}

// m rotations of one ciphertext from one decomposition of c1: ModUp once into
// ws_base; rotation j permutes every digit tower (and c0) into its own
// workspace, then runs the inner product with key j and ModDown there.
void PlanBuilder::rotate_hoisted(uint32_t l, uint32_t m, const uint64_t* ks, const uint64_t* c0,
                                 const uint64_t* c1, uint64_t* const* o0, uint64_t* const* o1,
                                 const DevKey* const* keys, uint64_t* ws_base, uint64_t* const* ws_rot) {
    const Ctx& C = C_;
    const uint32_t n = C.n;
    if (l < 1 || l > C.L) throw std::runtime_error("gpufhe: bad level");
    const LevelConst& lv = C.lv[l];
    for (uint32_t j = 0; j < m; ++j) {
        if (!(ks[j] & 1) || ks[j] >= 2ull * n)
            throw std::runtime_error("gpufhe: automorphism index must be odd and < 2n");
        if (keys[j]->nparts < lv.np) throw std::runtime_error("gpufhe: key has too few parts for this level");
    }
    const WsLayout w = WsLayout::make(C);
    modup(l, c1, ws_base, nullptr, nullptr, nullptr, nullptr);
    const uint64_t* dig = ws_base + w.dig * n;
    for (uint32_t j = 0; j < m; ++j) {
        uint64_t* ws = ws_rot[j];
        uint64_t* a0 = ws + w.t0 * n;
        for (uint32_t i = 0; i < l; ++i) perm_.push_back(PermJob{c0 + (size_t)i * n, a0 + (size_t)i * n, ks[j]});
        // the inner product reads the shared digits through sigma_k: no permuted copies
        inner_moddown_src(
            l, *keys[j], ws, [&](uint32_t pi, uint32_t i) { return digit_tower(l, c1, dig, pi, i); }, ks[j]);
        moddown_tail(l, ws, o0[j], o1[j], a0, nullptr);
    }
}

// sum_j pt_j * rot_kj(c0, c1): the rotations share one decomposition
// (rotate_hoisted, results left in each rotation's workspace), a term with
// no key and k = 1 is the ciphertext itself, and one accumulate launch per
// tower weights and sums all terms (kernels/stage_acc.cu:12-28 for every
// term at once).  pts[j]: plaintext in evaluation form, at least l towers.
void PlanBuilder::linear_transform(uint32_t l, uint32_t m, const uint64_t* ks, const uint64_t* c0,
                                   const uint64_t* c1, const uint64_t* const* pts, uint64_t* o0, uint64_t* o1,
                                   const DevKey* const* keys, uint64_t* ws_base, uint64_t* const* ws_rot) {
    const Ctx& C = C_;
    const uint32_t n = C.n;
    if (m == 0 || m > MAXP) throw std::runtime_error("gpufhe: a linear transform takes 1 to 16 terms");
    if (l < 1 || l > C.L) throw std::runtime_error("gpufhe: bad level");
    // the accumulate keeps sum_j pt_j x_j in 128 bits and reduces it with one
    // Montgomery step, which needs sum < q 2^64, i.e. m q < 2^64
    for (uint32_t i = 0; i < l; ++i)
        if (C.mods[i] > ~0ull / m) throw std::runtime_error("gpufhe: too many linear-transform terms for this modulus");
    const WsLayout w = WsLayout::make(C);
    std::vector<const uint64_t*> x0(m), x1(m);
    std::vector<uint64_t> rk;
    std::vector<uint64_t*> r0, r1, rws;
    std::vector<const DevKey*> rkey;
    for (uint32_t j = 0; j < m; ++j) {
        if (!keys[j]) {
            if (ks[j] != 1) throw std::runtime_error("gpufhe: a term without a rotation key must have k = 1");
            x0[j] = c0;
            x1[j] = c1;
            continue;
        }
        uint64_t* ws = ws_rot[j];
        r0.push_back(ws + w.t0 * n);   // sigma(c0) + ba0 overwrites sigma(c0) element by element
        r1.push_back(ws + w.t1 * n);
        rk.push_back(ks[j]);
        rkey.push_back(keys[j]);
        rws.push_back(ws);
        x0[j] = r0.back();
        x1[j] = r1.back();
    }
    if (!rk.empty())
        rotate_hoisted(l, (uint32_t)rk.size(), rk.data(), c0, c1, r0.data(), r1.data(), rkey.data(), ws_base,
                       rws.data());
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        PtAccJob J{};
        J.m = m;
        for (uint32_t j = 0; j < m; ++j) {
            J.pt[j] = pts[j] + o;
            J.x0[j] = x0[j] + o;
            J.x1[j] = x1[j] + o;
        }
        J.o0 = o0 + o;
        J.o1 = o1 + o;
        J.q = C.rc[i].q;
        J.qneg = C.rc[i].qneg;
        J.r2 = C.rc[i].r2;
        ptacc_.push_back(J);
    }
}

void PlanBuilder::add(uint32_t l, const uint64_t* a0, const uint64_t* a1, const uint64_t* b0,
                      const uint64_t* b1, uint64_t* o0, uint64_t* o1) {
    const uint32_t n = C_.n;
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        add_.push_back(AddJob{a0 + o, b0 + o, o0 + o, C_.mods[i]});
        add_.push_back(AddJob{a1 + o, b1 + o, o1 + o, C_.mods[i]});
    }
}

// (c0 pt, c1 pt) per tower: src/ptops.cpp:20-26, one accumulate job per tower
void PlanBuilder::mul_plain(uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                            uint64_t* o1) {
    if (l < 1 || l > C_.L) throw std::runtime_error("gpufhe: bad level");
    const uint32_t n = C_.n;
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        PtAccJob J{};
        J.m = 1;
        J.pt[0] = pt + o;
        J.x0[0] = c0 + o;
        J.x1[0] = c1 + o;
        J.o0 = o0 + o;
        J.o1 = o1 + o;
        J.q = C_.rc[i].q;
        J.qneg = C_.rc[i].qneg;
        J.r2 = C_.rc[i].r2;
        ptacc_.push_back(J);
    }
}

// (c0 + pt, c1): the ct_add_pt of src/ptops.cpp:1-5 (c1 copied when out is not in place)
void PlanBuilder::add_plain(uint32_t l, const uint64_t* c0, const uint64_t* c1, const uint64_t* pt, uint64_t* o0,
                            uint64_t* o1) {
    if (l < 1 || l > C_.L) throw std::runtime_error("gpufhe: bad level");
    const uint32_t n = C_.n;
    for (uint32_t i = 0; i < l; ++i) {
        const size_t o = (size_t)i * n;
        add_.push_back(AddJob{c0 + o, pt + o, o0 + o, C_.mods[i]});
        if (o1 != c1) copy_.push_back(CopyJob{c1 + o, o1 + o});
    }
}

// Jobs of one stage are independent; ordering them so that consecutive
// workgroups share a twiddle table (transforms), a key row (inner product) or a
// constant table (conversions) keeps those operands in L2 across the batch
// instead of re-reading them per ciphertext.
template <class T, class K>
static void group_by(std::vector<T>& v, K key) {
    std::stable_sort(v.begin(), v.end(), [&](const T& a, const T& b) { return key(a) < key(b); });
}

std::unique_ptr<Plan> PlanBuilder::finish() {
    std::vector<char> blob;
    auto P = layout(blob);
    P->dev.alloc(blob.size());
    if (!blob.empty())
        hip_check(hipMemcpy(P->dev.p, blob.data(), blob.size(), hipMemcpyHostToDevice), "upload plan");
    P->base = P->dev.p;
    return P;
}

std::unique_ptr<Plan> PlanBuilder::finish_into(void* d_arena, void* h_stage, size_t cap, hipStream_t s) {
    std::vector<char> blob;
    auto P = layout(blob);
    if (blob.size() > cap) throw std::runtime_error("gpufhe: batch plan larger than the job arena");
    if (!blob.empty()) {
        std::memcpy(h_stage, blob.data(), blob.size());
        hip_check(hipMemcpyAsync(d_arena, h_stage, blob.size(), hipMemcpyHostToDevice, s), "stage plan");
    }
    P->base = d_arena;
    return P;
}

std::unique_ptr<Plan> PlanBuilder::layout(std::vector<char>& blob) {
    // a fused ModUp stage too small to fill the chip runs unfused: the
    // inverse transforms whole, then the conversion and the column passes
    if (!modup_.empty() && modup_.size() * (C_.n / MODUP_TILE) < C_.modup_min_wgs) {
        for (NttJob& j : intt_upb_) intt_up_.push_back(j);
        intt_upb_.clear();
        for (const ModupSrc& m : modup_src_) modup_unfused(m.l, m.pi, m.x, m.dp, true);
        modup_.clear();
        modup_src_.clear();
    }
    // likewise a fused ModDown stage: the P towers' whole inverse transforms,
    // the prologue, the epilogue transforms whole
    if (!moddown_.empty() && moddown_.size() * (C_.n / MODUP_TILE) < C_.modup_min_wgs) {
        for (NttJob& j : intt_pb_) intt_p_.push_back(j);
        intt_pb_.clear();
        ntt_vb_.clear();
        moddown_.clear();
        std::vector<ModdownSrc> ms;
        ms.swap(moddown_src_);
        for (const ModdownSrc& m : ms) rescale_tail(m, false);
    }
    for (auto* v : {&intt_up_, &intt_upb_, &ntt_dig_, &ntt_diga_, &intt_p_, &intt_pb_, &ntt_conv_, &intt_drop_,
                    &ntt_sw_, &ntt_v_, &ntt_vb_}) {
        group_by(*v, [](const NttJob& j) { return (uintptr_t)j.tw; });
        // every transform of a plan uses the context's paired tables; its phase B
        // reads the matching row table of the current radix
        for (NttJob& j : *v) {
            j.twr = C_.row_table_of(j.tw);
            if (!j.twr && !C_.geom.single) throw std::runtime_error("gpufhe: transform table of another context");
        }
    }
    // forward transforms over moduli below 2^57 skip the butterflies'
    // conditional subtractions; the digits stay unreduced for the inner product
    for (auto* v : {&ntt_dig_, &ntt_diga_, &ntt_conv_, &ntt_sw_, &ntt_v_, &ntt_vb_})
        for (NttJob& j : *v) j.red = lazy_red_of(j.q);
    for (NttJob& j : ntt_dig_) j.lazy_out = j.red != 0;
    for (KsiJob& j : ksi_) {
        j.twr = C_.row_table_of(j.tw);
        if (!j.twr) throw std::runtime_error("gpufhe: transform table of another context");
    }
    // with the fused ModDown, a P row's inner product runs the inverse row
    // pass of its two output rows itself (jobs with at least two transformed
    // parts: two LDS regions to hand the rows over in), and those row-pass
    // jobs leave their stage
    if (C_.ksi_rowpass && !moddown_.empty() && !intt_pb_.empty()) {
        std::map<const uint64_t*, size_t> by_in;
        for (size_t i = 0; i < intt_pb_.size(); ++i) by_in[intt_pb_[i].in] = i;
        std::vector<char> used(intt_pb_.size(), 0);
        for (KsiJob& j : ksi_) {
            const auto a = by_in.find(j.o0), b = by_in.find(j.o1);
            if (a == by_in.end() || b == by_in.end() || __builtin_popcount(j.xf) < 2 || j.lazy_out) continue;
            if (used[a->second] || used[b->second]) continue;
            j.rp_on = 1;
            j.rp[0] = intt_pb_[a->second];
            j.rp[1] = intt_pb_[b->second];
            used[a->second] = used[b->second] = 1;
        }
        std::vector<NttJob> rest;
        for (size_t i = 0; i < intt_pb_.size(); ++i)
            if (!used[i]) rest.push_back(intt_pb_[i]);
        intt_pb_.swap(rest);
    }
    // With the fp64 fused ModUp in this plan, the inverse row pass in front
    // of it runs in fp64 too for moduli below 2^50 (its inputs, the tensor's
    // t2 or a ciphertext tower, are reduced residues), and with ksi_f64 so do
    // the row passes of the fused inner product (the ModUp leaves those
    // digits in [0, 2q)).
    if (!modup_.empty() && C_.modup_f64 && !C_.geom.single) {
        auto f64_job = [&](auto& j) {
            const uint64_t* td = C_.f64_table_of(j.tw);
            if (!td) return;
            j.twd = td;
            j.twrd = C_.row_table_of(td);
            j.fq = (double)j.q;
            j.fqinv = 1.0 / j.fq;
        };
        if (C_.ksi_f64)
            for (KsiJob& j : ksi_) f64_job(j);
        for (NttJob& j : intt_upb_) f64_job(j);
    }
    // KSI jobs by transformed-part count (one step each: that many LDS
    // regions and waves), by key within a step
    auto nxf = [](const KsiJob& j) { return (uint32_t)__builtin_popcount(j.xf); };
    group_by(ksi_, [](const KsiJob& j) { return (uintptr_t)j.kb[0]; });
    group_by(ksi_, nxf);
    group_by(inner_, [](const InnerJob& j) { return (uintptr_t)j.kb[0]; });
    split_chunks(conv_up_, C_.n, C_.rechunk_wgs);
    split_chunks(conv_dn_, C_.n, C_.rechunk_wgs);
    split_chunks(rs_, C_.n, C_.rechunk_wgs);
    group_by(conv_up_, [](const ConvJob& j) { return (uintptr_t)j.tab; });
    group_by(tensor_, [](const TensorJob& j) { return j.q; });
    auto P = std::make_unique<Plan>();
    P->n = C_.n;
    P->logn = C_.logn;
    P->geom = C_.geom;
    blob.clear();
    auto stage = [&](Step::Kind kind, const void* data, size_t cnt, size_t sz) {
        if (!cnt) return;
        size_t off = (blob.size() + 15) & ~(size_t)15;
        blob.resize(off + cnt * sz);
        std::memcpy(blob.data() + off, data, cnt * sz);
        P->steps.push_back(Step{kind, off, (uint32_t)cnt});
    };
#define GF_STAGE(k, v) stage(Step::k, v.data(), v.size(), sizeof(v[0]))
    // the tensor product fused with the inverse row pass of its t2 where every
    // tensor job has that row pass (the fused EvalMult+Rescale with the fused
    // ModUp): t2 goes to the row pass in LDS instead of through HBM
    if (C_.tensor_ib && !C_.geom.single && C_.geom.R * (1u << C_.geom.logN2) == NTT_TILE && !tensor_.empty() &&
        tensor_.size() == intt_upb_.size()) {
        std::map<const uint64_t*, size_t> by_in;
        for (size_t i = 0; i < intt_upb_.size(); ++i) by_in[intt_upb_[i].in] = i;
        bool all = true;
        for (const TensorJob& t : tensor_) all &= t.pairs && by_in.count(t.t2) == 1;
        if (all && by_in.size() == tensor_.size()) {
            for (const TensorJob& t : tensor_) tensor_ib_.push_back(TensorIbJob{t, intt_upb_[by_in[t.t2]]});
            tensor_.clear();
            intt_upb_.clear();
        }
    }
    GF_STAGE(TENSOR_IB, tensor_ib_);
    GF_STAGE(TENSOR, tensor_);
    GF_STAGE(PERM, perm_);
    GF_STAGE(NTT_IB, intt_upb_);
    GF_STAGE(NTT_I, intt_up_);
    // fused ModUp jobs by source count and conversion reduction (one launch each)
    auto mkey = [](const ModupJob& a) { return a.sp | a.c60 << 8; };
    std::stable_sort(modup_.begin(), modup_.end(),
                     [&](const ModupJob& a, const ModupJob& b) { return mkey(a) < mkey(b); });
    for (size_t i = 0; i < modup_.size();) {
        size_t e = i;
        while (e < modup_.size() && mkey(modup_[e]) == mkey(modup_[i])) ++e;
        stage(Step::MODUP, modup_.data() + i, e - i, sizeof(ModupJob));
        P->steps.back().arg = mkey(modup_[i]);
        i = e;
    }
    GF_STAGE(CONV, conv_up_);
    GF_STAGE(NTT_F, ntt_dig_);
    GF_STAGE(NTT_FA, ntt_diga_);
    GF_STAGE(INNER, inner_);
    for (size_t i = 0; i < ksi_.size();) {
        size_t e = i;
        while (e < ksi_.size() && nxf(ksi_[e]) == nxf(ksi_[i])) ++e;
        const uint32_t x = nxf(ksi_[i]), waves = std::max(1u, std::min<uint32_t>(x + C_.ksi_extra, KSI_WAVES));
        stage(Step::KSI, ksi_.data() + i, e - i, sizeof(KsiJob));
        P->steps.back().arg = x | waves << 8;
        i = e;
    }
    GF_STAGE(NTT_I, intt_p_);
    GF_STAGE(NTT_IB, intt_pb_);
    // fused ModDown jobs by source count (one launch each)
    std::stable_sort(moddown_.begin(), moddown_.end(),
                     [](const ModupJob& a, const ModupJob& b) { return a.sp < b.sp; });
    for (size_t i = 0; i < moddown_.size();) {
        size_t e = i;
        while (e < moddown_.size() && moddown_[e].sp == moddown_[i].sp) ++e;
        stage(Step::MODDOWN, moddown_.data() + i, e - i, sizeof(ModupJob));
        P->steps.back().arg = moddown_[i].sp;
        i = e;
    }
    GF_STAGE(CONV, conv_dn_);
    GF_STAGE(NTT_F, ntt_conv_);
    GF_STAGE(COMBINE, combine_);
    GF_STAGE(ADD, add_rot_);
    GF_STAGE(COPY, copy_rot_);
    GF_STAGE(NTT_I, intt_drop_);
    GF_STAGE(SWITCH, switch_);
    GF_STAGE(NTT_F, ntt_sw_);
    GF_STAGE(FUSE, fuse_);
    GF_STAGE(RSV, rs_);
    GF_STAGE(NTT_F, ntt_v_);
    GF_STAGE(NTT_FB, ntt_vb_);
    GF_STAGE(ADD, add_);
    GF_STAGE(COPY, copy_);
    GF_STAGE(PTACC, ptacc_);
#undef GF_STAGE
    return P;
}

}  // namespace gf
