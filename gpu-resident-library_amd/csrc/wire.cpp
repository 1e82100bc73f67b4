// Wire format for ciphertexts and switching keys (checkpointing resident
// state, moving it between processes).  The reference stores raw tower-major
// u64 arrays with no header (src/device_ciphertext.cpp:43-61); here the same
// arrays are framed by a header that names the ring, the moduli and a
// checksum, so a reader can refuse data for another parameter set.
//
// Layout, little-endian, every field 8-byte aligned:
//   ciphertext: "GFHECT01" | u32 n | u32 towers | f64 scale | u64 checksum |
//               u64 moduli[towers] | u64 c0[towers][n] | u64 c1[towers][n]
//   key:        "GFHEKY01" | u32 n | u32 nparts | u32 rows | u32 fullQ |
//               u32 L | u32 K | u64 checksum | u64 moduli[L+K] |
//               u64 av[nparts][rows][n] | u64 bv[nparts][rows][n]   (plain form)
// checksum = FNV-1a (64-bit) over everything after the header.
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/gpufhe_capi.h"
#include "handles.hpp"

using namespace gf;

namespace {
const char CT_MAGIC[8] = {'G', 'F', 'H', 'E', 'C', 'T', '0', '1'};
const char KY_MAGIC[8] = {'G', 'F', 'H', 'E', 'K', 'Y', '0', '1'};

struct CtHeader {
    char magic[8];
    uint32_t n, towers;
    double scale;
    uint64_t checksum;
};
struct KeyHeader {
    char magic[8];
    uint32_t n, nparts, rows, fullQ, L, K;
    uint64_t checksum;
};
static_assert(sizeof(CtHeader) == 32 && sizeof(KeyHeader) == 40, "wire header layout");

uint64_t fnv1a(const unsigned char* p, size_t len) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < len; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
// bounds a header must meet before any length is computed from it
constexpr uint32_t WIRE_MAX_N = 1u << 17, WIRE_MAX_TOWERS = 4096;
bool ct_shape_ok(uint32_t n, uint32_t towers) {
    return n >= 2 && n <= WIRE_MAX_N && (n & (n - 1)) == 0 && towers >= 1 && towers <= WIRE_MAX_TOWERS;
}
size_t ct_words(uint32_t n, uint32_t towers) { return (size_t)towers + 2 * (size_t)towers * n; }
// every residue of tower t below its modulus
void check_residues(const uint64_t* v, const uint64_t* mods, uint32_t towers, uint32_t n, size_t stride_towers) {
    for (uint32_t t = 0; t < towers; ++t) {
        const uint64_t q = mods[t % stride_towers];
        const uint64_t* x = v + (size_t)t * n;
        for (uint32_t k = 0; k < n; ++k)
            if (x[k] >= q) throw std::runtime_error("gpufhe: packed data holds a residue not below its modulus");
    }
}
}  // namespace

extern "C" {

size_t gf_ct_packed_bytes(uint32_t n, uint32_t towers) {
    return ct_shape_ok(n, towers) ? sizeof(CtHeader) + 8 * ct_words(n, towers) : 0;
}

int gf_ct_pack(const uint64_t* h_c0, const uint64_t* h_c1, uint32_t n, uint32_t towers, const uint64_t* h_moduli,
               double scale, void* out, size_t cap) {
    return guard([&] {
        const size_t need = gf_ct_packed_bytes(n, towers);
        if (!need) throw std::runtime_error("gpufhe: ciphertext shape outside the wire format (n a power of two <= 2^17, 1..4096 towers)");
        if (cap < need) throw std::runtime_error("gpufhe: output buffer smaller than gf_ct_packed_bytes");
        check_residues(h_c0, h_moduli, towers, n, towers);
        check_residues(h_c1, h_moduli, towers, n, towers);
        unsigned char* o = (unsigned char*)out;
        unsigned char* body = o + sizeof(CtHeader);
        const size_t T = (size_t)towers * n;
        std::memcpy(body, h_moduli, (size_t)towers * 8);
        std::memcpy(body + (size_t)towers * 8, h_c0, T * 8);
        std::memcpy(body + (size_t)towers * 8 + T * 8, h_c1, T * 8);
        CtHeader h{};
        std::memcpy(h.magic, CT_MAGIC, 8);
        h.n = n;
        h.towers = towers;
        h.scale = scale;
        h.checksum = fnv1a(body, need - sizeof(CtHeader));
        std::memcpy(o, &h, sizeof h);
    });
}

int gf_ct_unpack_header(const void* buf, size_t len, uint32_t* n, uint32_t* towers, double* scale) {
    return guard([&] {
        if (len < sizeof(CtHeader)) throw std::runtime_error("gpufhe: buffer shorter than a ciphertext header");
        CtHeader h;
        std::memcpy(&h, buf, sizeof h);
        if (std::memcmp(h.magic, CT_MAGIC, 8)) throw std::runtime_error("gpufhe: not a packed ciphertext");
        if (!ct_shape_ok(h.n, h.towers)) throw std::runtime_error("gpufhe: packed ciphertext header out of range");
        if (len != gf_ct_packed_bytes(h.n, h.towers)) throw std::runtime_error("gpufhe: packed ciphertext length mismatch");
        *n = h.n;
        *towers = h.towers;
        *scale = h.scale;
    });
}

int gf_ct_unpack(const void* buf, size_t len, uint64_t* h_moduli, uint64_t* h_c0, uint64_t* h_c1) {
    return guard([&] {
        uint32_t n, towers;
        double scale;
        if (gf_ct_unpack_header(buf, len, &n, &towers, &scale)) throw std::runtime_error(gf_last_error());
        CtHeader h;
        std::memcpy(&h, buf, sizeof h);
        const unsigned char* body = (const unsigned char*)buf + sizeof(CtHeader);
        if (fnv1a(body, len - sizeof(CtHeader)) != h.checksum) throw std::runtime_error("gpufhe: ciphertext checksum mismatch");
        const size_t T = (size_t)towers * n;
        std::memcpy(h_moduli, body, (size_t)towers * 8);
        std::memcpy(h_c0, body + (size_t)towers * 8, T * 8);
        std::memcpy(h_c1, body + (size_t)towers * 8 + T * 8, T * 8);
        check_residues(h_c0, h_moduli, towers, n, towers);
        check_residues(h_c1, h_moduli, towers, n, towers);
    });
}

size_t gf_key_packed_bytes(const gf_key* k) {
    const Ctx& C = *k->ctx->C;
    return sizeof(KeyHeader) + 8 * ((size_t)(C.L + C.K) + 2 * (size_t)k->K->nparts * k->K->rows * C.n);
}

int gf_key_pack(const gf_key* k, void* out, size_t cap, hipStream_t s) {
    return guard([&] {
        const Ctx& C = *k->ctx->C;
        const DevKey& K = *k->K;
        const size_t need = gf_key_packed_bytes(k);
        if (cap < need) throw std::runtime_error("gpufhe: output buffer smaller than gf_key_packed_bytes");
        if (K.rows != K.fullQ + C.K) throw std::runtime_error("gpufhe: key rows do not follow the fullQ + K layout");
        const size_t words = (size_t)K.nparts * K.rows * C.n;
        unsigned char* body = (unsigned char*)out + sizeof(KeyHeader);
        // rows between L and fullQ were never converted to Montgomery form and are
        // written as uploaded; every other row comes back in plain form
        std::vector<uint64_t> av(words), bv(words);
        download_key(C, K, av.data(), bv.data(), s);
        std::memcpy(body, C.mods.data(), (size_t)(C.L + C.K) * 8);
        std::memcpy(body + (size_t)(C.L + C.K) * 8, av.data(), words * 8);
        std::memcpy(body + (size_t)(C.L + C.K) * 8 + words * 8, bv.data(), words * 8);
        KeyHeader h{};
        std::memcpy(h.magic, KY_MAGIC, 8);
        h.n = C.n;
        h.nparts = K.nparts;
        h.rows = K.rows;
        h.fullQ = K.fullQ;
        h.L = C.L;
        h.K = C.K;
        h.checksum = fnv1a(body, need - sizeof(KeyHeader));
        std::memcpy(out, &h, sizeof h);
    });
}

gf_key* gf_key_unpack(gf_ctx* c, const void* buf, size_t len, hipStream_t s) {
    return guard_ptr<gf_key>([&] {
        const Ctx& C = *c->C;
        if (len < sizeof(KeyHeader)) throw std::runtime_error("gpufhe: buffer shorter than a key header");
        KeyHeader h;
        std::memcpy(&h, buf, sizeof h);
        if (std::memcmp(h.magic, KY_MAGIC, 8)) throw std::runtime_error("gpufhe: not a packed key");
        if (h.n != C.n || h.L != C.L || h.K != C.K)
            throw std::runtime_error("gpufhe: packed key belongs to another parameter set");
        if (h.nparts == 0 || h.nparts > C.L || h.fullQ < C.L || h.fullQ > WIRE_MAX_TOWERS || h.rows != h.fullQ + C.K)
            throw std::runtime_error("gpufhe: packed key shape out of range");
        const size_t words = (size_t)h.nparts * h.rows * h.n;
        const size_t need = sizeof(KeyHeader) + 8 * ((size_t)(h.L + h.K) + 2 * words);
        if (len != need) throw std::runtime_error("gpufhe: packed key length mismatch");
        const unsigned char* body = (const unsigned char*)buf + sizeof(KeyHeader);
        if (fnv1a(body, len - sizeof(KeyHeader)) != h.checksum) throw std::runtime_error("gpufhe: key checksum mismatch");
        if (std::memcmp(body, C.mods.data(), (size_t)(C.L + C.K) * 8))
            throw std::runtime_error("gpufhe: packed key moduli differ from the context's");
        std::vector<uint64_t> av(words), bv(words);
        std::memcpy(av.data(), body + (size_t)(h.L + h.K) * 8, words * 8);
        std::memcpy(bv.data(), body + (size_t)(h.L + h.K) * 8 + words * 8, words * 8);
        // row r of a part is mod q_r (r < L) or p_(r - fullQ) (r >= fullQ); rows
        // between L and fullQ belong to towers outside this context and are kept as they are
        for (uint32_t part = 0; part < h.nparts; ++part)
            for (uint32_t r = 0; r < h.rows; ++r) {
                if (r >= C.L && r < h.fullQ) continue;
                const uint64_t* m = &C.mods[r < C.L ? r : C.L + (r - h.fullQ)];
                const size_t o = ((size_t)part * h.rows + r) * h.n;
                check_residues(av.data() + o, m, 1, h.n, 1);
                check_residues(bv.data() + o, m, 1, h.n, 1);
            }
        auto k = new gf_key;
        k->ctx = c;
        try {
            k->K = upload_key(C, av.data(), bv.data(), h.nparts, h.rows, h.fullQ, s);
        } catch (...) {
            delete k;
            throw;
        }
        return k;
    });
}

}  // extern "C"
