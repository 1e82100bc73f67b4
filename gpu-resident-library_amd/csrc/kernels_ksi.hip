// The fused row pass of the keyswitch digits' forward transform and the key
// inner product (ksi.hpp), with its launch.
#include "ksi.hpp"
#include "ntt_kernels.hpp"

namespace gf {

// ------------------------------------------- fused digit row pass + inner --
// One wave's LDS region is private to it during the row passes.
struct WaveSync {
    __device__ void operator()() const { DevSync{}.wave(); }
    __device__ void wave() const { DevSync{}.wave(); }
};

// At most 96 VGPRs (5 waves per SIMD; 82 without spills): a 3-wave workgroup
// of the Q rows takes 26 KiB of LDS, so 6 of them fit a CU.
template <uint32_t S2, uint32_t RM>
__global__ __launch_bounds__(KSI_THREADS) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_ksi(const KsiJob* __restrict__ jobs, uint32_t off, uint32_t N1) {
    extern __shared__ uint64_t lds[];
    const KsiJob& j = jobs[off + blockIdx.y];
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    auto go = [&](auto L) {
        ksi_body<S2, RM, decltype(L)::value>(j, N1, blockIdx.x, lds, tid >> 6, nthr >> 6, tid & 63, 64, tid, nthr,
                                             WaveSync{}, DevSync{});
    };
    if (j.red) go(std::true_type{});
    else go(std::false_type{});
}

void launch_ksi(const KsiJob* d, uint32_t nj, const NttGeom& g, uint32_t regions, uint32_t waves, hipStream_t s) {
    if (g.single || g.n % KSI_TILE) throw std::runtime_error("gpufhe: fused inner product needs a two-phase ring");
    if (regions > KSI_MAXP || waves < 1 || waves > KSI_WAVES)
        throw std::runtime_error("gpufhe: fused inner product launch shape out of range");
    const size_t shm = (size_t)regions * KSI_PART_WORDS * 8;
    bool ok = with_radix(g.rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        const bool geom_ok = with_split(g.logN1, g.logN2, [&](auto, auto B) {
            constexpr uint32_t S2 = decltype(B)::value;
            for (uint32_t off = 0; off < nj; off += MAXY) {
                const uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;
                k_ksi<S2, RM><<<dim3(g.n / KSI_TILE, cnt), 64 * waves, shm, s>>>(d, off, 1u << g.logN1);
            }
        });
        if (!geom_ok) throw std::runtime_error("gpufhe: unsupported ring dimension for the NTT");
    });
    if (!ok) throw std::runtime_error("gpufhe: unsupported NTT radix");
}


}  // namespace gf
