// Inverse transforms over device job lists (library plans, paired tables),
// and the tensor product fused with the inverse row pass of its t2.
#include "ntt_kernels.hpp"
#include "tensor.hpp"

namespace gf {

// One workgroup per (job, row tile): the tensor product over the tile's
// NTT_TILE positions (t2 also left in LDS in the row pass's layout), one
// block barrier, then the inverse transform's row pass (phase B) of t2 from
// LDS, so t2 is written once and never read back from HBM for it.
template <uint32_t S2, uint32_t RM>
__global__ __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_tensor_ib(const TensorIbJob* __restrict__ jobs, uint32_t off, NttGeom g) {
    extern __shared__ uint64_t lds[];
    const TensorIbJob& J = jobs[off + blockIdx.y];
    tensor_tile_lds(J.t, blockIdx.x * NTT_TILE, lds, threadIdx.x, NTT_THREADS);
    __syncthreads();
    const NttJob j = J.r;
    if (j.twd) phaseB_inv_f64<S2, RM, DevSync, true>(j, g, blockIdx.x, lds, threadIdx.x, NTT_THREADS, DevSync{});
    else phaseB_body<false, false, true, S2, RM, DevSync, true>(j, g, blockIdx.x, lds, threadIdx.x, NTT_THREADS,
                                                                 DevSync{});
}

void launch_tensor_ib(const TensorIbJob* d, uint32_t nj, const NttGeom& g, hipStream_t s) {
    if (g.single || g.R * (1u << g.logN2) != NTT_TILE)
        throw std::runtime_error("gpufhe: the fused tensor + row pass needs a two-phase ring");
    bool ok = with_radix(g.rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        const bool geom_ok = with_split(g.logN1, g.logN2, [&](auto, auto B) {
            constexpr uint32_t S2 = decltype(B)::value;
            for (uint32_t off = 0; off < nj; off += MAXY) {
                const uint32_t cnt = nj - off < MAXY ? nj - off : MAXY;
                k_tensor_ib<S2, RM><<<dim3(g.n / NTT_TILE, cnt), NTT_THREADS, (size_t)NTT_TILE_PADDED * 8, s>>>(d, off, g);
            }
        });
        if (!geom_ok) throw std::runtime_error("gpufhe: unsupported ring dimension for the NTT");
    });
    if (!ok) throw std::runtime_error("gpufhe: unsupported NTT radix");
}

void launch_ntt_inv(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, hipStream_t s, uint32_t which) {
    for (uint32_t off = 0; off < njobs; off += MAXY) {
        const uint32_t cnt = njobs - off < MAXY ? njobs - off : MAXY;
        ntt_dispatch_dir<false, true>(JobsDev{d_jobs, off}, cnt, g, s, which);
    }
}

}  // namespace gf
