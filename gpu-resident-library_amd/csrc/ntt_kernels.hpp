// NTT kernel templates shared by the translation units that instantiate them
// (kernels_ntt_f.hip, kernels_ntt_i.hip, kernels_ntt_ref.hip, kernels_ksi.hip):
// the transform phases over job lists, and the launch dispatch over the ring
// split and register radix.  Each unit instantiates one direction / table
// layout, so the library's kernels compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace gf {

// ------------------------------------------------------------------ NTT ---
struct DevSync {
    __device__ void operator()() const { __syncthreads(); }
    // LDS words written by this wave become visible to this wave's lanes
    __device__ void wave() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};
struct JobsDev {
    const NttJob* p;
    uint32_t off;
    __device__ NttJob get(uint32_t i) const { return p[off + i]; }
};
struct JobsOne {
    NttJob j;
    __device__ NttJob get(uint32_t) const { return j; }
};

// Forward jobs with j.red take the butterflies without conditional
// subtractions (LZ); the choice is uniform per workgroup (one job each).
template <bool FWD, class F>
__device__ __forceinline__ void with_lz(const NttJob& j, F&& f) {
    if constexpr (FWD) {
        if (j.red) f(std::true_type{});
        else f(std::false_type{});
    } else {
        f(std::false_type{});
    }
}

template <bool FWD, bool PAIRED, uint32_t S1, uint32_t RM, class J>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_A(J jobs, NttGeom g) {
    extern __shared__ uint64_t lds[];
    const NttJob j = jobs.get(blockIdx.y);
    with_lz<FWD>(j, [&](auto L) {
        phaseA_body<FWD, decltype(L)::value, PAIRED, S1, RM>(j, g, blockIdx.x, lds, threadIdx.x, blockDim.x,
                                                             DevSync{});
    });
}
template <bool FWD, bool PAIRED, uint32_t S2, uint32_t RM, class J>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_B(J jobs, NttGeom g) {
    extern __shared__ uint64_t lds[];
    const NttJob j = jobs.get(blockIdx.y);
    if constexpr (!FWD && PAIRED) {
        // inverse jobs that carry fp64 tables (moduli below 2^50)
        if (j.twd) {
            phaseB_inv_f64<S2, RM>(j, g, blockIdx.x, lds, threadIdx.x, blockDim.x, DevSync{});
            return;
        }
    }
    with_lz<FWD>(j, [&](auto L) {
        phaseB_body<FWD, decltype(L)::value, PAIRED, S2, RM>(j, g, blockIdx.x, lds, threadIdx.x, blockDim.x,
                                                             DevSync{});
    });
}
template <bool FWD, bool PAIRED, uint32_t S, uint32_t RM, class J>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_S(J jobs) {
    extern __shared__ uint64_t lds[];
    const NttJob j = jobs.get(blockIdx.y);
    with_lz<FWD>(j, [&](auto L) {
        single_body<FWD, decltype(L)::value, PAIRED, S, RM>(j, lds, threadIdx.x, blockDim.x, DevSync{});
    });
}

// PAIRED: library plans use interleaved (w, Shoup precon) tables; the
// reference ABI launchers take the two arrays separately.  g.rmax selects the
// register radix (the tuning knob the autotuner sets).
template <bool FWD, bool PAIRED, class J>
inline void ntt_dispatch_dir(J jobs, uint32_t njobs, const NttGeom& g, hipStream_t s, uint32_t which) {
    const dim3 thr(NTT_THREADS);
    bool ok = with_radix(g.rmax, [&](auto Rc) {
        constexpr uint32_t RM = decltype(Rc)::value;
        bool geom_ok;
        if (g.single) {
            geom_ok = with_single(g.logn, [&](auto Sc) {
                constexpr uint32_t S = decltype(Sc)::value;
                if (which & 1)
                    k_ntt_S<FWD, PAIRED, S, RM><<<dim3(1, njobs), thr, (size_t)padded(1u << S) * 8, s>>>(jobs);
            });
        } else {
            geom_ok = with_split(g.logN1, g.logN2, [&](auto A, auto B) {
                constexpr uint32_t S1 = decltype(A)::value, S2 = decltype(B)::value;
                const dim3 gA((1u << S2) / g.C, njobs), gB((1u << S1) / g.R, njobs);
                const size_t shA = (size_t)NTT_TILE * 8, shB = (size_t)NTT_TILE_PADDED * 8;
                // which: bit 0 = first kernel of the transform, bit 1 = second
                if (FWD) {
                    if (which & 1) k_ntt_A<FWD, PAIRED, S1, RM><<<gA, thr, shA, s>>>(jobs, g);
                    if (which & 2) k_ntt_B<FWD, PAIRED, S2, RM><<<gB, thr, shB, s>>>(jobs, g);
                } else {
                    if (which & 1) k_ntt_B<FWD, PAIRED, S2, RM><<<gB, thr, shB, s>>>(jobs, g);
                    if (which & 2) k_ntt_A<FWD, PAIRED, S1, RM><<<gA, thr, shA, s>>>(jobs, g);
                }
            });
        }
        if (!geom_ok) throw std::runtime_error("gpufhe: unsupported ring dimension for the NTT");
    });
    if (!ok) throw std::runtime_error("gpufhe: unsupported NTT radix");
}

template <bool PAIRED, class J>
inline void ntt_dispatch(J jobs, uint32_t njobs, const NttGeom& g, bool fwd, hipStream_t s,
                         uint32_t which = 3) {
    if (fwd) ntt_dispatch_dir<true, PAIRED>(jobs, njobs, g, s, which);
    else ntt_dispatch_dir<false, PAIRED>(jobs, njobs, g, s, which);
}

static const uint32_t MAXY = 65535;

}  // namespace gf
