// Single transforms of the reference ABI (LaunchNTT_CT / LaunchINTT_GS: twiddles
// and Shoup precons as separate arrays), and the library's job-list entry point.
#include "ntt_kernels.hpp"

namespace gf {

void launch_ntt_fwd(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, hipStream_t s, uint32_t which);
void launch_ntt_inv(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, hipStream_t s, uint32_t which);

void launch_ntt(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, bool fwd, hipStream_t s, uint32_t which) {
    if (fwd) launch_ntt_fwd(d_jobs, njobs, g, s, which);
    else launch_ntt_inv(d_jobs, njobs, g, s, which);
}
void launch_ntt_one(const NttJob& job, const NttGeom& g, bool fwd, hipStream_t s) {
    ntt_dispatch<false>(JobsOne{job}, 1, g, fwd, s);
}

}  // namespace gf
