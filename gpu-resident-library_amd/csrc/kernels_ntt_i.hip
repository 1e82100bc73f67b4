// Inverse transforms over device job lists (library plans, paired tables).
#include "ntt_kernels.hpp"

namespace gf {

void launch_ntt_inv(const NttJob* d_jobs, uint32_t njobs, const NttGeom& g, hipStream_t s, uint32_t which) {
    for (uint32_t off = 0; off < njobs; off += MAXY) {
        const uint32_t cnt = njobs - off < MAXY ? njobs - off : MAXY;
        ntt_dispatch_dir<false, true>(JobsDev{d_jobs, off}, cnt, g, s, which);
    }
}

}  // namespace gf
