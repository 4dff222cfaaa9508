// The stage-fused BRUSS2D step (rkx_sweep_kernel.hpp) for DOPRI87, in a
// translation unit of its own: 13 stages of register ring take minutes to
// build, which then overlaps the build of the other tables' kernels.
#include "rkx_sweep_kernel.hpp"

namespace rkx {

hipError_t launch_sweep_dopri87(int K, bool gen, const sweep_inst *insts, const int2 *tiles,
                                int ntiles, hipStream_t stream)
{
  if (ntiles <= 0)
    return hipSuccess;
  return launch_pat<pat_dopri87>(K, gen, insts, tiles, ntiles, stream);
}

int resident_blocks_dopri87(int K) { return blocks_per_cu_k<pat_dopri87>(K); }

}  // namespace rkx
