// sg_tiled_bf16.hip -- tiled sweeps with bf16 storage and float arithmetic.
#include "sg_tiled.hpp"

namespace sg {

hipError_t sweep_bf16(int which, int ty, int ny, int dm, int comp, const void *shots, const WorkItem *items, int nitems,
                     int grid, hipStream_t st)
{
    return sweep_t<bf16, float>(which, ty, ny, dm, comp != 0, shots, items, nitems, grid, st);
}

hipError_t sweep_fused_bf16(int which, int comp, int waves, const void *shots, const WorkItem *items, const int *off,
                         int grid, hipStream_t st)
{
    return sweep_fused_t<bf16, float>(which, comp != 0, waves, shots, items, off, grid, st);
}

}  // namespace sg
