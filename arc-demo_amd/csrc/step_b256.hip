/* the step kernel with 256-thread workgroups (step_kernel.h) */
#include "step_kernel.h"

namespace xpbd {

hipError_t launch_step_b256(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                          int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s)
{
    return launch_step_block<256>(d, m_dev, targets, batch, total, tile, grid_blocks, work, leg, s);
}

int step_resident_blocks_b256(int tile, int leg) { return step_resident_blocks_of<256>(tile, leg); }

} /* namespace xpbd */
