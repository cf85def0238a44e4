/* the step kernel with 64-thread workgroups (step_kernel.h) */
#include "step_kernel.h"

namespace xpbd {

hipError_t launch_step_b64(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                          int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s)
{
    return launch_step_block<64>(d, m_dev, targets, batch, total, tile, grid_blocks, work, leg, s);
}

int step_resident_blocks_b64(int tile, int leg) { return step_resident_blocks_of<64>(tile, leg); }

} /* namespace xpbd */
