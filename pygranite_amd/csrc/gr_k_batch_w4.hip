// The batched kernel held to at most 128 VGPRs (four wavefronts per SIMD).
#include "gr_batch.h"

namespace gr
{

GR_BATCH_LAUNCHER(launch_batch_w4, 4, false)
GR_BATCH_OCCUPANCY(occupancy_batch_w4, 4)

} // namespace gr
