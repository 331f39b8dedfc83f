// The batched kernel with FP16 stencil reuse across RK4 stages, for narrow launches (every
// item's wavefront resident at once).
#include "gr_batch.h"

namespace gr
{

GR_BATCH_LAUNCHER(launch_batch_narrow, 1, true)

} // namespace gr
