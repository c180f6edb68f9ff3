// placeholder, replaced by the mixed-precision kernels
#include <hip/hip_runtime.h>
#include "kernels.h"
namespace icempm {
namespace mixed {
hipError_t launch_tiles(const StepArgs &, int, int, int, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_p2g_direct(const StepArgs &, int, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_g2p_direct(const StepArgs &, int, int, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_update(const StepArgs &, int, int, int, hipStream_t) { return hipErrorNotSupported; }
size_t tile_lds_bytes(const StepArgs &, int) { return 0; }
}
}
