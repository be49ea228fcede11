// Split-precision MFMA gridder / degridder (not built yet).
#include "device_common.h"

namespace p9idg {
hipError_t launch_grid_mfma(const KernelArgs&, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_degrid_mfma(const KernelArgs&, int, int, hipStream_t) { return hipErrorNotSupported; }
}  // namespace p9idg
