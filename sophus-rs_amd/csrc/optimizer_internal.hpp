// Access to an optimizer handle's device state for other translation units of the
// library (the preconditioner trainer reads the RCS the optimizer assembles).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sophus_gpu.h"
#include "pcg.hpp"

namespace sgpu {

// The iterative-solver state (valid for the PCG solvers; S is in .bsr after a step).
const DevPcg& optimizer_pcg(const sgpu_optimizer* opt);
hipStream_t optimizer_stream(const sgpu_optimizer* opt);
// Number of lower 6x6 blocks of the reduced camera system.
int64_t optimizer_nslot(const sgpu_optimizer* opt);

}  // namespace sgpu
