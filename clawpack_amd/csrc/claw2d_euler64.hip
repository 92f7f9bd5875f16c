// The FP64 Euler step kernel for the MC limiter at order 2 with transverse_waves = 2 (the bench's
// Euler kernel), in a translation unit of its own so that it can be built with LLVM's
// iterative-minreg machine scheduler, which fits it in 256 VGPRs without the spill the default
// scheduler leaves (clawpack_amd/build.py).  claw2d.hip declares it extern and launches it.
#define CLAW_KERNEL_UNIT 1
#include "claw2d.hip"

namespace claw {
template __global__ void step2_fused<double, EulerP<double>, 4, false, 2, false, 0, false>(
    StepArgs<double, EulerP<double>>);
}  // namespace claw
