// The FP32 Euler step kernel for the MC limiter at order 2 with transverse_waves = 2 (the bench's
// FP32 Euler kernel), in a translation unit of its own so that it can be built with LLVM's max-ilp
// machine scheduler (clawpack_amd/build.py).  claw2d.hip declares it extern and launches it.
#define CLAW_KERNEL_UNIT 1
#include "claw2d.hip"

namespace claw {
template __global__ void step2_fused<float, EulerP<float>, 4, false, 2, false, 0, false>(
    StepArgs<float, EulerP<float>>);
}  // namespace claw
