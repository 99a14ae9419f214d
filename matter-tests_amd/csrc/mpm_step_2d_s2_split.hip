// One translation unit per instantiation of the step, so that the engine's kernels compile in parallel.
#include "mpm_step.hpp"
MPM_INSTANTIATE_STEP(2, 2, float)
