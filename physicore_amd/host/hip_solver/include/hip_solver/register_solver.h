// Registration of the HIP backend under the name "hip_solver", the counterpart of
// reactions-diffusion/biofvm/kernels/openmp_solver/include/openmp_solver/register_solver.h:3-7.
#pragma once

namespace physicore::biofvm::kernels::hip_solver {

void attach_to_registry();

}
