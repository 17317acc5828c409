// Instantiations of the pipelined z-sweep (FP64 recurrences), lines of
// 65..128 voxels in buckets of 8.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_z_1(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<2, N, true, 0> : sweep_pipe_kernel<2, N, false, 0>;
	switch (nmax)
	{
		BFVM_BUCKET(72)
		BFVM_BUCKET(80)
		BFVM_BUCKET(88)
		BFVM_BUCKET(96)
		BFVM_BUCKET(104)
		BFVM_BUCKET(112)
		BFVM_BUCKET(120)
		BFVM_BUCKET(128)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
