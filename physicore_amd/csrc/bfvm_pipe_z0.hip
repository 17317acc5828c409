// Instantiations of the pipelined z-sweep (FP64 recurrences), lines of
// 1..64 voxels in buckets of 8.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_z_0(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<2, N, true, 0> : sweep_pipe_kernel<2, N, false, 0>;
	switch (nmax)
	{
		BFVM_BUCKET(8)
		BFVM_BUCKET(16)
		BFVM_BUCKET(24)
		BFVM_BUCKET(32)
		BFVM_BUCKET(40)
		BFVM_BUCKET(48)
		BFVM_BUCKET(56)
		BFVM_BUCKET(64)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
