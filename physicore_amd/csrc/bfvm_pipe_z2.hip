// Instantiations of the pipelined z-sweep (FP64 recurrences), lines of
// 129..168 voxels in buckets of 8.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_z_2(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<2, N, true, 0> : sweep_pipe_kernel<2, N, false, 0>;
	switch (nmax)
	{
		BFVM_BUCKET(136)
		BFVM_BUCKET(144)
		BFVM_BUCKET(152)
		BFVM_BUCKET(160)
		BFVM_BUCKET(168)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
