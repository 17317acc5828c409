// Instantiations of the pipelined z-sweep (FP64 recurrences), lines of
// 225..256 voxels in buckets of 8.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_z_5(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<2, N, true, 0> : sweep_pipe_kernel<2, N, false, 0>;
	switch (nmax)
	{
		BFVM_BUCKET(232)
		BFVM_BUCKET(240)
		BFVM_BUCKET(248)
		BFVM_BUCKET(256)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
