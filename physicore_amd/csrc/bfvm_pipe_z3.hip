// Instantiations of the pipelined z-sweep (FP64 recurrences), lines of
// 169..200 voxels in buckets of 8.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_z_3(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<2, N, true, 0> : sweep_pipe_kernel<2, N, false, 0>;
	switch (nmax)
	{
		BFVM_BUCKET(176)
		BFVM_BUCKET(184)
		BFVM_BUCKET(192)
		BFVM_BUCKET(200)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
