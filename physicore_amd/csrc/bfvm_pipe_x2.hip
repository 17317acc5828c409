// Instantiations of the pipelined x-sweep (FP64 recurrences) for S = 2.
#include "bfvm_pipe_kernels.h"

namespace bfvm {
namespace sweep {

pipe_fn pick_pipe_x_s2(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_pipe_kernel<0, N, true, 2> : sweep_pipe_kernel<0, N, false, 2>;
	switch (nmax)
	{
		BFVM_BUCKET(32)
		BFVM_BUCKET(64)
		BFVM_BUCKET(96)
		BFVM_BUCKET(128)
		BFVM_BUCKET(160)
		BFVM_BUCKET(192)
		BFVM_BUCKET(224)
		BFVM_BUCKET(256)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
