// Instantiations of the twisted two-lane y-sweep, lines of 257..320 voxels
// (half-line buckets of 8).
#include "bfvm_twist_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_twist_y_a(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_twist_kernel<1, N, true> : sweep_twist_kernel<1, N, false>;
	switch (nmax)
	{
		BFVM_BUCKET(136)
		BFVM_BUCKET(144)
		BFVM_BUCKET(152)
		BFVM_BUCKET(160)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
