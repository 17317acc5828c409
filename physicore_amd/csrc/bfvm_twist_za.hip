// Instantiations of the twisted two-lane z-sweep, lines of 257..320 voxels
// (half-line buckets of 8).
#include "bfvm_twist_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_twist_z_a(int nmax, int fin)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return fin == kFinBulk	? sweep_twist_kernel<2, N, true, true>  \
			   : fin == kFinFaces ? sweep_twist_kernel<2, N, true, false> \
								  : sweep_twist_kernel<2, N, false, false>;
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
