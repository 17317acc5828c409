// Instantiations of the twisted two-lane z-sweep, lines of 137..256 voxels
// (half-line buckets of 8).
#include "bfvm_twist_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_twist_z_e(int nmax, int fin)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return fin == kFinBulk	? sweep_twist_kernel<2, N, true, true>  \
			   : fin == kFinFaces ? sweep_twist_kernel<2, N, true, false> \
								  : sweep_twist_kernel<2, N, false, false>;
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
