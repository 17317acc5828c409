// Instantiations of the twisted two-lane z-sweep, lines of 449..512 voxels
// (half-line buckets of 8).
#include "bfvm_twist_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_twist_z_d(int nmax, int fin)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return fin == kFinBulk	? sweep_twist_kernel<2, N, true, true>  \
			   : fin == kFinFaces ? sweep_twist_kernel<2, N, true, false> \
								  : sweep_twist_kernel<2, N, false, false>;
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
