// Instantiations of the y-sweep register kernel, FP64 arithmetic,
// lines of 1..128 voxels in buckets of 16.
#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_cols_y64a(int nmax, int fin)
{
#define BFVM_BUCKET(N)                                                           \
	case N:                                                                      \
		return fin == kFinBulk	  ? sweep_cols_kernel<1, N, double, kFinBulk>  \
			   : fin == kFinFaces ? sweep_cols_kernel<1, N, double, kFinFaces> \
								  : sweep_cols_kernel<1, N, double, kFinNone>;
	switch (nmax)
	{
		BFVM_BUCKET(16)
		BFVM_BUCKET(32)
		BFVM_BUCKET(48)
		BFVM_BUCKET(64)
		BFVM_BUCKET(80)
		BFVM_BUCKET(96)
		BFVM_BUCKET(112)
		BFVM_BUCKET(128)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
