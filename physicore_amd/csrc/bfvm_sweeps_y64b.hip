// Instantiations of the y-sweep register kernel, FP64 arithmetic,
// lines of 129..256 voxels in buckets of 16.
#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_cols_y64b(int nmax, int fin)
{
#define BFVM_BUCKET(N)                                                           \
	case N:                                                                      \
		return fin == kFinBulk	  ? sweep_cols_kernel<1, N, double, kFinBulk>  \
			   : fin == kFinFaces ? sweep_cols_kernel<1, N, double, kFinFaces> \
								  : sweep_cols_kernel<1, N, double, kFinNone>;
	switch (nmax)
	{
		BFVM_BUCKET(144)
		BFVM_BUCKET(160)
		BFVM_BUCKET(176)
		BFVM_BUCKET(192)
		BFVM_BUCKET(208)
		BFVM_BUCKET(224)
		BFVM_BUCKET(240)
		BFVM_BUCKET(256)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
