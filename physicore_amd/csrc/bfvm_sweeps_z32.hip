// Instantiations of the z-sweep register kernel, FP32 arithmetic,
// lines of 1..256 voxels in buckets of 16.
#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_cols_z32(int nmax, bool faces)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return faces ? sweep_cols_kernel<2, N, float, true> : sweep_cols_kernel<2, N, float, false>;
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
