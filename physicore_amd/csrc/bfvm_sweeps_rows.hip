// Instantiations of the x-sweep register kernel (both precisions).
#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_rows(int nmax, bool fp64)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return fp64 ? sweep_rows_kernel<N, double> : sweep_rows_kernel<N, float>;
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
