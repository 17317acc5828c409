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
