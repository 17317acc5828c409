// Instantiations of the 4-byte chunked-transpose x sweep for 4 substrates,
// rows of 97..200 voxels.
#include "bfvm_xg_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_xg4b(int nmax)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return sweep_xg_kernel<N, 4>;
	switch (nmax)
	{
		BFVM_BUCKET(112)
		BFVM_BUCKET(128)
		BFVM_BUCKET(144)
		BFVM_BUCKET(160)
		BFVM_BUCKET(176)
		BFVM_BUCKET(192)
		BFVM_BUCKET(200)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
