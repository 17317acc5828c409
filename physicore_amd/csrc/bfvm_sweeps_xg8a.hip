// Instantiations of the 4-byte chunked-transpose x sweep for 8 substrates,
// rows of 1..96 voxels.
#include "bfvm_xg_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_xg8a(int nmax)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return sweep_xg_kernel<N, 8>;
	switch (nmax)
	{
		BFVM_BUCKET(16)
		BFVM_BUCKET(32)
		BFVM_BUCKET(48)
		BFVM_BUCKET(64)
		BFVM_BUCKET(80)
		BFVM_BUCKET(96)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
