// Instantiations of the two-lane chunked x sweep for 3 substrates, rows of
// 65..400 voxels (half-row buckets).
#include "bfvm_xgt_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_xgt3(int nmax)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return sweep_xgt_kernel<N, 3>;
	switch (nmax)
	{
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
		BFVM_BUCKET(200)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
