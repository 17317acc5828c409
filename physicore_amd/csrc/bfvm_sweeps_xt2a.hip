// Instantiations of the chunked-transpose x sweep for 2 substrates (rows of even length), rows of
// 1..96 voxels (row buckets: multiples of 8 that are not 24 mod 32).
#include "bfvm_xt_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_xt2a(int nmax)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return sweep_xt_kernel<N, 2>;
	switch (nmax)
	{
		BFVM_BUCKET(8)
		BFVM_BUCKET(16)
		BFVM_BUCKET(32)
		BFVM_BUCKET(40)
		BFVM_BUCKET(48)
		BFVM_BUCKET(64)
		BFVM_BUCKET(72)
		BFVM_BUCKET(80)
		BFVM_BUCKET(96)
	default:
		return nullptr;
	}
#undef BFVM_BUCKET
}

} // namespace sweep
} // namespace bfvm
