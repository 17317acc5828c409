// Instantiations of the chunked-transpose x sweep for 2 substrates (rows of even length), rows of
// 97..200 voxels (row buckets: multiples of 8 that are not 24 mod 32).
#include "bfvm_xt_kernels.h"

namespace bfvm {
namespace sweep {

kernel_fn pick_xt2b(int nmax)
{
#define BFVM_BUCKET(N) \
	case N:            \
		return sweep_xt_kernel<N, 2>;
	switch (nmax)
	{
		BFVM_BUCKET(104)
		BFVM_BUCKET(112)
		BFVM_BUCKET(128)
		BFVM_BUCKET(136)
		BFVM_BUCKET(144)
		BFVM_BUCKET(160)
		BFVM_BUCKET(168)
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
