// Instantiations of the x-sweep scan kernel: S = 1..8 substrates, VPL voxels
// per lane with S*VPL <= 32, both precisions.
#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

namespace {
template <int S, typename ACC>
kernel_fn pick_vpl(int vpl)
{
	if constexpr (S * 1 <= 32)
		if (vpl == 1)
			return sweep_xscan_kernel<S, 1, ACC>;
	if constexpr (S * 2 <= 32)
		if (vpl == 2)
			return sweep_xscan_kernel<S, 2, ACC>;
	if constexpr (S * 4 <= 32)
		if (vpl == 4)
			return sweep_xscan_kernel<S, 4, ACC>;
	if constexpr (S * 8 <= 32)
		if (vpl == 8)
			return sweep_xscan_kernel<S, 8, ACC>;
	return nullptr;
}

template <typename ACC>
kernel_fn pick_s(int S, int vpl)
{
	switch (S)
	{
	case 1: return pick_vpl<1, ACC>(vpl);
	case 2: return pick_vpl<2, ACC>(vpl);
	case 3: return pick_vpl<3, ACC>(vpl);
	case 4: return pick_vpl<4, ACC>(vpl);
	case 5: return pick_vpl<5, ACC>(vpl);
	case 6: return pick_vpl<6, ACC>(vpl);
	case 7: return pick_vpl<7, ACC>(vpl);
	case 8: return pick_vpl<8, ACC>(vpl);
	default: return nullptr;
	}
}
} // namespace

kernel_fn pick_xscan(int S, int vpl, bool fp64)
{
	return fp64 ? pick_s<double>(S, vpl) : pick_s<float>(S, vpl);
}

} // namespace sweep
} // namespace bfvm
