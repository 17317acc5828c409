// bfvm_sweeps.hip -- sweep launch dispatch: picks the kernel for (axis,
// variant, line bucket, precision, faces) and the LDS it needs.  The register
// kernels are instantiated in bfvm_sweeps_{y,z}{64,32}.hip and bfvm_sweeps_rows.hip
// (separate translation units so they compile in parallel).
#include "bfvm_sweep_kernels.h"

#include <algorithm>

namespace bfvm {
namespace sweep {
kernel_fn pick_cols_x64(int nmax, bool faces);
kernel_fn pick_cols_x32(int nmax, bool faces);
kernel_fn pick_cols_y64(int nmax, bool faces);
kernel_fn pick_cols_z64(int nmax, bool faces);
kernel_fn pick_cols_y32(int nmax, bool faces);
kernel_fn pick_cols_z32(int nmax, bool faces);
kernel_fn pick_rows(int nmax, bool fp64);
kernel_fn pick_xscan(int S, int vpl, bool fp64);

namespace {
template <typename ACC>
kernel_fn pick_lds(int axis)
{
	if (axis == 0)
		return sweep_lds_kernel<0, ACC>;
	if (axis == 1)
		return sweep_lds_kernel<1, ACC>;
	return sweep_lds_kernel<2, ACC>;
}

kernel_fn pick(int axis, int variant, int nmax, bool fp64, bool faces, int S)
{
	if (variant == kSweepScan)
		return axis == 0 ? pick_xscan(S, nmax, fp64) : nullptr;
	if (variant == kSweepLds)
		return fp64 ? pick_lds<double>(axis) : pick_lds<float>(axis);
	if (variant == kSweepRows)
		return axis == 0 ? pick_rows(nmax, fp64) : nullptr;
	if (axis == 0)
		return fp64 ? pick_cols_x64(nmax, faces) : pick_cols_x32(nmax, faces);
	if (axis == 1)
		return fp64 ? pick_cols_y64(nmax, faces) : pick_cols_y32(nmax, faces);
	return fp64 ? pick_cols_z64(nmax, faces) : pick_cols_z32(nmax, faces);
}
} // namespace
} // namespace sweep

int reg_bucket(int n)
{
	if (n > kRegMaxN)
		return 0;
	return ((n + 31) / 32) * 32;
}

size_t sweep_lds_bytes(int axis, int variant, int n, int S, int L, int fac_cap)
{
	const size_t fac = (size_t)(2 * fac_cap + 64) * sizeof(double);
	if (variant == kSweepScan)
		return 0;
	if (axis == 0 && (variant == kSweepRows || variant == kSweepLds))
		return fac + (size_t)std::max(1, x_rows_per_tile(S, L, fac_cap)) * x_row_pitch(L, S) * sizeof(float);
	if (variant == kSweepLds)
		return fac + (size_t)n * kLinesPerTile * sizeof(float);
	return fac;
}

hipError_t init_sweep_kernels()
{
	const int max_lds = kMaxLdsBytes;
	for (int axis = 0; axis < 3; axis++)
		for (int nmax = 0; nmax <= kRegMaxN; nmax += 32)
			for (int p = 0; p < 8; p++)
			{
				int variant = nmax == 0 ? kSweepLds : kSweepReg;
				if (p & 4)
				{
					if (axis != 0 || nmax == 0)
						continue;
					variant = kSweepRows;
				}
				sweep::kernel_fn k = sweep::pick(axis, variant, nmax, p & 1, p & 2, 1);
				if (!k)
					return hipErrorInvalidValue;
				hipError_t e =
					hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
				if (e != hipSuccess)
					return e;
			}
	return hipSuccess;
}

hipError_t launch_sweep(int axis, int variant, int nmax, bool fp64, bool faces, int S, const DomainDev* doms,
						const TileRef* tiles, int ntiles, int finish, int fac_cap, size_t lds_bytes, int grid,
						hipStream_t stream)
{
	if (ntiles <= 0)
		return hipSuccess;
	if (grid <= 0 || grid > ntiles)
		grid = ntiles;
	sweep::kernel_fn k = sweep::pick(axis, variant, nmax, fp64, faces, S);
	if (!k)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(k, dim3(grid), dim3(kSweepThreads), lds_bytes, stream, doms, tiles, ntiles, finish, fac_cap);
	return hipGetLastError();
}

} // namespace bfvm
