// bfvm_sweeps.hip -- sweep launch dispatch: picks the kernel for (axis,
// variant, line bucket, precision, faces) and the LDS it needs.  The register
// kernels are instantiated in bfvm_sweeps_{y,z}{64,32}.hip and bfvm_sweeps_rows.hip
// (separate translation units so they compile in parallel).
#include "bfvm_pipe_kernels.h"
#include "bfvm_twist_kernels.h"
#include "bfvm_xg_kernels.h"
#include "bfvm_xgt_kernels.h"
#include "bfvm_xt_kernels.h"

#include <algorithm>

namespace bfvm {
namespace sweep {
kernel_fn pick_cols_x64a(int nmax, bool faces);
kernel_fn pick_cols_x64b(int nmax, bool faces);
kernel_fn pick_cols_y64a(int nmax, int fin);
kernel_fn pick_cols_y64b(int nmax, int fin);
kernel_fn pick_cols_z64a(int nmax, int fin);
kernel_fn pick_cols_z64b(int nmax, int fin);
kernel_fn pick_cols_x32(int nmax, bool faces);
kernel_fn pick_cols_y32(int nmax, bool faces);
kernel_fn pick_cols_z32(int nmax, bool faces);
kernel_fn pick_rows(int nmax, bool fp64);
kernel_fn pick_xt2a(int nmax);
kernel_fn pick_xt2b(int nmax);
kernel_fn pick_xt4a(int nmax);
kernel_fn pick_xt4b(int nmax);
kernel_fn pick_xt8a(int nmax);
kernel_fn pick_xt8b(int nmax);
kernel_fn pick_xgt1(int nmax);
kernel_fn pick_xgt2(int nmax);
kernel_fn pick_xgt3(int nmax);
kernel_fn pick_xgt5(int nmax);
kernel_fn pick_xgt6(int nmax);
kernel_fn pick_xgt7(int nmax);
kernel_fn pick_xgt4(int nmax);
kernel_fn pick_xgt8(int nmax);
kernel_fn pick_xg1a(int nmax);
kernel_fn pick_xg1b(int nmax);
kernel_fn pick_xg2a(int nmax);
kernel_fn pick_xg2b(int nmax);
kernel_fn pick_xg3a(int nmax);
kernel_fn pick_xg3b(int nmax);
kernel_fn pick_xg5a(int nmax);
kernel_fn pick_xg5b(int nmax);
kernel_fn pick_xg6a(int nmax);
kernel_fn pick_xg6b(int nmax);
kernel_fn pick_xg7a(int nmax);
kernel_fn pick_xg7b(int nmax);
kernel_fn pick_xg4a(int nmax);
kernel_fn pick_xg4b(int nmax);
kernel_fn pick_xg8a(int nmax);
kernel_fn pick_xg8b(int nmax);
kernel_fn pick_twist_y_a(int nmax, bool faces);
kernel_fn pick_twist_y_b(int nmax, bool faces);
kernel_fn pick_twist_y_c(int nmax, bool faces);
kernel_fn pick_twist_y_d(int nmax, bool faces);
kernel_fn pick_twist_z_a(int nmax, int fin);
kernel_fn pick_twist_z_b(int nmax, int fin);
kernel_fn pick_twist_z_c(int nmax, int fin);
kernel_fn pick_twist_z_d(int nmax, int fin);
kernel_fn pick_twist_y_e(int nmax, bool faces);
kernel_fn pick_twist_z_e(int nmax, int fin);
kernel_fn pick_xscan(int S, int vpl, bool fp64);
pipe_fn pick_pipe_x_s1(int nmax, bool faces);
pipe_fn pick_pipe_x_s2(int nmax, bool faces);
pipe_fn pick_pipe_x_s4(int nmax, bool faces);
pipe_fn pick_pipe_x_s8(int nmax, bool faces);

// substrate counts the pipelined x kernel is built for (its LDS reads step by S)
bool pipe_x_supports(int S) { return S == 1 || S == 2 || S == 4 || S == 8; }

namespace {
pipe_fn pick_pipe_x(int nmax, bool faces, int S)
{
	switch (S)
	{
	case 1:
		return pick_pipe_x_s1(nmax, faces);
	case 2:
		return pick_pipe_x_s2(nmax, faces);
	case 4:
		return pick_pipe_x_s4(nmax, faces);
	case 8:
		return pick_pipe_x_s8(nmax, faces);
	default:
		return nullptr;
	}
}
} // namespace
pipe_fn pick_pipe_y_0(int nmax, bool faces);
pipe_fn pick_pipe_y_1(int nmax, bool faces);
pipe_fn pick_pipe_y_2(int nmax, bool faces);
pipe_fn pick_pipe_y_3(int nmax, bool faces);
pipe_fn pick_pipe_y_4(int nmax, bool faces);
pipe_fn pick_pipe_y_5(int nmax, bool faces);
pipe_fn pick_pipe_z_0(int nmax, bool faces);
pipe_fn pick_pipe_z_1(int nmax, bool faces);
pipe_fn pick_pipe_z_2(int nmax, bool faces);
pipe_fn pick_pipe_z_3(int nmax, bool faces);
pipe_fn pick_pipe_z_4(int nmax, bool faces);
pipe_fn pick_pipe_z_5(int nmax, bool faces);

namespace {
pipe_fn pick_pipe(int axis, int nmax, bool faces, int S)
{
	if (axis == 0)
		return pick_pipe_x(nmax, faces, S);
	// instantiation groups: 8..64, 72..128, 136..168, 176..200, 208..224, 232..256
	const int g = nmax <= 64 ? 0 : nmax <= 128 ? 1 : nmax <= 168 ? 2 : nmax <= 200 ? 3 : nmax <= 224 ? 4 : 5;
	static pipe_fn (*const y[6])(int, bool) = { pick_pipe_y_0, pick_pipe_y_1, pick_pipe_y_2,
												pick_pipe_y_3, pick_pipe_y_4, pick_pipe_y_5 };
	static pipe_fn (*const z[6])(int, bool) = { pick_pipe_z_0, pick_pipe_z_1, pick_pipe_z_2,
												pick_pipe_z_3, pick_pipe_z_4, pick_pipe_z_5 };
	return axis == 1 ? y[g](nmax, faces) : z[g](nmax, faces);
}
} // namespace

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

kernel_fn pick(int axis, int variant, int nmax, bool fp64, int fin, int S)
{
	const bool faces = fin != kFinNone;
	// the bulk-finishing stores exist for the FP64 register y/z kernels and
	// the twisted z kernels only
	if (fin == kFinBulk && ((variant != kSweepReg && !(variant == kSweepTwist && axis == 2)) || !fp64 || axis == 0))
		return nullptr;
	if (variant == kSweepScan)
		return axis == 0 ? pick_xscan(S, nmax, fp64) : nullptr;
	if (variant == kSweepLds)
		return fp64 ? pick_lds<double>(axis) : pick_lds<float>(axis);
	if (variant == kSweepRows)
		return axis == 0 ? pick_rows(nmax, fp64) : nullptr;
	if (variant == kSweepXT)
	{
		if (axis != 0 || !fp64 || (S != 2 && S != 4 && S != 8))
			return nullptr;
		const bool lo = nmax <= 96;
		if (S == 2)
			return lo ? pick_xt2a(nmax) : pick_xt2b(nmax);
		return S == 4 ? (lo ? pick_xt4a(nmax) : pick_xt4b(nmax)) : (lo ? pick_xt8a(nmax) : pick_xt8b(nmax));
	}
	if (variant == kSweepXGT)
	{
		if (axis != 0 || !fp64)
			return nullptr;
		switch (S)
		{
		case 1:
			return pick_xgt1(nmax);
		case 2:
			return pick_xgt2(nmax);
		case 3:
			return pick_xgt3(nmax);
		case 4:
			return pick_xgt4(nmax);
		case 5:
			return pick_xgt5(nmax);
		case 6:
			return pick_xgt6(nmax);
		case 7:
			return pick_xgt7(nmax);
		case 8:
			return pick_xgt8(nmax);
		default:
			return nullptr;
		}
	}
	if (variant == kSweepXG)
	{
		if (axis != 0 || !fp64)
			return nullptr;
		const bool lo = nmax <= 96;
		switch (S)
		{
		case 1:
			return lo ? pick_xg1a(nmax) : pick_xg1b(nmax);
		case 2:
			return lo ? pick_xg2a(nmax) : pick_xg2b(nmax);
		case 3:
			return lo ? pick_xg3a(nmax) : pick_xg3b(nmax);
		case 4:
			return lo ? pick_xg4a(nmax) : pick_xg4b(nmax);
		case 5:
			return lo ? pick_xg5a(nmax) : pick_xg5b(nmax);
		case 6:
			return lo ? pick_xg6a(nmax) : pick_xg6b(nmax);
		case 7:
			return lo ? pick_xg7a(nmax) : pick_xg7b(nmax);
		case 8:
			return lo ? pick_xg8a(nmax) : pick_xg8b(nmax);
		default:
			return nullptr;
		}
	}
	if (variant == kSweepTwist)
	{
		if (axis == 0 || !fp64)
			return nullptr;
		// instantiation groups of buckets: 136..160, 168..192, 200..224, 232..256
		// and, for lines of 209..256 voxels, 112..128
		const int g = nmax <= 128 ? 4 : nmax <= 160 ? 0 : nmax <= 192 ? 1 : nmax <= 224 ? 2 : 3;
		static kernel_fn (*const y[5])(int, bool) = { pick_twist_y_a, pick_twist_y_b, pick_twist_y_c, pick_twist_y_d,
													   pick_twist_y_e };
		static kernel_fn (*const z[5])(int, int) = { pick_twist_z_a, pick_twist_z_b, pick_twist_z_c, pick_twist_z_d,
													  pick_twist_z_e };
		// a uniform bulk only on the z lines (the last sweep of a 3-D domain)
		if (fin == kFinBulk && axis == 1)
			return nullptr;
		return axis == 1 ? y[g](nmax, faces) : z[g](nmax, fin);
	}
	// FP64 instantiations split at 128 lines (two translation units per axis)
	const bool lo = nmax <= 128;
	if (axis == 0)
		return fp64 ? (lo ? pick_cols_x64a(nmax, faces) : pick_cols_x64b(nmax, faces)) : pick_cols_x32(nmax, faces);
	if (axis == 1)
		return fp64 ? (lo ? pick_cols_y64a(nmax, fin) : pick_cols_y64b(nmax, fin)) : pick_cols_y32(nmax, faces);
	return fp64 ? (lo ? pick_cols_z64a(nmax, fin) : pick_cols_z64b(nmax, fin)) : pick_cols_z32(nmax, faces);
}
} // namespace
} // namespace sweep

int reg_bucket(int n)
{
	if (n > kRegMaxN)
		return 0;
	return ((n + 15) / 16) * 16;
}

int twist_bucket(int n)
{
	if (n <= kRegMaxN || n > 2 * kRegMaxN)
		return 0;
	return (((n + 1) / 2 + 7) / 8) * 8;
}

int twist_short_bucket(int n)
{
	if (n <= kTwistShortMin || n > kRegMaxN)
		return 0;
	return (((n + 1) / 2 + 7) / 8) * 8;
}

int xt_bucket(int n)
{
	if (n < 1 || n > kXtMaxN)
		return 0;
	const int b = ((n + 7) / 8) * 8;
	return b % 32 == 24 ? b + 8 : b;
}

int xg_bucket(int n, int S)
{
	if (S < 1 || S > 8 || n < 1 || n > xg_max_n(S))
		return 0;
	return n > 192 ? kXgMaxN : ((n + 15) / 16) * 16;
}

int xgt_bucket(int n, int S)
{
	if (S < 1 || S > 8 || n <= xg_max_n(S) || n > 2 * xg_max_n(S))
		return 0;
	const int h = (n + 1) / 2;
	const int nh = h > 192 ? kXgMaxN : ((h + 15) / 16) * 16;
	return nh <= xg_max_n(S) ? nh : 0;
}

int xgt_short_bucket(int n, int S)
{
	if (S < 1 || S > 8 || n < 65 || n > xg_max_n(S))
		return 0;
	return std::max(48, ((n + 1) / 2 + 15) / 16 * 16);
}

size_t sweep_lds_bytes(int axis, int variant, int n, int S, int L, int fac_cap, int threads)
{
	const size_t waves = (size_t)threads / 64;
	const size_t fac = (size_t)(2 * fac_cap + 64) * sizeof(double);
	if (variant == kSweepScan)
		return 0;
	if (axis == 0 && (variant == kSweepRows || variant == kSweepLds))
		return fac + (size_t)std::max(1, x_rows_per_tile(S, L, fac_cap)) * x_row_pitch(L, S) * sizeof(float);
	if (variant == kSweepLds)
		return fac + (size_t)n * kLinesPerTile * sizeof(float);
	if (variant == kSweepTwist)
		return fac + (size_t)sweep::kTwistExchange * sizeof(double);
	if (variant == kSweepXT)
		return fac + waves * sweep::xt_buffer_floats(S) * sizeof(float);
	if (variant == kSweepXG)
		return fac + waves * sweep::xg_buffer_floats(S) * sizeof(float);
	if (variant == kSweepXGT)
		return fac + (size_t)sweep::kXgtExchange * sizeof(double) +
			   (size_t)(kSweepThreads / 64) * sweep::xg_buffer_floats(S) * sizeof(float);
	return fac;
}

hipError_t init_sweep_kernels()
{
	const int max_lds = kMaxLdsBytes;
	for (int axis = 0; axis < 3; axis++)
		for (int nmax = 0; nmax <= kRegMaxN; nmax += 16)
			for (int p = 0; p < 8; p++)
			{
				int variant = nmax == 0 ? kSweepLds : kSweepReg;
				if (p & 4)
				{
					if (axis != 0 || nmax == 0)
						continue;
					variant = kSweepRows;
				}
				sweep::kernel_fn k = sweep::pick(axis, variant, nmax, p & 1, (p & 2) ? kFinFaces : kFinNone, 1);
				if (!k)
					return hipErrorInvalidValue;
				hipError_t e =
					hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
				if (e != hipSuccess)
					return e;
			}
	for (int axis = 1; axis < 3; axis++)
		for (int nmax = 16; nmax <= kRegMaxN; nmax += 16)
		{
			sweep::kernel_fn k = sweep::pick(axis, kSweepReg, nmax, true, kFinBulk, 1);
			if (!k)
				return hipErrorInvalidValue;
			hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
			if (e != hipSuccess)
				return e;
		}
	for (int S = 1; S <= 8; S++)
		for (int n = 65; n <= 2 * xg_max_n(S); n++)
		{
			const int nh = n > xg_max_n(S) ? xgt_bucket(n, S) : xgt_short_bucket(n, S);
			const int prev = n - 1 > xg_max_n(S) ? xgt_bucket(n - 1, S) : xgt_short_bucket(n - 1, S);
			if (nh == 0 || prev == nh)
				continue;
			sweep::kernel_fn k = sweep::pick(0, kSweepXGT, nh, true, kFinNone, S);
			if (!k)
				return hipErrorInvalidValue;
			hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
			if (e != hipSuccess)
				return e;
		}
	for (int S = 1; S <= 8; S++)
		for (int nmax = 16; nmax <= xg_max_n(S); nmax += 8)
		{
			if (xg_bucket(nmax, S) != nmax)
				continue;
			sweep::kernel_fn k = sweep::pick(0, kSweepXG, nmax, true, kFinNone, S);
			if (!k)
				return hipErrorInvalidValue;
			hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
			if (e != hipSuccess)
				return e;
		}
	for (int S : { 2, 4, 8 })
		for (int nmax = 8; nmax <= kXtMaxN; nmax += 8)
		{
			if (xt_bucket(nmax) != nmax)
				continue;
			sweep::kernel_fn k = sweep::pick(0, kSweepXT, nmax, true, kFinNone, S);
			if (!k)
				return hipErrorInvalidValue;
			hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
			if (e != hipSuccess)
				return e;
		}
	for (int axis = 1; axis < 3; axis++)
		for (int nmax = twist_short_bucket(kTwistShortMin + 1); nmax <= kRegMaxN; nmax += 8)
			for (int fin = 0; fin < (axis == 2 ? 3 : 2); fin++)
			{
				sweep::kernel_fn k = sweep::pick(axis, kSweepTwist, nmax, true, fin, 1);
				if (!k)
					return hipErrorInvalidValue;
				hipError_t e =
					hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
				if (e != hipSuccess)
					return e;
			}
	for (int axis = 0; axis < 3; axis++)
		for (int nmax = pipe_bucket(axis, 1); nmax <= kRegMaxN; nmax += pipe_bucket(axis, 1))
			for (int faces = 0; faces < 2; faces++)
				for (int S : { 1, 2, 4, 8 })
			{
				if (axis != 0 && S != 1)
					continue;
				sweep::pipe_fn k = sweep::pick_pipe(axis, nmax, faces, S);
				if (!k)
					return hipErrorInvalidValue;
				hipError_t e =
					hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
				if (e != hipSuccess)
					return e;
			}
	return hipSuccess;
}

int pipe_bucket(int axis, int n)
{
	const int q = axis == 0 ? 32 : 8;
	return n > kRegMaxN ? 0 : ((n + q - 1) / q) * q;
}

PipeGeometry pipe_geometry(int axis, int n, int nmax, int S, int L, int64_t lines_in_tiles)
{
	PipeGeometry g;
	g.S = S;
	if (nmax <= 0 || L % 4 != 0 || S > 64 || lines_in_tiles <= 0 || (axis == 0 && !sweep::pipe_x_supports(S)))
		return g;
	const size_t fac = (size_t)(fac_row(nmax) * S + 64) * sizeof(double);
	if (axis == 0)
	{
		g.x_rows = 64 / S;
		g.buf_floats = g.x_rows * pipe_x_pitch(L, S);
	}
	else
		g.buf_floats = ((n + 3) / 4) * 4 * 64;
	const size_t buf = (size_t)g.buf_floats * sizeof(float);
	if (fac + buf > (size_t)kMaxLdsBytes)
		return PipeGeometry {};
	// up to 4 waves; no more than there are tiles
	const int64_t tiles = (lines_in_tiles + 63) / 64;
	g.waves = (int)std::min<int64_t>({ (int64_t)((kMaxLdsBytes - fac) / buf), 4, tiles });
	g.lds_bytes = fac + (size_t)g.waves * buf;
	return g;
}

hipError_t launch_sweep_pipe(int axis, int nmax, bool faces, const DomainDev* doms, const TileRef* tiles, int ntiles,
							 int finish, int fac_cap, const PipeGeometry& g, int grid, hipStream_t stream)
{
	if (ntiles <= 0)
		return hipSuccess;
	sweep::pipe_fn k = sweep::pick_pipe(axis, nmax, faces, g.S);
	if (!k || g.waves <= 0)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(k, dim3(grid), dim3(64 * g.waves), g.lds_bytes, stream, doms, tiles, ntiles, finish, fac_cap,
					   g.buf_floats);
	return hipGetLastError();
}

hipError_t launch_sweep(int axis, int variant, int nmax, bool fp64, int fin, int S, const DomainDev* doms,
						const TileRef* tiles, int ntiles, int finish, int fac_cap, size_t lds_bytes, int grid,
						hipStream_t stream, int threads)
{
	if (ntiles <= 0)
		return hipSuccess;
	if (grid <= 0 || grid > ntiles)
		grid = ntiles;
	sweep::kernel_fn k = sweep::pick(axis, variant, nmax, fp64, fin, S);
	if (!k)
		return hipErrorInvalidValue;
	if (threads != kSweepThreads && (threads % 64 || threads < 64 || threads > kSweepThreads ||
									 sweep_threads(variant, threads) != threads))
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(k, dim3(grid), dim3(threads), lds_bytes, stream, doms, tiles, ntiles, finish, fac_cap);
	return hipGetLastError();
}

} // namespace bfvm
