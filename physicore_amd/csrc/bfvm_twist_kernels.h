// bfvm_twist_kernels.h -- y and z sweeps of lines of 257..512 voxels in
// registers, two lanes per line (device template; instantiated by
// bfvm_twist_{y,z}{a,b}.hip).
//
// One lane cannot hold such a line (256 VGPRs), so each line is split at
// m = ceil(n/2) and solved by a twisted factorization: the "top" lane
// eliminates rows 0..m-1 downwards, the "bottom" lane rows n-1..m upwards,
// both at once; the two lanes then meet in a 2x2 solve for x_{m-1}, x_m and
// substitute outwards.  The system is persymmetric (diagonal 1 + dt*decay/dims
// + 2c, less c at both ends; off-diagonals -c), so the pivots of the upward
// elimination are the downward pivots b'_k counted from the bottom end: both
// lanes walk the same factor table, indexed by the distance k from their own
// end.  In exact arithmetic this is the Thomas solution of the reference
// (diffusion_solver.h:11-33); in FP64 it differs from it by rounding only.
//
// Lane l < 32 of a wave holds the top half of column l, lane l + 32 the bottom
// half of the same column.  A half of `len` rows sits in register slots
// pad..NH-1 (pad = NH - len <= 16), so the interface row is always slot NH-1
// (a compile-time index) and the leading slots are padding: loaded as 0,
// factors 0 (the tables carry a 16-entry zero prefix), never stored.
#pragma once

#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

constexpr int kTwistPrefix = 16; // zero entries before each substrate's factor row
// recurrence steps per scheduling group: half the column kernels' group, as
// the per-lane offsets of the two directions cost registers the line needs
constexpr int kTwistChunk = 8;

template <int AXIS, int NH, bool FACES>
__global__ __launch_bounds__(kSweepThreads) void sweep_twist_kernel(const DomainDev* __restrict__ doms,
																	const TileRef* __restrict__ tiles, int ntiles,
																	int finish, int fac_cap)
{
	static_assert(AXIS == 1 || AXIS == 2, "twisted lines are y or z lines");
	static_assert(NH > kTwistPrefix && NH <= kRegMaxN, "half-line bucket");
	constexpr int W = NH + kTwistPrefix; // table row per substrate
	extern __shared__ double lds[];
	double* fb = lds;				// [s][W]: zeros, then b'_k
	double* fc = lds + 2 * fac_cap; // [s] (the second table of the layout stays unused:
									// e_k = c b'_k is formed in registers, the same FP64
									// product as the host's)
	const int lane = threadIdx.x & 63;
	const bool bottom = lane >= 32;
	int cur = -1;
	fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int n = AXIS == 1 ? d.ny : d.nz;
		if (tr.dom != cur)
		{
			__syncthreads();
			const double* gb = d.fb[AXIS];
			for (int k = threadIdx.x; k < W * d.S; k += blockDim.x)
			{
				const int s = k / W, i = k - s * W - kTwistPrefix;
				fb[k] = (i >= 0 && i < n) ? gb[i * d.S + s] : 0.0;
			}
			for (int k = threadIdx.x; k < d.S; k += blockDim.x)
				fc[k] = d.fc[AXIS][k];
			__syncthreads();
			cur = tr.dom;
		}
		const int S = d.S, L = d.L;
		const int outer_n = AXIS == 1 ? d.nz : d.ny;
		// 128 columns per tile: 32 per wave, two lanes per column
		const int f = tr.a + (int)(threadIdx.x >> 6) * 32 + (lane & 31);
		if (f >= outer_n * L)
			continue; // both lanes of the column leave together
		const int outer = f / L, c = f - outer * L;
		const int cs = d.coef_map ? d.coef_map[c] : c % S;
		const int m = (n + 1) / 2;
		const int len = bottom ? n - m : m;
		const int pad = NH - len;
		const int stride = AXIS == 1 ? L : d.ny * L;			   // elements from one row to the next
		const int colbase = AXIS == 1 ? outer * d.ny * L + c : f; // element of row 0
		// slot j holds row (top) j - pad, (bottom) n-1-(j-pad)
		const int row0 = bottom ? n - 1 + pad : -pad;
		const int rstep = bottom ? -1 : 1;
		// byte offsets in unsigned arithmetic: the virtual offsets of padding
		// slots may leave the int range, the offsets of real rows never do
		const uint32_t ostep = (uint32_t)(rstep * stride * (int)sizeof(float));
		const uint32_t first = (uint32_t)((colbase + (bottom ? n - 1 : 0) * stride) * (int)sizeof(float)); // slot pad
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const LineKey kf = line_key(d, 2 * AXIS, (uint32_t)(2 * f + (bottom ? 1 : 0)));
		const LineKey ko = line_key(d, 2 * AXIS + 1, (uint32_t)(2 * f + (bottom ? 1 : 0)));

		float v[NH];
		{
			// byte offset of slot j = first + (j - pad) * ostep, stepped once per slot
			uint32_t o = first - (uint32_t)pad * ostep;
			static_for<NH>([&](auto J) {
				constexpr int j = J;
				if constexpr (j <= kTwistPrefix)
				{
					const bool real = j >= pad;
					const float x = buf_load(rsrc, (int)(real ? o : first), 0);
					v[j] = real ? x : 0.0f;
				}
				else
					v[j] = buf_load(rsrc, (int)o, 0);
				o += ostep;
			});
		}
		__builtin_amdgcn_sched_barrier(0);

		line_faces lf {};
		if constexpr (FACES)
			if (finish && d.face_mask && AXIS == d.dims - 1)
			{
				const int x = c / S, s = c - (c / S) * S;
				lf = AXIS == 1 ? faces_of_line(d, 1, x, 0, outer, s) : faces_of_line(d, 2, x, outer, 0, s);
			}

		// lb[j] = b'_{j-pad} (zero in the padding slots)
		const double* lb = fb + cs * W + kTwistPrefix - pad;
		const double cc = fc[cs];
		constexpr int nchunks = (NH - 1 + kTwistChunk - 1) / kTwistChunk;

		// elimination from this lane's end towards the interface
		double prev = (double)v[0];
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = 1 + C * kTwistChunk + K;
				if constexpr (i < NH)
				{
					prev = (double)v[i] + (cc * lb[i - 1]) * prev;
					v[i] = field_round<double>(prev, kf, weyl29(i));
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});

		// the substitution reads the factors again from LDS: an opaque copy of
		// the pointer keeps the compiler from holding every b' of the
		// elimination in registers until then
		const double* lbs = lb;
		asm volatile("" : "+v"(lbs));

		// interface rows m-1 (top) and m (bottom):
		//   x_own = beta (d' + c x_other),  x_other = beta_o (d'_o + c x_own)
		const double beta = lbs[NH - 1];
		const double beta_o = __shfl_xor(beta, 32);
		const double dp_o = __shfl_xor(prev, 32);
		double nxt = beta * (prev + cc * beta_o * dp_o) / (1.0 - cc * cc * beta * beta_o);

		// substitution towards this lane's end: x_j = (d'_j + c x_{j+1}) b'_j
		uint32_t o = first + (uint32_t)(NH - 1 - pad) * ostep;
		auto out = [&](auto J, double x) {
			constexpr int j = J;
			float val = field_round<double>(x, ko, weyl29(j));
			if constexpr (FACES)
				val = face_value(lf, row0 + j * rstep, n, val);
			if constexpr (j <= kTwistPrefix)
			{
				if (j >= pad)
					buf_store(rsrc, (int)o, 0, val);
			}
			else
				buf_store(rsrc, (int)o, 0, val);
			o -= ostep;
		};
		out(std::integral_constant<int, NH - 1> {}, nxt);
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = NH - 2 - (C * kTwistChunk + K);
				if constexpr (i >= 0)
				{
					nxt = ((double)v[i] + cc * nxt) * lbs[i];
					out(std::integral_constant<int, i> {}, nxt);
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
	}
}

} // namespace sweep
} // namespace bfvm
