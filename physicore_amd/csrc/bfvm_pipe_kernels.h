// bfvm_pipe_kernels.h -- pipelined register sweeps: while a wave solves the
// 64 lines of one tile in registers, the next tile streams from HBM into the
// wave's own LDS buffer by buffer-to-LDS DMA (buffer_load ... lds), so memory
// traffic continues through the recurrences instead of stopping for them.
//
// Workgroup: W waves (W = blockDim.x / 64), each looping over tiles
// t = (blockIdx.x + k * gridDim.x) * W + wave.  LDS: the b' factor table of the
// current domain (shared, FP64, [s][i] with a row pitch of NMAX, zero past n)
// and the c values, then one tile buffer per wave.  The forward multipliers
// are recomputed as e[i-1] = c * b'[i-1], the same FP64 product the factor
// precompute forms (openmp_solver/src/diffusion_solver.cpp:29-86), so only b'
// takes LDS.  The host pads each domain's tile list to a multiple of W (tiles
// with a < 0 do no work) so that the W tiles of a round share one domain and
// the factor reload at a domain change is a workgroup barrier.
//
// Tiles:
//  y (AXIS 1) / z (AXIS 2): 64 consecutive columns of the flattened
//    (outer, x*S + s) space from tr.a; the buffer holds element i of column j
//    at [i * 64 + j], filled by 1 KiB DMA loads of 4 elements x 64 columns
//    (needs L % 4 == 0 so that 4 consecutive columns are contiguous).
//  x (AXIS 0): tr.b rows from row tr.a (contiguous in HBM), staged with the
//    row pitch pipe_x_pitch(L, S) so that the 64 lanes (row, substrate) read
//    their strided lines from distinct banks; DMA chunks of 256 floats are
//    split at row boundaries (needs L % 4 == 0).
// Buffers hold the real line length rounded up to 4; the register slots past
// it (at most the last 8 of a y/z bucket, 32 of an x bucket) read element n-1
// again, and the zero b' there decouples them.  Line lengths are bucketed by 8
// for y/z: register slots beyond 256 VGPRs would spill to AGPRs, which costs
// more than the memory time the pipeline hides.
#pragma once

#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

typedef __attribute__((address_space(3))) void* lds_void_t;
// recurrence steps per factor prefetch group (the next group's b' values are
// read from LDS while the current group computes)
constexpr int kPipeChunk = 8;
using pipe_fn = void (*)(const DomainDev*, const TileRef*, int, int, int, int);

__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, float* lds, int voff, int soff)
{
	__builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void_t)lds, 16, voff, soff, 0, 0);
}

// y/z: byte offset of flattened column f at element 0 (f clamped to the last
// column), and the line stride in bytes
template <int AXIS>
__device__ __forceinline__ int pipe_col_off(const DomainDev& d, int f, int total)
{
	f = min(f, total - 1);
	if constexpr (AXIS == 2)
		return f * (int)sizeof(float);
	const int o = f / d.L, c = f - o * d.L;
	return (o * d.ny * d.L + c) * (int)sizeof(float);
}

template <int AXIS>
__device__ __forceinline__ int pipe_total(const DomainDev& d)
{
	return AXIS == 0 ? d.ny * d.nz * d.S : (AXIS == 1 ? d.nz * d.L : d.ny * d.L);
}

// Issues the DMA loads of tile tr into buf (this wave's buffer).
template <int AXIS, int NMAX>
__device__ __forceinline__ void pipe_prefetch(__amdgpu_buffer_rsrc_t r, float* buf, const DomainDev& d,
											  const TileRef& tr, int lane)
{
	if constexpr (AXIS == 0)
	{
		const int L = d.L;
		const int pitch = pipe_x_pitch(L, d.S);
		const int total = tr.b * L; // floats, contiguous from row tr.a
		const int base4 = tr.a * L * (int)sizeof(float);
		int row = 0, row_start = 0; // the row holding float q0 of the tile
		for (int q0 = 0; q0 < total; q0 += 256)
		{
			while (q0 >= row_start + L)
			{
				row++;
				row_start += L;
			}
			// lanes of this chunk hold floats q0 + 4*lane .. +3; split at row ends
			const int my = 4 * lane;
			for (int rr = row, rs = row_start; rs < q0 + 256 && rr < tr.b; rr++, rs += L)
			{
				// lane-linear LDS destination: lane l lands at dst + 4*l floats
				float* dst = buf + rr * pitch + (q0 - rs);
				if (q0 + my >= rs && q0 + my < rs + L)
					dma16(r, dst, base4 + (q0 + my) * (int)sizeof(float), 0);
			}
		}
	}
	else
	{
		const int total = pipe_total<AXIS>(d);
		const int stride4 = (AXIS == 1 ? d.L : d.ny * d.L) * (int)sizeof(float);
		const int sub = lane >> 4;
		// groups of 4 columns beyond the end load a valid group instead; the
		// lanes of those columns read the last real column's slot
		const int g = min(tr.a + (lane & 15) * 4, total - 4);
		const int off = pipe_col_off<AXIS>(d, g, total);
		const int lane_off = off + sub * stride4;
		const int n = AXIS == 1 ? d.ny : d.nz;
		const int full = n / 4;
		for (int q = 0; q < full; q++)
			dma16(r, buf + q * 256, lane_off, 4 * q * stride4);
		// the last partial group repeats element n-1
		for (int q = full; q < (n + 3) / 4; q++)
			dma16(r, buf + q * 256, off + min(4 * q + sub, n - 1) * stride4, 0);
	}
}

// SX: the substrate count as a compile-time constant for the x kernel (its
// LDS reads step by S floats and need immediate offsets); 0 for y/z.
template <int AXIS, int NMAX, bool FACES, int SX>
__global__ __launch_bounds__(kSweepThreads) void sweep_pipe_kernel(const DomainDev* __restrict__ doms,
																   const TileRef* __restrict__ tiles, int ntiles,
																   int finish, int fac_cap, int buf_floats)
{
	extern __shared__ double lds[];
	double* fb = lds;
	double* fc = lds + fac_cap;
	const int W = blockDim.x >> 6;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	float* buf = reinterpret_cast<float*>(lds + fac_cap + 64) + wave * buf_floats;
	const int step = gridDim.x * W;
	int first = blockIdx.x * W;
	if (first >= ntiles)
		return;
	fp32_toward_zero();

	// stale buffer contents only ever feed padding steps (b' = 0 there), so
	// they must be finite: clear once
	for (int k = lane; k < buf_floats; k += 64)
		buf[k] = 0.0f;
	asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");

	{
		const TileRef t0 = tiles[first + wave];
		if (t0.a >= 0)
			pipe_prefetch<AXIS, NMAX>(field_rsrc(doms[t0.dom]), buf, doms[t0.dom], t0, lane);
	}
	int cur = -1;
	for (int base = first; base < ntiles; base += step)
	{
		const TileRef tr = tiles[base + wave];
		const DomainDev& d = doms[tr.dom];
		if (tr.dom != cur) // uniform across the workgroup: a round never spans two domains
		{
			__syncthreads();
			const int n = AXIS == 0 ? d.nx : (AXIS == 1 ? d.ny : d.nz);
			const double* gb = d.fb[AXIS];
			for (int k = threadIdx.x; k < fac_row(NMAX) * d.S; k += blockDim.x)
			{
				const int s = k / fac_row(NMAX), i = k - s * fac_row(NMAX);
				fb[k] = i < n ? gb[i * d.S + s] : 0.0;
			}
			for (int k = threadIdx.x; k < d.S; k += blockDim.x)
				fc[k] = d.fc[AXIS][k];
			__syncthreads();
			cur = tr.dom;
		}
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int S = AXIS == 0 ? SX : d.S, L = d.L;

		// this lane's line in the tile (line: its rounding-key index, as in the
		// register kernels)
		int slot, cs, voff, stride4;
		uint32_t line;
		line_faces lf {};
		if constexpr (AXIS == 0)
		{
			const int rows = max(tr.b, 1);
			const int r = min(lane / S, rows - 1);
			const int s = lane / S < rows ? lane - (lane / S) * S : S - 1;
			slot = r * pipe_x_pitch(L, S) + s;
			cs = s;
			line = (uint32_t)((tr.a + r) * S + s);
			voff = ((tr.a + r) * L + s) * (int)sizeof(float);
			stride4 = S * (int)sizeof(float);
			if constexpr (FACES)
				if (finish && d.face_mask && d.dims == 1)
					lf = faces_of_line(d, 0, 0, (tr.a + r) % d.ny, (tr.a + r) / d.ny, s);
		}
		else
		{
			const int total = pipe_total<AXIS>(d);
			const int f = min(tr.a + lane, total - 1);
			slot = f - tr.a;
			line = (uint32_t)f;
			const int outer = f / L, c = f - outer * L;
			cs = d.coef_map ? d.coef_map[c] : c % S;
			voff = pipe_col_off<AXIS>(d, f, total);
			stride4 = (AXIS == 1 ? L : d.ny * L) * (int)sizeof(float);
			if constexpr (FACES)
				if (finish && d.face_mask && AXIS == d.dims - 1)
				{
					const int x = c / S, s = c - (c / S) * S;
					lf = AXIS == 1 ? faces_of_line(d, 1, x, 0, outer, s) : faces_of_line(d, 2, x, outer, 0, s);
				}
		}

		// the tile has landed in this wave's buffer
		asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
		const int n = AXIS == 0 ? d.nx : (AXIS == 1 ? d.ny : d.nz);
		float v[NMAX];
		{
			// compile-time element stride: immediate LDS offsets
			constexpr int ls = AXIS == 0 ? SX : 64;
			const float* src = buf + slot;
			// n > NMAX - kPipeBucket: only the last slots of the bucket can pass the end
			constexpr int tail = AXIS == 0 ? 32 : 8;
			static_for<NMAX>([&](auto I) {
				if constexpr (I < NMAX - tail)
					v[I] = src[I * ls];
				else
					v[I] = src[min((int)I, n - 1) * ls];
			});
		}
		// every lane's reads are done before the next tile overwrites the buffer
		asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
		if (base + step < ntiles)
		{
			const TileRef nt = tiles[base + step + wave];
			if (nt.a >= 0)
				pipe_prefetch<AXIS, NMAX>(field_rsrc(doms[nt.dom]), buf, doms[nt.dom], nt, lane);
		}
		if (tr.a < 0)
			continue;
		const LineKey kf = line_key(d, 2 * AXIS, line);
		const LineKey ko = line_key(d, 2 * AXIS + 1, line);

		const double* lb = fb + cs * fac_row(NMAX);
		const double cc = fc[cs];
		int sn = n - 1, sstride4 = stride4;
		asm volatile("" : "+s"(sn), "+s"(sstride4));
		// forward: d_i += e_{i-1} d_{i-1}, e_{i-1} = c * b'_{i-1}.  The factors
		// of chunk C+1 are read from LDS while chunk C computes (one wave per
		// SIMD here, so LDS latency must be covered inside the wave).
		constexpr int kPC = kPipeChunk;
		constexpr int nchunks = (NMAX - 1 + kPC - 1) / kPC;
		double fa[2][kPC];
		auto load_fwd = [&](auto C, double(&dst)[kPC]) {
			static_for<kPC>([&](auto K) {
				constexpr int i = 1 + C * kPC + K;
				if constexpr (i < NMAX)
					dst[K] = lb[i - 1];
			});
		};
		load_fwd(std::integral_constant<int, 0> {}, fa[0]);
		__builtin_amdgcn_sched_barrier(0);
		double prev = (double)v[0];
		static_for<nchunks>([&](auto C) {
			if constexpr (C + 1 < nchunks)
				load_fwd(std::integral_constant<int, C + 1> {}, fa[(C + 1) & 1]);
			__builtin_amdgcn_sched_barrier(0);
			static_for<kPC>([&](auto K) {
				constexpr int i = 1 + C * kPC + K;
				if constexpr (i < NMAX)
				{
					prev = (double)v[i] + (cc * fa[C & 1][K]) * prev;
					v[i] = field_round<double>(prev, kf, weyl29(i));
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
		// backward: d_{n-1} *= b'_{n-1}; d_i = (d_i + c d_{i+1}) b'_i; stored
		// from the top, padding first (clamped to n-1, rewritten by the real value)
		// Faces of this sweep's own axis touch only element 0 (static i == 0)
		// and element n-1 (among the last slots of the bucket); lines lying in
		// a face of another axis are rewritten after the solve.
		constexpr int tail = AXIS == 0 ? 32 : 8;
		auto out = [&](auto I, float val) {
			constexpr int i = I;
			if constexpr (FACES && i == 0)
				val = lf.has_min ? lf.min_v : val;
			if constexpr (i >= NMAX - tail)
			{
				const int k = min(i, sn);
				if constexpr (FACES)
					val = (k == sn && lf.has_max) ? lf.max_v : val;
				buf_store(rsrc, voff, k * sstride4, val);
			}
			else
				buf_store(rsrc, voff, i * sstride4, val);
		};
		auto load_bwd = [&](auto C, double(&dst)[kPC]) {
			static_for<kPC>([&](auto K) {
				constexpr int i = NMAX - 2 - (C * kPC + K);
				if constexpr (i >= 0)
					dst[K] = lb[i];
			});
		};
		load_bwd(std::integral_constant<int, 0> {}, fa[0]);
		const double btop = lb[NMAX - 1];
		__builtin_amdgcn_sched_barrier(0);
		double nxt = (double)v[NMAX - 1] * btop;
		out(std::integral_constant<int, NMAX - 1> {}, field_round<double>(nxt, ko, weyl29(NMAX - 1)));
		static_for<nchunks>([&](auto C) {
			if constexpr (C + 1 < nchunks)
				load_bwd(std::integral_constant<int, C + 1> {}, fa[(C + 1) & 1]);
			__builtin_amdgcn_sched_barrier(0);
			static_for<kPC>([&](auto K) {
				constexpr int i = NMAX - 2 - (C * kPC + K);
				if constexpr (i >= 0)
				{
					nxt = ((double)v[i] + cc * nxt) * fa[C & 1][K];
					out(std::integral_constant<int, i> {}, field_round<double>(nxt, ko, weyl29(i)));
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
		if constexpr (FACES)
		{
			// lines inside a Dirichlet face of another axis: their values do not
			// depend on the solve (face_value ignores v when lo or hi is set);
			// later stores to the same address win
			const bool special = lf.lo || lf.hi;
			if (__builtin_amdgcn_ballot_w64(special) != 0)
				for (int k = 0; k <= sn; k++)
					if (special)
						buf_store(rsrc, voff, k * sstride4, face_value(lf, k, sn + 1, 0.0f));
		}
	}
}

} // namespace sweep
} // namespace bfvm
