// bfvm_sweep_kernels.h -- the three Thomas sweeps of the LOD step on gfx950
// (device templates; instantiated by the bfvm_sweeps*.hip translation units).
//
// Reference: diffusion_solver::solve and solve_slice_{x,y,z}
// (reactions-diffusion/biofvm/kernels/openmp_solver/src/diffusion_solver.cpp:116-202, 390-843),
// recurrences of diffusion_solver.h:11-33:
//   forward   d_i += e_{i-1} d_{i-1};  d_{n-1} *= b_{n-1}
//   backward  d_i = (d_i + c d_{i+1}) b_i
//
// Every sweep reads and writes each FP32 value once, so all of them are bound
// by HBM bandwidth.  Lines up to 256 long are held in VGPRs, one line per lane:
//  * y and z: a workgroup owns 256 consecutive columns (x*S + s, flattened with
//    the outer axis), so every load and store of a line step is one coalesced
//    256-lane access; no LDS is needed for the data.
//  * x: the lines of a row are interleaved by substrate, so a tile of 64/S rows
//    is staged through LDS with 16-byte coalesced accesses and each lane then
//    reads its strided line into registers (the transpose).
// The Thomas factors (FP64) sit in LDS, transposed to [substrate][i] so each
// lane reads them at an immediate offset from its own base.  Arithmetic is
// FP64 (or FP32 on request); the field stays FP32.  Longer lines fall back to
// a kernel that solves them inside an LDS tile.
#pragma once

#include "bfvm_internal.h"

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace bfvm {
namespace sweep {

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1.  Array
// indices inside f are constants from the front end on, so a per-lane line
// array is split into scalars (registers) before any other pass sees it.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
	(f(std::integral_constant<int, I> {}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
	static_for_impl(f, std::make_integer_sequence<int, N> {});
}

__device__ __forceinline__ float apply_faces(const DomainDev& d, int x, int y, int z, int s, float v)
{
	// Later faces win: x-min, x-max, y-min, y-max, z-min, z-max
	// (openmp_solver/src/dirichlet_solver.cpp:59-83), so test from the last.
	const int m = d.face_mask;
	const int S = d.S;
	if ((m & 32) && z == d.nz - 1 && d.face_cond[5 * S + s])
		return d.face_val[5 * S + s];
	if ((m & 16) && z == 0 && d.face_cond[4 * S + s])
		return d.face_val[4 * S + s];
	if ((m & 8) && y == d.ny - 1 && d.face_cond[3 * S + s])
		return d.face_val[3 * S + s];
	if ((m & 4) && y == 0 && d.face_cond[2 * S + s])
		return d.face_val[2 * S + s];
	if ((m & 2) && x == d.nx - 1 && d.face_cond[1 * S + s])
		return d.face_val[1 * S + s];
	if ((m & 1) && x == 0 && d.face_cond[0 * S + s])
		return d.face_val[0 * S + s];
	return v;
}

// Copies one axis' factors into LDS transposed to [s][i] with row stride npad:
// fb[s*npad + i], fe[s*npad + i], fc[s].  Past the end of the line the tables
// hold zeros (b_i = 0 for i >= n, e_i = 0 for i >= n-1), which decouples the
// padding lanes of the register kernels: a padded element then contributes
// exactly 0 to the last real one.
__device__ __forceinline__ void load_factors(const DomainDev& d, int axis, int n, int npad, double* fb, double* fe,
											 double* fc)
{
	const int S = d.S;
	const double* gb = d.fb[axis];
	const double* ge = d.fe[axis];
	for (int k = threadIdx.x; k < npad * S; k += blockDim.x)
	{
		const int s = k / npad, i = k - (k / npad) * npad;
		fb[k] = i < n ? gb[i * S + s] : 0.0;
		fe[k] = i < n - 1 ? ge[i * S + s] : 0.0;
	}
	for (int k = threadIdx.x; k < S; k += blockDim.x)
		fc[k] = d.fc[axis][k];
}

// Raw buffer access to the field: the resource holds the base address and the
// size, a lane passes its own byte offset, the line step is a scalar offset.
constexpr int kBufferWord3 = 0x00020000; // gfx9 raw buffer, 32-bit elements
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float u32_as_float(unsigned u) { return __builtin_bit_cast(float, u); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const DomainDev& d)
{
	const uint64_t bytes = (uint64_t)d.nx * d.ny * d.nz * d.S * sizeof(float);
	const int num = (int)(uint32_t)(bytes > 0xffffffffull ? 0xffffffffull : bytes);
	return __builtin_amdgcn_make_buffer_rsrc(d.field, (short)0, num, kBufferWord3);
}

__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, int voff, int soff, float v)
{
	__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// Thomas solve of a line held in registers v[0..NMAX-1] (entries past the real
// length are padding, decoupled by the zero factors).  No branches: the whole
// solve is one basic block, so the loads feeding it can all be in flight.
// `out(i, value)` receives the results from i = NMAX-1 down to 0; fb/fe point
// at this line's substrate row.  kf / ko: the line's rounding keys for the
// forward values kept in v and for the results (FP64: the caller has set the
// FP32 rounding mode to toward-zero).
// The recurrences run in chunks of kChunk steps separated by scheduling
// barriers, so the compiler neither hoists every factor read nor every store
// offset to the top of the (fully unrolled) solve, which would exhaust VGPRs
// and SGPRs; latency is hidden across the resident waves instead.
constexpr int kChunk = 16;

// Back substitution of thomas_regs (v holds the rounded forward values).  The
// factors are read from LDS one group of kBackGroup steps ahead: the backward
// steps have no memory loads to hide an LDS round trip behind.
constexpr int kBackGroup = 4;

template <int NMAX, typename ACC, typename Out>
__device__ __forceinline__ void thomas_back(float (&v)[NMAX], const double* __restrict__ fb, double c, LineKey ko,
											Out out)
{
	constexpr int ngroups = (NMAX - 1 + kBackGroup - 1) / kBackGroup;
	const ACC cc = (ACC)c;
	ACC nxt = (ACC)v[NMAX - 1] * (ACC)fb[NMAX - 1];
	out(NMAX - 1, field_round<ACC>(nxt, ko, weyl29(NMAX - 1)));
	// step g*kBackGroup + K solves element NMAX - 2 - (g*kBackGroup + K)
	auto fetch = [&](auto G, double (&f)[kBackGroup]) {
		static_for<kBackGroup>([&](auto K) {
			constexpr int i = NMAX - 2 - (G * kBackGroup + K);
			if constexpr (i >= 0)
				f[K] = fb[i];
		});
	};
	double fnext[kBackGroup];
	fetch(std::integral_constant<int, 0> {}, fnext);
	static_for<ngroups>([&](auto G) {
		double fcur[kBackGroup];
		static_for<kBackGroup>([&](auto K) { fcur[K] = fnext[K]; });
		if constexpr (G + 1 < ngroups)
			fetch(std::integral_constant<int, G + 1> {}, fnext);
		static_for<kBackGroup>([&](auto K) {
			constexpr int i = NMAX - 2 - (G * kBackGroup + K);
			if constexpr (i >= 0)
			{
				nxt = ((ACC)v[i] + cc * nxt) * (ACC)fcur[K];
				out(i, field_round<ACC>(nxt, ko, weyl29(i)));
			}
		});
		__builtin_amdgcn_sched_barrier(0);
	});
}

template <int NMAX, typename ACC, typename Out>
__device__ __forceinline__ void thomas_regs(float (&v)[NMAX], const double* __restrict__ fb,
											const double* __restrict__ fe, double c, LineKey kf, LineKey ko, Out out)
{
	constexpr int nchunks = (NMAX - 1 + kChunk - 1) / kChunk;
	ACC prev = (ACC)v[0];
	static_for<nchunks>([&](auto C) {
		static_for<kChunk>([&](auto K) {
			constexpr int i = 1 + C * kChunk + K;
			if constexpr (i < NMAX)
			{
				prev = (ACC)v[i] + (ACC)fe[i - 1] * prev;
				v[i] = field_round<ACC>(prev, kf, weyl29(i));
			}
		});
		__builtin_amdgcn_sched_barrier(0);
	});
	thomas_back<NMAX, ACC>(v, fb, c, ko, out);
}

// thomas_regs with the line's loads interleaved with the forward recurrence:
// `ld(i)` returns element i.  The first kLoadAhead loads are issued before the
// recurrence starts, then kChunk more before each chunk of kChunk steps, so the
// forward steps run as the data arrives instead of after the last load; at
// most kLoadAhead + kChunk = 56 loads are in flight, within the 63 a wave's
// vector-memory counter tracks.
constexpr int kLoadAhead = 40;

template <int NMAX, typename ACC, typename Load, typename Out>
__device__ __forceinline__ void thomas_regs_streamed(float (&v)[NMAX], Load ld, const double* __restrict__ fb,
													 const double* __restrict__ fe, double c, LineKey kf, LineKey ko,
													 Out out)
{
	constexpr int nchunks = (NMAX - 1 + kChunk - 1) / kChunk;
	constexpr int first = NMAX < kLoadAhead ? NMAX : kLoadAhead;
	static_for<first>([&](auto I) { v[I] = ld((int)I); });
	ACC prev = 0;
	static_for<nchunks>([&](auto C) {
		static_for<kChunk>([&](auto K) {
			constexpr int j = kLoadAhead + C * kChunk + K;
			if constexpr (j < NMAX)
				v[j] = ld(j);
		});
		if constexpr (C == 0)
			prev = (ACC)v[0];
		static_for<kChunk>([&](auto K) {
			constexpr int i = 1 + C * kChunk + K;
			if constexpr (i < NMAX)
			{
				prev = (ACC)v[i] + (ACC)fe[i - 1] * prev;
				v[i] = field_round<ACC>(prev, kf, weyl29(i));
			}
		});
		__builtin_amdgcn_sched_barrier(0);
	});
	thomas_back<NMAX, ACC>(v, fb, c, ko, out);
}

// Face Dirichlet values seen by one line of the y or z sweep.  Faces of the
// other axes that contain the line are constant along it; they split into
// those ranked below the line's own two faces in the reference's order
// (x-min < x-max < y-min < y-max < z-min < z-max, later wins;
// openmp_solver/src/dirichlet_solver.cpp:59-83) and those ranked above.
struct line_faces
{
	float lo_v, min_v, max_v, hi_v;
	bool lo, has_min, has_max, hi;
};

__device__ __forceinline__ line_faces faces_of_line(const DomainDev& d, int axis, int x, int y, int z, int s)
{
	line_faces r {};
	const int S = d.S, m = d.face_mask;
	const int coord[3] = { x, y, z };
	const int nn[3] = { d.nx, d.ny, d.nz };
	for (int f = 0; f < kFaces; f++)
	{
		if (!((m >> f) & 1) || !d.face_cond[f * S + s])
			continue;
		const int a = f / 2;
		const float v = d.face_val[f * S + s];
		if (a == axis)
		{
			if (f & 1)
			{
				r.has_max = true;
				r.max_v = v;
			}
			else
			{
				r.has_min = true;
				r.min_v = v;
			}
			continue;
		}
		const bool on = (f & 1) ? coord[a] == nn[a] - 1 : coord[a] == 0;
		if (!on)
			continue;
		if (a > axis)
		{
			r.hi = true;
			r.hi_v = v;
		}
		else
		{
			r.lo = true;
			r.lo_v = v;
		}
	}
	return r;
}

// Value stored at position k of a line of length n: precedence applied in rank order.
__device__ __forceinline__ float face_value(const line_faces& f, int k, int n, float v)
{
	float out = f.lo ? f.lo_v : v;
	out = (k == 0 && f.has_min) ? f.min_v : out;
	out = (k == n - 1 && f.has_max) ? f.max_v : out;
	return f.hi ? f.hi_v : out;
}

// Sweeps with one line per lane in registers.  A tile is 256 consecutive
// columns of a flattened (outer, column) space starting at tr.a: for y (AXIS 1)
// and z (AXIS 2) the columns are x*S + s and outer is z for the y-sweep and y
// for the z-sweep (contiguous within each z plane, resp. in memory); for x
// (AXIS 0) the columns are the S substrates of a row and outer is the row
// y + ny*z, so a wave reads 64/S rows with stride S and the caches turn the
// short stride into whole-line fetches.  Padding steps (i >= n) reload element n-1
// and store to it before the real value, which is stored last.  FIN (kFinFaces,
// kFinBulk): the launch finishes the step for domains whose last axis this is,
// with the face conditions and, for kFinBulk, the bulk stage of a bulk with
// uniform rates over one run of voxels per line (DomainDev::bulk_run): each
// stored value v in the run becomes (v + a_s) * r_s, rounded with the bulk
// stage's key of its line and position (bulk_round), the arithmetic and bits
// of bulk_kernel on the stored field.
template <int AXIS, int NMAX, typename ACC, int FIN>
__global__ __launch_bounds__(kSweepThreads) void sweep_cols_kernel(const DomainDev* __restrict__ doms,
																   const TileRef* __restrict__ tiles, int ntiles,
																   int finish, int fac_cap)
{
	extern __shared__ double lds[];
	double* fb = lds;
	double* fe = lds + fac_cap;
	double* fc = lds + 2 * fac_cap;
	int cur = -1;
	if constexpr (sizeof(ACC) == 8)
		fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int n = AXIS == 0 ? d.nx : (AXIS == 1 ? d.ny : d.nz);
		if (tr.dom != cur)
		{
			__syncthreads();
			load_factors(d, AXIS, n, fac_row(NMAX), fb, fe, fc);
			__syncthreads();
			cur = tr.dom;
		}
		const int S = d.S, L = d.L;
		// AXIS 0: columns are (row, s), row = y + ny*z, S of them per row
		const int width = AXIS == 0 ? S : L;
		const int outer_n = AXIS == 0 ? d.ny * d.nz : (AXIS == 1 ? d.nz : d.ny);
		const int f = tr.a + (int)threadIdx.x;
		if (f >= outer_n * width)
			continue;
		const int outer = f / width;
		const int c = f - outer * width;
		const int stride4 = (AXIS == 0 ? S : (AXIS == 1 ? L : d.ny * L)) * (int)sizeof(float);
		const int voff =
			(AXIS == 0 ? outer * L + c : (AXIS == 1 ? outer * d.ny * L + c : f)) * (int)sizeof(float);
		const int cs = AXIS == 0 ? c : (d.coef_map ? d.coef_map[c] : c % S);
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		// line f of this axis: (row, s) for x, (outer, x*S + s) for y and z
		const LineKey kf = line_key(d, 2 * AXIS, (uint32_t)f);
		const LineKey ko = line_key(d, 2 * AXIS + 1, (uint32_t)f);

		float v[NMAX];
		auto ld = [&](int i) { return buf_load(rsrc, voff, min(i, n - 1) * stride4); };

		// Opaque copies of the line geometry for the store phase: otherwise the
		// per-step store offsets are recognised as the load offsets and all of
		// them are kept live (in SGPRs) across the whole solve.
		int sn = n - 1, sstride4 = stride4;
		asm volatile("" : "+s"(sn), "+s"(sstride4));
		if constexpr (FIN != kFinNone)
		{
			const bool last = finish && AXIS == d.dims - 1;
			line_faces lf {};
			if (last && d.face_mask)
			{
				if constexpr (AXIS == 0)
					lf = faces_of_line(d, 0, 0, outer % d.ny, outer / d.ny, c);
				else
				{
					const int x = c / S, s = c - (c / S) * S;
					lf = AXIS == 1 ? faces_of_line(d, 1, x, 0, outer, s) : faces_of_line(d, 2, x, outer, 0, s);
				}
			}
			// face_value's precedence without its per-element chain: a line
			// lying in a face (lo, hi) takes that value everywhere (one select
			// per element), and the line's own faces are stored at its ends
			// afterwards (a lane's later store to a voxel wins; hi outranks them)
			const bool lohi = lf.lo || lf.hi;
			const float fv = lf.hi ? lf.hi_v : lf.lo_v;
			if constexpr (FIN == kFinBulk)
			{
				// the bulk stage on each final value, rounded with the bulk
				// key of its line and position (bulk_round's bits); the
				// padding stores are overwritten by the real value
				const double* bu = last ? d.bulk_u : nullptr;
				const int sub = AXIS == 0 ? c : c % S;
				const double ba = bu ? bu[sub] : 0.0, br = bu ? bu[S + sub] : 1.0;
				const int col = f / S;
				const int r0 = bu ? d.bulk_run[2 * col] : 0;
				const unsigned rlen = bu ? (unsigned)(d.bulk_run[2 * col + 1] - r0) : 0u;
				const LineKey kb = line_key(d, kRoundBulk, (uint32_t)f);
				// the bulk value is formed at every position and picked by a
				// bit mask: with a conditional the compiler branched per
				// element and kept one more VGPR per element live
				auto bulk = [&](int i, float x) -> float {
					const float r = field_round<double>(((double)x + ba) * br, kb, weyl29(i));
					const unsigned m = 0u - (unsigned)((unsigned)(i - r0) < rlen);
					return __builtin_bit_cast(float,
											  (__builtin_bit_cast(unsigned, r) & m) | (__builtin_bit_cast(unsigned, x) & ~m));
				};
				thomas_regs_streamed<NMAX, ACC>(v, ld, fb + cs * fac_row(NMAX), fe + cs * fac_row(NMAX), fc[cs], kf, ko,
												[&](int i, float val) {
													buf_store(rsrc, voff, min(i, sn) * sstride4, bulk(i, lohi ? fv : val));
												});
				if (!lf.hi && lf.has_min)
					buf_store(rsrc, voff, 0, bulk(0, lf.min_v));
				if (!lf.hi && lf.has_max)
					buf_store(rsrc, voff, sn * sstride4, bulk(sn, lf.max_v));
			}
			else
			{
				thomas_regs_streamed<NMAX, ACC>(v, ld, fb + cs * fac_row(NMAX), fe + cs * fac_row(NMAX), fc[cs], kf, ko,
												[&](int i, float val) {
													buf_store(rsrc, voff, min(i, sn) * sstride4, lohi ? fv : val);
												});
				if (!lf.hi && lf.has_min)
					buf_store(rsrc, voff, 0, lf.min_v);
				if (!lf.hi && lf.has_max)
					buf_store(rsrc, voff, sn * sstride4, lf.max_v);
			}
		}
		else
		{
			thomas_regs_streamed<NMAX, ACC>(v, ld, fb + cs * fac_row(NMAX), fe + cs * fac_row(NMAX), fc[cs], kf, ko,
											[&](int i, float val) { buf_store(rsrc, voff, min(i, sn) * sstride4, val); });
		}
	}
}

// x sweep: a tile of `rows` = 64/S consecutive rows (contiguous in HBM) is
// staged in LDS with row pitch Lp = S (mod 64), so the 64 lanes (r, s) read
// their strided lines from distinct banks; each lane solves in registers and
// writes back; the tile is stored coalesced.
template <int NMAX, typename ACC>
__global__ __launch_bounds__(kSweepThreads) void sweep_rows_kernel(const DomainDev* __restrict__ doms,
																   const TileRef* __restrict__ tiles, int ntiles,
																   int finish, int fac_cap)
{
	extern __shared__ double lds[];
	double* fb = lds;
	double* fe = lds + fac_cap;
	double* fc = lds + 2 * fac_cap;
	float* tile = reinterpret_cast<float*>(lds + 2 * fac_cap + 64);
	const int tid = threadIdx.x;
	int cur = -1;
	if constexpr (sizeof(ACC) == 8)
		fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int n = d.nx, S = d.S, L = d.L;
		if (tr.dom != cur)
		{
			load_factors(d, 0, n, fac_row(NMAX), fb, fe, fc); // the loop-end barrier of the previous tile orders this
			cur = tr.dom;
		}
		const int rows = tr.b;
		const int Lp = x_row_pitch(L, S);
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int base4 = tr.a * L * (int)sizeof(float);
		const int total = rows * L;
		const bool vec = (S % 4 == 0) && (L % 4 == 0);
		if (vec)
		{
			for (int q = tid; q < total / 4; q += kSweepThreads)
			{
				const int e = q * 4;
				const int r = e / L, j = e - r * L;
				*reinterpret_cast<u32x4*>(tile + r * Lp + j) = __builtin_amdgcn_raw_buffer_load_b128(rsrc, e * 4, base4, 0);
			}
		}
		else
		{
			for (int e = tid; e < total; e += kSweepThreads)
			{
				const int r = e / L, j = e - r * L;
				tile[r * Lp + j] = buf_load(rsrc, e * 4, base4);
			}
		}
		__syncthreads();
		if (tid < rows * S)
		{
			const int r = tid / S, s = tid - (tid / S) * S;
			float* line = tile + r * Lp + s;
			float v[NMAX];
			static_for<NMAX>([&](auto I) { v[I] = line[min((int)I, n - 1) * S]; });
			__builtin_amdgcn_sched_barrier(0);
			int sn = n - 1, sS = S;
			asm volatile("" : "+s"(sn), "+s"(sS));
			const uint32_t f = (uint32_t)((tr.a + r) * S + s); // line (row, s)
			thomas_regs<NMAX, ACC>(v, fb + s * fac_row(NMAX), fe + s * fac_row(NMAX), fc[s], line_key(d, kRoundFwdX, f),
								   line_key(d, kRoundOutX, f), [&](int i, float val) { line[min(i, sn) * sS] = val; });
		}
		__syncthreads();
		const bool faces = finish && d.face_mask && d.dims == 1;
		if (vec && !faces)
		{
			for (int q = tid; q < total / 4; q += kSweepThreads)
			{
				const int e = q * 4;
				const int r = e / L, j = e - r * L;
				__builtin_amdgcn_raw_buffer_store_b128(*reinterpret_cast<const u32x4*>(tile + r * Lp + j), rsrc, e * 4,
													   base4, 0);
			}
		}
		else
		{
			for (int e = tid; e < total; e += kSweepThreads)
			{
				const int r = e / L, j = e - r * L;
				float v = tile[r * Lp + j];
				if (faces)
				{
					const int row = tr.a + r;
					const int x = j / S, s = j - (j / S) * S;
					v = apply_faces(d, x, row % d.ny, row / d.ny, s, v);
				}
				buf_store(rsrc, e * 4, base4, v);
			}
		}
		__syncthreads();
	}
}

// Lines longer than the register budget: the line stays in the LDS tile and
// the recurrence walks it there (tile layouts as above; y/z tiles are 64
// columns of one plane: tr.a = outer, tr.b = first column).
template <int AXIS, typename ACC>
__device__ __forceinline__ void thomas_lds(float* line, int stride, int n, const double* fb, const double* fe,
										   double c, const DomainDev& d, uint32_t f)
{
	const LineKey kf = line_key(d, 2 * AXIS, f);
	const LineKey ko = line_key(d, 2 * AXIS + 1, f);
	ACC prev = (ACC)line[0];
	for (int i = 1; i < n; i++)
	{
		float* p = line + i * stride;
		const ACC v = (ACC)(*p) + (ACC)fe[i - 1] * prev;
		*p = field_round<ACC>(v, kf, weyl29(i));
		prev = v;
	}
	ACC nxt = prev * (ACC)fb[n - 1];
	line[(n - 1) * stride] = field_round<ACC>(nxt, ko, weyl29(n - 1));
	const ACC cc = (ACC)c;
	for (int i = n - 2; i >= 0; i--)
	{
		float* p = line + i * stride;
		nxt = ((ACC)(*p) + cc * nxt) * (ACC)fb[i];
		*p = field_round<ACC>(nxt, ko, weyl29(i));
	}
}

template <int AXIS, typename ACC>
__global__ __launch_bounds__(kSweepThreads) void sweep_lds_kernel(const DomainDev* __restrict__ doms,
																  const TileRef* __restrict__ tiles, int ntiles,
																  int finish, int fac_cap)
{
	extern __shared__ double lds[];
	double* fb = lds;
	double* fe = lds + fac_cap;
	double* fc = lds + 2 * fac_cap;
	float* tile = reinterpret_cast<float*>(lds + 2 * fac_cap + 64);
	const int tid = threadIdx.x;
	int cur = -1;
	if constexpr (sizeof(ACC) == 8)
		fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int S = d.S, L = d.L;
		const int n = AXIS == 0 ? d.nx : (AXIS == 1 ? d.ny : d.nz);
		if (tr.dom != cur)
		{
			load_factors(d, AXIS, n, n, fb, fe, fc);
			cur = tr.dom;
		}
		const bool faces = finish && d.face_mask && AXIS == d.dims - 1;
		if constexpr (AXIS == 0)
		{
			const int rows = tr.b;
			const int Lp = x_row_pitch(L, S);
			float* g = d.field + (size_t)tr.a * L;
			const int total = rows * L;
			for (int e = tid; e < total; e += kSweepThreads)
			{
				const int r = e / L, j = e - r * L;
				tile[r * Lp + j] = g[e];
			}
			__syncthreads();
			if (tid < rows * S)
			{
				const int r = tid / S, s = tid - (tid / S) * S;
				thomas_lds<AXIS, ACC>(tile + r * Lp + s, S, n, fb + s * n, fe + s * n, fc[s], d,
									  (uint32_t)((tr.a + r) * S + s));
			}
			__syncthreads();
			for (int e = tid; e < total; e += kSweepThreads)
			{
				const int r = e / L, j = e - r * L;
				float v = tile[r * Lp + j];
				if (faces)
				{
					const int row = tr.a + r;
					v = apply_faces(d, j / S, row % d.ny, row / d.ny, j % S, v);
				}
				g[e] = v;
			}
		}
		else
		{
			const int c0 = tr.b;
			const int w = min(kLinesPerTile, L - c0);
			const size_t line_stride = AXIS == 1 ? (size_t)L : (size_t)d.ny * L;
			const size_t base = AXIS == 1 ? (size_t)tr.a * d.ny * L : (size_t)tr.a * L;
			const int j = tid & (kLinesPerTile - 1);
			const int k0 = tid / kLinesPerTile;
			float* g = d.field + base + c0;
			if (j < w)
				for (int k = k0; k < n; k += kSweepThreads / kLinesPerTile)
					tile[k * kLinesPerTile + j] = g[k * line_stride + j];
			__syncthreads();
			if (tid < w)
			{
				const int c = c0 + tid;
				const int cs = d.coef_map ? d.coef_map[c] : c % S;
				thomas_lds<AXIS, ACC>(tile + tid, kLinesPerTile, n, fb + cs * n, fe + cs * n, fc[cs], d,
									  (uint32_t)(tr.a * L + c));
			}
			__syncthreads();
			if (j < w)
			{
				const int c = c0 + j;
				const int x = c / S, s = c - (c / S) * S;
				for (int k = k0; k < n; k += kSweepThreads / kLinesPerTile)
				{
					float v = tile[k * kLinesPerTile + j];
					if (faces)
						v = AXIS == 1 ? apply_faces(d, x, k, tr.a, s, v) : apply_faces(d, x, tr.a, k, s, v);
					g[k * line_stride + j] = v;
				}
			}
		}
		__syncthreads();
	}
}

using kernel_fn = void (*)(const DomainDev*, const TileRef*, int, int, int);

} // namespace sweep
} // namespace bfvm

namespace bfvm {
namespace sweep {

// ------------------------------------------------------------------------
// x sweep by affine scans (lanes along the row).
//
// Both Thomas recurrences are first-order linear recurrences,
//   forward   y_x = e_{x-1} y_{x-1} + d_x
//   backward  z_x = (c b_x) z_{x+1} + b_x y_x
// so a lane holding VPL consecutive voxels (all S substrates, contiguous in
// HBM) turns its chunk into one affine map per substrate, Y_out = A*Y_in + B.
// The maps are combined across the 64 lanes of the wave with a log-depth scan
// (DPP shifts inside rows of 16 lanes, lane reads across rows), after which
// each lane re-runs its chunk from the correct carry.  One wave solves one row
// (all S lines of it); loads and stores are fully coalesced, nothing goes
// through LDS, and the lane's factors stay in registers across rows.
// Padding (x >= nx) has e = b = 0, which makes it inert in both directions.

template <typename T>
__device__ __forceinline__ T dpp_identity_a()
{
	return (T)1;
}

template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ T dpp_mov(T old, T src)
{
	if constexpr (sizeof(T) == 8)
	{
		const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
		const unsigned long long v = __builtin_bit_cast(unsigned long long, src);
		const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)v, CTRL,
																   ROW_MASK, 0xf, false);
		const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(v >> 32),
																   CTRL, ROW_MASK, 0xf, false);
		return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
	}
	else
	{
		return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
																  __builtin_bit_cast(int, src), CTRL, ROW_MASK, 0xf,
																  false));
	}
}

template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane)
{
	if constexpr (sizeof(T) == 8)
	{
		const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
		const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
		const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
		return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
	}
	else
	{
		return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
	}
}

// (A, B) <- (A, B) o (A', B'): apply the other map first, then this one.
template <typename T>
__device__ __forceinline__ void compose(T& A, T& B, T Ao, T Bo)
{
	B = A * Bo + B;
	A = A * Ao;
}

// Inclusive scan of affine maps towards higher lanes (lane l ends with the
// composition of lanes 0..l).  DPP row_shr inside rows of 16, then
// row_bcast:15 / row_bcast:31 across rows (GFX9 DPP).
template <typename T>
__device__ __forceinline__ void scan_up(T& A, T& B)
{
	const T one = (T)1, zero = (T)0;
	compose(A, B, dpp_mov<0x111, 0xf>(one, A), dpp_mov<0x111, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x112, 0xf>(one, A), dpp_mov<0x112, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x114, 0xf>(one, A), dpp_mov<0x114, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x118, 0xf>(one, A), dpp_mov<0x118, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x142, 0xa>(one, A), dpp_mov<0x142, 0xa>(zero, B));
	compose(A, B, dpp_mov<0x143, 0xc>(one, A), dpp_mov<0x143, 0xc>(zero, B));
}

// Inclusive scan towards lower lanes (lane l ends with lanes l..63, the higher
// lane applied first): row_shl inside rows, then each row composes with the
// totals of the rows above it, read from their first lanes.
template <typename T>
__device__ __forceinline__ void scan_down(T& A, T& B, int lane)
{
	const T one = (T)1, zero = (T)0;
	compose(A, B, dpp_mov<0x101, 0xf>(one, A), dpp_mov<0x101, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x102, 0xf>(one, A), dpp_mov<0x102, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x104, 0xf>(one, A), dpp_mov<0x104, 0xf>(zero, B));
	compose(A, B, dpp_mov<0x108, 0xf>(one, A), dpp_mov<0x108, 0xf>(zero, B));
	T A3 = read_lane(A, 48), B3 = read_lane(B, 48);
	T A2 = read_lane(A, 32), B2 = read_lane(B, 32);
	T A1 = read_lane(A, 16), B1 = read_lane(B, 16);
	// totals of rows 3; 2..3; 1..3
	compose(A2, B2, A3, B3);
	compose(A1, B1, A2, B2);
	const int row = lane >> 4;
	const T Ar = row == 0 ? A1 : (row == 1 ? A2 : (row == 2 ? A3 : one));
	const T Br = row == 0 ? B1 : (row == 1 ? B2 : (row == 2 ? B3 : zero));
	compose(A, B, Ar, Br);
}

template <int S, int VPL, typename ACC>
__global__ __launch_bounds__(kSweepThreads) void sweep_xscan_kernel(const DomainDev* __restrict__ doms,
																	const TileRef* __restrict__ tiles, int ntiles,
																	int finish, int fac_cap)
{
	(void)finish;
	(void)fac_cap;
	constexpr int C = S * VPL; // floats per lane
	const int lane = threadIdx.x & 63;
	// uniform per wave: keeps the row offset in an SGPR (buffer soffset)
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	int cur = -1;
	int nx = 0;
	ACC fa[C], fbv[C], cs[S];
	if constexpr (sizeof(ACC) == 8)
		fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		if (tr.dom != cur)
		{
			nx = d.nx;
			const double* ge = d.fe[0];
			const double* gb = d.fb[0];
			static_for<C>([&](auto K) {
				constexpr int j = K / S, s = K % S;
				const int x = lane * VPL + j;
				fa[K] = (x >= 1 && x < nx) ? (ACC)ge[(x - 1) * S + s] : (ACC)0;
				fbv[K] = x < nx ? (ACC)gb[x * S + s] : (ACC)0;
			});
			static_for<S>([&](auto s) { cs[s] = (ACC)d.fc[0][s]; });
			cur = tr.dom;
		}
		if (wave >= tr.b)
			continue;
		const int row = tr.a + wave;
		const int L = nx * S;
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int rowoff = row * L * (int)sizeof(float);
		const int loff = lane * C * (int)sizeof(float);
		const bool full = lane * VPL + VPL <= nx;
		const bool vec = (C % 4 == 0) && (L % 4 == 0);

		float v[C];
		if (vec && full)
		{
			static_for<C / 4>([&](auto Q) {
				const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rsrc, loff + Q * 16, rowoff, 0);
				// element values, not element lvalues: __builtin_bit_cast of an
				// ext-vector element reads from the start of the vector
				v[4 * Q + 0] = u32_as_float(w.x);
				v[4 * Q + 1] = u32_as_float(w.y);
				v[4 * Q + 2] = u32_as_float(w.z);
				v[4 * Q + 3] = u32_as_float(w.w);
			});
		}
		else
		{
			static_for<C>([&](auto K) {
				constexpr int j = K / S;
				v[K] = lane * VPL + j < nx ? buf_load(rsrc, loff + K * 4, rowoff) : 0.0f;
			});
		}

		// forward: local maps, scan, re-run with the carry
		ACC A[S], B[S];
		static_for<S>([&](auto s) {
			A[s] = (ACC)1;
			B[s] = (ACC)0;
		});
		static_for<C>([&](auto K) {
			constexpr int s = K % S;
			B[s] = fa[K] * B[s] + (ACC)v[K];
			A[s] = fa[K] * A[s];
		});
		static_for<S>([&](auto s) {
			scan_up(A[s], B[s]);
			// carry into this lane = inclusive result of the lane before (0 for lane 0)
			B[s] = dpp_mov<0x138, 0xf>((ACC)0, B[s]);
		});
		// forward results kept in ACC (a lane holds only C of them), so only
		// the results are rounded to FP32
		ACC fw[C];
		static_for<C>([&](auto K) {
			constexpr int s = K % S;
			B[s] = fa[K] * B[s] + (ACC)v[K];
			fw[K] = B[s];
		});

		// backward
		static_for<S>([&](auto s) {
			A[s] = (ACC)1;
			B[s] = (ACC)0;
		});
		static_for<C>([&](auto K) {
			constexpr int j = VPL - 1 - K / S, s = K % S;
			constexpr int e = j * S + s;
			const ACC cb = cs[s] * fbv[e];
			B[s] = cb * B[s] + fbv[e] * fw[e];
			A[s] = cb * A[s];
		});
		static_for<S>([&](auto s) {
			scan_down(A[s], B[s], lane);
			// carry from the right = inclusive result of the next lane (0 past lane 63)
			B[s] = dpp_mov<0x130, 0xf>((ACC)0, B[s]);
		});
		// one key per (row, lane); element e of the lane adds weyl29(e)
		const LineKey ko = line_key(d, kRoundOutX, (uint32_t)(row * 64 + lane));
		static_for<C>([&](auto K) {
			constexpr int j = VPL - 1 - K / S, s = K % S;
			constexpr int e = j * S + s;
			B[s] = cs[s] * fbv[e] * B[s] + fbv[e] * fw[e];
			v[e] = field_round<ACC>(B[s], ko, weyl29(e));
		});

		if (vec && full)
		{
			static_for<C / 4>([&](auto Q) {
				u32x4 w;
				w.x = __builtin_bit_cast(unsigned, v[4 * Q + 0]);
				w.y = __builtin_bit_cast(unsigned, v[4 * Q + 1]);
				w.z = __builtin_bit_cast(unsigned, v[4 * Q + 2]);
				w.w = __builtin_bit_cast(unsigned, v[4 * Q + 3]);
				__builtin_amdgcn_raw_buffer_store_b128(w, rsrc, loff + Q * 16, rowoff, 0);
			});
		}
		else
		{
			static_for<C>([&](auto K) {
				constexpr int j = K / S;
				if (lane * VPL + j < nx)
					buf_store(rsrc, loff + K * 4, rowoff, v[K]);
			});
		}
	}
}

} // namespace sweep
} // namespace bfvm
