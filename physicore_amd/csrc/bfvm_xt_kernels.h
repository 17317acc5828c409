// bfvm_xt_kernels.h -- x sweep of rows of up to 200 voxels with 4 or 8
// substrates, or 2 when rows have an even length (16-byte aligned rows):
// register lines filled and emptied through a per-wave LDS transpose (device
// template; instantiated by bfvm_sweeps_xt{2,4,8}{a,b}.hip).
//
// The strided register lines (sweep_cols_kernel<0, ...>) read each line with
// stride S: one load instruction of a wave touches 64/S rows, 16 cache lines
// for 256 useful bytes, and the address/cache path, not HBM, bounds the
// sweep.  Here a wave (64 lines = 64/S rows) loads its rows with 16-byte
// accesses in chunk-major order (each instruction covers 512 contiguous bytes
// of two rows, or 1 KB of one), so the whole tile is in flight at once as in
// the strided kernel, and then, 32 voxels at a time, writes the staged values
// to an LDS buffer and reads back its own line (row pitch 33*S floats: the
// strided reads of a half-wave hit 32 distinct banks).  The registers of a
// staged chunk are free once it is in LDS, so the line and the staging share
// the same budget.  Recurrences, factors and rounding keys are those of the
// strided kernel, so the results are bit-identical to it.  The results leave
// the same way, chunk by chunk from the end of the row, with 16-byte stores.
#pragma once

#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

constexpr int kXtChunk = 32; // voxels per transposed chunk
constexpr int kXtGroup = 8;	 // recurrence steps per scheduling group (bounds the factor prefetch)

__host__ __device__ constexpr int xt_buffer_floats(int S) { return (64 / S) * 33 * S; } // one wave's buffer

// Lanes of one wave exchange values through its LDS buffer.  The wave's LDS
// accesses execute in issue order, so only the compiler has to be kept from
// moving a lane's access across another lane's: a wavefront-scope fence.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

template <int NMAX, int S>
__global__ __launch_bounds__(kSweepThreads) void sweep_xt_kernel(const DomainDev* __restrict__ doms,
																 const TileRef* __restrict__ tiles, int ntiles,
																 int finish, int fac_cap)
{
	(void)finish;
	static_assert(S == 2 || S == 4 || S == 8, "16-byte transposes of 2 (even rows), 4 or 8 substrates");
	static_assert(NMAX % 8 == 0 && NMAX % 32 != 24 && NMAX <= kXtMaxN, "row bucket");
	constexpr int R = 64 / S;	   // rows per wave
	constexpr int PITCH = 33 * S;  // floats per row in the LDS buffer
	constexpr int NCH = (NMAX + kXtChunk - 1) / kXtChunk;
	extern __shared__ double lds[];
	double* fb = lds;
	double* fe = lds + fac_cap;
	double* fc = lds + 2 * fac_cap;
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	float* buf = reinterpret_cast<float*>(lds + 2 * fac_cap + 64) + wave * xt_buffer_floats(S);
	const int rl = lane / S, sl = lane - (lane / S) * S; // this lane's line: row rl of the wave, substrate sl
	int cur = -1;
	fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int n = d.nx;
		if (tr.dom != cur)
		{
			__syncthreads();
			load_factors(d, 0, n, fac_row(NMAX), fb, fe, fc);
			__syncthreads();
			cur = tr.dom;
		}
		const int L = d.L; // floats per row (n*S, a multiple of 4)
		const int rows = d.ny * d.nz;
		// rows of this wave: R from row0.  The last wave of a domain that
		// would reach past the end moves back to the last R rows (at least R
		// exist) and stores only the rows from `own` on: the rows below are
		// another wave's, which may already hold their results, so what is
		// computed for them here is discarded
		const int nominal = tr.a / S + wave * R;
		if (nominal >= rows)
			continue;
		const int row0 = min(nominal, rows - R);
		const int own = nominal - row0;
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int nS = n * S; // real floats of a row

		// staged float4 q of chunk k (CV voxels, W4 = CV*S/4 float4 per row):
		// u = q*64 + lane -> row u / W4, float4 column u % W4 of the chunk.
		// Byte offset = per-lane part (row lane / W4, column) + uniform part
		// (row q*64/W4, chunk start).  Chunks hold 32 voxels except the last,
		// which holds TAIL = NMAX % 32 when that is not 0; chunks that can
		// reach past the row (padding voxels) clamp the column per lane.
		constexpr int TAIL = NMAX % kXtChunk;
		auto lane_part = [&](int W4) { return ((row0 + lane / W4) * L + (lane % W4) * 4) * (int)sizeof(float); };
		const int lane_full = lane_part(kXtChunk * S / 4);
		const int lane_tail = TAIL ? lane_part(TAIL * S / 4) : 0;
		auto chunk_voff = [&](auto K) -> int {
			constexpr int k = K;
			constexpr int CV = (NMAX - k * kXtChunk) < kXtChunk ? NMAX - k * kXtChunk : kXtChunk;
			constexpr int W4 = CV * S / 4;
			const int base = CV == kXtChunk ? lane_full : lane_tail;
			if constexpr ((k + 1) * kXtChunk > NMAX - 16)
			{
				const int e = k * kXtChunk * S + (lane % W4) * 4;
				return base + (min(e, nS - 4) - e) * (int)sizeof(float);
			}
			else
				return base;
		};
		auto chunk_soff = [&](auto K, auto Q) -> int {
			constexpr int k = K, q = Q;
			constexpr int CV = (NMAX - k * kXtChunk) < kXtChunk ? NMAX - k * kXtChunk : kXtChunk;
			constexpr int W4 = CV * S / 4;
			return (q * (64 / W4) * L + k * kXtChunk * S) * (int)sizeof(float);
		};

		// all loads of the tile, chunk-major
		float4 st[NMAX / 4];
		static_for<NCH>([&](auto K) {
			constexpr int k = K;
			constexpr int CV = (NMAX - k * kXtChunk) < kXtChunk ? NMAX - k * kXtChunk : kXtChunk;
			constexpr int base = k * (kXtChunk / 4); // staged index of the chunk's first float4
			const int vo = chunk_voff(K);
			static_for<CV / 4>([&](auto Q) {
				const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rsrc, vo, chunk_soff(K, Q), 0);
				st[base + Q].x = u32_as_float(w.x);
				st[base + Q].y = u32_as_float(w.y);
				st[base + Q].z = u32_as_float(w.z);
				st[base + Q].w = u32_as_float(w.w);
			});
		});

		// transpose chunk by chunk into the register line (the staged registers
		// of a chunk become its line registers), then eliminate: running the
		// recurrence while later chunks are still staged would need registers
		// the line does not have
		const int line = (row0 + rl) * S + sl;
		const LineKey kf = line_key(d, kRoundFwdX, (uint32_t)line);
		const LineKey ko = line_key(d, kRoundOutX, (uint32_t)line);
		const double* lb = fb + sl * fac_row(NMAX);
		const double* le = fe + sl * fac_row(NMAX);
		const double cc = fc[sl];
		float v[NMAX];
		static_for<NCH>([&](auto K) {
			constexpr int k = K;
			constexpr int CV = (NMAX - k * kXtChunk) < kXtChunk ? NMAX - k * kXtChunk : kXtChunk;
			constexpr int W4 = CV * S / 4;
			constexpr int base = k * (kXtChunk / 4);
			static_for<CV / 4>([&](auto Q) {
				const int r = Q * (64 / W4) + lane / W4;
				*reinterpret_cast<float4*>(buf + r * PITCH + (lane % W4) * 4) = st[base + Q];
			});
			wave_lds_fence();
			static_for<CV>([&](auto V) { v[k * kXtChunk + V] = buf[rl * PITCH + V * S + sl]; });
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
		});
		// the rounded forward values are pinned (opaque) at the end of each
		// group: otherwise the compiler defers the rounding to the
		// substitution and keeps every FP64 forward value alive until then
		double prev = (double)v[0];
		static_for<(NMAX - 1 + kXtGroup - 1) / kXtGroup>([&](auto G) {
			static_for<kXtGroup>([&](auto V) {
				constexpr int i = 1 + G * kXtGroup + V;
				if constexpr (i < NMAX)
				{
					prev = (double)v[i] + le[i - 1] * prev;
					v[i] = field_round<double>(prev, kf, weyl29(i));
				}
			});
			static_for<kXtGroup>([&](auto V) {
				constexpr int i = 1 + G * kXtGroup + V;
				if constexpr (i < NMAX)
				{
					float x = v[i];
					asm volatile("" : "+v"(x));
					v[i] = x;
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});

		// substitution from the end of the row; each chunk leaves through LDS
		// once its first slot is done.  Opaque copies of the offsets keep the
		// compiler from holding the load addresses until the stores
		int sfull = lane_full, stail = lane_tail, sL = L;
		asm volatile("" : "+v"(sfull), "+v"(stail), "+s"(sL));
		double nxt = (double)v[NMAX - 1] * lb[NMAX - 1];
		static_for<NCH>([&](auto KK) {
			constexpr int k = NCH - 1 - KK;
			constexpr int CV = (NMAX - k * kXtChunk) < kXtChunk ? NMAX - k * kXtChunk : kXtChunk;
			constexpr int W4 = CV * S / 4;
			static_for<CV / kXtGroup>([&](auto G) {
				static_for<kXtGroup>([&](auto VV) {
					constexpr int i = k * kXtChunk + CV - 1 - (G * kXtGroup + VV);
					if constexpr (i != NMAX - 1)
						nxt = ((double)v[i] + cc * nxt) * lb[i];
					buf[rl * PITCH + (i - k * kXtChunk) * S + sl] = field_round<double>(nxt, ko, weyl29(i));
				});
				__builtin_amdgcn_sched_barrier(0);
			});
			// lane-derived offsets recomputed here (an opaque lane index) rather
			// than kept alive since the loads.  All of the chunk's LDS reads are
			// issued before its stores (the line registers of the later chunks
			// are free by now): a read followed by its store would wait out
			// every LDS round trip in turn
			wave_lds_fence();
			int ln = lane;
			asm volatile("" : "+v"(ln));
			const int vo = CV == kXtChunk ? sfull : stail;
			u32x4 w[CV / 4];
			static_for<CV / 4>([&](auto Q) {
				const int r = Q * (64 / W4) + ln / W4;
				const float4 o = *reinterpret_cast<const float4*>(buf + r * PITCH + (ln % W4) * 4);
				w[Q].x = __builtin_bit_cast(unsigned, o.x);
				w[Q].y = __builtin_bit_cast(unsigned, o.y);
				w[Q].z = __builtin_bit_cast(unsigned, o.z);
				w[Q].w = __builtin_bit_cast(unsigned, o.w);
			});
			__builtin_amdgcn_sched_barrier(0);
			static_for<CV / 4>([&](auto Q) {
				const int r = Q * (64 / W4) + ln / W4;
				const int so = (Q * (64 / W4) * sL + k * kXtChunk * S) * (int)sizeof(float);
				// padding voxels (column >= n*S) exist only in the last chunks of the bucket
				bool keep = r >= own;
				if constexpr ((k + 1) * kXtChunk > NMAX - 16)
					keep = keep && k * kXtChunk * S + (ln % W4) * 4 < nS;
				if (keep)
					__builtin_amdgcn_raw_buffer_store_b128(w[Q], rsrc, vo, so, 0);
			});
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
		});
	}
}

} // namespace sweep
} // namespace bfvm
