// bfvm_xg_kernels.h -- x sweep of rows of up to 200 voxels with 1 to 8
// substrates: register lines filled and emptied through a per-wave LDS
// transpose with 4-byte accesses (device template; instantiated by
// bfvm_sweeps_xg{1..8}{a,b}.hip).
//
// The strided register lines read a line with stride S, so one load
// instruction of a wave touches 64/S rows: a cache line per row for 4*S useful
// bytes.  Here a wave of R rows (R*S lanes own lines; R even when S is odd)
// reads its rows in chunks of 32 voxels: a chunk of the R rows is R*32*S
// floats, a whole number of 64-float load instructions, and each instruction
// reads 64 consecutive floats of that chunk (the end of one row's chunk and the
// start of the next row's when S is odd: two contiguous runs).  The lane
// offsets of the instructions repeat with a period of P = 32*S / gcd(32*S, 64)
// instructions (S for odd S, 3 for S = 6, 1 for S = 2, 2 for S = 4, 4 for
// S = 8), so they are P per-lane registers
// plus scalar offsets.  The staged chunk goes to a per-wave LDS buffer (row
// pitch 33*S floats: the lines' strided reads of a half-wave hit 32 distinct
// banks) and each lane reads back its own line.  The last chunk (bucket
// padding, a partial last instruction) computes its offsets per lane.
// About 200 dword loads per lane exceed the 63 a wave can have in flight, so
// the loads of chunk k+2 are issued before chunk k is transposed and
// eliminated.  Recurrences, factors and rounding keys are those of the strided
// kernel, so the results are bit-identical to it.  The results leave the same
// way, chunk by chunk from the end of the row.
#pragma once

#include "bfvm_xt_kernels.h"

namespace bfvm {
namespace sweep {

constexpr int kXgChunk = 32; // voxels per transposed chunk
constexpr int kXgGroup = 4;	 // recurrence steps per scheduling group (bounds the factor prefetch)
constexpr int kXgStoreBatch = 8; // LDS reads in flight ahead of their stores

__host__ __device__ constexpr int xg_buffer_floats(int S) { return xg_rows(S) * 33 * S; } // one wave's buffer
__host__ __device__ constexpr int xg_gcd(int a, int b) { return b == 0 ? a : xg_gcd(b, a % b); }

template <int NMAX, int S>
__global__ __launch_bounds__(kSweepThreads) void sweep_xg_kernel(const DomainDev* __restrict__ doms,
																 const TileRef* __restrict__ tiles, int ntiles,
																 int finish, int fac_cap)
{
	(void)finish;
	static_assert(S >= 1 && S <= 8, "4-byte transposes of 1-8 substrates");
	static_assert(NMAX % 8 == 0 && NMAX <= xg_max_n(S), "row bucket");
	constexpr int R = xg_rows(S);
	constexpr int PITCH = 33 * S;				  // floats per row in the LDS buffer
	constexpr int CW = kXgChunk * S;			  // floats of a row in a full chunk
	constexpr int P = CW / xg_gcd(CW, 64);		  // instructions per period of the lane offsets
	constexpr int RP = 64 / xg_gcd(CW, 64);		  // rows per period
	constexpr int NQ = R * CW / 64;				  // load instructions of a full chunk
	constexpr int NCH = (NMAX + kXgChunk - 1) / kXgChunk;
	constexpr int CVL = NMAX - (NCH - 1) * kXgChunk; // voxels of the last chunk
	constexpr int CWL = CVL * S;
	constexpr int NQL = (R * CWL + 63) / 64; // load instructions of the last chunk
	static_assert(NQ % P == 0 && R % RP == 0, "whole periods per chunk");
	extern __shared__ double lds[];
	double* fb = lds;
	double* fe = lds + fac_cap;
	double* fc = lds + 2 * fac_cap;
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	float* buf = reinterpret_cast<float*>(lds + 2 * fac_cap + 64) + wave * xg_buffer_floats(S);
	// this lane's line: row rl of the wave, substrate sl; lanes past R*S solve
	// a copy of a line of the last row (same values, never stored twice over
	// different data)
	const int rl = min(lane / S, R - 1), sl = lane - (lane / S) * S;
	// per-lane part of instruction p of a period: row p*64+lane / CW of the
	// period, float (p*64+lane) % CW of that row's chunk; LDS offset of it
	int loff[P];
	static_for<P>([&](auto Pi) {
		constexpr int p = Pi;
		const int u = p * 64 + lane;
		loff[p] = (u / CW) * PITCH + u % CW;
	});
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
		const int L = d.L; // floats per row (n*S)
		const int rows = d.ny * d.nz;
		// rows of this wave: R from row0.  A tile holds the rows of 4 waves
		// from tr.a.  The last wave of a domain that would reach past the end
		// moves back to the last R rows (at least R exist) and stores only the
		// rows from `own` on: the rows below are another wave's, which may
		// already hold their results, so what is computed for them here is
		// discarded
		const int nominal = tr.a + wave * R;
		if (nominal >= rows)
			continue;
		const int row0 = min(nominal, rows - R);
		const int own = nominal - row0;
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int nS = n * S; // real floats of a row
		int voff[P];		  // bytes: per-lane part of instruction p of a period
		static_for<P>([&](auto Pi) {
			constexpr int p = Pi;
			const int u = p * 64 + lane;
			voff[p] = ((u / CW) * L + u % CW) * (int)sizeof(float);
		});
		// staged float q of a full chunk k: uniform part row0 + (q/P)*RP rows,
		// chunk start k*CW
		auto soff = [&](int k, int q) { return ((row0 + (q / P) * RP) * L + k * CW) * (int)sizeof(float); };
		// last chunk, instruction q: chunk-local float u of the R rows (valid
		// below R*CWL), its row and column; the column is clamped to the row's
		// last real float (padding voxels load a real value and are decoupled by
		// the zero factors)
		struct Tail
		{
			bool valid;
			int row, c;
		};
		// (an opaque lane index: the compiler would otherwise hoist these
		// per-lane values out of the tile loop and keep them all alive)
		auto tail = [&](int q) -> Tail {
			int ln = lane;
			asm volatile("" : "+v"(ln));
			const int u = q * 64 + ln;
			const int uu = min(u, R * CWL - 1);
			const int row = uu / CWL;
			return { u < R * CWL, row, uu - row * CWL };
		};

		float st[(NCH - 1) * NQ + NQL];
		auto load_chunk = [&](auto K) {
			constexpr int k = K;
			if constexpr (k < NCH - 1)
				static_for<NQ>([&](auto Q) {
					constexpr int q = Q;
					st[k * NQ + q] = buf_load(rsrc, voff[q % P], soff(k, q));
				});
			else if constexpr (k == NCH - 1)
				static_for<NQL>([&](auto Q) {
					constexpr int q = Q;
					const Tail tl = tail(q);
					const int col = min(k * CW + tl.c, nS - 1);
					st[k * NQ + q] = buf_load(rsrc, ((row0 + tl.row) * L + col) * (int)sizeof(float), 0);
				});
		};

		const int line = (row0 + rl) * S + sl;
		const LineKey kf = line_key(d, kRoundFwdX, (uint32_t)line);
		const LineKey ko = line_key(d, kRoundOutX, (uint32_t)line);
		const double* lb = fb + sl * fac_row(NMAX);
		const double* le = fe + sl * fac_row(NMAX);
		const double cc = fc[sl];
		float v[NMAX];
		double prev = 0;
		load_chunk(std::integral_constant<int, 0> {});
		load_chunk(std::integral_constant<int, 1> {});
		static_for<NCH>([&](auto K) {
			constexpr int k = K;
			constexpr int CV = k == NCH - 1 ? CVL : kXgChunk;
			load_chunk(std::integral_constant<int, k + 2> {});
			// chunk k into the LDS buffer, then each lane's line
			if constexpr (k < NCH - 1)
				static_for<NQ>([&](auto Q) {
					constexpr int q = Q;
					buf[loff[q % P] + (q / P) * RP * PITCH] = st[k * NQ + q];
				});
			else
				static_for<NQL>([&](auto Q) {
					constexpr int q = Q;
					const Tail tl = tail(q);
					if (tl.valid)
						buf[tl.row * PITCH + tl.c] = st[k * NQ + q];
				});
			wave_lds_fence();
			static_for<CV>([&](auto V) { v[k * kXgChunk + V] = buf[rl * PITCH + V * S + sl]; });
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
			if constexpr (k == 0)
				prev = (double)v[0];
			// forward recurrence over the chunk; the rounded values are pinned
			// (opaque) at the end of each group, or the compiler defers the
			// rounding and keeps the FP64 forward values alive
			static_for<CV / kXgGroup>([&](auto G) {
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = k * kXgChunk + G * kXgGroup + V;
					if constexpr (i > 0)
					{
						prev = (double)v[i] + le[i - 1] * prev;
						v[i] = field_round<double>(prev, kf, weyl29(i));
					}
				});
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = k * kXgChunk + G * kXgGroup + V;
					if constexpr (i > 0)
					{
						float x = v[i];
						asm volatile("" : "+v"(x));
						v[i] = x;
					}
				});
				__builtin_amdgcn_sched_barrier(0);
			});
		});

		// substitution from the end of the row; each chunk leaves through LDS
		// once its first slot is done
		int sL = L, srow0 = row0;
		asm volatile("" : "+s"(sL), "+s"(srow0));
		double nxt = (double)v[NMAX - 1] * lb[NMAX - 1];
		static_for<NCH>([&](auto KK) {
			constexpr int k = NCH - 1 - KK;
			constexpr int CV = k == NCH - 1 ? CVL : kXgChunk;
			static_for<CV / kXgGroup>([&](auto G) {
				static_for<kXgGroup>([&](auto VV) {
					constexpr int i = k * kXgChunk + CV - 1 - (G * kXgGroup + VV);
					if constexpr (i != NMAX - 1)
						nxt = ((double)v[i] + cc * nxt) * lb[i];
					buf[rl * PITCH + (i - k * kXgChunk) * S + sl] = field_round<double>(nxt, ko, weyl29(i));
				});
				__builtin_amdgcn_sched_barrier(0);
			});
			wave_lds_fence();
			// the chunk's LDS reads are issued kXgStoreBatch at a time before
			// their stores: a read followed by its store would wait out every LDS
			// round trip in turn
			if constexpr (k < NCH - 1)
			{
				static_for<(NQ + kXgStoreBatch - 1) / kXgStoreBatch>([&](auto B) {
					constexpr int q0 = B * kXgStoreBatch;
					constexpr int nb = NQ - q0 < kXgStoreBatch ? NQ - q0 : kXgStoreBatch;
					float o[nb];
					static_for<nb>([&](auto J) {
						constexpr int q = q0 + J;
						o[J] = buf[loff[q % P] + (q / P) * RP * PITCH];
					});
					if (own == 0)
						static_for<nb>([&](auto J) {
							constexpr int q = q0 + J;
							buf_store(rsrc, voff[q % P], ((srow0 + (q / P) * RP) * sL + k * CW) * (int)sizeof(float),
									  o[J]);
						});
					else
						// the last wave of the domain: only rows from `own` on
						static_for<nb>([&](auto J) {
							constexpr int q = q0 + J;
							const int u = (q % P) * 64 + lane;
							if ((q / P) * RP + u / CW >= own)
								buf_store(rsrc, voff[q % P],
										  ((srow0 + (q / P) * RP) * sL + k * CW) * (int)sizeof(float), o[J]);
						});
					__builtin_amdgcn_sched_barrier(0);
				});
			}
			else
			{
				float o[NQL];
				static_for<NQL>([&](auto Q) {
					constexpr int q = Q;
					const Tail tl = tail(q);
					o[q] = buf[tl.row * PITCH + tl.c];
				});
				__builtin_amdgcn_sched_barrier(0);
				static_for<NQL>([&](auto Q) {
					constexpr int q = Q;
					const Tail tl = tail(q);
					const int col = k * CW + tl.c;
					if (tl.valid && col < nS && tl.row >= own)
						buf_store(rsrc, ((srow0 + tl.row) * sL + col) * (int)sizeof(float), 0, o[q]);
				});
			}
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
		});
	}
}

} // namespace sweep
} // namespace bfvm
