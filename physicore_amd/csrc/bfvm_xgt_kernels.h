// bfvm_xgt_kernels.h -- x sweep of rows of 201..400 voxels with 1 to 8
// substrates: two lanes per line (device template; instantiated by
// bfvm_sweeps_xgt*.hip).
//
// The 4-byte chunked x kernel (bfvm_xg_kernels.h) holds a whole row in one
// lane's registers, 200 voxels at most.  Here two waves share R rows: the top
// wave holds voxels 0..NH-1 of each line, the bottom wave voxels NH..n-1
// (n - NH >= NH - 31 of them, padding slots after them).  Each loads and
// stores its half with the chunked transposes of the one-lane kernel (a half
// is a row segment starting at voxel 0 or NH).  The elimination is the
// twisted one of the long y/z lines (bfvm_twist_kernels.h): the top half
// eliminates downwards with b'_i, the bottom half upwards with the same table
// read from its end (the matrix is persymmetric), the two interface values are
// exchanged through LDS (one barrier per tile, which every wave reaches) and
// combined with per-substrate coefficients, and each half substitutes back
// from the interface.  The factor table carries a zero prefix of kXgtPrefix
// entries, which decouples the bottom half's padding slots.
#pragma once

#include "bfvm_xg_kernels.h"

namespace bfvm {
namespace sweep {

// LDS after the factor table and c: interface coefficients [s][4], then two
// exchange buffers of one double per lane
constexpr int kXgtIface = 4 * 64;
constexpr int kXgtExchange = kXgtIface + 2 * kSweepThreads;
constexpr int kXgtStoreBatch = 4; // LDS reads in flight ahead of their stores

template <int NH, int S>
__global__ __launch_bounds__(kSweepThreads, 2) void sweep_xgt_kernel(const DomainDev* __restrict__ doms,
																  const TileRef* __restrict__ tiles, int ntiles,
																  int finish, int fac_cap)
{
	(void)finish;
	static_assert(S >= 1 && S <= 8, "4-byte transposes");
	static_assert(NH % 16 == 0 || NH == kXgMaxN, "half-row bucket");
	static_assert(NH > kXgtPrefix && NH <= xg_max_n(S), "half-row bucket");
	constexpr int R = xg_rows(S);
	constexpr int PITCH = 33 * S;
	constexpr int CW = kXgChunk * S;
	constexpr int P = CW / xg_gcd(CW, 64);
	constexpr int RP = 64 / xg_gcd(CW, 64);
	constexpr int NQ = R * CW / 64;
	constexpr int NCH = (NH + kXgChunk - 1) / kXgChunk;
	constexpr int W = fac_row(NH + kXgtPrefix); // factor table row per substrate (padded: see fac_row)
	static_assert(NQ % P == 0 && R % RP == 0, "whole periods per chunk");
	// chunk k: voxels of the half, and whether its loads/stores take the
	// per-lane path (a partial chunk, or one that can reach the bottom half's
	// padding: the bottom half has at least NH - 31 real voxels)
	auto cvox = [](int k) __attribute__((always_inline)) { return NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk; };
	auto nq_of = [&](int k) __attribute__((always_inline)) { return (R * cvox(k) * S + 63) / 64; };
	extern __shared__ double lds[];
	double* fb = lds;				// [s][W]: zeros, then b'_k
	double* fc = lds + 2 * fac_cap; // [s]
	double* ifc = fc + 64;			// [s][4]: top (A, B), bottom (A, B); then exchange [2][wave][lane]
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const bool bottom = (wave & 1) != 0;
	float* buf = reinterpret_cast<float*>(lds + 2 * fac_cap + 64 + kXgtExchange) + wave * xg_buffer_floats(S);
	const int rl = min(lane / S, R - 1), sl = lane - (lane / S) * S;
	int loff[P];
	static_for<P>([&](auto Pi) {
		constexpr int p = Pi;
		const int u = p * 64 + lane;
		loff[p] = (u / CW) * PITCH + u % CW;
	});
	int cur = -1;
	int parity = 0;
	fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x, parity ^= 1)
	{
		const TileRef tr = tiles[t];
		const DomainDev& d = doms[tr.dom];
		const int n = d.nx;
		if (tr.dom != cur)
		{
			__syncthreads();
			const double* gb = d.fb[0];
			for (int k = threadIdx.x; k < W * S; k += blockDim.x)
			{
				const int s = k / W, i = k - s * W - kXgtPrefix;
				fb[k] = (i >= 0 && i < n) ? gb[i * S + s] : 0.0;
			}
			for (int k = threadIdx.x; k < S; k += blockDim.x)
			{
				// x_T = bT (d'_T + c x_B), x_B = bB (d'_B + c x_T), bT = b'_{NH-1},
				// bB = b'_{nb-1}: x_T = (bT d'_T + bT c bB d'_B) / (1 - c^2 bT bB)
				const double c = d.fc[0][k];
				const int nb = n - NH;
				const double bT = gb[(NH - 1) * S + k], bB = gb[(nb - 1) * S + k];
				const double den = 1.0 - c * c * bT * bB;
				fc[k] = c;
				ifc[4 * k + 0] = bT / den;
				ifc[4 * k + 1] = bT * c * bB / den;
				ifc[4 * k + 2] = bB / den;
				ifc[4 * k + 3] = bB * c * bT / den;
			}
			__syncthreads();
			cur = tr.dom;
		}
		const int L = d.L;
		const int rows = d.ny * d.nz;
		const int nb = n - NH; // real voxels of the bottom half
		// rows of this wave pair: R from row0 (a tile holds two pairs).  Every
		// wave reaches the exchange barrier: a pair past the end, or the last
		// pair of a domain reaching past it, works on the last R rows and
		// stores only the rows from `own` on (none when past the end)
		const int nominal = tr.a + (wave >> 1) * R;
		const int row0 = min(nominal, rows - R);
		const int own = nominal - row0;
		const __amdgpu_buffer_rsrc_t rsrc = field_rsrc(d);
		const int nS = n * S;
		// each half is its own code path (a runtime branch between the halves
		// inside one path would keep both halves' state alive)
		auto half = [&](auto BOT) __attribute__((always_inline)) {
		constexpr bool bottom = BOT;
		const int hoff = bottom ? NH * S : 0; // first float of this half in a row
		int voff[P];
		static_for<P>([&](auto Pi) {
			constexpr int p = Pi;
			const int u = p * 64 + lane;
			voff[p] = ((u / CW) * L + u % CW) * (int)sizeof(float);
		});
		// per-lane path of chunk k (CV voxels): float u of the chunk's R rows,
		// its row and column (the column clamped to the row's last real float)
		struct Slot
		{
			bool valid;
			int row, c;
		};
		auto slot = [&](int q, int cwk) __attribute__((always_inline)) -> Slot {
			int ln = lane;
			asm volatile("" : "+v"(ln));
			const int u = q * 64 + ln;
			const int uu = min(u, R * cwk - 1);
			const int row = uu / cwk;
			return { u < R * cwk, row, uu - row * cwk };
		};
		// the per-lane path is taken by partial chunks and, in the bottom half,
		// by chunks that can reach its padding
		auto generic = [&](int k) __attribute__((always_inline)) { return cvox(k) < kXgChunk || (k + 1) * kXgChunk > NH - 31; };

		constexpr int NST = [] {
			int s = 0;
			for (int k = 0; k < NCH; k++)
				s += (R * (NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk) * S + 63) / 64;
			return s;
		}();
		float st[NST];
		auto base_of = [&](int k) __attribute__((always_inline)) {
			int s = 0;
			for (int j = 0; j < k; j++)
				s += nq_of(j);
			return s;
		};
		auto load_chunk = [&](auto K) __attribute__((always_inline)) {
			constexpr int k = K;
			if constexpr (k < NCH)
			{
				constexpr int CV = NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk;
				constexpr int NQK = (R * CV * S + 63) / 64;
				constexpr int b0 = [] {
					int s = 0;
					for (int j = 0; j < k; j++)
						s += (R * (NH - j * kXgChunk < kXgChunk ? NH - j * kXgChunk : kXgChunk) * S + 63) / 64;
					return s;
				}();
				if (!bottom || !generic(k))
				{
					if constexpr (CV == kXgChunk)
						static_for<NQK>([&](auto Q) {
							constexpr int q = Q;
							st[b0 + q] = buf_load(rsrc, voff[q % P],
												  ((row0 + (q / P) * RP) * L + hoff + k * CW) * (int)sizeof(float));
						});
					else
						static_for<NQK>([&](auto Q) {
							constexpr int q = Q;
							const Slot sl2 = slot(q, CV * S);
							st[b0 + q] =
								buf_load(rsrc, ((row0 + sl2.row) * L + hoff + k * CW + sl2.c) * (int)sizeof(float), 0);
						});
				}
				else
					static_for<NQK>([&](auto Q) {
						constexpr int q = Q;
						const Slot sl2 = slot(q, CV * S);
						const int col = min(hoff + k * CW + sl2.c, nS - 1);
						st[b0 + q] = buf_load(rsrc, ((row0 + sl2.row) * L + col) * (int)sizeof(float), 0);
					});
			}
		};
		(void)base_of;

		const int line = (row0 + rl) * S + sl;
		const LineKey kf = line_key(d, kRoundFwdX, (uint32_t)(2 * line + (bottom ? 1 : 0)));
		const LineKey ko = line_key(d, kRoundOutX, (uint32_t)(2 * line + (bottom ? 1 : 0)));
		const double cc = fc[sl];
		const int tb = sl * W + kXgtPrefix; // lds[tb + i] = b'_i, zeros below i = 0
		float v[NH];
		load_chunk(std::integral_constant<int, 0> {});
		load_chunk(std::integral_constant<int, 1> {});
		static_for<NCH>([&](auto K) {
			constexpr int k = K;
			constexpr int CV = NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk;
			constexpr int NQK = (R * CV * S + 63) / 64;
			constexpr int b0 = [] {
				int s = 0;
				for (int j = 0; j < k; j++)
					s += (R * (NH - j * kXgChunk < kXgChunk ? NH - j * kXgChunk : kXgChunk) * S + 63) / 64;
				return s;
			}();
			load_chunk(std::integral_constant<int, k + 2> {});
			if constexpr (CV == kXgChunk)
				static_for<NQK>([&](auto Q) {
					constexpr int q = Q;
					buf[loff[q % P] + (q / P) * RP * PITCH] = st[b0 + q];
				});
			else
				static_for<NQK>([&](auto Q) {
					constexpr int q = Q;
					const Slot sl2 = slot(q, CV * S);
					if (sl2.valid)
						buf[sl2.row * PITCH + sl2.c] = st[b0 + q];
				});
			wave_lds_fence();
			static_for<CV>([&](auto V) { v[k * kXgChunk + V] = buf[rl * PITCH + V * S + sl]; });
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
		});

		// elimination towards the interface (top: slot NH-1, bottom: slot 0);
		// rounded values pinned per group, as in the one-lane kernel
		double prev;
		auto pin = [&](auto I) __attribute__((always_inline)) {
			constexpr int i = I;
			float x = v[i];
			asm volatile("" : "+v"(x));
			v[i] = x;
		};
		if constexpr (!bottom)
		{
			prev = (double)v[0];
			static_for<(NH - 1 + kXgGroup - 1) / kXgGroup>([&](auto G) {
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = 1 + G * kXgGroup + V;
					if constexpr (i < NH)
					{
						prev = (double)v[i] + (cc * lds[tb + i - 1]) * prev;
						v[i] = field_round<double>(prev, kf, weyl29(i));
					}
				});
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = 1 + G * kXgGroup + V;
					if constexpr (i < NH)
						pin(std::integral_constant<int, i> {});
				});
				__builtin_amdgcn_sched_barrier(0);
			});
		}
		else
		{
			// slot i couples to slot i+1 through e_{k-1}, k = nb-1-i:
			// ub[NH-2-i] = b'_{nb-2-i}; slots >= nb meet the zero prefix
			const int ub = tb - (NH - nb);
			prev = (double)v[NH - 1];
			static_for<(NH - 1 + kXgGroup - 1) / kXgGroup>([&](auto G) {
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = NH - 2 - (G * kXgGroup + V);
					if constexpr (i >= 0)
					{
						prev = (double)v[i] + (cc * lds[ub + NH - 2 - i]) * prev;
						v[i] = field_round<double>(prev, kf, weyl29(i));
					}
				});
				static_for<kXgGroup>([&](auto V) {
					constexpr int i = NH - 2 - (G * kXgGroup + V);
					if constexpr (i >= 0)
						pin(std::integral_constant<int, i> {});
				});
				__builtin_amdgcn_sched_barrier(0);
			});
		}
		// interface: x_own = A d'_own + B d'_other
		const int xbase = 2 * fac_cap + 64 + kXgtIface + parity * kSweepThreads;
		lds[xbase + (int)threadIdx.x] = prev;
		__syncthreads();
		const int fi = 2 * fac_cap + 64 + 4 * sl + (bottom ? 2 : 0);
		double nxt = lds[fi] * prev + lds[fi + 1] * lds[xbase + ((int)threadIdx.x ^ 64)];

		int sL = L, srow0 = row0;
		asm volatile("" : "+s"(sL), "+s"(srow0));
		// stores of chunk k from the LDS buffer (reads ahead of their stores)
		auto store_chunk = [&](auto K) __attribute__((always_inline)) {
			constexpr int k = K;
			constexpr int CV = NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk;
			constexpr int NQK = (R * CV * S + 63) / 64;
			wave_lds_fence();
			if (CV == kXgChunk && (!bottom || !generic(k)))
			{
				static_for<(NQK + kXgtStoreBatch - 1) / kXgtStoreBatch>([&](auto B) {
					constexpr int q0 = B * kXgtStoreBatch;
					constexpr int nbt = NQK - q0 < kXgtStoreBatch ? NQK - q0 : kXgtStoreBatch;
					float o[nbt];
					static_for<nbt>([&](auto J) {
						constexpr int q = q0 + J;
						o[J] = buf[loff[q % P] + (q / P) * RP * PITCH];
					});
					static_for<nbt>([&](auto J) {
						constexpr int q = q0 + J;
						const int u = (q % P) * 64 + lane;
						if ((q / P) * RP + u / CW >= own)
							buf_store(rsrc, voff[q % P],
									  ((srow0 + (q / P) * RP) * sL + hoff + k * CW) * (int)sizeof(float), o[J]);
					});
					__builtin_amdgcn_sched_barrier(0);
				});
			}
			else
			{
				static_for<(NQK + kXgtStoreBatch - 1) / kXgtStoreBatch>([&](auto B) {
					constexpr int q0 = B * kXgtStoreBatch;
					constexpr int nbt = NQK - q0 < kXgtStoreBatch ? NQK - q0 : kXgtStoreBatch;
					float o[nbt];
					static_for<nbt>([&](auto J) {
						constexpr int q = q0 + J;
						const Slot sl2 = slot(q, CV * S);
						o[J] = buf[sl2.row * PITCH + sl2.c];
					});
					static_for<nbt>([&](auto J) {
						constexpr int q = q0 + J;
						const Slot sl2 = slot(q, CV * S);
						const int col = hoff + k * CW + sl2.c;
						if (sl2.valid && col < nS && sl2.row >= own)
							buf_store(rsrc, ((srow0 + sl2.row) * sL + col) * (int)sizeof(float), 0, o[J]);
					});
					__builtin_amdgcn_sched_barrier(0);
				});
			}
			wave_lds_fence();
			__builtin_amdgcn_sched_barrier(0);
		};

		if constexpr (!bottom)
		{
			// x = (d' + c x_next) b', from slot NH-1 down; each chunk leaves once
			// its first slot is done
			static_for<NCH>([&](auto KK) {
				constexpr int k = NCH - 1 - KK;
				constexpr int CV = NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk;
				static_for<CV / kXgGroup>([&](auto G) {
					static_for<kXgGroup>([&](auto VV) {
						constexpr int i = k * kXgChunk + CV - 1 - (G * kXgGroup + VV);
						if constexpr (i != NH - 1)
							nxt = ((double)v[i] + cc * nxt) * lds[tb + i];
						buf[rl * PITCH + (i - k * kXgChunk) * S + sl] = field_round<double>(nxt, ko, weyl29(i));
					});
					__builtin_amdgcn_sched_barrier(0);
				});
				store_chunk(std::integral_constant<int, k> {});
			});
		}
		else
		{
			// x = (d'' + c x_prev) b'_{nb-1-i}, from slot 0 up; each chunk leaves
			// once its last slot is done
			const int db = tb - (NH - nb); // db[NH-1-i] = b'_{nb-1-i}
			static_for<NCH>([&](auto K) {
				constexpr int k = K;
				constexpr int CV = NH - k * kXgChunk < kXgChunk ? NH - k * kXgChunk : kXgChunk;
				static_for<CV / kXgGroup>([&](auto G) {
					static_for<kXgGroup>([&](auto VV) {
						constexpr int i = k * kXgChunk + G * kXgGroup + VV;
						if constexpr (i != 0)
							nxt = ((double)v[i] + cc * nxt) * lds[db + NH - 1 - i];
						buf[rl * PITCH + (i - k * kXgChunk) * S + sl] = field_round<double>(nxt, ko, weyl29(i));
					});
					__builtin_amdgcn_sched_barrier(0);
				});
				store_chunk(std::integral_constant<int, k> {});
			});
		}
		};
		if (bottom)
			half(std::true_type {});
		else
			half(std::false_type {});
	}
}

} // namespace sweep
} // namespace bfvm
