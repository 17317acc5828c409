// bfvm_twist_kernels.h -- y and z sweeps of lines of 257..512 voxels in
// registers, two lanes per line (device template; instantiated by
// bfvm_twist_{y,z}{a,b,c,d}.hip), and of lines of 137..256 voxels, whose
// one-lane kernels need more than 256 registers (209..256) or hold fewer waves
// per SIMD than half lines do (bfvm_twist_{y,z}e.hip).
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
// A workgroup of four waves solves 128 columns: waves 0 and 2 hold the top
// halves of 64 columns each, waves 1 and 3 the bottom halves of the same
// columns, one line half per lane.  Each wave thus walks memory in one
// direction, so the row offsets are scalar.  The top half of m rows sits in
// register slots pad..NH-1 (row j - pad), the bottom half of n - m rows in
// slots 0..n-m-1 (row m + j), so both interface rows are compile-time slots
// (NH-1 and 0); the padding slots (fewer than 16) load a real row of the column and are
// never stored, and the factor table carries a 16-entry zero prefix so that
// they decouple.  The
// partner halves exchange their interface values through LDS.
#pragma once

#include "bfvm_sweep_kernels.h"

namespace bfvm {
namespace sweep {

constexpr int kTwistPrefix = 16; // zero entries before each substrate's factor row
// LDS after the factor table and c: the interface coefficients [s][4] (64
// substrates at most), then two exchange buffers of one double per lane
constexpr int kTwistIface = 4 * 64;
constexpr int kTwistExchange = kTwistIface + 2 * kSweepThreads;
// recurrence steps per scheduling group (half the column kernels' group: the
// exchange and key state of a half-line lane cost registers the line needs)
constexpr int kTwistChunk = 8;

// Everything one half-line lane needs beyond its register slots.
struct TwistLine
{
	__amdgpu_buffer_rsrc_t rsrc;
	int voff;		   // byte offset of the column's row 0
	int stride4;	   // bytes from one row to the next
	int n, m, nb;	   // line length, top rows, bottom rows
	double cc;		 // c of the substrate
	LineKey kf, ko;	 // rounding keys (forward values, results)
	// the last sweep's uniform bulk (BULK): (v + ba) * br over rows r0 ..
	// r0 + rlen - 1, rounded with the bulk stage's key of the line (kb) and row
	double ba, br;
	int r0;
	unsigned rlen;
	LineKey kb;
	// LDS element offsets (integers, so that the LDS address space stays visible)
	int tb;		// tb[k] = b'_k, tb[-16..-1] = 0
	int ifc;	// this half's interface coefficients (own d', other d')
	int mine;	// this lane's exchange slot
	int theirs; // the partner lane's
};

// The bulk stage on a final value of row r (bulk_round's bits): formed at
// every row and picked by a mask, as in the one-lane kernels.
__device__ __forceinline__ float twist_bulk(const TwistLine& L, int r, float x)
{
	const float v = field_round<double>(((double)x + L.ba) * L.br, L.kb, weyl29((uint32_t)r));
	const unsigned m = 0u - (unsigned)((unsigned)(r - L.r0) < L.rlen);
	return __builtin_bit_cast(float, (__builtin_bit_cast(unsigned, v) & m) | (__builtin_bit_cast(unsigned, x) & ~m));
}

// One half of a line: load, eliminate towards the interface, exchange, solve
// the 2x2 interface, substitute back.  Every wave reaches the one barrier.
template <bool BOTTOM, int NH, bool BULK = false>
__device__ __forceinline__ void twist_half(const TwistLine& L, double* lds)
{
	constexpr int nchunks = (NH - 1 + kTwistChunk - 1) / kTwistChunk;
	const int n = L.n, m = L.m, nb = L.nb, stride4 = L.stride4, voff = L.voff;
	const double cc = L.cc;
	float v[NH];
	double prev;
	if constexpr (!BOTTOM)
	{
		// rows 0..m-1 in slots pad..NH-1; padding slots load row 0: their
		// values never reach a real row (the factors that couple them are the
		// table's zero prefix, and their results are not stored)
		const int pad = NH - m;
		auto ld = [&](auto J) {
			constexpr int j = J;
			if constexpr (j < kTwistPrefix)
				v[j] = buf_load(L.rsrc, voff, max(j - pad, 0) * stride4);
			else
				v[j] = buf_load(L.rsrc, voff, (j - pad) * stride4);
		};
		// loads interleaved with the elimination (as in thomas_regs_streamed):
		// kLoadAhead slots first, then kTwistChunk more before each group
		static_for<(NH < kLoadAhead ? NH : kLoadAhead)>([&](auto J) { ld(J); });
		const double* lb = lds + L.tb - pad; // lb[j] = b'_{j-pad}
		prev = 0;
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int j = kLoadAhead + C * kTwistChunk + K;
				if constexpr (j < NH)
					ld(std::integral_constant<int, j> {});
			});
			if constexpr (C == 0)
				prev = (double)v[0];
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = 1 + C * kTwistChunk + K;
				if constexpr (i < NH)
				{
					prev = (double)v[i] + (cc * lb[i - 1]) * prev;
					v[i] = field_round<double>(prev, L.kf, weyl29(i));
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
	}
	else
	{
		// rows m..n-1 in slots 0..nb-1, eliminated from the last row upwards;
		// padding slots (>= nb) load row n-1 (decoupled as above)
		auto ld = [&](auto J) {
			constexpr int j = J;
			if constexpr (j >= NH - kTwistPrefix - 1)
				v[j] = buf_load(L.rsrc, voff, min(m + j, n - 1) * stride4);
			else
				v[j] = buf_load(L.rsrc, voff, (m + j) * stride4);
		};
		// loads interleaved with the elimination, from the last slot down
		static_for<(NH < kLoadAhead ? NH : kLoadAhead)>([&](auto J) { ld(std::integral_constant<int, NH - 1 - J> {}); });
		// slot i couples to slot i+1 through e_{k-1}, k = nb-1-i:
		// ub[NH-2-i] = b'_{nb-2-i}
		const double* ub = lds + L.tb - (NH - nb);
		prev = 0;
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int j = NH - 1 - (kLoadAhead + C * kTwistChunk + K);
				if constexpr (j >= 0)
					ld(std::integral_constant<int, j> {});
			});
			if constexpr (C == 0)
				prev = (double)v[NH - 1];
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = NH - 2 - (C * kTwistChunk + K);
				if constexpr (i >= 0)
				{
					prev = (double)v[i] + (cc * ub[NH - 2 - i]) * prev;
					v[i] = field_round<double>(prev, L.kf, weyl29(i));
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
	}

	// interface rows m-1 (top) and m (bottom): x_own = beta (d' + c x_other)
	// and x_other = beta_o (d'_o + c x_own) give x_own = A d' + B d'_o with
	// per-substrate A, B (see sweep_twist_kernel)
	lds[L.mine] = prev;
	__syncthreads();
	double nxt = lds[L.ifc] * prev + lds[L.ifc + 1] * lds[L.theirs];

	// substitution towards this half's end: x = (d' + c x_prev) b'.  Opaque
	// copies of the row geometry keep the compiler from computing every store
	// offset up front (they would fill the SGPRs)
	int sm = m, snb = nb, sstride4 = stride4;
	asm volatile("" : "+s"(sm), "+s"(snb), "+s"(sstride4));
	if constexpr (!BOTTOM)
	{
		const int pad = NH - sm;
		// an opaque offset (not pointer: that would turn the reads into flat
		// loads) makes the compiler read the factors again rather than keep them
		int tbo = L.tb - pad;
		asm volatile("" : "+v"(tbo));
		const double* lb = lds + tbo;
		auto out = [&](auto J, double x) {
			constexpr int j = J;
			float val = field_round<double>(x, L.ko, weyl29(j));
			if constexpr (BULK)
				val = twist_bulk(L, j - pad, val);
			if constexpr (j < kTwistPrefix) // pad <= 16: later slots are real rows
			{
				if (j >= pad)
					buf_store(L.rsrc, voff, (j - pad) * sstride4, val);
			}
			else
				buf_store(L.rsrc, voff, (j - pad) * sstride4, val);
		};
		out(std::integral_constant<int, NH - 1> {}, nxt);
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = NH - 2 - (C * kTwistChunk + K);
				if constexpr (i >= 0)
				{
					nxt = ((double)v[i] + cc * nxt) * lb[i];
					out(std::integral_constant<int, i> {}, nxt);
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
	}
	else
	{
		// slot i: x = (d'' + c x_{i-1}) b'_{nb-1-i} = db[NH-1-i]
		int tbo = L.tb - (NH - snb);
		asm volatile("" : "+v"(tbo));
		const double* db = lds + tbo;
		auto out = [&](auto J, double x) {
			constexpr int j = J;
			float val = field_round<double>(x, L.ko, weyl29(j));
			if constexpr (BULK)
				val = twist_bulk(L, sm + j, val);
			if constexpr (j >= NH - kTwistPrefix - 1) // nb >= NH - 16: earlier slots are real rows
			{
				if (j < snb)
					buf_store(L.rsrc, voff, (sm + j) * sstride4, val);
			}
			else
				buf_store(L.rsrc, voff, (sm + j) * sstride4, val);
		};
		out(std::integral_constant<int, 0> {}, nxt);
		static_for<nchunks>([&](auto C) {
			static_for<kTwistChunk>([&](auto K) {
				constexpr int i = 1 + C * kTwistChunk + K;
				if constexpr (i < NH)
				{
					nxt = ((double)v[i] + cc * nxt) * db[NH - 1 - i];
					out(std::integral_constant<int, i> {}, nxt);
				}
			});
			__builtin_amdgcn_sched_barrier(0);
		});
	}
}

// BULK (with FACES): the launch also applies the last sweep's uniform bulk
// as it stores (DomainDev::bulk_run), after the faces, with the arithmetic and
// bits of the bulk stage on the stored field.
template <int AXIS, int NH, bool FACES, bool BULK = false>
__global__ __launch_bounds__(kSweepThreads) void sweep_twist_kernel(const DomainDev* __restrict__ doms,
																	const TileRef* __restrict__ tiles, int ntiles,
																	int finish, int fac_cap)
{
	static_assert(AXIS == 1 || AXIS == 2, "twisted lines are y or z lines");
	static_assert(NH > kTwistPrefix && NH <= kRegMaxN, "half-line bucket");
	constexpr int W = fac_row(NH + kTwistPrefix); // table row per substrate (padded: see fac_row)
	extern __shared__ double lds[];
	double* fb = lds;				// [s][W]: zeros, then b'_k (the second table of the
									// layout is unused: e_k = c b'_k is formed in registers,
									// the same FP64 product as the host's)
	double* fc = lds + 2 * fac_cap; // [s]
	double* ifc = fc + 64;			// [s][4]: top (A, B), bottom (A, B); then the
									// exchange buffers [2][wave][lane]
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const bool bottom = (wave & 1) != 0;
	int cur = -1;
	int parity = 0;
	fp32_toward_zero();

	for (int t = blockIdx.x; t < ntiles; t += gridDim.x, parity ^= 1)
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
			{
				// x_T = bT (d'_T + c x_B), x_B = bB (d'_B + c x_T), bT = b'_{m-1},
				// bB = b'_{nb-1}: x_T = (bT d'_T + bT c bB d'_B) / (1 - c^2 bT bB)
				const double c = d.fc[AXIS][k];
				const int m = (n + 1) / 2, nb = n - m;
				const double bT = gb[(m - 1) * d.S + k], bB = gb[(nb - 1) * d.S + k];
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
		const int S = d.S, L = d.L;
		const int outer_n = AXIS == 1 ? d.nz : d.ny;
		// every wave takes part in the exchange barrier: columns past the end
		// repeat the last column (same inputs and keys, so the same values are
		// stored twice)
		const int f = min(tr.a + (wave >> 1) * 64 + lane, outer_n * L - 1);
		const int outer = f / L, c = f - outer * L;
		const int cs = d.coef_map ? d.coef_map[c] : c % S;
		TwistLine T;
		T.rsrc = field_rsrc(d);
		T.n = n;
		T.m = (n + 1) / 2;
		T.nb = n - T.m;
		T.stride4 = (AXIS == 1 ? L : d.ny * L) * (int)sizeof(float);
		T.voff = (AXIS == 1 ? outer * d.ny * L + c : f) * (int)sizeof(float);
		T.cc = fc[cs];
		T.tb = cs * W + kTwistPrefix; // fb = lds
		T.kf = line_key(d, 2 * AXIS, (uint32_t)(2 * f + (bottom ? 1 : 0)));
		T.ko = line_key(d, 2 * AXIS + 1, (uint32_t)(2 * f + (bottom ? 1 : 0)));
		const int xbase = 2 * fac_cap + 64 + kTwistIface + parity * kSweepThreads;
		T.ifc = 2 * fac_cap + 64 + 4 * cs + (bottom ? 2 : 0);
		T.mine = xbase + (int)threadIdx.x;
		T.theirs = xbase + ((int)threadIdx.x ^ 64);
		if constexpr (BULK)
		{
			const double* bu = finish && AXIS == d.dims - 1 ? d.bulk_u : nullptr;
			const int sub = c % S;
			const int col = f / S;
			T.ba = bu ? bu[sub] : 0.0;
			T.br = bu ? bu[S + sub] : 1.0;
			T.r0 = bu ? d.bulk_run[2 * col] : 0;
			T.rlen = bu ? (unsigned)(d.bulk_run[2 * col + 1] - T.r0) : 0u;
			T.kb = line_key(d, kRoundBulk, (uint32_t)f);
		}
		if (bottom)
			twist_half<true, NH, BULK>(T, lds);
		else
			twist_half<false, NH, BULK>(T, lds);
		auto fin = [&](int r, float x) -> float {
			if constexpr (BULK)
				return twist_bulk(T, r, x);
			else
				return x;
		};
		// Dirichlet faces after the solve (face values replace solved ones, and
		// this lane's later stores win), so the solve carries no face state:
		// row 0 / row n-1 of this half, or every row of a line lying in a face
		// of another axis
		if constexpr (FACES)
			if (finish && d.face_mask && AXIS == d.dims - 1)
			{
				const int x = c / S, s = c - (c / S) * S;
				const line_faces lf =
					AXIS == 1 ? faces_of_line(d, 1, x, 0, outer, s) : faces_of_line(d, 2, x, outer, 0, s);
				const int r0 = bottom ? T.m : 0, r1 = bottom ? n : T.m;
				if (lf.lo || lf.hi)
				{
					for (int r = r0; r < r1; r++)
						buf_store(T.rsrc, T.voff, r * T.stride4, fin(r, face_value(lf, r, n, 0.0f)));
				}
				else if (!bottom && lf.has_min)
					buf_store(T.rsrc, T.voff, 0, fin(0, lf.min_v));
				else if (bottom && lf.has_max)
					buf_store(T.rsrc, T.voff, (n - 1) * T.stride4, fin(n - 1, lf.max_v));
			}
	}
}

} // namespace sweep
} // namespace bfvm
