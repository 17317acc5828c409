// Internal interface between the kernels (bfvm_kernels.hip) and the context
// logic (bfvm_context.hip) of libbfvm_hip.so.  Not installed, not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bfvm {

constexpr int kSweepThreads = 256; // workgroup of every sweep kernel (4 wave64s)
constexpr int kLinesPerTile = 64;  // lines solved per tile: one per lane of one wave
constexpr int kFaces = 6;		   // x-min, x-max, y-min, y-max, z-min, z-max

// Everything a kernel needs to know about one domain; lives in device memory
// (one entry per domain of a launch).
struct DomainDev
{
	float* field; // ((z*ny + y)*nx + x)*S + s
	uint32_t* step; // [1] steps taken: seeds the rounding keys (field_round)
	int32_t nx, ny, nz, S, dims, L; // L = nx*S, one row
	int32_t face_mask;				// bit f: face f has a condition for some substrate
	int32_t round_nearest;			// 1: FP64 results rounded to nearest, not stochastically
	uint32_t step_offset;			// this step's place in a run of steps captured as one graph
	double dt, voxel_volume;
	int64_t bbox_min[3];
	int64_t voxel[3];

	// FP64 Thomas factors per axis: b[n*S], e[(n-1)*S], c[S]
	const double* fb[3];
	const double* fe[3];
	const double* fc[3];
	const int32_t* coef_map; // [L] coefficient substrate of a y/z column, or null (c % S)

	const float* face_val;	  // [6*S]
	const uint8_t* face_cond; // [6*S]

	// interior Dirichlet: linear voxel, values, conditions
	const int64_t* dir_vox;
	const float* dir_val;
	const uint8_t* dir_cond;
	// bulk: linear voxel, a = dt*S*T, r = 1 / (1 + dt*(U+S)) (FP64)
	const int64_t* bulk_vox;
	const double* bulk_a;
	const double* bulk_r;
	// a bulk list with the same rates at every listed voxel (per substrate)
	// whose voxels form at most one run along each line of the last axis,
	// when the last sweep applies the bulk stage as it stores its results
	// (null: the bulk stage runs over the list): a and r per substrate [2*S],
	// and each column's run [first, end) along the last axis [2*columns],
	// the column of line f being f / S
	const double* bulk_u;
	const int32_t* bulk_run;
	// bulk and cell lists sharing voxels (rebuilt with either list): the apply
	// list entry at a bulk voxel or -1, and whether an apply entry's voxel is
	// in the bulk list (stages_kernel applies such voxels from the bulk side)
	int32_t* bulk_cell; // [bulk count]
	uint8_t* cell_in_bulk; // [n_owners]

	// agents (FP64, as the reference)
	int64_t n_agents;
	const double* positions;
	const double* secretion;
	const double* saturation;
	const double* uptake;
	const double* net_export;
	double* internalized;
	const double* fraction_released;
	const double* volumes;
	const uint8_t* is_active;
	// cell stage state, rebuilt at recompute
	int32_t* cell_voxel; // [n] linear voxel, -1 inactive
	int32_t* owner_of;	 // [n] owner agent of the cell's voxel; once the owner list is packed, the owner's slot in it
	int32_t* ballot;	 // [V] owner agent per voxel (atomic mode)
	double* num;		 // [n*S] per-agent terms
	double* den;
	double* fac;
	double* rnum; // [n*S] reduced sums, indexed by owner
	double* rden;
	double* rfac;
	int32_t* owners; // [n] owner agents (the first n_owners entries)
	// apply list, packed at recompute in owner-list order k: voxel, sums, and
	// the new density at that voxel (FP64) for the per-agent update
	int32_t* pvox;	  // [n_owners]
	int32_t* slot_of; // [n] list slot of an owner agent
	double* pnum;	  // [n_owners*S]
	double* pden;
	double* pfac;
	double* dnew;
	// with a shared voxel somewhere and tracking on: per slot, whether the
	// owner's voxel holds the owner only (its agent update then runs in the
	// apply); per agent (atomic mode), whether another agent joined its voxel
	uint8_t* slot_alone; // [n]
	uint8_t* joined;	 // [n]
	// (tracking on) the agents whose voxel is shared, in agent order: the
	// per-agent update's work list, flags[3] entries
	int32_t* shared_list; // [n]
	int32_t* flags; // kCellFlags: [0] conflict, [1] agent outside grid, [2] owner count,
					// [3] shared_list entries, [4] owner list entries below their predecessor's agent
	int32_t compute_internalized;
	int32_t pad1;
};

// ---------------------------------------------------------------- rounding
// Every FP64 result written to the FP32 field (sweep outputs, the FP32-stored
// forward values of the register and LDS sweeps, bulk and cell updates) is
// rounded stochastically: up with probability equal to the discarded fraction
// of an FP32 ulp, so the rounding error has mean zero.  Round-to-nearest makes
// the same error at a voxel every step once the field sits near a fixed point
// of the step, and that bias accumulates as error / (per-step relaxation
// rate): with decay 0.23 and bulk supply 0.5 (relaxation 0.0073 per step at
// dt = 0.01) it reaches 1e-5 of the field after 1000 steps, against 2.5e-6 with
// stochastic rounding (C2-like 64^3 emulation, substrate of D = 1.66).  This is
// the stagnation of low-precision time stepping analysed by Croci and Giles,
// "Effects of round-to-nearest and stochastic rounding in the numerical solution
// of the heat equation in low precision" (IMA J. Numer. Anal., 2022).
//
// The random bits are a pure function of (domain step count, stage, line,
// position), so results are reproducible run to run and identical whether a
// domain is stepped alone or in a batch.
constexpr uint32_t kWeyl = 0x9e3779b9u;

// Rounding stages (distinct key streams within a step).
enum RoundStage : uint32_t
{
	kRoundFwdX = 0,
	kRoundOutX = 1,
	kRoundFwdY = 2,
	kRoundOutY = 3,
	kRoundFwdZ = 4,
	kRoundOutZ = 5,
	kRoundBulk = 6,
	kRoundCells = 7
};

// 32-bit integer finaliser ("lowbias32", C. Wellons' hash-prospector).
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

__device__ __forceinline__ uint32_t stage_seed(const DomainDev& d, uint32_t stage)
{
	return mix32((*d.step + d.step_offset) * 8u + stage);
}

// FP64 -> FP32 with stochastic rounding, exact in any FP32 rounding mode: the
// top 29 bits of `key` are added to the 29 mantissa bits that FP32 drops and
// the value is truncated to FP32 precision (the conversion is then exact;
// FP32 subnormals round to nearest).  Infinities and NaNs pass unchanged;
// `nearest` selects round-to-nearest.  For the sparse stages.
__device__ __forceinline__ float sr_round(double v, uint32_t key, bool nearest)
{
	const uint64_t b = __builtin_bit_cast(uint64_t, v);
	const uint64_t up = (b + (uint64_t)(key >> 3)) & ~(uint64_t)0x1fffffff;
	const bool special = ((b >> 52) & 0x7ff) == 0x7ff;
	return (float)__builtin_bit_cast(double, (nearest || special) ? b : up);
}

// The sweeps' form, three VALU instructions per value: the kernel sets the
// wave's FP32 rounding mode to round-toward-zero once (fp32_toward_zero), so
// the conversion itself truncates after the 29 random bits r are added.  The
// bits of element i of a line are k ^ (weyl29(i) & m): k is the line's hashed
// key (29 bits), weyl29(i) a compile-time Weyl sequence, m a wave-uniform
// mask, so the AND runs on the scalar unit.  Round-to-nearest (ties away from
// zero) is k = 2^28, m = 0.  An infinite result becomes a NaN.
__device__ __forceinline__ void fp32_toward_zero()
{
	__builtin_amdgcn_s_setreg(0x801, 3); // MODE[1:0] (FP32 round mode) = 3, toward zero
}

struct LineKey
{
	uint32_t k, m;
};

__device__ __forceinline__ LineKey line_key(const DomainDev& d, uint32_t stage, uint32_t line)
{
	if (d.round_nearest)
		return { 1u << 28, 0u };
	return { mix32(line ^ stage_seed(d, stage)) >> 3, 0x1fffffffu };
}

__host__ __device__ constexpr uint32_t weyl29(uint32_t i) { return (i * kWeyl) >> 3; }

// Rounding of the bulk stage's results, keyed as the last sweep keys its own:
// element e lies on line f of the last axis at position p, and its 29 random
// bits are line_key(d, kRoundBulk, f) ^ weyl29(p), so a last sweep that
// applies the bulk as it stores (sweep_cols_kernel, kFinBulk) rounds with
// field_round and gets the same bits.  Round-to-nearest is ties away from
// zero, as in the sweeps.  The truncated value converts exactly in any FP32
// rounding mode (FP32 subnormals aside).
__device__ __forceinline__ float bulk_round(const DomainDev& d, double v, int64_t e)
{
	const int64_t stride = d.dims == 3 ? (int64_t)d.ny * d.L : d.dims == 2 ? (int64_t)d.L : (int64_t)d.S;
	const uint32_t f = (uint32_t)(e % stride), p = (uint32_t)(e / stride);
	const uint32_t k = d.round_nearest ? (1u << 28) : ((mix32(f ^ stage_seed(d, kRoundBulk)) >> 3) ^ weyl29(p));
	const uint64_t b = __builtin_bit_cast(uint64_t, v);
	const bool special = ((b >> 52) & 0x7ff) == 0x7ff;
	return (float)__builtin_bit_cast(double, special ? b : (b + k) & ~(uint64_t)0x1fffffff);
}

template <typename ACC>
__device__ __forceinline__ float field_round(ACC v, LineKey key, uint32_t w)
{
	if constexpr (sizeof(ACC) == 8)
	{
		// w & m on the scalar unit: the VALU then needs one VOP2 XOR with an
		// SGPR operand (no VGPR for the constant)
		uint32_t wm;
		asm("s_and_b32 %0, %1, %2" : "=s"(wm) : "s"(w), "s"(key.m) : "scc");
		return (float)__builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) + (uint64_t)(key.k ^ wm));
	}
	else
		return (float)v; // an FP32 recurrence has rounded already
}

// One tile of a sweep.
//   x (both variants): a = first row (row = y + ny*z), b = rows in the tile (64/S)
//   y/z register variant: a = first flattened column (outer*L + x*S + s), 256 columns
//   y/z LDS variant: a = outer (z for y, y for z), b = first column, 64 columns
struct TileRef
{
	int32_t dom, a, b, pad;
};

// An entry of a sparse per-voxel stage: (domain, index in that domain's list).
struct SparseRef
{
	int32_t dom, idx;
};

// Sweep kernel variants: lines held in VGPRs (n <= kRegMaxN, templated on a
// bucket of 16) or solved inside an LDS tile (any n the LDS can hold).
enum SweepVariant
{
	kSweepLds = 0,
	kSweepReg = 1,	// one line per lane in VGPRs, read straight from HBM (x: stride S)
	kSweepScan = 2, // x only: lanes along the row, affine scans; nmax carries voxels per lane
	kSweepRows = 3, // x only: rows staged in LDS, one line per lane in VGPRs
	kSweepPipe = 4,	// register lines fed by buffer-to-LDS prefetch of the next tile (FP64)
	kSweepTwist = 5, // y/z lines of 257..512: two lanes per line, twisted elimination (FP64)
	kSweepXT = 6,	 // x, S in {4, 8}: register lines through a chunked LDS transpose (FP64)
	kSweepXG = 7,	 // x, S <= 7 but 4: the same with 4-byte accesses (FP64)
	kSweepXGT = 8	 // x rows past xg_max_n: two lanes per line, twisted, 4-byte chunks (FP64)
};

// Geometry of the pipelined sweeps (kSweepPipe): the per-wave tile buffer in
// floats, the waves per workgroup that fit the LDS beside the b' table, and
// the x rows per tile.  waves == 0: the shape does not fit (or L % 4 != 0).
struct PipeGeometry
{
	int buf_floats = 0;
	int waves = 0;
	int x_rows = 0;
	int S = 0;
	size_t lds_bytes = 0;
};
// Register bucket of the pipelined kernels: 8 lines for y/z, 32 for x.
int pipe_bucket(int axis, int n);
PipeGeometry pipe_geometry(int axis, int n, int nmax, int S, int L, int64_t lines_in_tiles);
constexpr int kScanRowsPerTile = kSweepThreads / 64; // one row per wave
// voxels per lane of the x-scan kernel for a row of nx voxels and S substrates, 0 if unsupported
__host__ inline int xscan_vpl(int nx, int S)
{
	if (S < 1 || S > 8)
		return 0;
	for (int v = 1; v <= 8; v *= 2)
		if (64 * v >= nx && S * v <= 32)
			return v;
	return 0;
}
constexpr int kRegMaxN = 256;
constexpr int kColsPerTile = kSweepThreads; // y/z register tiles: 256 flattened columns
// Register bucket: n rounded up to 16 (16..256), or 0 when n > kRegMaxN.
int reg_bucket(int n);
// Half-line bucket of the twisted kernels (lines of kRegMaxN+1 .. 2*kRegMaxN):
// ceil(n/2) rounded up to 8, or 0 outside that range.
int twist_bucket(int n);
// The same for y/z lines of kTwistShortMin+1 .. kRegMaxN: one-lane buckets
// of 224..256 need more than 256 registers (one wave per SIMD); from 144 the
// half lines fit three waves per SIMD against two.
constexpr int kTwistShortMin = 136;
int twist_short_bucket(int n);
constexpr int kTwistColsPerTile = kSweepThreads / 2; // two lanes per column
// Row bucket of the chunked x kernel: n rounded up to 8, and up to 32 when
// that leaves a last chunk of 24 voxels (whose rows do not split into 16-byte
// accesses evenly over a wave); 0 beyond kXtMaxN, where a line, its staging
// and the recurrence state exceed 256 VGPRs.
constexpr int kXtMaxN = 200;
int xt_bucket(int n);
// Row bucket of the 4-byte chunked x kernel (S = 1..8): n rounded up
// to 16, and 200 for rows of 193..200 unless S is 1, 5 or 7; 0 beyond
// xg_max_n(S) (the line, its staging, the per-lane offsets and the recurrence
// state exceed 256 VGPRs).
constexpr int kXgMaxN = 200;
__host__ __device__ constexpr int xg_max_n(int S) { return S == 1 ? 176 : (S == 5 || S == 7 ? 192 : kXgMaxN); }
int xg_bucket(int n, int S);
// Rows per wave of that kernel: as many as fit 64 lanes, even for odd S (a
// 32-voxel chunk of the rows is then a whole number of 64-float loads); a
// tile holds the rows of kSweepThreads/64 waves.
__host__ __device__ constexpr int xg_rows(int S) { return S % 2 ? (64 / S) & ~1 : 64 / S; }
// Half-row bucket of the two-lane chunked x kernel (rows of xg_max_n(S)+1 ..
// 2*xg_max_n(S) voxels): the top half holds voxels 0..NH-1, NH = ceil(n/2)
// rounded up to 16 (200 above 192), the bottom half the remaining n - NH >=
// NH - 31; 0 outside that range.  The factor table carries kXgtPrefix zeros.
int xgt_bucket(int n, int S);
constexpr int kXgtPrefix = 32;
// The same kernel on rows one lane would hold (65..xg_max_n(S) voxels), on
// request: NH = ceil(n/2) rounded up to 16 (at least 48), 0 outside that range.
int xgt_short_bucket(int n, int S);

// Row stride (doubles) of the factor tables the one-lane register kernels
// keep in LDS, [s][fac_row(NMAX)]: the bucket plus 2.  A wave's lanes read the
// rows of up to S substrates at one step; with a stride that is a multiple of
// 32 doubles (buckets of 32, 64, ..., 160, 192) every row starts in the same
// LDS bank and the reads conflict S ways, while NMAX + 2 (NMAX a multiple of
// 8) puts 8 rows in distinct banks and keeps 16-byte alignment.
__host__ __device__ constexpr int fac_row(int nmax) { return nmax + 2; }

// LDS bytes of a sweep launch of `threads` per workgroup: factor tables
// (fac_cap doubles each) plus the tile.
size_t sweep_lds_bytes(int axis, int variant, int n, int S, int L, int fac_cap, int threads = kSweepThreads);
// Workgroup size of a sweep variant: `lines_per_block` (64, 128, 256; 0 = 256)
// for the one-lane register kernels (column lines, chunked x lines), 256 for
// the others.
inline int sweep_threads(int variant, int lines_per_block)
{
	const bool one_lane = variant == kSweepReg || variant == kSweepXT || variant == kSweepXG;
	return one_lane && lines_per_block > 0 ? lines_per_block : kSweepThreads;
}
// x tiles: 64/S rows; LDS row pitch Lp = L + pad with Lp = S (mod 64), so the
// 64 lanes (r, s) reading element i of their lines hit distinct banks.
__host__ __device__ inline int x_row_pitch(int L, int S) { return L + ((S - L % 64) % 64 + 64) % 64; }
constexpr int kMaxLdsBytes = 160 * 1024;
// Row pitch of the pipelined x tiles: a multiple of 4 floats (the buffer-to-LDS
// loads write 16 bytes per lane) congruent to 4*ceil(S/4) mod 64, so the lanes
// (row, substrate) reading their lines hit distinct banks when S % 4 == 0.
__host__ __device__ inline int pipe_x_pitch(int L, int S)
{
	const int T = 4 * ((S + 3) / 4) % 64;
	return L + ((T - L % 64) % 64 + 64) % 64;
}
// Rows of an x tile staged in LDS: 64/S (one lane per (row, substrate)), fewer
// when the rows would not fit beside the factor tables (fac_cap entries each of
// b', e, plus 64 c values, all FP64); 0 when not even one row fits.
inline int x_rows_per_tile(int S, int L, int fac_cap)
{
	const long long fac = (2LL * fac_cap + 64) * (long long)sizeof(double);
	const long long row = (long long)x_row_pitch(L, S) * (long long)sizeof(float);
	const long long fit = (kMaxLdsBytes - fac) / row;
	const int want = S >= kLinesPerTile ? 1 : kLinesPerTile / S;
	return fit <= 0 ? 0 : (int)(fit < want ? fit : want);
}

// Raises the dynamic-LDS limit of the sweep kernels to 160 KiB (once per process).
hipError_t init_sweep_kernels();
hipError_t launch_sweep_pipe(int axis, int nmax, bool faces, const DomainDev* doms, const TileRef* tiles, int ntiles,
							 int finish, int fac_cap, const PipeGeometry& g, int grid, hipStream_t stream);

// ---- launchers (enqueue on `stream`, return hipError_t) ----
// fac_cap: doubles reserved per factor table in LDS (max n*S over the domains).
// finish: apply face Dirichlet conditions in the last sweep of each domain.
// fin: the instantiation's finishing work in the last sweep: kFinNone,
// kFinFaces (face conditions), kFinBulk (face conditions, then the bulk of
// DomainDev::bulk_u on each stored value; FP64 register y/z kernels only).
// S: substrate count (selects the x-scan instantiation; all its domains share it).
constexpr int kFinNone = 0, kFinFaces = 1, kFinBulk = 2;
hipError_t launch_sweep(int axis, int variant, int nmax, bool fp64, int fin, int S, const DomainDev* doms,
						const TileRef* tiles, int ntiles, int finish, int fac_cap, size_t lds_bytes, int grid,
						hipStream_t stream, int threads = kSweepThreads);
hipError_t launch_dirichlet_interior(const DomainDev* doms, const SparseRef* refs, int64_t n, hipStream_t stream);
// The per-step sparse stages run one thread per (entry, substrate); `S` is the
// substrate count of every referenced domain (batched launches group by S).
// Sparse stages over tiles (dom, first entry, count <= kSparseTile / S), one
// 256-thread workgroup a tile.
constexpr int kSparseTile = 256;
hipError_t launch_bulk(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream);
// Ends a step, or a captured run of `by` steps, of `ndoms` domains: advances
// each one's step count by `by`.
hipError_t launch_step_tick(const DomainDev* doms, int ndoms, uint32_t by, hipStream_t stream);
hipError_t launch_cells_apply(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream);
hipError_t launch_cells_internalized(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream);
// Bulk and cell apply of one domain in one launch (bulk tiles first), once its
// overlap maps are current (launch_bulk_overlap).
hipError_t launch_stages(const DomainDev* dom, const TileRef* bulk_tiles, int n_bulk_tiles, const TileRef* cell_tiles,
						 int n_cell_tiles, hipStream_t stream);
// Voxel -> bulk entry map ([V], -1 elsewhere) of the bulk list.
hipError_t launch_bulk_map(const int64_t* bulk_vox, int64_t count, int32_t* map, int64_t V, hipStream_t stream);
// The domain's overlap maps (bulk_cell: the domain's, on the device) from the
// voxel map and its apply list.
// n: the agent count (the owner list's capacity; the owner count is read on the device)
hipError_t launch_bulk_overlap(const DomainDev* dom, const int32_t* map, int32_t* bulk_cell, int64_t bulk_count,
							   int64_t n, hipStream_t stream);
hipError_t launch_fill_field(float* field, const float* init, int S, int64_t nvox, hipStream_t stream);
hipError_t launch_copy_field(const float* src, float* dst, int64_t n, int num_cu, hipStream_t stream);
hipError_t launch_convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t stream);
hipError_t launch_convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t stream);
hipError_t launch_release_flush(const DomainDev* dom, const int64_t* groups, int64_t n_groups, const int64_t* agents,
								int S, hipStream_t stream);

// cell recompute (single domain)
hipError_t launch_cells_bin(const DomainDev* dom, int64_t n, int S, bool clear_ballots, hipStream_t stream);
hipError_t launch_cells_ballot(const DomainDev* dom, int64_t n, hipStream_t stream);
// packs the apply list from owners[0..flags[2]) (both modes); n: the agent
// count; own_terms: the sums hold the other agents' terms only (atomic mode)
hipError_t launch_cells_pack(const DomainDev* dom, int64_t n, int S, bool own_terms, hipStream_t stream);
// After the pack: owner_of[i] becomes the slot of agent i's owner (-1 stays).
hipError_t launch_cells_agent_slots(const DomainDev* dom, int64_t n, hipStream_t stream);
constexpr int kCellFlags = 8; // entries of DomainDev::flags
hipError_t launch_cells_owner_order(const DomainDev* dom, int64_t n, hipStream_t stream);
// pos: n entries of scratch (the scan), tmp: scan_tmp_entries(n)
hipError_t launch_cells_shared_list(const DomainDev* dom, int64_t n, int32_t* pos, int32_t* tmp, hipStream_t stream);
// atomic mode: keys/vals = (voxel, agent) of each owner, key V for the other
// agents, to sort the owners into voxel order (radix_sort_pairs); then the
// owner count (flags[2]) from the sorted keys
hipError_t launch_cells_owner_keys(const DomainDev* dom, uint32_t* keys, uint32_t* vals, int64_t n,
								   hipStream_t stream);
hipError_t launch_cells_count_owners(const uint32_t* keys, int64_t n, uint32_t V, int32_t* flags, hipStream_t stream);
// sorted mode: keys/vals are scratch of n entries (sorted in place by radix_sort_pairs)
hipError_t launch_cells_make_keys(const DomainDev* dom, uint32_t* keys, uint32_t* vals, int64_t n,
								  hipStream_t stream);
hipError_t launch_cells_heads(const uint32_t* keys, int32_t* heads, int64_t n, uint32_t V, hipStream_t stream);
hipError_t launch_cells_segment_reduce(const DomainDev* dom, const uint32_t* keys, const uint32_t* vals,
									   const int32_t* heads, const int32_t* head_scan, int64_t n,
									   hipStream_t stream);

// LSD radix sort of (key, value) pairs on the low `bits` bits of the keys; stable.
// tmp_keys/tmp_vals: n entries; hist: radix_hist_entries(n) ints.
size_t radix_hist_entries(int64_t n);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys, uint32_t* tmp_vals, int32_t* hist,
							int32_t* scan_tmp, int64_t n, int bits, hipStream_t stream);
// Exclusive scan of n int32 in place; tmp needs scan_tmp_entries(n) ints.
size_t scan_tmp_entries(int64_t n);
hipError_t exclusive_scan(int32_t* data, int32_t* tmp, int64_t n, hipStream_t stream);

// FP64 per-substrate sum of the field (fixed order); partial: S*mass_blocks(nvox) doubles.
int mass_blocks(int64_t nvox);
hipError_t launch_mass(const float* field, int S, int64_t nvox, double* partial, double* out, double scale,
					   hipStream_t stream);

} // namespace bfvm
