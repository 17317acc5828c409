// bfvm_kernels.hip -- the non-sweep kernels of the BioFVM backend on gfx950:
// interior Dirichlet, bulk, the cell stage (recompute by CAS + wavefront-
// aggregated atomics or by radix sort + segmented reduce, per-step apply,
// internalized bookkeeping, release), field fill/convert, the FP64 mass
// reduction, and the scan / radix sort they use.  The sweeps are in
// bfvm_sweeps.hip.  Step order of the reference
// (reactions-diffusion/biofvm/kernels/openmp_solver/src/openmp_solver.cpp:23-40):
// diffusion -> Dirichlet -> bulk -> cells.
#include "bfvm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace bfvm {

namespace {

constexpr int kWave = 64;

// Interior Dirichlet voxels (dirichlet_solver.cpp:22-42).
__global__ void dirichlet_interior_kernel(const DomainDev* __restrict__ doms, const SparseRef* __restrict__ refs,
										  int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const SparseRef r = refs[k];
	const DomainDev& d = doms[r.dom];
	const int S = d.S;
	float* D = d.field + d.dir_vox[r.idx] * S;
	for (int s = 0; s < S; s++)
		if (d.dir_cond[(int64_t)r.idx * S + s])
			D[s] = d.dir_val[(int64_t)r.idx * S + s];
}

// Bulk supply/uptake (bulk_solver.cpp:12-41): D = (D + dt*S*T) / (1 + dt*(U+S)),
// as D = (D + a) * r with a = dt*S*T and r = 1 / (1 + dt*(U+S)) precomputed in
// FP64 on the host (within an FP64 ulp of the quotient; the division leaves the
// per-step path, where its latency chain showed).  One thread per (listed
// voxel, substrate).
// Sparse stages (bulk, cell apply, per-agent internalized) run over tiles of
// up to kSparseTile / S entries of one domain: tile = (dom, first entry, count),
// one workgroup per tile, so the domain's descriptor is read once into SGPRs.
__global__ void bulk_kernel(const DomainDev* __restrict__ doms, const TileRef* __restrict__ tiles)
{
	const TileRef tr = tiles[blockIdx.x];
	const DomainDev& d = doms[tr.dom];
	const int S = d.S;
	for (int t = threadIdx.x; t < tr.b * S; t += blockDim.x)
	{
		const int k = tr.a + t / S, s = t % S;
		const int64_t q = (int64_t)k * S + s;
		const int64_t e = d.bulk_vox[k] * S + s;
		float* D = d.field + e;
		*D = bulk_round(d, ((double)*D + d.bulk_a[q]) * d.bulk_r[q], e);
	}
}

// Advances the step count of every domain of a launch by `by` (the last kernel
// of a step, or of a captured run of `by` steps).
__global__ void step_tick_kernel(const DomainDev* __restrict__ doms, int ndoms, uint32_t by)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ndoms)
		doms[i].step[0] += by;
}

// ---------------------------------------------------------------- cells

#pragma clang fp contract(off)

// compute_intermediates + the clearing half of clear_ballots
// (openmp_solver/src/cell_solver.cpp:22-77); voxel of the agent as
// cartesian_mesh::voxel_position (reactions-diffusion/biofvm/src/mesh.cpp:37-61).
// One thread per (agent, substrate), so the per-agent arrays [i*S + s] are read
// and written coalesced; the substrate-0 thread of each agent bins it.
__global__ void cells_bin_kernel(const DomainDev* __restrict__ dp, int64_t n, int clear_ballots)
{
	const DomainDev& d = *dp;
	const int S = d.S;
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n * S)
		return;
	const int64_t i = k / S;
	const double V = d.voxel_volume, dt = d.dt, vol = d.volumes[i];
	d.num[k] = d.secretion[k] * d.saturation[k] * dt * vol / V;
	d.den[k] = (d.uptake[k] + d.secretion[k]) * dt * vol / V;
	d.fac[k] = d.net_export[k] * dt / V;
	d.rnum[k] = 0;
	d.rden[k] = 0;
	d.rfac[k] = 0;
	if (k != i * S)
		return;
	d.joined[i] = 0;
	if (!d.is_active[i])
	{
		d.cell_voxel[i] = -1;
		d.owner_of[i] = -1;
		return;
	}
	const int64_t nn[3] = { d.nx, d.ny, d.nz };
	int64_t idx[3] = { 0, 0, 0 };
	for (int a = 0; a < d.dims; a++)
	{
		const double q = (d.positions[i * d.dims + a] - (double)d.bbox_min[a]) / (double)d.voxel[a];
		if (!(q >= 0.0) || q >= (double)nn[a])
		{
			d.flags[1] = 1;
			d.cell_voxel[i] = -1;
			d.owner_of[i] = -1;
			return;
		}
		idx[a] = (int64_t)q;
	}
	const int64_t v = idx[0] + nn[0] * (idx[1] + nn[1] * idx[2]);
	d.cell_voxel[i] = (int32_t)v;
	if (clear_ballots)
		d.ballot[v] = -1;
}

// ballot_and_sum (cell_solver.cpp:79-129) with wavefront aggregation: lanes
// holding the same voxel in a run of consecutive lanes elect one CAS and
// reduce their terms with a segmented scan, so one FP64 atomic per run and
// substrate reaches memory.  The owner's own terms (and the 1 of the
// denominator) are not added here: the pack adds them to what the other
// agents of the voxel accumulated, so a voxel held by one agent costs no
// atomic at all.
__global__ __launch_bounds__(256) void cells_ballot_kernel(const DomainDev* __restrict__ dp, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const DomainDev& d = *dp;
	const int lane = threadIdx.x & (kWave - 1);
	const int key = (i < n) ? d.cell_voxel[i] : -1;
	const int prev_key = __shfl_up(key, 1);
	const bool head = key < 0 || lane == 0 || prev_key != key;
	int hp = head ? lane : 0;
	for (int off = 1; off < kWave; off <<= 1)
	{
		const int o = __shfl_up(hp, off);
		if (lane >= off)
			hp = max(hp, o);
	}
	const int seg = hp; // first lane of my run
	const int next_seg = __shfl_down(hp, 1);
	const bool tail = lane == kWave - 1 || next_seg != seg;

	int owner = -1;
	int won = 0;
	if (key >= 0 && head)
	{
		const int old = atomicCAS(&d.ballot[key], -1, (int)i);
		owner = old == -1 ? (int)i : old;
		won = old == -1;
		if (!won)
			d.flags[0] = 1;
	}
	owner = __shfl(owner, seg);
	won = __shfl(won, seg);
	if (key >= 0)
	{
		d.owner_of[i] = owner;
		if (!head)
			d.flags[0] = 1;
		if (owner != (int)i)
			d.joined[owner] = 1; // every writer stores 1
	}
	const int S = d.S;
	const bool self = key >= 0 && head && won; // this lane is the owner
	const bool alone = won && lane == seg;	   // (at the tail) a run of the owner only
	for (int s = 0; s < S; s++)
	{
		double vn = 0, vd = 0, vf = 0;
		if (key >= 0 && !self)
		{
			vn = d.num[i * S + s];
			vd = d.den[i * S + s];
			vf = d.fac[i * S + s];
		}
		for (int off = 1; off < kWave; off <<= 1)
		{
			const double on = __shfl_up(vn, off), od = __shfl_up(vd, off), of = __shfl_up(vf, off);
			if (lane - off >= seg)
			{
				vn += on;
				vd += od;
				vf += of;
			}
		}
		if (key >= 0 && tail && !alone)
		{
			atomicAdd(&d.rnum[(int64_t)owner * S + s], vn);
			atomicAdd(&d.rden[(int64_t)owner * S + s], vd);
			atomicAdd(&d.rfac[(int64_t)owner * S + s], vf);
		}
	}
}

__global__ void cells_make_keys_kernel(const DomainDev* __restrict__ dp, uint32_t* keys, uint32_t* vals, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const DomainDev& d = *dp;
	const int v = d.cell_voxel[i];
	const uint32_t V = (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz;
	keys[i] = v >= 0 ? (uint32_t)v : V;
	vals[i] = (uint32_t)i;
}

__global__ void cells_heads_kernel(const uint32_t* __restrict__ keys, int32_t* heads, int64_t n, uint32_t V)
{
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	const uint32_t k = keys[p];
	heads[p] = (k < V && (p == 0 || keys[p - 1] != k)) ? 1 : 0;
}

// Sorted reduce: one thread per voxel segment sums its agents in ascending
// agent order, the owner (lowest index) first with its den + 1 -- the order
// in which a sequential run of ballot_and_sum adds them.  Bit-reproducible.
__global__ void cells_segment_reduce_kernel(const DomainDev* __restrict__ dp, const uint32_t* __restrict__ keys,
											const uint32_t* __restrict__ vals, const int32_t* __restrict__ heads,
											const int32_t* __restrict__ head_scan, int64_t n)
{
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	const DomainDev& d = *dp;
	if (p == n - 1)
		d.flags[2] = head_scan[p] + heads[p];
	if (!heads[p])
		return;
	const uint32_t key = keys[p];
	const int owner = (int)vals[p];
	d.owners[head_scan[p]] = owner;
	const int S = d.S;
	int64_t q1 = p + 1;
	while (q1 < n && keys[q1] == key)
		q1++;
	if (q1 - p > 1)
		d.flags[0] = 1;
	d.slot_alone[head_scan[p]] = q1 - p == 1;
	for (int64_t q = p; q < q1; q++)
		d.owner_of[vals[q]] = owner;
	for (int s = 0; s < S; s++)
	{
		double sn = d.num[(int64_t)owner * S + s];
		double sd = d.den[(int64_t)owner * S + s] + 1;
		double sf = d.fac[(int64_t)owner * S + s];
		for (int64_t q = p + 1; q < q1; q++)
		{
			const int64_t c = vals[q];
			sn += d.num[c * S + s];
			sd += d.den[c * S + s];
			sf += d.fac[c * S + s];
		}
		d.rnum[(int64_t)owner * S + s] = sn;
		d.rden[(int64_t)owner * S + s] = sd;
		d.rfac[(int64_t)owner * S + s] = sf;
	}
}

// Atomic mode: (voxel, agent) for the owners and key V for every other agent,
// so that sorting the n pairs puts the owners first, in voxel order.
__global__ void cells_owner_keys_kernel(const DomainDev* __restrict__ dp, uint32_t* keys, uint32_t* vals, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const DomainDev& d = *dp;
	const int v = d.cell_voxel[i];
	const bool own = v >= 0 && d.owner_of[i] == (int)i;
	keys[i] = own ? (uint32_t)v : (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz;
	vals[i] = (uint32_t)i;
}

// The owner count of sorted keys (owners before key V): where the keys below V end.
__global__ void cells_count_owners_kernel(const uint32_t* __restrict__ keys, int64_t n, uint32_t V, int32_t* flags)
{
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	const bool mine = keys[p] < V;
	const bool next = p + 1 < n && keys[p + 1] < V;
	if (mine && !next)
		flags[2] = (int32_t)(p + 1);
	else if (p == 0 && !mine)
		flags[2] = 0;
}

// Packs the apply list in owner-list order: the owner's voxel and its FP64
// sums, so the per-step kernels read them coalesced.  One thread per (slot,
// substrate).  own_terms: the sums hold the other agents' terms only (atomic
// mode), so the owner's own terms and the 1 of the denominator are added here;
// otherwise (sorted mode) they are the full sums.
__global__ void cells_pack_kernel(const DomainDev* __restrict__ dp, int64_t n, int own_terms)
{
	const DomainDev& d = *dp;
	const int S = d.S;
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t k = t / S;
	if (t >= n * S || k >= d.flags[2])
		return;
	const int s = (int)(t - k * S);
	const int o = d.owners[k];
	const int64_t q = (int64_t)o * S + s;
	if (s == 0)
	{
		d.pvox[k] = d.cell_voxel[o];
		d.slot_of[o] = (int32_t)k;
		if (own_terms)
			d.slot_alone[k] = !d.joined[o];
	}
	if (own_terms)
	{
		d.pnum[t] = d.num[q] + d.rnum[q];
		d.pden[t] = (d.den[q] + 1.0) + d.rden[q];
		d.pfac[t] = d.fac[q] + d.rfac[q];
	}
	else
	{
		d.pnum[t] = d.rnum[q];
		d.pden[t] = d.rden[q];
		d.pfac[t] = d.rfac[q];
	}
}

// Replaces each agent's owner by the owner's slot in the packed list, so that
// the per-step per-agent update reads its owner's results with one gather.
// Thread i reads and writes only owner_of[i]; slot_of is read only.
__global__ void cells_agent_slots_kernel(const DomainDev* __restrict__ dp, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const DomainDev& d = *dp;
	const int o = d.owner_of[i];
	if (o >= 0)
		d.owner_of[i] = d.slot_of[o];
}

// compute_result, owner part (cell_solver.cpp:187-227, compute_densities :146-163,
// compute_fused :165-185): D = (D + sum num) / sum den + sum fac at each occupied
// voxel.  Without a conflict and with tracking on, the owner's internalized
// amount takes V*(D_old - D_new); with tracking on and a conflict, D_new is
// kept in FP64 for the per-agent update (without tracking nothing is kept).  One thread per (occupied voxel, substrate).  Agent
// activity is fixed between recomputes (new agent arrays always trigger one),
// so every listed owner is active.
// The cell update of apply entry k, substrate s, from the density d0 at its
// voxel; returns the FP32 value to store there.
// The per-agent internalized update with a shared voxel somewhere
// (cell_solver.cpp:229-243): I - V*(num - D_new*den + fac), the agent's own
// terms.  One function for both callers, so both compile to the same
// operations.
__device__ __forceinline__ double agent_internalized(double I, double V, double num, double dnew, double den,
													 double fac)
{
	return I - V * (num - dnew * den + fac);
}

// Whether the owner list (in voxel order) is also nearly in agent order, at
// most one entry in 16 below its predecessor's agent: then the owners' agent
// arrays read in list order are contiguous, and the owners alone in their
// voxel take their per-agent update in the apply; otherwise those reads would
// scatter, and the per-agent pass (in agent order) takes them.
__device__ __forceinline__ bool owners_in_agent_order(const DomainDev& d)
{
	return d.flags[4] * 16 <= d.flags[2];
}

__device__ __forceinline__ float cell_update(const DomainDev& d, int k, int s, double d0, uint32_t seed, bool fused,
											 bool nearest)
{
	const int S = d.S;
	const int64_t q = (int64_t)k * S + s;
	const int64_t e = (int64_t)d.pvox[k] * S + s;
	const double d1 = (d0 + d.pnum[q]) / d.pden[q] + d.pfac[q];
	if (fused)
		d.internalized[(int64_t)d.owners[k] * S + s] += d.voxel_volume * (d0 - d1);
	else if (d.compute_internalized)
	{
		if (d.slot_alone[k] && owners_in_agent_order(d))
		{
			// the owner alone in its voxel: its own num and fac are the
			// slot's sums (num + 0 or num itself), its den is read back
			const int64_t qa = (int64_t)d.owners[k] * S + s;
			d.internalized[qa] = agent_internalized(d.internalized[qa], d.voxel_volume, d.pnum[q], d1, d.den[qa],
													d.pfac[q]);
		}
		else
			d.dnew[q] = d1; // read by the per-agent update
	}
	return sr_round(d1, mix32((uint32_t)e ^ seed), nearest);
}

// Entries of tile tr that are owners: the tiles cover every agent slot of the
// list, and the owner count of the last recompute stays on the device.
__device__ __forceinline__ int owners_in_tile(const DomainDev& d, const TileRef& tr)
{
	return max(0, min(tr.b, d.flags[2] - tr.a));
}

__global__ void cells_apply_kernel(const DomainDev* __restrict__ doms, const TileRef* __restrict__ tiles)
{
	const TileRef tr = tiles[blockIdx.x];
	const DomainDev& d = doms[tr.dom];
	const int S = d.S;
	const bool fused = d.compute_internalized && !d.flags[0];
	const uint32_t seed = stage_seed(d, kRoundCells);
	const bool nearest = d.round_nearest != 0;
	const int cnt = owners_in_tile(d, tr);
	for (int t = threadIdx.x; t < cnt * S; t += blockDim.x)
	{
		const int k = tr.a + t / S, s = t % S;
		float* D = d.field + (int64_t)d.pvox[k] * S + s;
		*D = cell_update(d, k, s, (double)*D, seed, fused, nearest);
	}
}

// Bulk then cell apply of one domain in one launch: blocks below nb take bulk
// tiles, the rest apply tiles.  A voxel in both lists is done by its bulk
// thread: the bulk update rounded to FP32, then the cell update of that value,
// the two stages' arithmetic in the reference's order; its apply thread skips
// it.  Results equal bulk_kernel followed by cells_apply_kernel.
__global__ void stages_kernel(const DomainDev* __restrict__ dom, const TileRef* __restrict__ btiles, int nb,
							  const TileRef* __restrict__ ctiles)
{
	const DomainDev& d = *dom;
	const int S = d.S;
	const bool nearest = d.round_nearest != 0;
	const uint32_t cseed = stage_seed(d, kRoundCells);
	const bool fused = d.compute_internalized && !d.flags[0];
	if ((int)blockIdx.x < nb)
	{
		const TileRef tr = btiles[blockIdx.x];
		for (int t = threadIdx.x; t < tr.b * S; t += blockDim.x)
		{
			const int k = tr.a + t / S, s = t % S;
			const int64_t q = (int64_t)k * S + s;
			const int64_t e = d.bulk_vox[k] * S + s;
			float* D = d.field + e;
			float v = bulk_round(d, ((double)*D + d.bulk_a[q]) * d.bulk_r[q], e);
			const int j = d.bulk_cell[k];
			if (j >= 0)
				v = cell_update(d, j, s, (double)v, cseed, fused, nearest);
			*D = v;
		}
	}
	else
	{
		const TileRef tr = ctiles[blockIdx.x - nb];
		const int cnt = owners_in_tile(d, tr);
		for (int t = threadIdx.x; t < cnt * S; t += blockDim.x)
		{
			const int k = tr.a + t / S, s = t % S;
			if (d.cell_in_bulk[k])
				continue;
			float* D = d.field + (int64_t)d.pvox[k] * S + s;
			*D = cell_update(d, k, s, (double)*D, cseed, fused, nearest);
		}
	}
}

__global__ void bulk_map_kernel(const int64_t* __restrict__ vox, int64_t count, int32_t* map)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < count)
		map[vox[k]] = (int32_t)k;
}

__global__ void bulk_cell_clear_kernel(int32_t* bulk_cell, int64_t count)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < count)
		bulk_cell[k] = -1;
}

__global__ void bulk_overlap_kernel(const DomainDev* __restrict__ dom, const int32_t* __restrict__ map,
									int64_t n)
{
	const DomainDev& d = *dom;
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n || k >= d.flags[2])
		return;
	const int32_t b = map[d.pvox[k]];
	d.cell_in_bulk[k] = b >= 0;
	if (b >= 0)
		d.bulk_cell[b] = (int32_t)k;
}

// The per-agent update's work list (after the owner slots are known): every
// agent whose voxel holds more than one agent, in agent order (the update
// then reads the agent arrays in order); the owners alone in theirs take the
// update in the apply.  Built as a compaction: the membership of each agent,
// an exclusive scan of it, then the scatter, which also stores the length.
__device__ __forceinline__ int in_shared_voxel(const DomainDev& d, int64_t i)
{
	const int slot = d.owner_of[i];
	return slot >= 0 && !(d.slot_alone[slot] && owners_in_agent_order(d));
}

// Counts the owner list entries whose agent precedes the previous entry's.
__global__ void cells_owner_order_kernel(const DomainDev* __restrict__ dp, int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const DomainDev& d = *dp;
	const bool down = k >= 1 && k < n && k < d.flags[2] && d.owners[k] < d.owners[k - 1];
	const unsigned long long m = __ballot(down);
	if ((threadIdx.x & (kWave - 1)) == 0 && m)
		atomicAdd(&d.flags[4], __popcll(m));
}

__global__ void cells_shared_mark_kernel(const DomainDev* __restrict__ dp, int64_t n, int32_t* mark)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		mark[i] = in_shared_voxel(*dp, i);
}

__global__ void cells_shared_scatter_kernel(const DomainDev* __restrict__ dp, int64_t n, const int32_t* __restrict__ pos)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const DomainDev& d = *dp;
	const int m = in_shared_voxel(d, i);
	if (m)
		d.shared_list[pos[i]] = (int32_t)i;
	if (i == n - 1)
		d.flags[3] = pos[i] + m;
}

// compute_internalized (cell_solver.cpp:131-144, :229-243):
// I -= V*(num - D_new*den + fac) for every active agent, when voxels are
// shared: the agents of the shared voxels from the work list (tiles over list
// positions, stopping at its length on the device); the owners alone in their
// voxels were updated by the apply.
// The workgroups loop over the tiles (the list is usually far shorter than
// the agent slots the tiles cover, so most tiles are empty).
__global__ void cells_internalized_kernel(const DomainDev* __restrict__ doms, const TileRef* __restrict__ tiles,
										  int ntiles)
{
	for (int b = blockIdx.x; b < ntiles; b += gridDim.x)
	{
		const TileRef tr = tiles[b];
		const DomainDev& d = doms[tr.dom];
		if (!d.flags[0]) // no shared voxel: the apply updated the owners' amounts
			continue;
		const int S = d.S;
		const int cnt = max(0, min(tr.b, d.flags[3] - tr.a));
		for (int t = threadIdx.x; t < cnt * S; t += blockDim.x)
		{
			const int i = d.shared_list[tr.a + t / S], s = t % S;
			const int64_t slot = d.owner_of[i];
			const int64_t q = (int64_t)i * S + s;
			d.internalized[q] = agent_internalized(d.internalized[q], d.voxel_volume, d.num[q],
												   d.dnew[slot * S + s], d.den[q], d.fac[q]);
		}
	}
}

// release_internal (cell_solver.cpp:274-289): D += I*F/V, I = 0, for a batch
// of releases grouped by voxel.  groups[3g..3g+2] = (voxel, first, count) into
// `agents` (call order within a voxel).  The sum runs in FP64 from the stored
// density, as the reference's FP64 field does, and is rounded to the FP32 field
// once.  One thread per (voxel, substrate).
__global__ void release_flush_kernel(const DomainDev* __restrict__ dp, const int64_t* __restrict__ groups,
									 int64_t n_groups, const int64_t* __restrict__ agents)
{
	const DomainDev& d = *dp;
	const int S = d.S;
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n_groups * S)
		return;
	const int64_t g = t / S;
	const int s = (int)(t - g * S);
	const int64_t v = groups[3 * g], first = groups[3 * g + 1], count = groups[3 * g + 2];
	float* D = d.field + v * S + s;
	double acc = (double)*D;
	for (int64_t j = first; j < first + count; j++)
	{
		const int64_t q = agents[j] * S + s;
		acc += d.internalized[q] * d.fraction_released[q] / d.voxel_volume;
		d.internalized[q] = 0;
	}
	*D = (float)acc;
}

#pragma clang fp contract(on)

// ------------------------------------------------------------- utilities

__global__ void fill_field_kernel(float* __restrict__ f, const float* __restrict__ init, int S, int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n)
		f[k] = init[k % S];
}

// Field snapshot: grid-stride copy in 16-byte pieces (both buffers come from
// hipMalloc, so they are aligned), the last n % 4 floats by block 0.
__global__ void copy_field_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4,
								  const float* __restrict__ src_tail, float* __restrict__ dst_tail, int tail)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += stride)
		dst[k] = src[k];
	if (blockIdx.x == 0 && (int)threadIdx.x < tail)
		dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n)
		b[k] = (float)a[k];
}

__global__ void f32_to_f64_kernel(const float* __restrict__ a, double* __restrict__ b, int64_t n)
{
	const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n)
		b[k] = (double)a[k];
}

constexpr int kMassThreads = 256;
constexpr int64_t kMassChunk = 65536;

__global__ __launch_bounds__(kMassThreads) void mass_partial_kernel(const float* __restrict__ f, int S, int64_t nvox,
																	double* __restrict__ partial)
{
	__shared__ double red[kMassThreads];
	const int s = blockIdx.y;
	const int64_t v0 = (int64_t)blockIdx.x * kMassChunk;
	const int64_t v1 = min(v0 + kMassChunk, nvox);
	double acc = 0;
	for (int64_t v = v0 + threadIdx.x; v < v1; v += kMassThreads)
		acc += (double)f[v * S + s];
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int w = kMassThreads / 2; w > 0; w >>= 1)
	{
		if ((int)threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		partial[(int64_t)s * gridDim.x + blockIdx.x] = red[0];
}

__global__ void mass_final_kernel(const double* __restrict__ partial, int nb, int S, double scale,
								  double* __restrict__ out)
{
	const int s = threadIdx.x;
	if (s >= S)
		return;
	double acc = 0;
	for (int b = 0; b < nb; b++)
		acc += partial[(int64_t)s * nb + b];
	out[s] = acc * scale;
}

// ------------------------------------------------------------------ scan

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

// exclusive scan of kScanThreads values held one per thread; returns the total
__device__ int block_exclusive_scan(int v, int* sh)
{
	const int lane = threadIdx.x & (kWave - 1);
	const int wave = threadIdx.x / kWave;
	int incl = v;
	for (int off = 1; off < kWave; off <<= 1)
	{
		const int o = __shfl_up(incl, off);
		if (lane >= off)
			incl += o;
	}
	if (lane == kWave - 1)
		sh[wave] = incl;
	__syncthreads();
	int wave_off = 0, total = 0;
	for (int w = 0; w < kScanThreads / kWave; w++)
	{
		if (w < wave)
			wave_off += sh[w];
		total += sh[w];
	}
	__syncthreads();
	sh[kScanThreads / kWave] = total;
	return wave_off + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_blocks_kernel(int32_t* data, int32_t* block_sums, int64_t n)
{
	__shared__ int sh[kScanThreads / kWave + 1];
	const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
	int v[kScanItems];
	int sum = 0;
	for (int k = 0; k < kScanItems; k++)
	{
		v[k] = (base + k < n) ? data[base + k] : 0;
		sum += v[k];
	}
	int off = block_exclusive_scan(sum, sh);
	for (int k = 0; k < kScanItems; k++)
	{
		if (base + k < n)
			data[base + k] = off;
		off += v[k];
	}
	if (threadIdx.x == kScanThreads - 1)
		block_sums[blockIdx.x] = off;
}

__global__ __launch_bounds__(kScanThreads) void scan_single_kernel(int32_t* data, int64_t n)
{
	__shared__ int sh[kScanThreads / kWave + 1];
	int carry = 0;
	for (int64_t c0 = 0; c0 < n; c0 += kScanThreads)
	{
		const int64_t k = c0 + threadIdx.x;
		const int v = k < n ? data[k] : 0;
		const int ex = block_exclusive_scan(v, sh);
		if (k < n)
			data[k] = carry + ex;
		__syncthreads();
		carry += sh[kScanThreads / kWave];
		__syncthreads();
	}
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(int32_t* data, const int32_t* block_sums, int64_t n)
{
	const int add = block_sums[blockIdx.x];
	const int64_t base = (int64_t)blockIdx.x * kScanBlock;
	for (int k = threadIdx.x; k < kScanBlock; k += kScanThreads)
		if (base + k < n)
			data[base + k] += add;
}

// ----------------------------------------------------------- radix sort

constexpr int kRadixThreads = 256;
constexpr int kRadixChunk = 4096; // elements per block and pass
constexpr int kRadixBins = 256;

__global__ __launch_bounds__(kRadixThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t n,
																   int shift, int32_t* __restrict__ hist, int nblocks)
{
	__shared__ int h[kRadixBins];
	h[threadIdx.x] = 0;
	__syncthreads();
	const int64_t b0 = (int64_t)blockIdx.x * kRadixChunk;
	const int64_t b1 = min(b0 + kRadixChunk, n);
	for (int64_t k = b0 + threadIdx.x; k < b1; k += kRadixThreads)
		atomicAdd(&h[(keys[k] >> shift) & 255], 1);
	__syncthreads();
	hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: within a block, elements keep their order (sub-tile, wave,
// lane), so equal digits leave in input order.
__global__ __launch_bounds__(kRadixThreads) void radix_scatter_kernel(const uint32_t* __restrict__ keys,
																	  const uint32_t* __restrict__ vals,
																	  uint32_t* __restrict__ okeys,
																	  uint32_t* __restrict__ ovals, int64_t n,
																	  int shift, const int32_t* __restrict__ offsets,
																	  int nblocks)
{
	__shared__ int base[kRadixBins];
	__shared__ int wcount[kRadixThreads / kWave][kRadixBins];
	const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
	base[tid] = offsets[(int64_t)tid * nblocks + blockIdx.x];
	const int64_t b0 = (int64_t)blockIdx.x * kRadixChunk;
	const int64_t b1 = min(b0 + kRadixChunk, n);
	for (int64_t t0 = b0; t0 < b1; t0 += kRadixThreads)
	{
		for (int w = 0; w < kRadixThreads / kWave; w++)
			wcount[w][tid] = 0;
		__syncthreads();
		const int64_t k = t0 + tid;
		const bool valid = k < b1;
		const uint32_t key = valid ? keys[k] : 0u;
		const uint32_t val = valid ? vals[k] : 0u;
		const int digit = (key >> shift) & 255;
		unsigned long long m = __ballot(valid);
		for (int bit = 0; bit < 8; bit++)
		{
			const unsigned long long bb = __ballot((digit >> bit) & 1);
			m &= ((digit >> bit) & 1) ? bb : ~bb;
		}
		const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
		const int rank = __popcll(m & below);
		if (valid && rank == 0)
			wcount[wave][digit] = __popcll(m);
		__syncthreads();
		if (valid)
		{
			int pre = 0;
			for (int w = 0; w < wave; w++)
				pre += wcount[w][digit];
			const int pos = base[digit] + pre + rank;
			okeys[pos] = key;
			ovals[pos] = val;
		}
		__syncthreads();
		int add = 0;
		for (int w = 0; w < kRadixThreads / kWave; w++)
			add += wcount[w][tid];
		base[tid] += add;
		__syncthreads();
	}
}

template <typename K, typename... Args>
hipError_t launch1d(K kernel, int64_t n, int threads, hipStream_t stream, Args... args)
{
	if (n <= 0)
		return hipSuccess;
	const int64_t blocks = (n + threads - 1) / threads;
	hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, args...);
	return hipGetLastError();
}

} // namespace

// ------------------------------------------------------------ launchers

hipError_t launch_dirichlet_interior(const DomainDev* doms, const SparseRef* refs, int64_t n, hipStream_t stream)
{
	return launch1d(dirichlet_interior_kernel, n, 256, stream, doms, refs, n);
}

hipError_t launch_bulk(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream)
{
	if (ntiles <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(bulk_kernel, dim3(ntiles), dim3(256), 0, stream, doms, tiles);
	return hipGetLastError();
}

hipError_t launch_step_tick(const DomainDev* doms, int ndoms, uint32_t by, hipStream_t stream)
{
	if (ndoms <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(step_tick_kernel, dim3((ndoms + 63) / 64), dim3(64), 0, stream, doms, ndoms, by);
	return hipGetLastError();
}

hipError_t launch_cells_apply(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream)
{
	if (ntiles <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(cells_apply_kernel, dim3(ntiles), dim3(256), 0, stream, doms, tiles);
	return hipGetLastError();
}

hipError_t launch_stages(const DomainDev* dom, const TileRef* bulk_tiles, int n_bulk_tiles, const TileRef* cell_tiles,
						 int n_cell_tiles, hipStream_t stream)
{
	if (n_bulk_tiles + n_cell_tiles <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(stages_kernel, dim3(n_bulk_tiles + n_cell_tiles), dim3(256), 0, stream, dom, bulk_tiles,
					   n_bulk_tiles, cell_tiles);
	return hipGetLastError();
}

hipError_t launch_bulk_map(const int64_t* bulk_vox, int64_t count, int32_t* map, int64_t V, hipStream_t stream)
{
	if (hipError_t e = hipMemsetAsync(map, 0xff, (size_t)V * sizeof(int32_t), stream); e != hipSuccess)
		return e;
	return launch1d(bulk_map_kernel, count, 256, stream, bulk_vox, count, map);
}

hipError_t launch_bulk_overlap(const DomainDev* dom, const int32_t* map, int32_t* bulk_cell, int64_t bulk_count,
							   int64_t n, hipStream_t stream)
{
	if (hipError_t e = launch1d(bulk_cell_clear_kernel, bulk_count, 256, stream, bulk_cell, bulk_count);
		e != hipSuccess)
		return e;
	return launch1d(bulk_overlap_kernel, n, 256, stream, dom, map, n);
}

hipError_t launch_cells_internalized(const DomainDev* doms, const TileRef* tiles, int ntiles, hipStream_t stream)
{
	if (ntiles <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(cells_internalized_kernel, dim3(std::min(ntiles, 2048)), dim3(256), 0, stream, doms, tiles,
					   ntiles);
	return hipGetLastError();
}

hipError_t launch_cells_owner_keys(const DomainDev* dom, uint32_t* keys, uint32_t* vals, int64_t n,
								   hipStream_t stream)
{
	return launch1d(cells_owner_keys_kernel, n, 256, stream, dom, keys, vals, n);
}

hipError_t launch_cells_count_owners(const uint32_t* keys, int64_t n, uint32_t V, int32_t* flags, hipStream_t stream)
{
	return launch1d(cells_count_owners_kernel, n, 256, stream, keys, n, V, flags);
}

hipError_t launch_cells_pack(const DomainDev* dom, int64_t n, int S, bool own_terms, hipStream_t stream)
{
	return launch1d(cells_pack_kernel, n * S, 256, stream, dom, n, own_terms ? 1 : 0);
}

hipError_t launch_cells_agent_slots(const DomainDev* dom, int64_t n, hipStream_t stream)
{
	return launch1d(cells_agent_slots_kernel, n, 256, stream, dom, n);
}

hipError_t launch_cells_owner_order(const DomainDev* dom, int64_t n, hipStream_t stream)
{
	return launch1d(cells_owner_order_kernel, n, 256, stream, dom, n);
}

hipError_t launch_cells_shared_list(const DomainDev* dom, int64_t n, int32_t* pos, int32_t* tmp, hipStream_t stream)
{
	if (n <= 0)
		return hipSuccess;
	hipError_t e = launch1d(cells_shared_mark_kernel, n, 256, stream, dom, n, pos);
	if (e == hipSuccess)
		e = exclusive_scan(pos, tmp, n, stream);
	if (e == hipSuccess)
		e = launch1d(cells_shared_scatter_kernel, n, 256, stream, dom, n, (const int32_t*)pos);
	return e;
}

hipError_t launch_copy_field(const float* src, float* dst, int64_t n, int num_cu, hipStream_t stream)
{
	if (n <= 0)
		return hipSuccess;
	const int64_t n4 = n / 4;
	const int tail = (int)(n - n4 * 4);
	const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n4 + 255) / 256, (int64_t)num_cu * 8));
	hipLaunchKernelGGL(copy_field_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
					   reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4, src + n4 * 4,
					   dst + n4 * 4, tail);
	return hipGetLastError();
}

hipError_t launch_fill_field(float* field, const float* init, int S, int64_t nvox, hipStream_t stream)
{
	const int64_t n = nvox * S;
	return launch1d(fill_field_kernel, n, 256, stream, field, init, S, n);
}

hipError_t launch_convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t stream)
{
	return launch1d(f64_to_f32_kernel, n, 256, stream, src, dst, n);
}

hipError_t launch_convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t stream)
{
	return launch1d(f32_to_f64_kernel, n, 256, stream, src, dst, n);
}

hipError_t launch_release_flush(const DomainDev* dom, const int64_t* groups, int64_t n_groups, const int64_t* agents,
								int S, hipStream_t stream)
{
	return launch1d(release_flush_kernel, n_groups * S, 256, stream, dom, groups, n_groups, agents);
}

hipError_t launch_cells_bin(const DomainDev* dom, int64_t n, int S, bool clear_ballots, hipStream_t stream)
{
	return launch1d(cells_bin_kernel, n * S, 256, stream, dom, n, clear_ballots ? 1 : 0);
}

hipError_t launch_cells_ballot(const DomainDev* dom, int64_t n, hipStream_t stream)
{
	return launch1d(cells_ballot_kernel, n, 256, stream, dom, n);
}

hipError_t launch_cells_make_keys(const DomainDev* dom, uint32_t* keys, uint32_t* vals, int64_t n, hipStream_t stream)
{
	return launch1d(cells_make_keys_kernel, n, 256, stream, dom, keys, vals, n);
}

hipError_t launch_cells_heads(const uint32_t* keys, int32_t* heads, int64_t n, uint32_t V, hipStream_t stream)
{
	return launch1d(cells_heads_kernel, n, 256, stream, keys, heads, n, V);
}

hipError_t launch_cells_segment_reduce(const DomainDev* dom, const uint32_t* keys, const uint32_t* vals,
									   const int32_t* heads, const int32_t* head_scan, int64_t n,
									   hipStream_t stream)
{
	return launch1d(cells_segment_reduce_kernel, n, 256, stream, dom, keys, vals, heads, head_scan, n);
}

size_t scan_tmp_entries(int64_t n) { return (size_t)((n + kScanBlock - 1) / kScanBlock) + 1; }

hipError_t exclusive_scan(int32_t* data, int32_t* tmp, int64_t n, hipStream_t stream)
{
	if (n <= 0)
		return hipSuccess;
	const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
	hipLaunchKernelGGL(scan_blocks_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, data, tmp, n);
	hipLaunchKernelGGL(scan_single_kernel, dim3(1), dim3(kScanThreads), 0, stream, tmp, nb);
	hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, data, tmp, n);
	return hipGetLastError();
}

size_t radix_hist_entries(int64_t n)
{
	const int64_t nb = (n + kRadixChunk - 1) / kRadixChunk;
	return (size_t)(nb * kRadixBins);
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys, uint32_t* tmp_vals, int32_t* hist,
							int32_t* scan_tmp, int64_t n, int bits, hipStream_t stream)
{
	if (n <= 1)
		return hipSuccess;
	const int nb = (int)((n + kRadixChunk - 1) / kRadixChunk);
	uint32_t *ik = keys, *iv = vals, *ok = tmp_keys, *ov = tmp_vals;
	int passes = 0;
	for (int shift = 0; shift < bits; shift += 8, passes++)
	{
		hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(kRadixThreads), 0, stream, ik, n, shift, hist, nb);
		hipError_t e = exclusive_scan(hist, scan_tmp, (int64_t)nb * kRadixBins, stream);
		if (e != hipSuccess)
			return e;
		hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kRadixThreads), 0, stream, ik, iv, ok, ov, n, shift,
						   hist, nb);
		std::swap(ik, ok);
		std::swap(iv, ov);
	}
	if (passes & 1)
	{
		hipError_t e = hipMemcpyAsync(keys, ik, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
		if (e != hipSuccess)
			return e;
		e = hipMemcpyAsync(vals, iv, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
		if (e != hipSuccess)
			return e;
	}
	return hipGetLastError();
}

int mass_blocks(int64_t nvox) { return (int)((nvox + kMassChunk - 1) / kMassChunk); }

hipError_t launch_mass(const float* field, int S, int64_t nvox, double* partial, double* out, double scale,
					   hipStream_t stream)
{
	const int nb = mass_blocks(nvox);
	hipLaunchKernelGGL(mass_partial_kernel, dim3(nb, S), dim3(kMassThreads), 0, stream, field, S, nvox, partial);
	hipLaunchKernelGGL(mass_final_kernel, dim3(1), dim3(64 * ((S + 63) / 64)), 0, stream, partial, nb, S, scale,
					   out);
	return hipGetLastError();
}

} // namespace bfvm
