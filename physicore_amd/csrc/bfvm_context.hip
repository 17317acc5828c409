// bfvm_context.hip -- the C ABI of libbfvm_hip.so (include/bfvm_hip.h): one
// BioFVM microenvironment resident in HBM, stepped on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "bfvm_hip.h"
#include "bfvm_internal.h"

namespace {
// shortest y/z line the pipelined kernel is used for by default
constexpr int kPipeMinLine = 192;

// x kernel of BFVM_X_AUTO.  The row-parallel scan keeps a line's forward
// values in FP64 registers and handles rows up to 64*32/S voxels; the strided
// register lines (lines up to 256) store theirs in FP32, stochastically
// rounded.  x-sweep times on MI355X (nx x 128 x 128, FP64 recurrences,
// strided / LDS transpose / scan, us): S = 1, nx = 200: 64.6 / 38.4 / 21.1;
// S = 2, nx = 64: 23.5 / 18.9 / 22.9; S = 2, nx = 200: 58.9 / 40.0 / 25.9;
// S = 4, nx = 64: 16.0 / 20.3 / 33.2; S = 4, nx = 200: 38.8 / 50.5 / 41.0;
// S = 8, nx = 200: 63.9 / 111 / 98.6.  The chunked transpose (S = 4, 8,
// rows up to 200) reads the strided kernel's lines with 16-byte accesses
// (strided / scan / chunked, us): 200^3 x 4: 75.9 / 80.5 / 60.8; 200^3 x 8:
// 151 / 203 / 121; 160^3 x 4: 41.2 / - / 40.2; 100^3 x 8: 37.1 / 39.5 / 28.7;
// 32 x 256^2 x 4: 23.2 / - / 22.4; 8 x 512^2 x 4: 35.5 / - / 22.6; and slower
// for 64 x 128^2 x 4: 16.3 / - / 17.3, 64 x 128^2 x 8: 25.6 / - / 31.0,
// 16 x 256^2 x 8: 23.8 / - / 32.1.  With 4-byte accesses (S = 3, 5, 6, 7;
// strided / chunked, us): 200^3 x 3: 83.7 / 39.9; 128^3 x 3: 31.8 / 25.5;
// 32 x 256^2 x 3: 27.7 / 23.2; 200^3 x 6: 120 / 76.7; 192^3 x 5: 96.4 / 76.3;
// 96 x 128^2 x 5: 36.3 / 31.4; and slower for 16 x 256^2 x 6: 21.9 / 25.0,
// 48 x 160^2 x 5: 23.3 / 24.6, 192^3 x 7: 108 / 128, 64 x 128^2 x 7:
// 25.3 / 28.4.  For S = 1, 2 (scan or transpose / chunked, us): 200^3 x 2:
// 48.9 / 37.0; 128^3 x 2: 26.6 / 26.3; 64 x 128^2 x 2: 19.0 / 20.8; S = 1
// (64 rows per wave: too few waves below ~10^5 rows): 176 x 200^2: 36.4 /
// 28.7; 160^3: 27.4 / 29.8; 128^3: 18.9 / 28.3.  Shapes the chosen kernel does
// not take fall back to the strided lines.  Rows longer than the one-lane
// chunked kernels hold (strided / scan / two-lane chunked, us): 256 x 160^2 x 3:
// 108 / 43.4 / 54.6; 240 x 160^2 x 5: 84.7 / 65.9 / 84.8; 256 x 128^2 x 6: 82.0
// / 54.8 / 76.6; 400 x 400 x 100 x 2: 407 / 74.1 / 74.1; and where the scan does
// not fit, 300 x 128^2 x 7: 507 (LDS tiles) / - / 101.
int default_x_sweep(int nx, int S)
{
	// S = 2: the 16-byte chunked kernel for even rows below 176 voxels, the
	// 4-byte one from 176 on (scan or transpose / chunked, us): 128^3: 26.1 /
	// 22.0; 64 x 128^2: 18.8 / 15.5; 32 x 256^2: 25.9 / 19.6; 96 x 160^2: 32.8 /
	// 22.2; 200^3: 48.9 / 37.0 (4-byte; 46.0 with 16-byte accesses)
	if (S == 2 && nx <= 200 && (nx % 2 == 0 || nx >= 160) && (bfvm::xt_bucket(nx) || bfvm::xg_bucket(nx, S)))
		return BFVM_X_CHUNKED;
	// S = 2 past one-lane rows: the two-lane chunked kernel (scan / chunked,
	// us, back-to-back launches): 400^2 x 100: 74.9 / 69.8; 300 x 200 x 100:
	// 40.7 / 34.1
	if (S == 2 && nx > 200 && bfvm::xgt_bucket(nx, S))
		return BFVM_X_CHUNKED;
	// S = 4, 8 past 176 voxels: 4-byte chunks, one lane per line up to 200
	// voxels, two up to 400 (16-byte chunks or scan / 4-byte chunks, us,
	// back-to-back launches): 200^3 x 4: 56.2 / 52.9; 200^3 x 8: 108.2 / 100.6;
	// 250^3 x 4: 130.1 / 117.2 (scan); 256^3 x 8: 340.5 / 312.1 (scan).  Below,
	// the 16-byte kernel: 160^3 x 4: 37.0 / 38.3; 96^3 x 4: 13.1 / 18.2
	if ((S == 4 || S == 8) && nx > 176 && (bfvm::xg_bucket(nx, S) || bfvm::xgt_bucket(nx, S)))
		return BFVM_X_CHUNKED4;
	if (nx > 200 || (S != 4 && S != 8 && nx > bfvm::xg_max_n(S)))
		return bfvm::xscan_vpl(nx, S) ? BFVM_X_SCAN : BFVM_X_CHUNKED;
	if (S <= 2 || nx > bfvm::kRegMaxN)
		return (S == 2 && nx < 96) ? BFVM_X_TRANSPOSE : BFVM_X_SCAN;
	if ((S == 4 || S == 8) && bfvm::xt_bucket(nx) && (nx >= 96 || (S == 4 && nx <= 32)))
		return BFVM_X_CHUNKED;
	if (bfvm::xg_bucket(nx, S) && (S == 3 || ((S == 5 || S == 6) && nx >= 96)))
		return BFVM_X_CHUNKED;
	return BFVM_X_STRIDED;
}
} // namespace

using namespace bfvm;

namespace {

thread_local std::string g_error;

// Device blocks of destroyed contexts and batches.  hipFree waits for the
// whole device, so a server that frees a finished domain while others step
// would stall for as long as the queued work takes, once per buffer.  The
// destroy calls first synchronize the object's streams (after which no queued
// work reads its buffers) and then hand the blocks to this cache; later
// allocations on the same device take a cached block of a fitting size.
// Buffers released any other way still go back through hipFree.
class BlockCache
{
  public:
	static constexpr size_t kLimit = size_t(48) << 30; // bytes held at most; beyond, hipFree

	// a block of at least `bytes` (already rounded by block_bytes), or nullptr
	void* take(int dev, size_t bytes)
	{
		std::lock_guard<std::mutex> g(m_);
		const size_t most = bytes >= kLarge ? bytes + bytes / 4 : bytes;
		for (auto it = blocks_.lower_bound({ dev, bytes }); it != blocks_.end(); ++it)
		{
			if (it->first.first != dev || it->first.second > most)
				break;
			void* p = it->second;
			total_ -= it->first.second;
			blocks_.erase(it);
			return p;
		}
		return nullptr;
	}
	void put(int dev, size_t bytes, void* p)
	{
		{
			std::lock_guard<std::mutex> g(m_);
			if (total_ + bytes <= kLimit)
			{
				blocks_.insert({ { dev, bytes }, p });
				total_ += bytes;
				return;
			}
		}
		(void)hipFree(p);
	}
	// returns every cached block of `dev` to the runtime (after an allocation failed)
	bool trim(int dev)
	{
		std::vector<void*> out;
		{
			std::lock_guard<std::mutex> g(m_);
			for (auto it = blocks_.begin(); it != blocks_.end();)
				if (it->first.first == dev)
				{
					total_ -= it->first.second;
					out.push_back(it->second);
					it = blocks_.erase(it);
				}
				else
					++it;
		}
		for (void* p : out)
			(void)hipFree(p);
		return !out.empty();
	}
	// allocation granularity: powers of two below 2 MiB, 2 MiB multiples above
	static size_t block_bytes(size_t bytes)
	{
		if (bytes >= kLarge)
			return (bytes + kLarge - 1) / kLarge * kLarge;
		size_t b = 256;
		while (b < bytes)
			b <<= 1;
		return b;
	}

  private:
	static constexpr size_t kLarge = size_t(2) << 20;
	std::mutex m_;
	std::multimap<std::pair<int, size_t>, void*> blocks_;
	size_t total_ = 0;
};

// never destroyed: cached blocks are not freed after the runtime shuts down
BlockCache& block_cache()
{
	static BlockCache* c = new BlockCache;
	return *c;
}

// set while a destroy call releases the buffers of an object whose streams
// it has synchronized
thread_local bool t_recycle = false;

struct RecycleScope
{
	bool prev;
	explicit RecycleScope(bool on) : prev(t_recycle) { t_recycle = on; }
	~RecycleScope() { t_recycle = prev; }
};

template <typename T>
struct dvec
{
	T* p = nullptr;
	size_t n = 0;
	size_t cap = 0; // bytes of the block
	int dev = 0;
	dvec() = default;
	dvec(const dvec&) = delete;
	dvec& operator=(const dvec&) = delete;
	~dvec() { release(); }
	void release()
	{
		if (p)
		{
			if (t_recycle)
				block_cache().put(dev, cap, p);
			else
				(void)hipFree(p);
		}
		p = nullptr;
		n = 0;
		cap = 0;
	}
	hipError_t alloc(size_t count)
	{
		if (count == n && p)
			return hipSuccess;
		release();
		if (count == 0)
			return hipSuccess;
		hipError_t e = hipGetDevice(&dev);
		if (e != hipSuccess)
			return e;
		const size_t bytes = BlockCache::block_bytes(count * sizeof(T));
		p = (T*)block_cache().take(dev, bytes);
		if (!p)
		{
			e = hipMalloc((void**)&p, bytes);
			if (e == hipErrorOutOfMemory && block_cache().trim(dev))
			{
				(void)hipGetLastError();
				e = hipMalloc((void**)&p, bytes);
			}
			if (e != hipSuccess)
			{
				p = nullptr;
				return e;
			}
		}
		n = count;
		cap = bytes;
		return hipSuccess;
	}
	hipError_t upload(const T* h, size_t count, hipStream_t s)
	{
		hipError_t e = alloc(count);
		if (e != hipSuccess || count == 0)
			return e;
		e = hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
		if (e != hipSuccess)
			return e;
		return hipStreamSynchronize(s);
	}
};

// Thomas factors of one axis in FP64, openmp_solver/src/diffusion_solver.cpp:29-86
// (copies = 1); a single-voxel line keeps only the decay share on its diagonal
// (thrust_solver/src/diffusion_solver.cpp:89-100) unless the OpenMP formula is asked for.
void precompute_factors(int n, int S, int64_t shape, int dims, double dt, const double* D, const double* lam,
						bool openmp_quirks, std::vector<double>& b, std::vector<double>& e, std::vector<double>& c)
{
	const double h2 = (double)(shape * shape);
	b.assign((size_t)n * S, 0.0);
	e.assign((size_t)std::max(n - 1, 1) * S, 0.0);
	c.assign(S, 0.0);
	for (int s = 0; s < S; s++)
		c[s] = dt * D[s] / h2;
	if (n == 1 && !openmp_quirks)
	{
		for (int s = 0; s < S; s++)
			b[s] = 1 / (1 + lam[s] * dt / (double)dims);
		return;
	}
	for (int i = 0; i < n; i++)
		for (int s = 0; s < S; s++)
		{
			double v = 1 + lam[s] * dt / (double)dims + 2 * dt * D[s] / h2;
			if (i == 0 || i == n - 1)
				v -= dt * D[s] / h2;
			b[(size_t)i * S + s] = v;
		}
	for (int s = 0; s < S; s++)
		b[s] = 1 / b[s];
	for (int i = 1; i < n; i++)
		for (int s = 0; s < S; s++)
		{
			b[(size_t)i * S + s] = 1 / (b[(size_t)i * S + s] - c[s] * c[s] * b[(size_t)(i - 1) * S + s]);
			e[(size_t)(i - 1) * S + s] = c[s] * b[(size_t)(i - 1) * S + s];
		}
}

// The OpenMP backend's y/z coefficient index (openmp_solver/src/diffusion_solver.cpp:390-401, :427):
// blocks of S*ceil(48/S) columns, tiles of 48, diagonals indexed tile-relative.
int32_t openmp_coef_sub(int64_t c, int S, int64_t L)
{
	const int64_t copies = (48 + S - 1) / S;
	const int64_t B = S * copies;
	const int64_t nb = L / B;
	const int64_t r = (c < nb * B) ? c % B : c - nb * B;
	return (int32_t)((r % 48) % S);
}

bool fp_ok(double v) { return std::isfinite(v); }

} // namespace

struct bfvm_ctx
{
	std::string err;
	int dims = 3, S = 1;
	int64_t n[3] = { 1, 1, 1 };
	int64_t voxel[3] = { 1, 1, 1 };
	int64_t bbox_min[3] = { 0, 0, 0 };
	int64_t V = 1;
	int L = 1;
	double dt = 0.01;
	bool internalized = false;
	bool fp64 = true;
	int reduce_mode = BFVM_REDUCE_ATOMIC;
	bool quirks = false;
	std::vector<double> D, lam, init;

	hipStream_t stream = nullptr;
	int device = 0; // the HIP device current at bfvm_create
	int num_cu = 256;

	dvec<float> field;
	dvec<uint32_t> step_count; // [1] steps taken (keys of the stochastic rounding)
	dvec<double> fb[3], fe[3], fc[3];
	dvec<int32_t> coef_map;
	dvec<float> face_val;
	dvec<uint8_t> face_cond;
	int face_mask = 0;

	int64_t dir_count = 0;
	dvec<int64_t> dir_vox;
	dvec<float> dir_val;
	dvec<uint8_t> dir_cond;
	dvec<SparseRef> dir_refs;

	int64_t bulk_count = 0;
	dvec<int64_t> bulk_vox;
	dvec<double> bulk_a, bulk_r;
	dvec<TileRef> bulk_tiles; // sparse tiles of the bulk list
	// a bulk list with the same rates per substrate at every listed voxel and
	// at most one run of listed voxels along each line of the last axis: its
	// a and r ([2*S]) and the runs ([2*columns]); the last sweep applies it
	// while storing when that sweep is a register y/z kernel and no interior
	// Dirichlet stage sits between them
	bool bulk_runs = false;
	// the list's voxels alone (whatever the rates) have that shape: the last
	// sweep then stays on one lane per line, so that a domain keeps its
	// kernels when only the rates change
	bool bulk_layout = false;
	bool bulk_fused = false;
	dvec<double> bulk_u;
	dvec<int32_t> bulk_run;
	int n_bulk_tiles = 0;
	// voxels in both the bulk and the cell apply lists: voxel -> bulk entry
	// ([V]), apply entry per bulk entry, bulk flag per apply entry; `overlap`
	// is true while they match both lists (bulk and cells then run as one
	// launch, stages_kernel)
	dvec<int32_t> bulk_map, bulk_cell;
	dvec<uint8_t> cell_in_bulk;
	bool overlap = false;

	int64_t n_agents = 0;
	dvec<double> positions, secretion, saturation, uptake, net_export, internalized_d, fraction_released, volumes;
	dvec<uint8_t> is_active;
	std::vector<double> h_positions;
	dvec<int32_t> cell_voxel, owner_of, ballot, owners, flags;
	dvec<double> num, den, fac, rnum, rden, rfac, dnew, pnum, pden, pfac;
	dvec<int32_t> pvox, slot_of;
	dvec<uint8_t> slot_alone, joined;
	dvec<int32_t> shared_list;
	dvec<TileRef> cell_tiles, int_tiles; // sparse tiles of the owner list / of all agents
	int n_cell_tiles = 0, n_int_tiles = 0;
	dvec<uint32_t> keys, vals, tkeys, tvals;
	dvec<int32_t> hist, scan_tmp, heads, head_scan;
	bool recompute_pending = false;
	// a recompute was enqueued: the next call that waits for the stream reads
	// its out-of-grid flag
	bool grid_check_pending = false;
	// positions were written on the device since the host copy was taken
	bool h_positions_stale = false;
	// owner count and shared-voxel flag of the last recompute, read back only
	// for bfvm_cell_sums (the step kernels read them on the device)
	int64_t n_owners = 0;
	bool conflict = false;
	// releases not yet applied to the field: (agent, voxel) in call order
	std::vector<int64_t> rel_agent, rel_voxel;
	dvec<int64_t> rel_dev;

	dvec<TileRef> tiles[3];
	int ntiles[3] = { 0, 0, 0 };
	// the x sweep in two halves, the second on `side`; the y tiles of the
	// planes the first half covers whole follow it on the context stream, the
	// others follow the second half (tile counts of the first halves; 0: one
	// launch per axis)
	int split_x = 0, split_y = 0;
	hipStream_t side = nullptr;
	hipEvent_t ev_split[3] = {};
	int fac_cap[3] = { 0, 0, 0 };
	int variant[3] = { 0, 0, 0 };
	int nmax[3] = { 0, 0, 0 };
	int threads[3] = { kSweepThreads, kSweepThreads, kSweepThreads }; // workgroup size of each sweep
	size_t lds[3] = { 0, 0, 0 };
	PipeGeometry pipe[3];
	bfvm_tuning tuning {};

	DomainDev hdom {};
	dvec<DomainDev> ddom;
	dvec<DomainDev> dslots; // the captured graph's steps: ddom with step_offset 0, 1, ...

	dvec<double> mass_partial, mass_out;

	// field snapshots for output: device copy, then host copy on copy_stream;
	// ticket t completes with snap_done[t % kSnapRing]
	static constexpr int kSnapRing = 4;
	dvec<float> snap;
	hipStream_t copy_stream = nullptr;
	hipEvent_t snap_taken = nullptr;
	hipEvent_t snap_done[kSnapRing] = {};
	std::atomic<int64_t> snap_next { 0 };

	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	int64_t graph_steps = 0;
	uint64_t graph_version = ~0ull;
	uint64_t version = 0;

	int fail(int code, const std::string& m)
	{
		err = m;
		return code;
	}
	int hip_fail(hipError_t e, const char* what)
	{
		err = std::string(what) + ": " + hipGetErrorString(e);
		return e == hipErrorOutOfMemory ? BFVM_ERR_NO_MEMORY : BFVM_ERR_HIP;
	}
	void drop_graph()
	{
		if (gexec)
			(void)hipGraphExecDestroy(gexec);
		if (graph)
			(void)hipGraphDestroy(graph);
		gexec = nullptr;
		graph = nullptr;
		graph_version = ~0ull;
	}
	~bfvm_ctx()
	{
		if (stream)
			(void)hipStreamSynchronize(stream);
		if (copy_stream)
			(void)hipStreamSynchronize(copy_stream);
		drop_graph();
		for (hipEvent_t e : snap_done)
			if (e)
				(void)hipEventDestroy(e);
		if (snap_taken)
			(void)hipEventDestroy(snap_taken);
		if (side)
			(void)hipStreamSynchronize(side);
		for (hipEvent_t e : ev_split)
			if (e)
				(void)hipEventDestroy(e);
		if (side)
			(void)hipStreamDestroy(side);
		if (copy_stream)
			(void)hipStreamDestroy(copy_stream);
		if (stream)
			(void)hipStreamDestroy(stream);
	}
};

#define TRY_HIP(ctx, expr)                               \
	do                                                   \
	{                                                    \
		hipError_t e_ = (expr);                          \
		if (e_ != hipSuccess)                            \
			return (ctx)->hip_fail(e_, #expr);           \
	} while (0)

namespace {

// Tiles (dom, first, count) over `count` entries of one domain with S
// substrates: at most kSparseTile / S entries a tile, so that each thread of a
// 256-thread workgroup takes one (entry, substrate) and all of a tile's loads
// are in flight at once.
std::vector<TileRef> sparse_tiles(int dom, int64_t count, int S)
{
	const int per = std::max(1, kSparseTile / std::max(1, S));
	std::vector<TileRef> t;
	for (int64_t k = 0; k < count; k += per)
		t.push_back({ dom, (int32_t)k, (int32_t)std::min<int64_t>(per, count - k), 0 });
	return t;
}

int sync_domain(bfvm_ctx* c)
{
	DomainDev& d = c->hdom;
	d.field = c->field.p;
	d.step = c->step_count.p;
	d.round_nearest = c->tuning.round_nearest;
	d.nx = (int32_t)c->n[0];
	d.ny = (int32_t)c->n[1];
	d.nz = (int32_t)c->n[2];
	d.S = c->S;
	d.dims = c->dims;
	d.L = c->L;
	d.face_mask = c->face_mask;
	d.dt = c->dt;
	d.voxel_volume = (double)(c->voxel[0] * c->voxel[1] * c->voxel[2]);
	for (int a = 0; a < 3; a++)
	{
		d.bbox_min[a] = c->bbox_min[a];
		d.voxel[a] = c->voxel[a];
		d.fb[a] = c->fb[a].p;
		d.fe[a] = c->fe[a].p;
		d.fc[a] = c->fc[a].p;
	}
	d.coef_map = c->coef_map.p;
	d.face_val = c->face_val.p;
	d.face_cond = c->face_cond.p;
	d.dir_vox = c->dir_vox.p;
	d.dir_val = c->dir_val.p;
	d.dir_cond = c->dir_cond.p;
	d.bulk_vox = c->bulk_vox.p;
	d.bulk_a = c->bulk_a.p;
	d.bulk_r = c->bulk_r.p;
	c->bulk_fused = c->bulk_runs && c->dims >= 2 && c->fp64 && c->dir_count == 0 &&
					(c->variant[c->dims - 1] == kSweepReg || (c->variant[c->dims - 1] == kSweepTwist && c->dims == 3));
	d.bulk_u = c->bulk_fused ? c->bulk_u.p : nullptr;
	d.bulk_run = c->bulk_fused ? c->bulk_run.p : nullptr;
	d.bulk_cell = c->bulk_cell.p;
	d.cell_in_bulk = c->cell_in_bulk.p;
	d.n_agents = c->n_agents;
	d.positions = c->positions.p;
	d.secretion = c->secretion.p;
	d.saturation = c->saturation.p;
	d.uptake = c->uptake.p;
	d.net_export = c->net_export.p;
	d.internalized = c->internalized_d.p;
	d.fraction_released = c->fraction_released.p;
	d.volumes = c->volumes.p;
	d.is_active = c->is_active.p;
	d.cell_voxel = c->cell_voxel.p;
	d.owner_of = c->owner_of.p;
	d.ballot = c->ballot.p;
	d.num = c->num.p;
	d.den = c->den.p;
	d.fac = c->fac.p;
	d.rnum = c->rnum.p;
	d.rden = c->rden.p;
	d.rfac = c->rfac.p;
	d.dnew = c->dnew.p;
	d.slot_alone = c->slot_alone.p;
	d.joined = c->joined.p;
	d.shared_list = c->shared_list.p;
	d.owners = c->owners.p;
	d.pvox = c->pvox.p;
	d.slot_of = c->slot_of.p;
	d.pnum = c->pnum.p;
	d.pden = c->pden.p;
	d.pfac = c->pfac.p;
	d.flags = c->flags.p;
	d.compute_internalized = c->internalized ? 1 : 0;
	TRY_HIP(c, c->ddom.upload(&d, 1, c->stream));
	c->version++;
	return BFVM_OK;
}

// Tiles of one axis for one domain (dom index `dom` of a launch).
// Tiles of the pipelined sweeps for one domain, padded with idle tiles (a < 0)
// to a multiple of the workgroup's waves so that no round spans two domains.
void domain_pipe_tiles(int axis, int dom, int nx, int ny, int nz, int S, const PipeGeometry& g,
					   std::vector<TileRef>& t)
{
	const int L = nx * S;
	const size_t first = t.size();
	if (axis == 0)
	{
		const int rows = ny * nz;
		for (int r0 = 0; r0 < rows; r0 += g.x_rows)
			t.push_back({ dom, r0, std::min(g.x_rows, rows - r0), 0 });
	}
	else
	{
		const int64_t total = (int64_t)(axis == 1 ? nz : ny) * L;
		for (int64_t f0 = 0; f0 < total; f0 += 64)
			t.push_back({ dom, (int)f0, 0, 0 });
	}
	while ((t.size() - first) % (size_t)g.waves)
		t.push_back({ dom, -1, 0, 0 });
}

// `threads`: the launch's workgroup size (sweep_threads), which sets the
// lines of a one-lane register tile.
void domain_tiles(int axis, int variant, int dom, int nx, int ny, int nz, int S, int fac_cap, std::vector<TileRef>& t,
				  int threads = kSweepThreads)
{
	const int L = nx * S;
	if (axis == 0 && variant == kSweepScan)
	{
		const int rows = ny * nz;
		for (int r0 = 0; r0 < rows; r0 += kScanRowsPerTile)
			t.push_back({ dom, r0, std::min(kScanRowsPerTile, rows - r0), 0 });
		return;
	}
	// chunked 4-byte x kernels: the rows of the workgroup's waves (of 2 wave
	// pairs for the two-lane kernel) from the first row
	if (axis == 0 && (variant == kSweepXG || variant == kSweepXGT))
	{
		const int rows = ny * nz;
		const int per_tile = variant == kSweepXGT ? xg_rows(S) * (kSweepThreads / 64) / 2 : xg_rows(S) * (threads / 64);
		for (int r0 = 0; r0 < rows; r0 += per_tile)
			t.push_back({ dom, r0, 0, 0 });
		return;
	}
	// x tiles of whole rows (LDS-staged kernels): first row, row count
	if (axis == 0 && variant != kSweepReg && variant != kSweepXT)
	{
		const int R = std::max(1, x_rows_per_tile(S, L, fac_cap));
		const int rows = ny * nz;
		for (int r0 = 0; r0 < rows; r0 += R)
			t.push_back({ dom, r0, std::min(R, rows - r0), 0 });
		return;
	}
	const int outer = axis == 1 ? nz : ny;
	if (variant == kSweepReg || variant == kSweepTwist || variant == kSweepXT)
	{
		// flattened (outer, column) space: x columns are the S lines of a row
		const int64_t total = axis == 0 ? (int64_t)ny * nz * S : (int64_t)outer * L;
		const int per_tile = variant == kSweepTwist ? kTwistColsPerTile : threads;
		for (int64_t f0 = 0; f0 < total; f0 += per_tile)
			t.push_back({ dom, (int)f0, 0, 0 });
		return;
	}
	for (int o = 0; o < outer; o++)
		for (int c0 = 0; c0 < L; c0 += kLinesPerTile)
			t.push_back({ dom, o, c0, 0 });
}

int build_tiles(bfvm_ctx* c)
{
	std::vector<TileRef> host_tiles[3];
	for (int axis = 0; axis < 3; axis++)
	{
		std::vector<TileRef> t;
		c->variant[axis] = kSweepLds;
		c->nmax[axis] = 0;
		if (axis < c->dims)
		{
			const int n = (int)c->n[axis];
			const int bucket = reg_bucket(n);
			const int vpl = xscan_vpl(n, c->S);
			const int xs = c->tuning.x_sweep == BFVM_X_AUTO ? default_x_sweep(n, c->S) : c->tuning.x_sweep;
			const int64_t lines = axis == 0 ? c->n[1] * c->n[2] * c->S : c->V / c->n[axis] * c->S;
			const int pbucket = pipe_bucket(axis, n);
			const PipeGeometry pg = pipe_geometry(axis, n, pbucket, c->S, c->L, lines);
			c->pipe[axis] = PipeGeometry {};
			if (c->tuning.force_lds_sweeps)
				;
			// pipelined register lines (FP64) on request: y/z lines of at least
			// kPipeMinLine voxels (shorter tiles are too short to hide the next
			// tile's load), x by its own x_sweep value.  At 200^3 x 4 on MI355X
			// the plain register lines are faster (y 58 us, z 57 us against 61
			// and 66 us pipelined), so neither is the default.
			else if (c->fp64 && pbucket && pg.waves > 0 &&
					 (axis == 0 ? c->tuning.x_sweep == BFVM_X_PIPELINED
								: c->tuning.pipeline && n >= kPipeMinLine))
			{
				c->variant[axis] = kSweepPipe;
				c->nmax[axis] = pbucket;
				c->pipe[axis] = pg;
			}
			// the scan and chunked kernels do not apply faces, which 1-D domains need in this sweep
			// (the chunked kernel also needs a full wave of rows, 64/S)
			else if (axis == 0 && xs == BFVM_X_CHUNKED && xt_bucket(n) && c->fp64 &&
					 (c->S == 4 || c->S == 8 || (c->S == 2 && n % 2 == 0 && n < 176)) && c->n[1] * c->n[2] >= 64 / c->S &&
					 !(c->dims == 1 && c->face_mask))
			{
				c->variant[axis] = kSweepXT;
				c->nmax[axis] = xt_bucket(n);
			}
			else if (axis == 0 && (xs == BFVM_X_CHUNKED || xs == BFVM_X_CHUNKED4) && xg_bucket(n, c->S) &&
					 c->fp64 && c->n[1] * c->n[2] >= xg_rows(c->S) && !(c->dims == 1 && c->face_mask))
			{
				c->variant[axis] = kSweepXG;
				c->nmax[axis] = xg_bucket(n, c->S);
			}
			// rows too long for one lane: two lanes per line (the top and bottom
			// halves in two waves)
			else if (axis == 0 &&
					 (((xs == BFVM_X_CHUNKED || xs == BFVM_X_CHUNKED4 || xs == BFVM_X_TWO_LANE) &&
					   xgt_bucket(n, c->S)) ||
					  (xs == BFVM_X_TWO_LANE && xgt_short_bucket(n, c->S))) &&
					 c->fp64 && c->n[1] * c->n[2] >= xg_rows(c->S) && !(c->dims == 1 && c->face_mask))
			{
				c->variant[axis] = kSweepXGT;
				c->nmax[axis] = xgt_bucket(n, c->S) ? xgt_bucket(n, c->S) : xgt_short_bucket(n, c->S);
			}
			// the scan kernel: asked for, or the default's chunked choice did not
			// take the shape (too few rows for a wave, or 1-D with faces)
			else if (axis == 0 && vpl && !(c->dims == 1 && c->face_mask) &&
					 (xs == BFVM_X_SCAN ||
					  (c->tuning.x_sweep == BFVM_X_AUTO && (xs == BFVM_X_CHUNKED || xs == BFVM_X_CHUNKED4))))
			{
				c->variant[axis] = kSweepScan;
				c->nmax[axis] = vpl;
			}
			else if (axis == 0 && xs == BFVM_X_TRANSPOSE && bucket)
			{
				c->variant[axis] = kSweepRows;
				c->nmax[axis] = bucket;
			}
			// y/z lines of 209..256 voxels: two lanes per line, two waves per
			// SIMD instead of one (ms/step, one lane / two, y lines: 224^3 x 4
			// 0.239 / 0.221, 240^3 x 4 0.265 / 0.251, 250^3 x 4 0.292 / 0.285,
			// 216^3 x 3 0.181 / 0.172); lines of 137..208 when one lane per
			// line needs more than one and at most two rounds of the register
			// file at two waves per SIMD (three with half lines: 192^3 x 4
			// 0.134 / 0.128, 144^3 x 8 0.115 / 0.110, 160^3 x 8 0.144 / 0.138;
			// within one round the extra waves only cost, 176^3 x 4 0.089 /
			// 0.096, and at 200^3 x 8, 2.4 rounds, 0.249 / 0.252).  Results
			// differ by rounding only (4e-7 after 10 steps).  The last sweep
			// (with the faces) only for lines of 209..256 (200^3 x 4 with an
			// x-min face: z 46.4 us back to back on one lane, 47.7 on two); z
			// also applies a uniform bulk as it stores, a 2-D domain's y only
			// one lane per line does, so there a bulk list of that shape
			// (judged on its voxels, whatever its rates, so that the kernels
			// do not change with them) keeps one lane per line
			else if (axis > 0 && c->fp64 && twist_short_bucket(n) &&
					 (axis < c->dims - 1 ? n > 208 || (lines > (int64_t)c->num_cu * 512 && lines <= (int64_t)c->num_cu * 1024)
										 : n > 208 && (axis == 2 || !c->bulk_layout)))
			{
				c->variant[axis] = kSweepTwist;
				c->nmax[axis] = twist_short_bucket(n);
			}
			else if (bucket)
			{
				c->variant[axis] = kSweepReg;
				c->nmax[axis] = bucket;
			}
			// y/z lines of 257..512: two lanes per line in registers
			else if (axis > 0 && c->fp64 && twist_bucket(n))
			{
				c->variant[axis] = kSweepTwist;
				c->nmax[axis] = twist_bucket(n);
			}
			// register kernels pad the factor rows to their bucket (the twisted
			// ones also carry a 16-entry zero prefix per substrate)
			const bool padded = c->variant[axis] == kSweepReg || c->variant[axis] == kSweepRows ||
								c->variant[axis] == kSweepPipe || c->variant[axis] == kSweepXT ||
								c->variant[axis] == kSweepXG;
			c->fac_cap[axis] = c->variant[axis] == kSweepTwist  ? fac_row(c->nmax[axis] + 16) * c->S
							   : c->variant[axis] == kSweepXGT ? fac_row(c->nmax[axis] + kXgtPrefix) * c->S
							   : padded						   ? fac_row(c->nmax[axis]) * c->S
															   : n * c->S;
			if (c->variant[axis] == kSweepPipe)
			{
				domain_pipe_tiles(axis, 0, (int)c->n[0], (int)c->n[1], (int)c->n[2], c->S, c->pipe[axis], t);
				c->lds[axis] = c->pipe[axis].lds_bytes;
				c->threads[axis] = kSweepThreads;
			}
			else
			{
				c->threads[axis] = sweep_threads(c->variant[axis], c->tuning.lines_per_block);
				domain_tiles(axis, c->variant[axis], 0, (int)c->n[0], (int)c->n[1], (int)c->n[2], c->S,
							 c->fac_cap[axis], t, c->threads[axis]);
				c->lds[axis] =
					sweep_lds_bytes(axis, c->variant[axis], n, c->S, c->L, c->fac_cap[axis], c->threads[axis]);
			}
			if (c->lds[axis] > (size_t)kMaxLdsBytes)
				return c->fail(BFVM_ERR_UNSUPPORTED, "grid axis " + std::to_string(axis) + " of length " +
														 std::to_string(n) + " needs more than 160 KiB of LDS per tile");
		}
		c->ntiles[axis] = (int)t.size();
		TRY_HIP(c, c->tiles[axis].upload(t.data(), t.size(), c->stream));
		host_tiles[axis] = std::move(t);
	}
	// x in two halves (3-D domains; x tiles of rows or of lines, y tiles of
	// columns): the last round of x waves then overlaps the y tiles of the
	// first half's planes instead of running alone.  By default where that
	// was measured to pay: the x lanes (one or two per line; the scan kernel
	// excepted, which loops over its rows) and the y lanes each fill more
	// than one round of the register file at two waves per SIMD, and the y
	// lines run two waves per SIMD (ms/step, one launch / halves: 200^3 x 4
	// 0.159 / 0.151, 400x400x100 x 2 0.158 / 0.144, 300^3 x 1 (scan) 0.174 /
	// 0.161, 256^3 x 8 0.674 / 0.662, 250^3 x 4 0.286 / 0.285).  Below one
	// round the halves only add launches (176^3 x 4 0.090 / 0.095,
	// 400x300x100 x 2 0.115 / 0.119, 300x120x120 x 1 0.053 / 0.059); y lines
	// in one-lane buckets of 224..256 (one wave per SIMD) crowd out the second
	// half (FP64 takes two lanes per line there)
	c->split_x = c->split_y = 0;
	const int vx = c->variant[0], vy = c->variant[1];
	const int64_t lanes = c->n[1] * c->n[2] * c->S * (vx == kSweepXGT ? 2 : 1);
	const int64_t ylanes = c->n[0] * c->n[2] * c->S * (vy == kSweepTwist ? 2 : 1);
	const int64_t round = (int64_t)c->num_cu * 4 * 2 * 64;
	bool want = c->tuning.split_xy == 1;
	if (c->tuning.split_xy == 0)
		want = (vx == kSweepScan || lanes > round) && ylanes > round &&
			   (vy == kSweepTwist || (vy == kSweepReg && c->nmax[1] <= 208));
	if (want && c->dims == 3 && host_tiles[0].size() >= 2 &&
		(vx == kSweepXG || vx == kSweepXGT || vx == kSweepXT || vx == kSweepReg || vx == kSweepScan) &&
		(vy == kSweepReg || vy == kSweepTwist))
	{
		const std::vector<TileRef>& tx = host_tiles[0];
		const std::vector<TileRef>& ty = host_tiles[1];
		// the first half's share of the x tiles, by the rounds the x lanes
		// need (measured at 30-75%, ms/step: 200x164x200 x 4, 1.0 rounds,
		// 0.1235 at 70% against 0.1246 at 60%; 200^3 x 4, 1.22 rounds, 0.138
		// at 60% against 0.141 at 50% and 70%; 192^3 x 6, 1.69 rounds, 0.175
		// at 50% against 0.184 at 60%; 200^3 x 8, 2.44 rounds, 0.246 at 50%
		// against 0.251 at 60%)
		const double r = (double)lanes / (double)round;
		const int tenths = r < 1.1 ? 7 : r < 1.4 ? 6 : 5;
		const int kx = std::clamp((int)(tx.size() * tenths / 10), 1, (int)tx.size() - 1);
		// lines (row*S + s) of the first half: those below tile kx's first
		const int64_t first = vx == kSweepXT || vx == kSweepReg ? (int64_t)tx[kx].a : (int64_t)tx[kx].a * c->S;
		const int64_t planes = first / c->S / c->n[1]; // planes of whole rows in the first half
		const int per = vy == kSweepTwist ? kTwistColsPerTile : c->threads[1];
		int ky = 0;
		while (ky < (int)ty.size() && (int64_t)ty[ky].a + per <= planes * c->L)
			ky++;
		if (ky > 0 && ky < (int)ty.size())
		{
			if (!c->side)
				TRY_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
			for (hipEvent_t& e : c->ev_split)
				if (!e)
					TRY_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
			c->split_x = kx;
			c->split_y = ky;
		}
	}
	return BFVM_OK;
}

int sweep_grid(const bfvm_ctx* c, int axis)
{
	int per_cu = c->tuning.persistent_wgs_per_cu;
	// default: the scan kernel loops (2 workgroups per CU), which keeps its
	// factors in registers across rows; the others take one workgroup per tile
	if (c->variant[axis] == kSweepPipe)
	{
		// persistent: as many workgroups per CU as the LDS holds (short lines
		// leave room for several), each looping over rounds of tiles
		const int rounds = c->ntiles[axis] / c->pipe[axis].waves;
		const int fit = std::max(1, (int)((size_t)kMaxLdsBytes / c->pipe[axis].lds_bytes));
		return std::min(rounds, c->num_cu * (per_cu > 0 ? per_cu : fit));
	}
	if (per_cu == 0)
		per_cu = c->variant[axis] == kSweepScan ? 2 : -1;
	if (per_cu > 0)
		return std::min(c->ntiles[axis], c->num_cu * per_cu);
	return c->ntiles[axis];
}

// The sweep of one axis on the context stream (faces fused into the last axis).
hipError_t launch_axis_sweep(bfvm_ctx* c, int axis, const DomainDev* dd)
{
	const bool faces = axis == c->dims - 1 && c->face_mask != 0;
	if (c->variant[axis] == kSweepPipe)
		return launch_sweep_pipe(axis, c->nmax[axis], faces, dd, c->tiles[axis].p, c->ntiles[axis], 1,
								 c->fac_cap[axis], c->pipe[axis], sweep_grid(c, axis), c->stream);
	const int fin = (axis == c->dims - 1 && c->bulk_fused) ? kFinBulk : faces ? kFinFaces : kFinNone;
	return launch_sweep(axis, c->variant[axis], c->nmax[axis], c->fp64, fin, c->S, dd, c->tiles[axis].p,
						c->ntiles[axis], 1, c->fac_cap[axis], c->lds[axis], sweep_grid(c, axis), c->stream,
						c->threads[axis]);
}

// Tiles [first, first + count) of one axis' sweep on `stream` (an axis that
// is not the last: no face or bulk stage).
hipError_t launch_axis_part(bfvm_ctx* c, int axis, const DomainDev* dd, int first, int count, hipStream_t stream)
{
	const int per_cu = c->tuning.persistent_wgs_per_cu > 0 ? c->tuning.persistent_wgs_per_cu
					   : c->variant[axis] == kSweepScan	   ? 2
														   : 0;
	const int grid = per_cu > 0 ? std::min(count, c->num_cu * per_cu) : count;
	return launch_sweep(axis, c->variant[axis], c->nmax[axis], c->fp64, kFinNone, c->S, dd, c->tiles[axis].p + first,
						count, 1, c->fac_cap[axis], c->lds[axis], grid, stream, c->threads[axis]);
}

// One step on the stream: sweeps x, y, z (faces fused into the last), interior
// Dirichlet, bulk, cells.  Stage order of openmp_solver.cpp:27-37.  With
// split_x the x sweep and the y sweep run as halves forked onto the side
// stream and joined back before z, so the step still starts and ends on the
// context stream.  `ev` (7 events or null) brackets the stages for timing
// (one launch per sweep then).  `dd` is the domain table
// of the step (its step_offset places it in a captured run of steps), and
// `tick` the steps the step count advances by at its end (0: none).
int enqueue_step(bfvm_ctx* c, hipEvent_t* ev, const DomainDev* dd = nullptr, uint32_t tick = 1)
{
	if (!dd)
		dd = c->ddom.p;
	auto mark = [&](int k) -> hipError_t { return ev ? hipEventRecord(ev[k], c->stream) : hipSuccess; };
	TRY_HIP(c, mark(0));
	if (c->split_x && !ev)
	{
		// x halves A (context stream) and B (side stream); y of A's planes
		// after A, the other y tiles after both; z after all of y
		const int nx = c->ntiles[0] - c->split_x, ny = c->ntiles[1] - c->split_y;
		TRY_HIP(c, hipEventRecord(c->ev_split[0], c->stream));
		TRY_HIP(c, hipStreamWaitEvent(c->side, c->ev_split[0], 0));
		TRY_HIP(c, launch_axis_part(c, 0, dd, 0, c->split_x, c->stream));
		TRY_HIP(c, hipEventRecord(c->ev_split[1], c->stream));
		TRY_HIP(c, launch_axis_part(c, 0, dd, c->split_x, nx, c->side));
		TRY_HIP(c, launch_axis_part(c, 1, dd, 0, c->split_y, c->stream));
		TRY_HIP(c, hipStreamWaitEvent(c->side, c->ev_split[1], 0));
		TRY_HIP(c, launch_axis_part(c, 1, dd, c->split_y, ny, c->side));
		TRY_HIP(c, hipEventRecord(c->ev_split[2], c->side));
		TRY_HIP(c, hipStreamWaitEvent(c->stream, c->ev_split[2], 0));
		TRY_HIP(c, launch_axis_sweep(c, 2, dd));
	}
	else
		for (int axis = 0; axis < 3; axis++)
		{
			if (axis < c->dims)
				TRY_HIP(c, launch_axis_sweep(c, axis, dd));
			TRY_HIP(c, mark(axis + 1));
		}
	if (c->dir_count)
		TRY_HIP(c, launch_dirichlet_interior(dd, c->dir_refs.p, c->dir_count, c->stream));
	TRY_HIP(c, mark(4));
	// bulk and cell apply as one launch when both run (voxels in both lists
	// are applied in order by the bulk side); the stage timer then counts the
	// cell apply with bulk.  The cell tiles cover every agent slot; the kernels
	// stop at the owner count of the last recompute, which stays on the device
	// (as does the shared-voxel flag the internalized update checks), so the
	// launches do not change when the agents move
	// (a uniform bulk over every voxel may have been applied by the last sweep)
	const bool cells = c->n_agents > 0 && c->n_cell_tiles > 0;
	const bool bulk = c->bulk_count && !c->bulk_fused;
	const bool both = bulk && c->n_bulk_tiles > 0 && cells && c->overlap;
	if (both)
		TRY_HIP(c, launch_stages(dd, c->bulk_tiles.p, c->n_bulk_tiles, c->cell_tiles.p, c->n_cell_tiles,
								 c->stream));
	else if (bulk)
		TRY_HIP(c, launch_bulk(dd, c->bulk_tiles.p, c->n_bulk_tiles, c->stream));
	TRY_HIP(c, mark(5));
	if (cells)
	{
		if (!both)
			TRY_HIP(c, launch_cells_apply(dd, c->cell_tiles.p, c->n_cell_tiles, c->stream));
		if (c->internalized)
			TRY_HIP(c, launch_cells_internalized(dd, c->int_tiles.p, c->n_int_tiles, c->stream));
	}
	if (tick)
		TRY_HIP(c, launch_step_tick(dd, 1, tick, c->stream));
	TRY_HIP(c, mark(6));
	return BFVM_OK;
}

// The bulk/apply overlap maps for the current lists (after a recompute or a
// new bulk list); without both lists the stages run as separate launches.
int rebuild_overlap(bfvm_ctx* c)
{
	c->overlap = false;
	if (c->bulk_count == 0 || c->n_agents == 0 || c->recompute_pending || !c->bulk_map.p)
		return BFVM_OK;
	TRY_HIP(c, launch_bulk_overlap(c->ddom.p, c->bulk_map.p, c->bulk_cell.p, c->bulk_count, c->n_agents, c->stream));
	c->overlap = true;
	return BFVM_OK;
}

int bits_for(uint64_t v)
{
	int b = 0;
	while (b < 32 && (v >> b) != 0)
		b++;
	return std::max(b, 1);
}

// Re-bins the agents and rebuilds the per-voxel sums (the recompute branch of
// cell_solver::simulate_secretion_and_uptake, openmp_solver/src/cell_solver.cpp:302-311, :253-267).
int recompute(bfvm_ctx* c)
{
	c->recompute_pending = false;
	c->overlap = false;
	const int64_t n = c->n_agents;
	if (n == 0)
		return BFVM_OK;
	// everything below is enqueued on the stream: the owner count and the
	// flags stay on the device (the sort buffers were sized by bfvm_set_agents)
	hipStream_t s = c->stream;
	TRY_HIP(c, hipMemsetAsync(c->flags.p, 0, kCellFlags * sizeof(int32_t), s));
	if (c->reduce_mode == BFVM_REDUCE_ATOMIC)
	{
		TRY_HIP(c, launch_cells_bin(c->ddom.p, n, c->S, true, s));
		TRY_HIP(c, launch_cells_ballot(c->ddom.p, n, s));
		// the owner list in voxel order (the per-step apply then walks the
		// field in order; the order changes no result): the agents sorted by
		// (owner voxel, or V for the others), the owners counted on the device
		TRY_HIP(c, launch_cells_owner_keys(c->ddom.p, c->keys.p, c->vals.p, n, s));
		TRY_HIP(c, radix_sort_pairs(c->keys.p, c->vals.p, c->tkeys.p, c->tvals.p, c->hist.p, c->scan_tmp.p, n,
									bits_for((uint64_t)c->V), s));
		TRY_HIP(c, launch_cells_count_owners(c->keys.p, n, (uint32_t)c->V, c->flags.p, s));
		TRY_HIP(c, hipMemcpyAsync(c->owners.p, c->vals.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
	}
	else
	{
		TRY_HIP(c, launch_cells_bin(c->ddom.p, n, c->S, false, s));
		TRY_HIP(c, launch_cells_make_keys(c->ddom.p, c->keys.p, c->vals.p, n, s));
		TRY_HIP(c, radix_sort_pairs(c->keys.p, c->vals.p, c->tkeys.p, c->tvals.p, c->hist.p, c->scan_tmp.p, n,
									bits_for((uint64_t)c->V), s));
		TRY_HIP(c, launch_cells_heads(c->keys.p, c->heads.p, n, (uint32_t)c->V, s));
		TRY_HIP(c, hipMemcpyAsync(c->head_scan.p, c->heads.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
		TRY_HIP(c, exclusive_scan(c->head_scan.p, c->scan_tmp.p, n, s));
		TRY_HIP(c, launch_cells_segment_reduce(c->ddom.p, c->keys.p, c->vals.p, c->heads.p, c->head_scan.p, n, s));
	}
	TRY_HIP(c, launch_cells_pack(c->ddom.p, n, c->S, c->reduce_mode == BFVM_REDUCE_ATOMIC, s));
	TRY_HIP(c, launch_cells_agent_slots(c->ddom.p, n, s));
	// (the scan reuses the sorted path's head scan buffer, free by now)
	if (c->internalized)
	{
		TRY_HIP(c, launch_cells_owner_order(c->ddom.p, n, s));
		TRY_HIP(c, launch_cells_shared_list(c->ddom.p, n, c->head_scan.p, c->scan_tmp.p, s));
	}
	c->grid_check_pending = true;
	return rebuild_overlap(c);
}

// After the stream has drained: reports an active agent outside the mesh
// found by the last recompute (the agent took no part in the steps since; the
// positional data stays marked for another recompute).
int check_recompute(bfvm_ctx* c)
{
	if (!c->grid_check_pending)
		return BFVM_OK;
	c->grid_check_pending = false;
	int32_t f[4] = { 0, 0, 0, 0 };
	TRY_HIP(c, hipMemcpyAsync(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	c->n_owners = f[2];
	c->conflict = f[0] != 0;
	if (f[1])
	{
		c->recompute_pending = true;
		return c->fail(BFVM_ERR_OUT_OF_GRID, "an active agent lies outside the mesh");
	}
	return BFVM_OK;
}

int valid_desc(const bfvm_domain_desc* d, std::string& why)
{
	if (!d)
		return why = "null descriptor", 0;
	if (d->dims < 1 || d->dims > 3)
		return why = "dims must be 1, 2 or 3", 0;
	if (d->substrates < 1 || d->substrates > 64)
		return why = "substrates must be in [1, 64]", 0;
	for (int a = 0; a < 3; a++)
	{
		if (d->grid[a] < 1)
			return why = "grid sizes must be >= 1", 0;
		if (a >= d->dims && d->grid[a] != 1)
			return why = "grid size of an axis beyond dims must be 1", 0;
		if (a < d->dims && d->voxel[a] < 1)
			return why = "voxel sizes must be >= 1", 0;
	}
	if ((double)d->grid[0] * d->grid[1] * d->grid[2] > 2147483647.0)
		return why = "more than 2^31-1 voxels", 0;
	if ((double)d->grid[0] * d->grid[1] * d->grid[2] * d->substrates * 4.0 > 2147483647.0)
		return why = "field larger than 2 GiB (the sweeps address it with 32-bit byte offsets)", 0;
	if (!(d->dt > 0) || !fp_ok(d->dt))
		return why = "dt must be positive", 0;
	if (!d->diffusion_coefficients || !d->decay_rates || !d->initial_conditions)
		return why = "null coefficient arrays", 0;
	for (int s = 0; s < d->substrates; s++)
		if (!fp_ok(d->diffusion_coefficients[s]) || !fp_ok(d->decay_rates[s]) || !fp_ok(d->initial_conditions[s]) ||
			d->diffusion_coefficients[s] < 0 || d->decay_rates[s] < 0)
			return why = "coefficients must be finite and non-negative", 0;
	if (d->precision != BFVM_SWEEP_FP64 && d->precision != BFVM_SWEEP_FP32)
		return why = "unknown precision", 0;
	if (d->reduce_mode != BFVM_REDUCE_ATOMIC && d->reduce_mode != BFVM_REDUCE_SORTED)
		return why = "unknown reduce mode", 0;
	return 1;
}

bool voxel_of(const bfvm_ctx* c, const double* pos, int64_t* v)
{
	int64_t idx[3] = { 0, 0, 0 };
	for (int a = 0; a < c->dims; a++)
	{
		const double q = (pos[a] - (double)c->bbox_min[a]) / (double)c->voxel[a];
		if (!(q >= 0.0) || q >= (double)c->n[a])
			return false;
		idx[a] = (int64_t)q;
	}
	*v = idx[0] + c->n[0] * (idx[1] + c->n[1] * idx[2]);
	return true;
}

std::vector<SparseRef> refs_for(int64_t count)
{
	std::vector<SparseRef> r((size_t)count);
	for (int64_t k = 0; k < count; k++)
		r[(size_t)k] = { 0, (int32_t)k };
	return r;
}

// Applies the logged releases: grouped by voxel (call order kept within a
// voxel), summed in FP64 and rounded to the field once per voxel.
int flush_releases(bfvm_ctx* c)
{
	const size_t n = c->rel_agent.size();
	if (n == 0)
		return BFVM_OK;
	std::vector<size_t> order(n);
	for (size_t k = 0; k < n; k++)
		order[k] = k;
	std::stable_sort(order.begin(), order.end(),
					 [&](size_t a, size_t b) { return c->rel_voxel[a] < c->rel_voxel[b]; });
	std::vector<int64_t> groups, agents(n);
	for (size_t k = 0; k < n; k++)
	{
		agents[k] = c->rel_agent[order[k]];
		const int64_t v = c->rel_voxel[order[k]];
		if (k == 0 || v != groups[groups.size() - 3])
			groups.insert(groups.end(), { v, (int64_t)k, 0 });
		groups.back()++;
	}
	const int64_t n_groups = (int64_t)groups.size() / 3;
	std::vector<int64_t> packed(groups);
	packed.insert(packed.end(), agents.begin(), agents.end());
	c->rel_agent.clear();
	c->rel_voxel.clear();
	TRY_HIP(c, c->rel_dev.upload(packed.data(), packed.size(), c->stream));
	TRY_HIP(c, launch_release_flush(c->ddom.p, c->rel_dev.p, n_groups, c->rel_dev.p + 3 * n_groups, c->S, c->stream));
	return BFVM_OK;
}

// The sweep kernels' LDS limits, set once per device (contexts may be
// created from several threads).
hipError_t sweep_kernels_ready(int dev)
{
	static std::mutex mu;
	static std::map<int, hipError_t> done;
	std::lock_guard<std::mutex> lock(mu);
	auto it = done.find(dev);
	if (it == done.end())
		it = done.emplace(dev, init_sweep_kernels()).first;
	return it->second;
}

} // namespace

#define FLUSH_RELEASES(c)               \
	do                                  \
	{                                   \
		int rc_ = flush_releases(c);    \
		if (rc_)                        \
			return rc_;                 \
	} while (0)

// ---------------------------------------------------------------- batches
//
// A batch steps several contexts together: each sweep is one launch per group
// of domains that share the kernel (axis, variant, register bucket; for the
// x kernels templated on S also S) over the concatenated tile lists, so many
// small domains fill the GPU that one of them alone would leave idle.  The
// boundary, bulk and cell stages launch once per substrate count.  Every
// domain runs the same stage order as alone, so its arithmetic is identical.
struct bfvm_batch
{
	std::string err;
	std::vector<bfvm_ctx*> ctx;
	hipStream_t stream = nullptr;
	int num_cu = 256;
	hipEvent_t ev_in = nullptr, ev_out = nullptr;
	// the groups of one axis touch disjoint domains: they run on `lanes`
	// streams (BFVM_BATCH_LANES, 1..kMaxBatchLanes, default 4) so that the
	// tail of one launch overlaps the next
	static constexpr int kMaxBatchLanes = 16;
	int lanes = 4;
	hipStream_t lane[kMaxBatchLanes - 1] = {};
	hipEvent_t ev_fork = nullptr, ev_join[kMaxBatchLanes - 1] = {};
	// bfvm_batch_mark / bfvm_batch_wait: ticket t completes with
	// marks[t % kMarkRing] (created on first use)
	static constexpr int kMarkRing = 8;
	hipEvent_t marks[kMarkRing] = {};
	int64_t next_mark = 0;

	struct SweepLaunch
	{
		int axis = 0, variant = 0, nmax = 0, S = 0;
		bool fp64 = true, faces = false;
		bool bulk = false; // the domains' last sweep, applying their uniform bulk
		int fac_cap = 0, grid = 0, ntiles = 0;
		int threads = kSweepThreads; // workgroup size
		size_t lds = 0;
		int64_t units = 0; // voxel*substrates swept: the launch's share of the axis
		int lane = 0;	   // stream of the launch among the axis' lanes
		PipeGeometry pg;
		std::vector<int> doms;
		dvec<TileRef> tiles;
	};
	struct RefLaunch
	{
		int64_t n = 0;
		dvec<SparseRef> refs;
	};
	struct TileLaunch
	{
		int n = 0;
		dvec<TileRef> tiles;
	};
	std::vector<std::unique_ptr<SweepLaunch>> sweeps;
	// the stages after the diffusion of the domains whose last sweep runs on
	// one lane of one axis, queued on that lane behind its sweeps: they
	// overlap the sweeps of the other lanes
	struct StageSet
	{
		std::unique_ptr<RefLaunch> dirichlet;
		std::unique_ptr<TileLaunch> bulk, cells, internalized;
		int launches() const { return (dirichlet ? 1 : 0) + (bulk ? 1 : 0) + (cells ? 1 : 0) + (internalized ? 1 : 0); }
	};
	StageSet stages[3][kMaxBatchLanes];
	dvec<DomainDev> ddoms;
	std::vector<DomainDev> hdoms;
	dvec<DomainDev> dslots; // the captured graph's steps: ddoms with step_offset 0, 1, ...
	std::vector<uint64_t> signature;

	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	int64_t graph_steps = 0;
	bool graph_valid = false;

	int fail(int code, const std::string& m)
	{
		err = m;
		return code;
	}
	int hip_fail(hipError_t e, const char* what)
	{
		err = std::string(what) + ": " + hipGetErrorString(e);
		return e == hipErrorOutOfMemory ? BFVM_ERR_NO_MEMORY : BFVM_ERR_HIP;
	}
	void drop_graph()
	{
		if (gexec)
			(void)hipGraphExecDestroy(gexec);
		if (graph)
			(void)hipGraphDestroy(graph);
		gexec = nullptr;
		graph = nullptr;
		graph_valid = false;
	}
	~bfvm_batch()
	{
		// the launches in flight read the graph, the tables and the tiles
		if (stream)
			(void)hipStreamSynchronize(stream);
		drop_graph();
		for (hipEvent_t e : marks)
			if (e)
				(void)hipEventDestroy(e);
		if (ev_in)
			(void)hipEventDestroy(ev_in);
		if (ev_out)
			(void)hipEventDestroy(ev_out);
		if (ev_fork)
			(void)hipEventDestroy(ev_fork);
		for (int k = 0; k < kMaxBatchLanes - 1; k++)
		{
			if (ev_join[k])
				(void)hipEventDestroy(ev_join[k]);
			if (lane[k])
				(void)hipStreamDestroy(lane[k]);
		}
		if (stream)
			(void)hipStreamDestroy(stream);
	}
};

namespace {

// what the batch layout depends on, per context
std::vector<uint64_t> batch_signature(const bfvm_batch* b)
{
	std::vector<uint64_t> sig;
	for (const bfvm_ctx* c : b->ctx)
	{
		sig.push_back(c->version);
		sig.push_back((uint64_t)c->n_agents);
		sig.push_back((uint64_t)c->dir_count);
		sig.push_back((uint64_t)c->bulk_count);
	}
	return sig;
}

// Brings every context up to date (pending releases and recomputes, on its
// own stream) and rebuilds the launch lists when any of them changed.
int batch_prepare(bfvm_batch* b)
{
	for (size_t i = 0; i < b->ctx.size(); i++)
	{
		bfvm_ctx* c = b->ctx[i];
		int rc = flush_releases(c);
		if (!rc && c->recompute_pending)
			rc = recompute(c);
		if (rc)
			return b->fail(rc, "domain " + std::to_string(i) + ": " + c->err);
	}
	std::vector<uint64_t> sig = batch_signature(b);
	if (sig == b->signature && b->ddoms.p)
		return BFVM_OK;
	b->drop_graph();
	// launches in flight still read the tables about to be replaced (a new
	// batch has none: its first step is ordered after the contexts' work by
	// events, so it need not wait for them here)
	if (b->ddoms.p)
	{
		if (hipError_t e = hipStreamSynchronize(b->stream); e != hipSuccess)
			return b->hip_fail(e, "hipStreamSynchronize");
		for (bfvm_ctx* c : b->ctx)
			if (hipError_t e = hipStreamSynchronize(c->stream); e != hipSuccess)
				return b->hip_fail(e, "hipStreamSynchronize");
	}

	std::vector<DomainDev> h(b->ctx.size());
	for (size_t i = 0; i < b->ctx.size(); i++)
		h[i] = b->ctx[i]->hdom;
	if (hipError_t e = b->ddoms.upload(h.data(), h.size(), b->stream); e != hipSuccess)
		return b->hip_fail(e, "domain table upload");
	b->hdoms = h;

	// sweeps: group domains by kernel
	b->sweeps.clear();
	for (int axis = 0; axis < 3; axis++)
	{
		std::map<std::tuple<int, int, int, int, int, int>, std::vector<int>> groups;
		for (size_t i = 0; i < b->ctx.size(); i++)
		{
			const bfvm_ctx* c = b->ctx[i];
			if (axis >= c->dims)
				continue;
			const int v = c->variant[axis];
			// kernels templated on S: the x scan, the pipelined x kernel
			const int keyS =
				(v == kSweepScan || v == kSweepXT || v == kSweepXG || v == kSweepXGT || (v == kSweepPipe && axis == 0))
					? c->S
					: 0;
			const int bulk = axis == c->dims - 1 && c->bulk_fused ? 1 : 0;
			groups[{ v, c->nmax[axis], c->fp64 ? 1 : 0, keyS, bulk, c->threads[axis] }].push_back((int)i);
		}
		for (auto& [key, doms] : groups)
		{
			auto L = std::make_unique<bfvm_batch::SweepLaunch>();
			L->axis = axis;
			L->variant = std::get<0>(key);
			L->nmax = std::get<1>(key);
			L->fp64 = std::get<2>(key) != 0;
			L->bulk = std::get<4>(key) != 0;
			L->threads = std::get<5>(key);
			L->doms = doms;
			int maxS = 1;
			int64_t lines = 0;
			for (int i : doms)
			{
				const bfvm_ctx* c = b->ctx[i];
				L->fac_cap = std::max(L->fac_cap, c->fac_cap[axis]);
				L->faces |= axis == c->dims - 1 && c->face_mask != 0;
				maxS = std::max(maxS, c->S);
				lines += axis == 0 ? c->n[1] * c->n[2] * c->S : c->V / c->n[axis] * c->S;
				L->units += c->V * c->S;
			}
			L->S = maxS;
			std::vector<TileRef> t;
			if (L->variant == kSweepPipe)
			{
				// one geometry for the group: the largest tile buffer of its
				// domains, the factor table of the most substrates
				PipeGeometry g;
				g.S = maxS;
				for (int i : doms)
				{
					const bfvm_ctx* c = b->ctx[i];
					g.buf_floats = std::max(g.buf_floats, c->pipe[axis].buf_floats);
					g.x_rows = c->pipe[axis].x_rows;
				}
				const size_t fac = (size_t)(fac_row(L->nmax) * maxS + 64) * sizeof(double);
				const size_t buf = (size_t)g.buf_floats * sizeof(float);
				const int64_t tiles = (lines + 63) / 64;
				g.waves = fac + buf > (size_t)kMaxLdsBytes
							  ? 0
							  : (int)std::min<int64_t>({ (int64_t)((kMaxLdsBytes - fac) / buf), 4, tiles });
				if (g.waves <= 0)
					return b->fail(BFVM_ERR_UNSUPPORTED, "pipelined sweep group does not fit the LDS");
				g.lds_bytes = fac + (size_t)g.waves * buf;
				L->pg = g;
				L->fac_cap = fac_row(L->nmax) * maxS;
				L->lds = g.lds_bytes;
				for (int i : doms)
				{
					const bfvm_ctx* c = b->ctx[i];
					domain_pipe_tiles(axis, i, (int)c->n[0], (int)c->n[1], (int)c->n[2], c->S, g, t);
				}
				L->ntiles = (int)t.size();
				const int fit = std::max(1, (int)((size_t)kMaxLdsBytes / g.lds_bytes));
				L->grid = std::min(L->ntiles / g.waves, b->num_cu * fit);
			}
			else
			{
				for (int i : doms)
				{
					const bfvm_ctx* c = b->ctx[i];
					domain_tiles(axis, L->variant, i, (int)c->n[0], (int)c->n[1], (int)c->n[2], c->S, L->fac_cap, t,
								 L->threads);
					L->lds = std::max(L->lds, sweep_lds_bytes(axis, L->variant, (int)c->n[axis], c->S, c->L,
															  L->fac_cap, L->threads));
				}
				if (L->lds > (size_t)kMaxLdsBytes)
					return b->fail(BFVM_ERR_UNSUPPORTED, "sweep group needs more than 160 KiB of LDS");
				L->ntiles = (int)t.size();
				L->grid = L->variant == kSweepScan ? std::min(L->ntiles, 2 * b->num_cu) : L->ntiles;
			}
			if (hipError_t e = L->tiles.upload(t.data(), t.size(), b->stream); e != hipSuccess)
				return b->hip_fail(e, "tile upload");
			b->sweeps.push_back(std::move(L));
		}
	}

	// lanes: per axis, the biggest launches first, each on the lane with the
	// least work so far, so that the lanes finish the axis together.  A
	// launch's work is the data it sweeps (the sweeps are memory-bound) plus
	// the stages of the domains whose last sweep it holds, which follow it on
	// its lane (a sparse entry costs about 10 swept voxel*substrates)
	const int K = b->lanes;
	auto stage_work = [&](int i) -> int64_t {
		const bfvm_ctx* c = b->ctx[i];
		// owners are at most the agents (their count stays on the device)
		const int64_t entries = c->dir_count + (c->bulk_fused ? 0 : c->bulk_count) + c->n_agents + (c->internalized ? c->n_agents : 0);
		return 10 * entries * c->S;
	};
	std::vector<int64_t> work(b->sweeps.size());
	for (size_t j = 0; j < b->sweeps.size(); j++)
	{
		const auto& L = b->sweeps[j];
		work[j] = L->units;
		for (int i : L->doms)
			if (L->axis == b->ctx[i]->dims - 1)
				work[j] += stage_work(i);
	}
	{
		std::vector<size_t> order(b->sweeps.size());
		for (size_t j = 0; j < order.size(); j++)
			order[j] = j;
		std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
			return b->sweeps[x]->axis != b->sweeps[y]->axis ? b->sweeps[x]->axis < b->sweeps[y]->axis
															: work[x] > work[y];
		});
		std::vector<std::unique_ptr<bfvm_batch::SweepLaunch>> sorted;
		for (size_t j : order)
			sorted.push_back(std::move(b->sweeps[j]));
		b->sweeps = std::move(sorted);
		std::vector<int64_t> w2(order.size());
		for (size_t j = 0; j < order.size(); j++)
			w2[j] = work[order[j]];
		work = std::move(w2);
	}
	for (int axis = 0; axis < 3; axis++)
	{
		int count = 0;
		for (auto& L : b->sweeps)
			count += L->axis == axis;
		const int lanes = std::min(K, count);
		int64_t load[bfvm_batch::kMaxBatchLanes] = {};
		for (size_t j = 0; j < b->sweeps.size(); j++)
			if (b->sweeps[j]->axis == axis)
			{
				const int k = (int)(std::min_element(load, load + lanes) - load);
				load[k] += work[j];
				b->sweeps[j]->lane = k;
			}
	}

	// stages per (last axis, lane): boundary refs (domain, entry); bulk,
	// cells and internalized: sparse tiles of each domain
	std::vector<int> lane_of(b->ctx.size(), 0);
	for (auto& L : b->sweeps)
		for (int i : L->doms)
			if (L->axis == b->ctx[i]->dims - 1)
				lane_of[i] = L->lane;
	auto make_tiles = [&](int axis, int lane, auto count_of, std::unique_ptr<bfvm_batch::TileLaunch>& out) -> int {
		std::vector<TileRef> t;
		for (size_t i = 0; i < b->ctx.size(); i++)
			if (b->ctx[i]->dims - 1 == axis && lane_of[i] == lane)
			{
				auto ti = sparse_tiles((int)i, count_of(b->ctx[i]), b->ctx[i]->S);
				t.insert(t.end(), ti.begin(), ti.end());
			}
		out.reset();
		if (t.empty())
			return BFVM_OK;
		out = std::make_unique<bfvm_batch::TileLaunch>();
		out->n = (int)t.size();
		if (hipError_t e = out->tiles.upload(t.data(), t.size(), b->stream); e != hipSuccess)
			return b->hip_fail(e, "tile upload");
		return BFVM_OK;
	};
	for (int axis = 0; axis < 3; axis++)
		for (int lane = 0; lane < K; lane++)
		{
			bfvm_batch::StageSet& st = b->stages[axis][lane];
			std::vector<SparseRef> v;
			for (size_t i = 0; i < b->ctx.size(); i++)
				if (b->ctx[i]->dims - 1 == axis && lane_of[i] == lane)
					for (int64_t k = 0; k < b->ctx[i]->dir_count; k++)
						v.push_back({ (int32_t)i, (int32_t)k });
			st.dirichlet.reset();
			if (!v.empty())
			{
				st.dirichlet = std::make_unique<bfvm_batch::RefLaunch>();
				st.dirichlet->n = (int64_t)v.size();
				if (hipError_t e = st.dirichlet->refs.upload(v.data(), v.size(), b->stream); e != hipSuccess)
					return b->hip_fail(e, "ref upload");
			}
			int rc = make_tiles(
				axis, lane, [](const bfvm_ctx* c) { return c->bulk_fused ? (int64_t)0 : c->bulk_count; }, st.bulk);
			if (!rc)
				rc = make_tiles(axis, lane, [](const bfvm_ctx* c) { return c->n_agents; }, st.cells);
			if (!rc)
				rc = make_tiles(axis, lane,
								[](const bfvm_ctx* c) { return c->internalized ? c->n_agents : 0; },
								st.internalized);
			if (rc)
				return rc;
		}
	b->signature = sig;
	return BFVM_OK;
}

int batch_enqueue_stages(bfvm_batch* b, const DomainDev* dd, const bfvm_batch::StageSet& st, hipStream_t stream)
{
	if (st.dirichlet)
		if (hipError_t e = launch_dirichlet_interior(dd, st.dirichlet->refs.p, st.dirichlet->n, stream);
			e != hipSuccess)
			return b->hip_fail(e, "dirichlet launch");
	if (st.bulk)
		if (hipError_t e = launch_bulk(dd, st.bulk->tiles.p, st.bulk->n, stream); e != hipSuccess)
			return b->hip_fail(e, "bulk launch");
	if (st.cells)
		if (hipError_t e = launch_cells_apply(dd, st.cells->tiles.p, st.cells->n, stream); e != hipSuccess)
			return b->hip_fail(e, "cells launch");
	if (st.internalized)
		if (hipError_t e = launch_cells_internalized(dd, st.internalized->tiles.p, st.internalized->n, stream);
			e != hipSuccess)
			return b->hip_fail(e, "internalized launch");
	return BFVM_OK;
}

// One step of every domain: per axis, the groups on their lanes (forked from
// and joined back into the batch stream), each lane followed by the stages of
// the domains whose last sweep it ran; then the step counts advance by `tick`
// (0: none).  `dd` is the domain table of the step (its step_offset places it
// in a captured run of steps).
int batch_enqueue_step(bfvm_batch* b, const DomainDev* dd = nullptr, uint32_t tick = 1)
{
	if (!dd)
		dd = b->ddoms.p;
	for (int axis = 0; axis < 3; axis++)
	{
		int lanes = 0;
		for (auto& L : b->sweeps)
			if (L->axis == axis)
				lanes = std::max(lanes, L->lane + 1);
		if (!lanes)
			continue;
		auto lane_stream = [&](int k) { return k == 0 ? b->stream : b->lane[k - 1]; };
		if (lanes > 1)
		{
			if (hipError_t e = hipEventRecord(b->ev_fork, b->stream); e != hipSuccess)
				return b->hip_fail(e, "hipEventRecord");
			for (int k = 0; k < lanes - 1; k++)
				if (hipError_t e = hipStreamWaitEvent(b->lane[k], b->ev_fork, 0); e != hipSuccess)
					return b->hip_fail(e, "hipStreamWaitEvent");
		}
		for (auto& Lp : b->sweeps)
		{
			const bfvm_batch::SweepLaunch* L = Lp.get();
			if (L->axis != axis)
				continue;
			hipStream_t st = lane_stream(L->lane);
			hipError_t e;
			if (L->variant == kSweepPipe)
				e = launch_sweep_pipe(L->axis, L->nmax, L->faces, dd, L->tiles.p, L->ntiles, 1, L->fac_cap,
									  L->pg, L->grid, st);
			else
				e = launch_sweep(L->axis, L->variant, L->nmax, L->fp64,
								 L->bulk ? kFinBulk : L->faces ? kFinFaces : kFinNone, L->S, dd, L->tiles.p,
								 L->ntiles, 1, L->fac_cap, L->lds, L->grid, st, L->threads);
			if (e != hipSuccess)
				return b->hip_fail(e, "sweep launch");
		}
		for (int k = 0; k < lanes; k++)
			if (int rc = batch_enqueue_stages(b, dd, b->stages[axis][k], lane_stream(k)))
				return rc;
		// the next axis starts when every group of this one is done
		for (int k = 0; k < lanes - 1; k++)
		{
			if (hipError_t e = hipEventRecord(b->ev_join[k], b->lane[k]); e != hipSuccess)
				return b->hip_fail(e, "hipEventRecord");
			if (hipError_t e = hipStreamWaitEvent(b->stream, b->ev_join[k], 0); e != hipSuccess)
				return b->hip_fail(e, "hipStreamWaitEvent");
		}
	}
	// the tick ends every step, also with nothing to add (tick 0, inside a
	// captured run of steps): replayed, a graph whose steps end in a kernel on
	// the batch stream starts the next step's lanes together (C3, 10-step
	// graphs: 1.624 against 1.650 ms/step without it, where two of the four
	// lanes' x chains started 150-460 us late)
	if (hipError_t e = launch_step_tick(dd, (int)b->ctx.size(), tick, b->stream); e != hipSuccess)
		return b->hip_fail(e, "step tick launch");
	return BFVM_OK;
}

// the batch stream waits for each context's queued work; afterwards each
// context stream waits for the batch
int batch_join_in(bfvm_batch* b)
{
	for (bfvm_ctx* c : b->ctx)
	{
		if (hipError_t e = hipEventRecord(b->ev_in, c->stream); e != hipSuccess)
			return b->hip_fail(e, "hipEventRecord");
		if (hipError_t e = hipStreamWaitEvent(b->stream, b->ev_in, 0); e != hipSuccess)
			return b->hip_fail(e, "hipStreamWaitEvent");
	}
	return BFVM_OK;
}

int batch_join_out(bfvm_batch* b)
{
	if (hipError_t e = hipEventRecord(b->ev_out, b->stream); e != hipSuccess)
		return b->hip_fail(e, "hipEventRecord");
	for (bfvm_ctx* c : b->ctx)
		if (hipError_t e = hipStreamWaitEvent(c->stream, b->ev_out, 0); e != hipSuccess)
			return b->hip_fail(e, "hipStreamWaitEvent");
	return BFVM_OK;
}

int batch_run(bfvm_batch* b, int64_t iterations, int64_t graph_steps)
{
	int rc = batch_prepare(b);
	if (rc)
		return rc;
	if ((rc = batch_join_in(b)))
		return rc;
	int64_t done = 0;
	if (graph_steps > 0 && iterations >= graph_steps)
	{
		if (!b->graph_valid || b->graph_steps != graph_steps)
		{
			b->drop_graph();
			// one domain table per captured step, the step counts advanced once
			// at the end of the run
			const size_t nd = b->hdoms.size();
			std::vector<DomainDev> slots((size_t)graph_steps * nd);
			for (int64_t k = 0; k < graph_steps; k++)
				for (size_t i = 0; i < nd; i++)
				{
					slots[k * nd + i] = b->hdoms[i];
					slots[k * nd + i].step_offset = (uint32_t)k;
				}
			if (hipError_t e = b->dslots.upload(slots.data(), slots.size(), b->stream); e != hipSuccess)
				return b->hip_fail(e, "domain table upload");
			if (hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal); e != hipSuccess)
				return b->hip_fail(e, "hipStreamBeginCapture");
			for (int64_t k = 0; k < graph_steps && rc == BFVM_OK; k++)
				rc = batch_enqueue_step(b, b->dslots.p + k * nd, k == graph_steps - 1 ? (uint32_t)graph_steps : 0);
			hipGraph_t g = nullptr;
			hipError_t e = hipStreamEndCapture(b->stream, &g);
			if (!rc && e == hipSuccess)
			{
				b->graph = g;
				e = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
			}
			else if (g)
				(void)hipGraphDestroy(g);
			if (rc || e != hipSuccess)
			{
				// nothing captured ran; a capture can also be spoiled by work
				// of another thread, so the steps go out one by one instead
				// (an error of their own is reported from there)
				b->drop_graph();
				(void)hipGetLastError();
				rc = BFVM_OK;
				graph_steps = 0;
			}
			else
			{
				b->graph_steps = graph_steps;
				b->graph_valid = true;
			}
		}
		for (; graph_steps > 0 && done + graph_steps <= iterations; done += graph_steps)
			if (hipError_t e = hipGraphLaunch(b->gexec, b->stream); e != hipSuccess)
				return b->hip_fail(e, "hipGraphLaunch");
	}
	for (; done < iterations; done++)
		if ((rc = batch_enqueue_step(b)))
			return rc;
	return batch_join_out(b);
}

} // namespace

namespace {
// Device memory of bfvm_total_mass_static: the mass reduction's per-block
// partial sums and its result, in a file-scope array instead of buffers the
// context allocates.  Every GPU holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for that
// device (the array's name is not a device address on the host).
constexpr int64_t kStaticScratchDoubles = 16384; // 128 KiB
__device__ __align__(256) double scratch[kStaticScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr; // per device id; null until resolved

// The scratch array's address on device `dev`, resolved once per device with
// that device current (the caller's current device is restored).
hipError_t static_scratch(int dev, double** out)
{
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (dev < 0)
		return hipErrorInvalidDevice;
	if ((size_t)dev >= g_scratch_ptr.size())
		g_scratch_ptr.resize((size_t)dev + 1, nullptr);
	if (!g_scratch_ptr[(size_t)dev])
	{
		int cur = 0;
		hipError_t e = hipGetDevice(&cur);
		if (e == hipSuccess && cur != dev)
			e = hipSetDevice(dev);
		double* p = nullptr;
		if (e == hipSuccess)
			e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
		if (cur != dev)
		{
			const hipError_t r = hipSetDevice(cur);
			if (e == hipSuccess)
				e = r;
		}
		if (e != hipSuccess)
			return e;
		if (!p)
			return hipErrorInvalidSymbol;
		g_scratch_ptr[(size_t)dev] = p;
	}
	*out = g_scratch_ptr[(size_t)dev];
	return hipSuccess;
}
} // namespace

extern "C" {

const char* bfvm_last_error(const bfvm_ctx* ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

int bfvm_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

int bfvm_create(const bfvm_domain_desc* desc, bfvm_ctx** out)
{
	if (!out)
		return g_error = "null output pointer", BFVM_ERR_INVALID;
	*out = nullptr;
	std::string why;
	if (!valid_desc(desc, why))
		return g_error = why, BFVM_ERR_INVALID;

	auto* c = new bfvm_ctx;
	c->dims = desc->dims;
	c->S = desc->substrates;
	for (int a = 0; a < 3; a++)
	{
		c->n[a] = desc->grid[a];
		c->voxel[a] = a < desc->dims ? desc->voxel[a] : std::max<int64_t>(desc->voxel[a], 1);
		c->bbox_min[a] = desc->bbox_min[a];
	}
	c->V = c->n[0] * c->n[1] * c->n[2];
	c->L = (int)(c->n[0] * c->S);
	c->dt = desc->dt;
	c->internalized = desc->compute_internalized != 0;
	c->fp64 = desc->precision == BFVM_SWEEP_FP64;
	c->reduce_mode = desc->reduce_mode;
	c->quirks = desc->openmp_quirks != 0;
	c->D.assign(desc->diffusion_coefficients, desc->diffusion_coefficients + c->S);
	c->lam.assign(desc->decay_rates, desc->decay_rates + c->S);
	c->init.assign(desc->initial_conditions, desc->initial_conditions + c->S);
	c->tuning.persistent_wgs_per_cu = 0;
	c->tuning.force_lds_sweeps = 0;
	c->tuning.x_sweep = BFVM_X_AUTO;
	c->tuning.pipeline = 0;
	c->tuning.round_nearest = 0;

	auto bail = [&](int rc) {
		g_error = c->err;
		delete c;
		return rc;
	};

	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	c->device = dev;
	if (e == hipSuccess)
		e = sweep_kernels_ready(dev);
	if (e == hipSuccess)
		e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev);
	if (e == hipSuccess)
		e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
	if (e != hipSuccess)
		return bail(c->hip_fail(e, "device setup"));

	if ((e = c->field.alloc((size_t)c->V * c->S)) != hipSuccess)
		return bail(c->hip_fail(e, "field allocation"));
	{
		const uint32_t zero = 0;
		if ((e = c->step_count.upload(&zero, 1, c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "step counter"));
	}
	{
		std::vector<float> init(c->S);
		for (int s = 0; s < c->S; s++)
			init[s] = (float)c->init[s];
		dvec<float> dinit;
		if ((e = dinit.upload(init.data(), init.size(), c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "initial conditions"));
		if ((e = launch_fill_field(c->field.p, dinit.p, c->S, c->V, c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "field fill"));
		if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "field fill"));
	}
	for (int a = 0; a < c->dims; a++)
	{
		std::vector<double> b, ee, cc;
		precompute_factors((int)c->n[a], c->S, c->voxel[a], c->dims, c->dt, c->D.data(), c->lam.data(), c->quirks, b,
						   ee, cc);
		if ((e = c->fb[a].upload(b.data(), b.size(), c->stream)) != hipSuccess ||
			(e = c->fe[a].upload(ee.data(), ee.size(), c->stream)) != hipSuccess ||
			(e = c->fc[a].upload(cc.data(), cc.size(), c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "factor upload"));
	}
	if (c->quirks && c->dims >= 2)
	{
		std::vector<int32_t> m(c->L);
		for (int col = 0; col < c->L; col++)
			m[col] = openmp_coef_sub(col, c->S, c->L);
		if ((e = c->coef_map.upload(m.data(), m.size(), c->stream)) != hipSuccess)
			return bail(c->hip_fail(e, "coefficient map"));
	}
	if ((e = c->flags.alloc(kCellFlags)) != hipSuccess ||
		(e = hipMemsetAsync(c->flags.p, 0, kCellFlags * sizeof(int32_t), c->stream)) != hipSuccess ||
		(e = hipStreamSynchronize(c->stream)) != hipSuccess)
		return bail(c->hip_fail(e, "flags"));
	int rc = build_tiles(c);
	if (rc)
		return bail(rc);
	rc = sync_domain(c);
	if (rc)
		return bail(rc);
	*out = c;
	return BFVM_OK;
}

void bfvm_destroy(bfvm_ctx* ctx)
{
	if (!ctx)
		return;
	// once both streams are idle no queued work reads the buffers (batch steps
	// are ordered before the context's stream), so they go to the block cache
	const bool idle = (!ctx->stream || hipStreamSynchronize(ctx->stream) == hipSuccess) &&
					  (!ctx->copy_stream || hipStreamSynchronize(ctx->copy_stream) == hipSuccess);
	RecycleScope r(idle);
	delete ctx;
}

int bfvm_set_dirichlet_faces(bfvm_ctx* c, const double* values, const uint8_t* conds)
{
	if (!c)
		return BFVM_ERR_INVALID;
	const int S = c->S;
	std::vector<float> v(6 * S, 0.0f);
	std::vector<uint8_t> k(6 * S, 0);
	int mask = 0;
	if (values && conds)
		for (int f = 0; f < 6; f++)
		{
			if (f / 2 >= c->dims)
				continue; // faces of axes beyond dims do not exist
			for (int s = 0; s < S; s++)
			{
				if (conds[f * S + s])
				{
					if (!fp_ok(values[f * S + s]))
						return c->fail(BFVM_ERR_INVALID, "non-finite Dirichlet value");
					k[f * S + s] = 1;
					mask |= 1 << f;
				}
				v[f * S + s] = (float)values[f * S + s];
			}
		}
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	TRY_HIP(c, c->face_val.upload(v.data(), v.size(), c->stream));
	TRY_HIP(c, c->face_cond.upload(k.data(), k.size(), c->stream));
	const bool rebuild = c->dims == 1 && ((c->face_mask != 0) != (mask != 0));
	c->face_mask = mask;
	if (rebuild)
	{
		int rc = build_tiles(c);
		if (rc)
			return rc;
	}
	return sync_domain(c);
}

int bfvm_set_dirichlet_interior(bfvm_ctx* c, int64_t count, const int64_t* voxels, const double* values,
								const uint8_t* conds)
{
	if (!c || count < 0 || (count > 0 && (!voxels || !values || !conds)))
		return c ? c->fail(BFVM_ERR_INVALID, "bad interior Dirichlet arrays") : BFVM_ERR_INVALID;
	const int S = c->S;
	std::vector<int64_t> lin;
	std::vector<float> val;
	std::vector<uint8_t> cond;
	std::vector<int64_t> slot_of; // voxel -> entry, to merge repeats
	for (int64_t k = 0; k < count; k++)
	{
		int64_t idx[3] = { 0, 0, 0 };
		for (int a = 0; a < c->dims; a++)
		{
			idx[a] = voxels[k * c->dims + a];
			if (idx[a] < 0 || idx[a] >= c->n[a])
				return c->fail(BFVM_ERR_INVALID, "interior Dirichlet voxel outside the mesh");
		}
		const int64_t v = idx[0] + c->n[0] * (idx[1] + c->n[1] * idx[2]);
		int64_t e = -1;
		for (size_t q = 0; q < lin.size(); q++)
			if (lin[q] == v)
				e = (int64_t)q;
		if (e < 0)
		{
			e = (int64_t)lin.size();
			lin.push_back(v);
			val.resize(val.size() + S, 0.0f);
			cond.resize(cond.size() + S, 0);
		}
		for (int s = 0; s < S; s++)
			if (conds[k * S + s])
			{
				if (!fp_ok(values[k * S + s]))
					return c->fail(BFVM_ERR_INVALID, "non-finite Dirichlet value");
				val[e * S + s] = (float)values[k * S + s];
				cond[e * S + s] = 1;
			}
	}
	(void)slot_of;
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	c->dir_count = (int64_t)lin.size();
	TRY_HIP(c, c->dir_vox.upload(lin.data(), lin.size(), c->stream));
	TRY_HIP(c, c->dir_val.upload(val.data(), val.size(), c->stream));
	TRY_HIP(c, c->dir_cond.upload(cond.data(), cond.size(), c->stream));
	auto r = refs_for(c->dir_count);
	TRY_HIP(c, c->dir_refs.upload(r.data(), r.size(), c->stream));
	return sync_domain(c);
}

int bfvm_set_bulk(bfvm_ctx* c, int64_t count, const int64_t* voxels, const double* supply, const double* uptake,
				  const double* target)
{
	if (!c || count < 0 || (count > 0 && (!voxels || !supply || !uptake || !target)))
		return c ? c->fail(BFVM_ERR_INVALID, "bad bulk arrays") : BFVM_ERR_INVALID;
	const int S = c->S;
	std::vector<int64_t> vox(voxels, voxels + count);
	std::vector<double> a((size_t)count * S), b((size_t)count * S);
	{
		std::vector<int64_t> sorted = vox;
		std::sort(sorted.begin(), sorted.end());
		if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
			return c->fail(BFVM_ERR_INVALID, "bulk voxel listed twice");
		if (count && (sorted.front() < 0 || sorted.back() >= c->V))
			return c->fail(BFVM_ERR_INVALID, "bulk voxel outside the mesh");
	}
	for (int64_t k = 0; k < count * S; k++)
	{
		if (!fp_ok(supply[k]) || !fp_ok(uptake[k]) || !fp_ok(target[k]))
			return c->fail(BFVM_ERR_INVALID, "non-finite bulk rate");
		// bulk_solver.cpp:29: D = (D + time_step*S*T) / (1 + time_step*(U+S)),
		// applied as (D + a) * b with b the reciprocal
		a[k] = c->dt * supply[k] * target[k];
		b[k] = 1 / (1 + c->dt * (uptake[k] + supply[k]));
	}
	// the same rates at every listed voxel, the listed voxels of each line of
	// the last axis one run, and at least an eighth of the voxels listed: then
	// the last sweep applies the bulk.  It forms the bulk value at every
	// position (branch-free), which costs the z sweep of 200^3 x 4 about 7 us:
	// less than a sparse launch over a dense list (250^3 x 4, every voxel:
	// +35 us against 480 us), more than one over a small ball (200^3 x 4,
	// 0.8% of the voxels: 5.7 us)
	// (the layout alone, whatever the rates, also decides the last sweep's
	// kernel: one lane per line where such a bulk could ride on it)
	bool layout = count > 0 && c->dims >= 2 && count * 8 >= c->V;
	std::vector<int32_t> run;
	if (layout)
	{
		const int la = c->dims - 1;
		const int64_t plane = c->n[0] * (la == 2 ? c->n[1] : 1); // voxels per position of the last axis
		const int64_t cols = c->V / c->n[la];
		std::vector<int32_t> lo(cols, INT32_MAX), hi(cols, -1), cnt(cols, 0);
		for (int64_t v : vox)
		{
			const int64_t pos = la == 2 ? v / plane : (v / c->n[0]) % c->n[1];
			const int64_t col = la == 2 ? v % plane : (v / (c->n[0] * c->n[1])) * c->n[0] + v % c->n[0];
			lo[col] = std::min(lo[col], (int32_t)pos);
			hi[col] = std::max(hi[col], (int32_t)pos);
			cnt[col]++;
		}
		run.assign(2 * cols, 0);
		for (int64_t q = 0; layout && q < cols; q++)
			if (cnt[q])
			{
				layout = hi[q] - lo[q] + 1 == cnt[q];
				run[2 * q] = lo[q];
				run[2 * q + 1] = hi[q] + 1;
			}
	}
	bool runs = layout;
	for (int64_t k = S; runs && k < count * S; k++)
		runs = a[k] == a[k % S] && b[k] == b[k % S];
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	const bool had_layout = c->bulk_layout;
	c->bulk_count = count;
	c->bulk_runs = runs;
	c->bulk_layout = layout;
	if (runs)
	{
		std::vector<double> u(2 * (size_t)S);
		for (int s = 0; s < S; s++)
		{
			u[s] = a[s];
			u[S + s] = b[s];
		}
		TRY_HIP(c, c->bulk_u.upload(u.data(), u.size(), c->stream));
		TRY_HIP(c, c->bulk_run.upload(run.data(), run.size(), c->stream));
	}
	TRY_HIP(c, c->bulk_vox.upload(vox.data(), vox.size(), c->stream));
	TRY_HIP(c, c->bulk_a.upload(a.data(), a.size(), c->stream));
	TRY_HIP(c, c->bulk_r.upload(b.data(), b.size(), c->stream));
	auto t = sparse_tiles(0, count, c->S);
	c->n_bulk_tiles = (int)t.size();
	TRY_HIP(c, c->bulk_tiles.upload(t.data(), t.size(), c->stream));
	if (count > 0)
	{
		TRY_HIP(c, c->bulk_map.alloc((size_t)c->V));
		TRY_HIP(c, c->bulk_cell.alloc((size_t)count));
		TRY_HIP(c, launch_bulk_map(c->bulk_vox.p, count, c->bulk_map.p, c->V, c->stream));
	}
	// the last sweep's kernel depends on the bulk list's layout
	int rc = had_layout != layout ? build_tiles(c) : BFVM_OK;
	if (!rc)
		rc = sync_domain(c);
	return rc ? rc : rebuild_overlap(c);
}

int bfvm_set_agents(bfvm_ctx* c, int64_t n, const double* positions, const double* secretion,
					const double* saturation, const double* uptake, const double* net_export,
					const double* internalized, const double* fraction_released, const double* volumes,
					const uint8_t* is_active)
{
	if (!c || n < 0 || n > 2147483647LL)
		return c ? c->fail(BFVM_ERR_INVALID, "bad agent count") : BFVM_ERR_INVALID;
	if (n > 0 && (!positions || !secretion || !saturation || !uptake || !net_export || !fraction_released ||
				  !volumes || !is_active))
		return c->fail(BFVM_ERR_INVALID, "null agent array");
	FLUSH_RELEASES(c); // pending releases belong to the old agent arrays
	const size_t nS = (size_t)n * c->S;
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	c->n_agents = n;
	c->h_positions.assign(positions, positions + (size_t)n * c->dims);
	TRY_HIP(c, c->positions.upload(positions, (size_t)n * c->dims, c->stream));
	TRY_HIP(c, c->secretion.upload(secretion, nS, c->stream));
	TRY_HIP(c, c->saturation.upload(saturation, nS, c->stream));
	TRY_HIP(c, c->uptake.upload(uptake, nS, c->stream));
	TRY_HIP(c, c->net_export.upload(net_export, nS, c->stream));
	if (internalized)
		TRY_HIP(c, c->internalized_d.upload(internalized, nS, c->stream));
	else
	{
		TRY_HIP(c, c->internalized_d.alloc(nS));
		if (nS)
		{
			TRY_HIP(c, hipMemsetAsync(c->internalized_d.p, 0, nS * sizeof(double), c->stream));
			TRY_HIP(c, hipStreamSynchronize(c->stream));
		}
	}
	TRY_HIP(c, c->fraction_released.upload(fraction_released, nS, c->stream));
	TRY_HIP(c, c->volumes.upload(volumes, (size_t)n, c->stream));
	TRY_HIP(c, c->is_active.upload(is_active, (size_t)n, c->stream));
	TRY_HIP(c, c->cell_voxel.alloc(n));
	TRY_HIP(c, c->owner_of.alloc(n));
	TRY_HIP(c, c->owners.alloc(n));
	TRY_HIP(c, c->num.alloc(nS));
	TRY_HIP(c, c->den.alloc(nS));
	TRY_HIP(c, c->fac.alloc(nS));
	TRY_HIP(c, c->rnum.alloc(nS));
	TRY_HIP(c, c->rden.alloc(nS));
	TRY_HIP(c, c->rfac.alloc(nS));
	TRY_HIP(c, c->dnew.alloc(nS));
	TRY_HIP(c, c->pnum.alloc(nS));
	TRY_HIP(c, c->pden.alloc(nS));
	TRY_HIP(c, c->pfac.alloc(nS));
	TRY_HIP(c, c->pvox.alloc(n));
	TRY_HIP(c, c->cell_in_bulk.alloc(n));
	TRY_HIP(c, c->slot_of.alloc(n));
	TRY_HIP(c, c->slot_alone.alloc(n));
	TRY_HIP(c, c->joined.alloc(n));
	TRY_HIP(c, c->shared_list.alloc(n));
	// the recompute's sort buffers (the owner list has a slot per agent)
	TRY_HIP(c, c->keys.alloc(n));
	TRY_HIP(c, c->vals.alloc(n));
	TRY_HIP(c, c->tkeys.alloc(n));
	TRY_HIP(c, c->tvals.alloc(n));
	TRY_HIP(c, c->heads.alloc(n));
	TRY_HIP(c, c->head_scan.alloc(n));
	TRY_HIP(c, c->hist.alloc(radix_hist_entries(n)));
	TRY_HIP(c, c->scan_tmp.alloc(std::max(scan_tmp_entries(n), scan_tmp_entries((int64_t)radix_hist_entries(n)))));
	if (c->reduce_mode == BFVM_REDUCE_ATOMIC && n > 0)
		TRY_HIP(c, c->ballot.alloc((size_t)c->V));
	// per-agent tiles (internalized update) and tiles over the owner list's
	// slots (cell apply; the kernels stop at the owner count)
	auto t = sparse_tiles(0, n, c->S);
	c->n_int_tiles = (int)t.size();
	TRY_HIP(c, c->int_tiles.upload(t.data(), t.size(), c->stream));
	c->n_cell_tiles = (int)t.size();
	TRY_HIP(c, c->cell_tiles.upload(t.data(), t.size(), c->stream));
	c->recompute_pending = true;
	c->grid_check_pending = false;
	c->h_positions_stale = false;
	c->n_owners = 0;
	c->overlap = false;
	return sync_domain(c);
}

int bfvm_recompute_positional_data(bfvm_ctx* c)
{
	if (!c)
		return BFVM_ERR_INVALID;
	// the recompute runs on the stream before the next step; the step launches
	// (and a captured graph) stay as they are
	c->recompute_pending = true;
	return BFVM_OK;
}

int bfvm_agent_positions_device_ptr(bfvm_ctx* c, double** ptr)
{
	if (!c || !ptr)
		return BFVM_ERR_INVALID;
	if (c->n_agents == 0)
		return c->fail(BFVM_ERR_STATE, "no agents set");
	FLUSH_RELEASES(c); // releases resolve voxels from the positions they were asked at
	*ptr = c->positions.p;
	c->h_positions_stale = true;
	return BFVM_OK;
}

int bfvm_step(bfvm_ctx* c, int64_t iterations)
{
	if (!c || iterations < 0)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	for (int64_t it = 0; it < iterations; it++)
	{
		if (c->recompute_pending)
		{
			int rc = recompute(c);
			if (rc)
				return rc;
		}
		int rc = enqueue_step(c, nullptr);
		if (rc)
			return rc;
	}
	return BFVM_OK;
}

int bfvm_step_graph(bfvm_ctx* c, int64_t iterations, int64_t graph_steps)
{
	if (!c || iterations < 0 || graph_steps < 1)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	if (iterations == 0)
		return BFVM_OK;
	if (c->recompute_pending)
	{
		int rc = recompute(c);
		if (rc)
			return rc;
	}
	if (!c->gexec || c->graph_version != c->version || c->graph_steps != graph_steps)
	{
		c->drop_graph();
		// one domain table per captured step, the step count advanced once at
		// the end of the run: no step-count update between the steps
		std::vector<DomainDev> slots((size_t)graph_steps, c->hdom);
		for (int64_t k = 0; k < graph_steps; k++)
			slots[k].step_offset = (uint32_t)k;
		TRY_HIP(c, c->dslots.upload(slots.data(), slots.size(), c->stream));
		TRY_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
		int rc = BFVM_OK;
		for (int64_t k = 0; k < graph_steps && rc == BFVM_OK; k++)
			rc = enqueue_step(c, nullptr, c->dslots.p + k, k == graph_steps - 1 ? (uint32_t)graph_steps : 0);
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(c->stream, &g);
		if (!rc && e == hipSuccess)
		{
			c->graph = g;
			e = hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0);
		}
		else if (g)
			(void)hipGraphDestroy(g);
		if (rc || e != hipSuccess)
		{
			// nothing captured ran; a capture can also be spoiled by work of
			// another thread, so the steps go out one by one instead (an
			// error of their own is reported from there)
			c->drop_graph();
			(void)hipGetLastError();
			for (int64_t k = 0; k < iterations; k++)
				if ((rc = enqueue_step(c, nullptr)))
					return rc;
			return BFVM_OK;
		}
		c->graph_steps = graph_steps;
		c->graph_version = c->version;
	}
	const int64_t full = iterations / graph_steps;
	for (int64_t k = 0; k < full; k++)
		TRY_HIP(c, hipGraphLaunch(c->gexec, c->stream));
	for (int64_t k = 0; k < iterations - full * graph_steps; k++)
	{
		int rc = enqueue_step(c, nullptr);
		if (rc)
			return rc;
	}
	return BFVM_OK;
}

int bfvm_sync(bfvm_ctx* c)
{
	if (!c)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return check_recompute(c);
}

int bfvm_download_field_f32(bfvm_ctx* c, float* host)
{
	if (!c || !host)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	TRY_HIP(c, hipMemcpyAsync(host, c->field.p, c->field.n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return check_recompute(c);
}

int bfvm_snapshot_begin(bfvm_ctx* c, float* host, int64_t* ticket)
{
	if (!c || !host || !ticket)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	if (!c->copy_stream)
	{
		TRY_HIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
		TRY_HIP(c, hipEventCreateWithFlags(&c->snap_taken, hipEventDisableTiming));
		for (hipEvent_t& e : c->snap_done)
			TRY_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
	}
	if (!c->snap.p)
		TRY_HIP(c, c->snap.alloc(c->field.n));
	const int64_t t = c->snap_next.load();
	const size_t bytes = c->field.n * sizeof(float);
	// the snapshot buffer is reused: the previous host copy must have read it
	if (t > 0)
		TRY_HIP(c, hipStreamWaitEvent(c->stream, c->snap_done[(t - 1) % bfvm_ctx::kSnapRing], 0));
	// a kernel copy at HBM speed (a DMA-engine copy would hold the stream far longer)
	TRY_HIP(c, launch_copy_field(c->field.p, c->snap.p, (int64_t)c->field.n, c->num_cu, c->stream));
	TRY_HIP(c, hipEventRecord(c->snap_taken, c->stream));
	TRY_HIP(c, hipStreamWaitEvent(c->copy_stream, c->snap_taken, 0));
	TRY_HIP(c, hipMemcpyAsync(host, c->snap.p, bytes, hipMemcpyDeviceToHost, c->copy_stream));
	TRY_HIP(c, hipEventRecord(c->snap_done[t % bfvm_ctx::kSnapRing], c->copy_stream));
	c->snap_next.store(t + 1);
	*ticket = t;
	return BFVM_OK;
}

int bfvm_snapshot_wait(bfvm_ctx* c, int64_t ticket)
{
	if (!c || ticket < 0 || ticket >= c->snap_next.load())
		return BFVM_ERR_INVALID;
	// a later snapshot may have re-recorded the slot: waiting for it is later still
	if (hipError_t e = hipEventSynchronize(c->snap_done[ticket % bfvm_ctx::kSnapRing]); e != hipSuccess)
	{
		g_error = std::string("hipEventSynchronize: ") + hipGetErrorString(e);
		return BFVM_ERR_HIP;
	}
	return BFVM_OK;
}

int bfvm_host_alloc(size_t bytes, void** out)
{
	if (!out || bytes == 0)
		return BFVM_ERR_INVALID;
	*out = nullptr;
	if (hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault); e != hipSuccess)
	{
		g_error = std::string("hipHostMalloc: ") + hipGetErrorString(e);
		return e == hipErrorOutOfMemory ? BFVM_ERR_NO_MEMORY : BFVM_ERR_HIP;
	}
	return BFVM_OK;
}

int bfvm_host_free(void* p)
{
	if (!p)
		return BFVM_OK;
	if (hipError_t e = hipHostFree(p); e != hipSuccess)
	{
		g_error = std::string("hipHostFree: ") + hipGetErrorString(e);
		return BFVM_ERR_HIP;
	}
	return BFVM_OK;
}

int bfvm_upload_field_f32(bfvm_ctx* c, const float* host)
{
	if (!c || !host)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	TRY_HIP(c, hipMemcpyAsync(c->field.p, host, c->field.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return BFVM_OK;
}

int bfvm_download_field_f64(bfvm_ctx* c, double* host)
{
	if (!c || !host)
		return BFVM_ERR_INVALID;
	std::vector<float> tmp(c->field.n);
	int rc = bfvm_download_field_f32(c, tmp.data());
	if (rc)
		return rc;
	for (size_t k = 0; k < tmp.size(); k++)
		host[k] = (double)tmp[k];
	return BFVM_OK;
}

int bfvm_upload_field_f64(bfvm_ctx* c, const double* host)
{
	if (!c || !host)
		return BFVM_ERR_INVALID;
	std::vector<float> tmp(c->field.n);
	for (size_t k = 0; k < tmp.size(); k++)
		tmp[k] = (float)host[k];
	return bfvm_upload_field_f32(c, tmp.data());
}

int bfvm_download_internalized(bfvm_ctx* c, double* host)
{
	if (!c || (!host && c->n_agents))
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	if (c->n_agents == 0)
		return BFVM_OK;
	TRY_HIP(c, hipMemcpyAsync(host, c->internalized_d.p, (size_t)c->n_agents * c->S * sizeof(double),
							  hipMemcpyDeviceToHost, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return check_recompute(c);
}

int bfvm_release_internalized(bfvm_ctx* c, int64_t agent)
{
	if (!c)
		return BFVM_ERR_INVALID;
	if (!c->internalized)
		return BFVM_OK; // cell_solver.cpp:353-354
	if (agent < 0 || agent >= c->n_agents)
		return c->fail(BFVM_ERR_INVALID, "agent index out of range");
	if (c->h_positions_stale)
	{
		// positions were written on the device (bfvm_agent_positions_device_ptr)
		TRY_HIP(c, hipStreamSynchronize(c->stream));
		TRY_HIP(c, hipMemcpyAsync(c->h_positions.data(), c->positions.p, c->h_positions.size() * sizeof(double),
								  hipMemcpyDeviceToHost, c->stream));
		TRY_HIP(c, hipStreamSynchronize(c->stream));
		c->h_positions_stale = false;
	}
	int64_t v;
	if (!voxel_of(c, c->h_positions.data() + agent * c->dims, &v))
		return c->fail(BFVM_ERR_OUT_OF_GRID, "agent outside the mesh");
	// applied in FP64 at the next call that reads or writes the field
	c->rel_agent.push_back(agent);
	c->rel_voxel.push_back(v);
	return BFVM_OK;
}

int bfvm_total_mass(bfvm_ctx* c, double* out)
{
	if (!c || !out)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	TRY_HIP(c, c->mass_partial.alloc((size_t)mass_blocks(c->V) * c->S));
	TRY_HIP(c, c->mass_out.alloc(c->S));
	TRY_HIP(c, launch_mass(c->field.p, c->S, c->V, c->mass_partial.p, c->mass_out.p,
						   (double)(c->voxel[0] * c->voxel[1] * c->voxel[2]), c->stream));
	TRY_HIP(c, hipMemcpyAsync(out, c->mass_out.p, c->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return check_recompute(c);
}

int bfvm_total_mass_static(bfvm_ctx* c, double* out)
{
	if (!c || !out)
		return c ? c->fail(BFVM_ERR_INVALID, "null output") : BFVM_ERR_INVALID;
	if (!c->rel_agent.empty())
		return c->fail(BFVM_ERR_STATE, "releases are pending: applying them uploads their list (device memory); "
									   "call bfvm_sync or bfvm_total_mass first");
	const int64_t nb = mass_blocks(c->V);
	const int64_t need = nb * c->S + c->S;
	if (need > kStaticScratchDoubles)
		return c->fail(BFVM_ERR_UNSUPPORTED, "the mass reduction needs " + std::to_string(need) +
												 " doubles of scratch, more than the " +
												 std::to_string(kStaticScratchDoubles) + " of the static array");
	double* base = nullptr;
	TRY_HIP(c, static_scratch(c->device, &base));
	// every partial and the result are written before they are read: nothing to clear
	TRY_HIP(c, launch_mass(c->field.p, c->S, c->V, base, base + nb * c->S,
						   (double)(c->voxel[0] * c->voxel[1] * c->voxel[2]), c->stream));
	TRY_HIP(c, hipMemcpyAsync(out, base + nb * c->S, c->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	return check_recompute(c);
}

int bfvm_cell_sums(bfvm_ctx* c, int64_t* n_owners, int64_t* owners, double* num, double* den, double* fac,
				   int32_t* is_conflict)
{
	if (!c || !n_owners)
		return BFVM_ERR_INVALID;
	if (c->recompute_pending)
	{
		int rc = recompute(c);
		if (rc)
			return rc;
	}
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	c->grid_check_pending = c->n_agents > 0; // read the count and flags of the last recompute
	if (int rc = check_recompute(c))
		return rc;
	*n_owners = c->n_owners;
	if (is_conflict)
		*is_conflict = c->conflict ? 1 : 0;
	if (!owners)
		return BFVM_OK;
	// the packed apply list: owners in voxel order and their full sums
	const int S = c->S;
	const size_t m = (size_t)c->n_owners;
	std::vector<int32_t> o(m);
	std::vector<double> pn(m * S), pd(m * S), pf(m * S);
	if (m)
	{
		TRY_HIP(c, hipMemcpyAsync(o.data(), c->owners.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
		TRY_HIP(c, hipMemcpyAsync(pn.data(), c->pnum.p, pn.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
		TRY_HIP(c, hipMemcpyAsync(pd.data(), c->pden.p, pd.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
		TRY_HIP(c, hipMemcpyAsync(pf.data(), c->pfac.p, pf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
		TRY_HIP(c, hipStreamSynchronize(c->stream));
	}
	std::vector<size_t> order(m);
	for (size_t k = 0; k < m; k++)
		order[k] = k;
	std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return o[a] < o[b]; });
	for (size_t k = 0; k < m; k++)
	{
		const size_t j = order[k];
		owners[k] = o[j];
		for (int s = 0; s < S; s++)
		{
			if (num)
				num[k * S + s] = pn[j * S + s];
			if (den)
				den[k * S + s] = pd[j * S + s];
			if (fac)
				fac[k * S + s] = pf[j * S + s];
		}
	}
	return BFVM_OK;
}

int bfvm_field_device_ptr(bfvm_ctx* c, void** ptr)
{
	if (!c || !ptr)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	*ptr = c->field.p;
	return BFVM_OK;
}

int bfvm_stream(bfvm_ctx* c, void** stream)
{
	if (!c || !stream)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	*stream = (void*)c->stream;
	return BFVM_OK;
}

int bfvm_time_sweep(bfvm_ctx* c, int32_t axis, int64_t launches, double* out_us)
{
	if (!c || !out_us || launches < 1 || axis < 0 || axis >= c->dims)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	if (c->recompute_pending)
	{
		int rc = recompute(c);
		if (rc)
			return rc;
	}
	hipEvent_t e0 = nullptr, e1 = nullptr;
	TRY_HIP(c, hipEventCreate(&e0));
	if (hipError_t e = hipEventCreate(&e1); e != hipSuccess)
	{
		(void)hipEventDestroy(e0);
		return c->hip_fail(e, "hipEventCreate");
	}
	hipError_t e = hipEventRecord(e0, c->stream);
	for (int64_t k = 0; k < launches && e == hipSuccess; k++)
		e = launch_axis_sweep(c, axis, c->ddom.p);
	if (e == hipSuccess)
		e = hipEventRecord(e1, c->stream);
	if (e == hipSuccess)
		e = hipEventSynchronize(e1);
	float ms = 0;
	if (e == hipSuccess)
		e = hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (e != hipSuccess)
		return c->hip_fail(e, "sweep timing");
	*out_us = (double)ms * 1000.0 / (double)launches;
	return BFVM_OK;
}

int bfvm_time_stages(bfvm_ctx* c, int64_t steps, double* out_us)
{
	if (!c || !out_us || steps < 1)
		return BFVM_ERR_INVALID;
	FLUSH_RELEASES(c);
	if (c->recompute_pending)
	{
		int rc = recompute(c);
		if (rc)
			return rc;
	}
	// the steps are queued back to back, as a run queues them, with events
	// between the launches; read once the last step has run
	constexpr int64_t kMaxTimed = 64;
	steps = std::min(steps, kMaxTimed);
	std::vector<hipEvent_t> ev((size_t)steps * 7, nullptr);
	int rc = BFVM_OK;
	for (auto& e : ev)
		if (hipError_t err = hipEventCreate(&e); err != hipSuccess)
		{
			rc = c->hip_fail(err, "hipEventCreate");
			break;
		}
	for (int64_t k = 0; k < steps && rc == BFVM_OK; k++)
		rc = enqueue_step(c, ev.data() + k * 7);
	double acc[6] = { 0, 0, 0, 0, 0, 0 };
	if (rc == BFVM_OK)
	{
		if (hipError_t e = hipEventSynchronize(ev.back()); e != hipSuccess)
			rc = c->hip_fail(e, "hipEventSynchronize");
		for (int64_t k = 0; k < steps && rc == BFVM_OK; k++)
			for (int s = 0; s < 6; s++)
			{
				float ms = 0;
				(void)hipEventElapsedTime(&ms, ev[k * 7 + s], ev[k * 7 + s + 1]);
				acc[s] += ms * 1000.0;
			}
	}
	for (auto& e : ev)
		if (e)
			(void)hipEventDestroy(e);
	if (rc)
		return rc;
	for (int s = 0; s < 6; s++)
		out_us[s] = acc[s] / (double)steps;
	return BFVM_OK;
}

int bfvm_sweep_plan(const bfvm_ctx* c, int32_t* kernels, int32_t* split)
{
	if (!c || !kernels)
		return BFVM_ERR_INVALID;
	for (int axis = 0; axis < 3; axis++)
		kernels[axis] = axis < c->dims ? c->variant[axis] : 0;
	if (split)
	{
		split[0] = c->split_x;
		split[1] = c->split_y;
		split[2] = c->ntiles[0];
		split[3] = c->dims > 1 ? c->ntiles[1] : 0;
	}
	return BFVM_OK;
}

int bfvm_get_tuning(const bfvm_ctx* c, bfvm_tuning* out)
{
	if (!c || !out)
		return BFVM_ERR_INVALID;
	*out = c->tuning;
	return BFVM_OK;
}

int bfvm_set_tuning(bfvm_ctx* c, const bfvm_tuning* t)
{
	if (!c || !t)
		return BFVM_ERR_INVALID;
	if (t->persistent_wgs_per_cu < -1 || t->persistent_wgs_per_cu > 64)
		return c->fail(BFVM_ERR_INVALID, "persistent_wgs_per_cu must be in [-1, 64]");
	if (t->force_lds_sweeps != 0 && t->force_lds_sweeps != 1)
		return c->fail(BFVM_ERR_INVALID, "force_lds_sweeps must be 0 or 1");
	if (t->x_sweep < BFVM_X_AUTO || t->x_sweep > BFVM_X_TWO_LANE)
		return c->fail(BFVM_ERR_INVALID, "x_sweep must be one of the bfvm_x_sweep values");
	if (t->pipeline != 0 && t->pipeline != 1)
		return c->fail(BFVM_ERR_INVALID, "pipeline must be 0 or 1");
	if (t->round_nearest != 0 && t->round_nearest != 1)
		return c->fail(BFVM_ERR_INVALID, "round_nearest must be 0 or 1");
	if (t->lines_per_block != 0 && t->lines_per_block != 64 && t->lines_per_block != 128 &&
		t->lines_per_block != 256)
		return c->fail(BFVM_ERR_INVALID, "lines_per_block must be 0, 64, 128 or 256");
	if (t->split_xy < 0 || t->split_xy > 2)
		return c->fail(BFVM_ERR_INVALID, "split_xy must be 0, 1 or 2");
	TRY_HIP(c, hipStreamSynchronize(c->stream));
	c->tuning = *t;
	// the kernel variants decide whether a uniform bulk is applied by the last
	// sweep, which the domain descriptor records
	int rc = build_tiles(c);
	if (!rc)
		rc = sync_domain(c);
	if (rc)
		return rc;
	c->version++;
	return BFVM_OK;
}


int bfvm_batch_create(bfvm_ctx* const* ctxs, int32_t count, bfvm_batch** out)
{
	if (!out)
		return BFVM_ERR_INVALID;
	*out = nullptr;
	if (!ctxs || count < 1)
	{
		g_error = "a batch needs at least one context";
		return BFVM_ERR_INVALID;
	}
	auto b = std::make_unique<bfvm_batch>();
	for (int32_t i = 0; i < count; i++)
	{
		if (!ctxs[i])
		{
			g_error = "null context in batch";
			return BFVM_ERR_INVALID;
		}
		for (int32_t j = 0; j < i; j++)
			if (ctxs[j] == ctxs[i])
			{
				g_error = "a context appears twice in the batch";
				return BFVM_ERR_INVALID;
			}
		b->ctx.push_back(ctxs[i]);
	}
	b->num_cu = ctxs[0]->num_cu;
	hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
	if (e == hipSuccess)
		e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
	if (e == hipSuccess)
		e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
	if (e == hipSuccess)
		e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
	if (const char* env = std::getenv("BFVM_BATCH_LANES"))
		b->lanes = std::clamp(std::atoi(env), 1, bfvm_batch::kMaxBatchLanes);
	for (int k = 0; k < b->lanes - 1 && e == hipSuccess; k++)
	{
		e = hipStreamCreateWithFlags(&b->lane[k], hipStreamNonBlocking);
		if (e == hipSuccess)
			e = hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming);
	}
	if (e != hipSuccess)
	{
		g_error = std::string("batch stream/events: ") + hipGetErrorString(e);
		return BFVM_ERR_HIP;
	}
	*out = b.release();
	return BFVM_OK;
}

void bfvm_batch_destroy(bfvm_batch* b)
{
	if (!b)
		return;
	bool idle = !b->stream || hipStreamSynchronize(b->stream) == hipSuccess;
	for (hipStream_t l : b->lane)
		idle = idle && (!l || hipStreamSynchronize(l) == hipSuccess);
	RecycleScope r(idle);
	delete b;
}

const char* bfvm_batch_last_error(const bfvm_batch* b) { return b ? b->err.c_str() : g_error.c_str(); }

int bfvm_batch_step(bfvm_batch* b, int64_t iterations, int64_t graph_steps)
{
	if (!b || iterations < 0 || graph_steps < 0)
		return BFVM_ERR_INVALID;
	if (iterations == 0)
		return BFVM_OK;
	return batch_run(b, iterations, graph_steps);
}

int bfvm_batch_sync(bfvm_batch* b)
{
	if (!b)
		return BFVM_ERR_INVALID;
	if (hipError_t e = hipStreamSynchronize(b->stream); e != hipSuccess)
		return b->hip_fail(e, "hipStreamSynchronize");
	for (size_t i = 0; i < b->ctx.size(); i++)
		if (int rc = check_recompute(b->ctx[i]))
			return b->fail(rc, "domain " + std::to_string(i) + ": " + b->ctx[i]->err);
	return BFVM_OK;
}

int bfvm_batch_mark(bfvm_batch* b, int64_t* ticket)
{
	if (!b || !ticket)
		return BFVM_ERR_INVALID;
	const int k = (int)(b->next_mark % bfvm_batch::kMarkRing);
	if (!b->marks[k])
		if (hipError_t e = hipEventCreateWithFlags(&b->marks[k], hipEventDisableTiming); e != hipSuccess)
			return b->hip_fail(e, "hipEventCreate");
	if (hipError_t e = hipEventRecord(b->marks[k], b->stream); e != hipSuccess)
		return b->hip_fail(e, "hipEventRecord");
	*ticket = b->next_mark++;
	return BFVM_OK;
}

namespace {
int check_ticket(bfvm_batch* b, int64_t ticket)
{
	if (!b || ticket < 0 || ticket >= b->next_mark)
		return b ? b->fail(BFVM_ERR_INVALID, "unknown ticket") : BFVM_ERR_INVALID;
	if (ticket + bfvm_batch::kMarkRing <= b->next_mark - 1)
		return b->fail(BFVM_ERR_INVALID, "ticket too old (its event was reused)");
	return BFVM_OK;
}
} // namespace

int bfvm_batch_wait(bfvm_batch* b, int64_t ticket)
{
	if (int rc = check_ticket(b, ticket))
		return rc;
	if (hipError_t e = hipEventSynchronize(b->marks[ticket % bfvm_batch::kMarkRing]); e != hipSuccess)
		return b->hip_fail(e, "hipEventSynchronize");
	return BFVM_OK;
}

int bfvm_batch_query(bfvm_batch* b, int64_t ticket, int32_t* done)
{
	if (!done)
		return BFVM_ERR_INVALID;
	if (int rc = check_ticket(b, ticket))
		return rc;
	const hipError_t e = hipEventQuery(b->marks[ticket % bfvm_batch::kMarkRing]);
	if (e == hipErrorNotReady)
	{
		*done = 0;
		return BFVM_OK;
	}
	if (e != hipSuccess)
		return b->hip_fail(e, "hipEventQuery");
	*done = 1;
	return BFVM_OK;
}

int bfvm_batch_info(bfvm_batch* b, int32_t* sweep_launches, int32_t* stage_launches)
{
	if (!b)
		return BFVM_ERR_INVALID;
	int rc = batch_prepare(b);
	if (rc)
		return rc;
	if (sweep_launches)
		*sweep_launches = (int32_t)b->sweeps.size();
	if (stage_launches)
	{
		int n = 0;
		for (const auto& axis : b->stages)
			for (const auto& st : axis)
				n += st.launches();
		*stage_launches = (int32_t)n;
	}
	return BFVM_OK;
}

int bfvm_batch_autotune(bfvm_batch* b, int64_t steps, bfvm_tuning* chosen)
{
	if (!b || steps < 1)
		return BFVM_ERR_INVALID;
	// candidate tunings, applied to every context of the batch: the x kernel,
	// pipelining of y/z, persistence and lines per workgroup of the register
	// kernels (each context keeps its own rounding mode)
	std::vector<bfvm_tuning> cand;
	auto T = [](int x, int pipeline, int per_cu, int lines = 0) {
		bfvm_tuning t {};
		t.x_sweep = x;
		t.pipeline = pipeline;
		t.persistent_wgs_per_cu = per_cu;
		t.lines_per_block = lines;
		return t;
	};
	cand.push_back(T(BFVM_X_AUTO, 0, 0));
	cand.push_back(T(BFVM_X_SCAN, 0, 0));
	cand.push_back(T(BFVM_X_TRANSPOSE, 0, 0));
	cand.push_back(T(BFVM_X_PIPELINED, 0, 0));
	cand.push_back(T(BFVM_X_CHUNKED, 0, 0));
	cand.push_back(T(BFVM_X_TWO_LANE, 0, 0));
	cand.push_back(T(BFVM_X_AUTO, 1, 0));
	cand.push_back(T(BFVM_X_AUTO, 0, 1));
	cand.push_back(T(BFVM_X_AUTO, 0, 2));
	cand.push_back(T(BFVM_X_AUTO, 0, 4));
	cand.push_back(T(BFVM_X_TRANSPOSE, 1, 0));
	// lines per workgroup of the one-lane register kernels
	cand.push_back(T(BFVM_X_AUTO, 0, 0, 64));
	cand.push_back(T(BFVM_X_AUTO, 0, 0, 128));
	cand.push_back(T(BFVM_X_AUTO, 1, 0, 128));
	auto apply = [&](bfvm_ctx* c, const bfvm_tuning& t0) {
		bfvm_tuning t = t0;
		t.round_nearest = c->tuning.round_nearest;
		t.split_xy = c->tuning.split_xy;
		return bfvm_set_tuning(c, &t);
	};
	// the candidates step the domains' own buffers: what a step changes (the
	// field, the step count, the internalized amounts) is copied aside first
	// and put back afterwards, so the domains leave as they came
	if (int rc = batch_prepare(b))
		return rc;
	struct Saved
	{
		dvec<float> field;
		dvec<uint32_t> step;
		dvec<double> internalized;
	};
	std::vector<std::unique_ptr<Saved>> saved;
	auto copy = [&](void* dst, const void* src, size_t bytes, bfvm_ctx* c) -> hipError_t {
		if (!bytes)
			return hipSuccess;
		hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
		return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
	};
	for (bfvm_ctx* c : b->ctx)
	{
		auto sv = std::make_unique<Saved>();
		hipError_t e = hipStreamSynchronize(c->stream);
		if (e == hipSuccess)
			e = sv->field.alloc(c->field.n);
		if (e == hipSuccess)
			e = sv->step.alloc(1);
		if (e == hipSuccess)
			e = sv->internalized.alloc(c->internalized_d.n);
		if (e == hipSuccess)
			e = copy(sv->field.p, c->field.p, c->field.n * sizeof(float), c);
		if (e == hipSuccess)
			e = copy(sv->step.p, c->step_count.p, sizeof(uint32_t), c);
		if (e == hipSuccess)
			e = copy(sv->internalized.p, c->internalized_d.p, c->internalized_d.n * sizeof(double), c);
		if (e != hipSuccess)
			return b->hip_fail(e, "autotune state copy");
		saved.push_back(std::move(sv));
	}
	auto restore = [&]() -> int {
		if (hipError_t e = hipStreamSynchronize(b->stream); e != hipSuccess)
			return b->hip_fail(e, "hipStreamSynchronize");
		for (size_t i = 0; i < b->ctx.size(); i++)
		{
			bfvm_ctx* c = b->ctx[i];
			const Saved& sv = *saved[i];
			hipError_t e = hipStreamSynchronize(c->stream);
			if (e == hipSuccess)
				e = copy(c->field.p, sv.field.p, c->field.n * sizeof(float), c);
			if (e == hipSuccess)
				e = copy(c->step_count.p, sv.step.p, sizeof(uint32_t), c);
			if (e == hipSuccess)
				e = copy(c->internalized_d.p, sv.internalized.p, c->internalized_d.n * sizeof(double), c);
			if (e != hipSuccess)
				return b->hip_fail(e, "autotune state restore");
		}
		return BFVM_OK;
	};
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipError_t e = hipEventCreate(&e0); e != hipSuccess)
		return b->hip_fail(e, "hipEventCreate");
	if (hipError_t e = hipEventCreate(&e1); e != hipSuccess)
	{
		(void)hipEventDestroy(e0);
		return b->hip_fail(e, "hipEventCreate");
	}
	int best = -1, rc = BFVM_OK;
	float best_ms = 0;
	for (size_t k = 0; k < cand.size() && rc == BFVM_OK; k++)
	{
		for (bfvm_ctx* c : b->ctx)
			if ((rc = apply(c, cand[k])))
			{
				b->err = c->err;
				break;
			}
		if (rc)
			break;
		if ((rc = batch_run(b, 2, 0)))
			break;
		(void)hipEventRecord(e0, b->stream);
		if ((rc = batch_run(b, steps, 0)))
			break;
		(void)hipEventRecord(e1, b->stream);
		float ms = 0;
		if (hipError_t e = hipEventSynchronize(e1); e != hipSuccess)
		{
			rc = b->hip_fail(e, "hipEventSynchronize");
			break;
		}
		(void)hipEventElapsedTime(&ms, e0, e1);
		if (best < 0 || ms < best_ms)
		{
			best = (int)k;
			best_ms = ms;
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (int r2 = restore(); !rc)
		rc = r2;
	if (rc)
		return rc;
	for (bfvm_ctx* c : b->ctx)
		if ((rc = apply(c, cand[best])))
			return b->fail(rc, c->err);
	if (chosen)
		*chosen = cand[best];
	return batch_prepare(b);
}

} // extern "C"
