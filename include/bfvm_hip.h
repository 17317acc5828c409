/*
 * bfvm_hip.h -- C ABI of the MI355X BioFVM backend (libbfvm_hip.so).
 *
 * One bfvm_ctx is one BioFVM microenvironment on one GPU: the substrate field,
 * the Thomas factors, the Dirichlet and bulk hooks and the agent arrays, all
 * resident in HBM, with one HIP stream.  The C++20 hip_solver
 * (physicore_amd/host) is the reference-side caller; Python tests and the
 * benchmark call the same symbols through ctypes.  Plain pointers and sizes
 * only, no exceptions across the ABI: every call returns BFVM_OK (0) or a
 * negative bfvm_status, and bfvm_last_error() holds the message.  A context is
 * not thread-safe, as the reference solvers are not.
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference checkout; BFVM = reactions-diffusion/biofvm):
 *   biofvm::solver                     BFVM/include/biofvm/solver.h:12-37
 *   openmp_solver::initialize/solve    BFVM/kernels/openmp_solver/src/openmp_solver.cpp:8-40
 *   thrust_solver data_manager         BFVM/kernels/thrust_solver/src/data_manager.cpp:11-109
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bfvm_status
{
	BFVM_OK = 0,
	BFVM_ERR_INVALID = -1,	   /* bad argument or descriptor */
	BFVM_ERR_HIP = -2,		   /* a HIP runtime call failed */
	BFVM_ERR_NO_MEMORY = -3,   /* device allocation failed or budget exceeded */
	BFVM_ERR_OUT_OF_GRID = -4, /* an active agent lies outside the mesh */
	BFVM_ERR_UNSUPPORTED = -5, /* shape beyond what the kernels handle */
	BFVM_ERR_STATE = -6		   /* call out of order (e.g. no agents set) */
} bfvm_status;

/* Arithmetic of the Thomas recurrences.  The field is FP32 in HBM in both.
 * With FP64 recurrences every FP64 result written to FP32 (sweep outputs,
 * forward values stored in FP32, bulk and cell updates) is rounded
 * stochastically, from random bits that are a function of the domain's step
 * count, the stage and the position: the error has mean zero, so it does not
 * pile up near slowly converging states as round-to-nearest does, and results
 * are reproducible and independent of batching.  bfvm_tuning.round_nearest
 * selects round-to-nearest instead. */
typedef enum bfvm_precision
{
	BFVM_SWEEP_FP64 = 0, /* FP64 recurrences on FP32 storage (default) */
	BFVM_SWEEP_FP32 = 1	 /* FP32 recurrences on FP32 factors */
} bfvm_precision;

/* How the cell stage accumulates the per-voxel sums at recompute. */
typedef enum bfvm_reduce_mode
{
	BFVM_REDUCE_ATOMIC = 0, /* owner election by CAS, wavefront-aggregated FP64 atomics */
	BFVM_REDUCE_SORTED = 1	/* sort by voxel, segmented reduce: bit-reproducible */
} bfvm_reduce_mode;

/*
 * One microenvironment.  Mirrors the fields solver backends read from
 * biofvm::microenvironment (BFVM/include/biofvm/microenvironment.h:46-81) and
 * cartesian_mesh (BFVM/include/biofvm/mesh.h:11-32).
 */
typedef struct bfvm_domain_desc
{
	int32_t dims;		 /* 1, 2 or 3 */
	int32_t substrates;	 /* S >= 1 */
	int64_t grid[3];	 /* nx, ny, nz (1 for axes >= dims) */
	int64_t voxel[3];	 /* dx, dy, dz (integers, as cartesian_mesh::voxel_shape) */
	int64_t bbox_min[3]; /* mesh bounding box minimum */
	double dt;			 /* diffusion_timestep */
	const double* diffusion_coefficients; /* [S] */
	const double* decay_rates;			  /* [S] */
	const double* initial_conditions;	  /* [S] */
	int32_t compute_internalized;		  /* microenvironment::compute_internalized_substrates */
	int32_t precision;					  /* bfvm_precision */
	int32_t reduce_mode;				  /* bfvm_reduce_mode */
	/* 1: reproduce the OpenMP backend's tile-relative y/z factor index and its
	 * single-voxel-line factor (BFVM/kernels/openmp_solver/src/diffusion_solver.cpp:29-113,390-794);
	 * 0: per-substrate factors as BFVM/kernels/thrust_solver/src/diffusion_solver.cpp:85-140. */
	int32_t openmp_quirks;
} bfvm_domain_desc;

typedef struct bfvm_ctx bfvm_ctx;

/* Message of the last failed call on ctx (or of the last failed bfvm_create
 * on this thread when ctx is NULL).  Never NULL. */
const char* bfvm_last_error(const bfvm_ctx* ctx);

/* Number of HIP devices visible (0 when none); never fails. */
int bfvm_device_count(void);

/* openmp_solver::initialize (openmp_solver.cpp:8-21) + problem_t::construct
 * (openmp_solver/src/problem.h:27-46): allocates the field, fills it with the
 * initial conditions, precomputes the Thomas factors in FP64. */
int bfvm_create(const bfvm_domain_desc* desc, bfvm_ctx** out);
void bfvm_destroy(bfvm_ctx* ctx);

/* Face Dirichlet conditions, faces ordered x-min, x-max, y-min, y-max, z-min,
 * z-max; values[6*S], conds[6*S] (0/1).  NULL values or conds turn every face
 * off.  Replaces dirichlet_solver::initialize / reinitialize_dirichlet
 * (thrust_solver/src/dirichlet_solver.cpp:192-218; openmp_solver.cpp:58-62)
 * for the face arrays of microenvironment.h:72-75. */
int bfvm_set_dirichlet_faces(bfvm_ctx* ctx, const double* values, const uint8_t* conds);

/* Interior Dirichlet voxels (microenvironment.h:65-69): voxels[count*dims]
 * grid indices, values[count*S], conds[count*S].  Later entries of a repeated
 * voxel win, as the reference's sequential order does. */
int bfvm_set_dirichlet_interior(bfvm_ctx* ctx, int64_t count, const int64_t* voxels, const double* values,
								const uint8_t* conds);

/* Bulk supply/uptake hook (bulk_functor, BFVM/include/biofvm/bulk_functor.h:8-14;
 * bulk_solver, openmp_solver/src/bulk_solver.cpp:12-49) as a sparse list of
 * voxels: voxels[count] linear indices x + nx*(y + ny*z), supply/uptake/target
 * [count*S].  Voxels not listed have S = U = 0, for which the bulk update is
 * the identity. */
int bfvm_set_bulk(bfvm_ctx* ctx, int64_t count, const int64_t* voxels, const double* supply, const double* uptake,
				  const double* target);

/* Agent arrays (agent_data_generic_storage, BFVM/include/biofvm/agent_data_generic_storage.h:16-31),
 * FP64 as in the reference: positions[n*dims], secretion/saturation/uptake/
 * net_export/internalized/fraction_released[n*S], volumes[n], is_active[n].
 * Copies them to HBM (data_manager::transfer_to_device,
 * thrust_solver/src/data_manager.cpp:61-93) and marks the positional data for
 * recompute.  internalized may be NULL (zeros). */
int bfvm_set_agents(bfvm_ctx* ctx, int64_t n, const double* positions, const double* secretion,
					const double* saturation, const double* uptake, const double* net_export,
					const double* internalized, const double* fraction_released, const double* volumes,
					const uint8_t* is_active);

/* solver::recompute_positional_data (openmp_solver.cpp:64): the next step
 * re-bins the agents and rebuilds the per-voxel sums.  The recompute runs on
 * the context stream before that step with no host wait: the owner count and
 * the shared-voxel flag stay on the device, the step launches (and a captured
 * graph) are kept.  An active agent outside the mesh is left out of the steps
 * and reported (BFVM_ERR_OUT_OF_GRID) by the next call that waits for the
 * stream (bfvm_sync, downloads, bfvm_total_mass, bfvm_cell_sums; for a batch,
 * bfvm_batch_sync); the positional data then stays marked for recompute. */
int bfvm_recompute_positional_data(bfvm_ctx* ctx);

/* Device address of the agent positions (n*dims FP64, the layout of
 * bfvm_set_agents) for callers that move agents on the GPU: write new
 * positions there, ordered on the context stream (bfvm_stream), then call
 * bfvm_recompute_positional_data.  Nothing is copied from the host.  No
 * counterpart in the reference, whose agent arrays live in host memory
 * (data_manager::transfer_to_device copies them, thrust_solver/src/data_manager.cpp:61-93). */
int bfvm_agent_positions_device_ptr(bfvm_ctx* ctx, double** ptr);

/* openmp_solver::solve (openmp_solver.cpp:23-40): `iterations` steps of
 * diffusion -> Dirichlet -> bulk -> cells, enqueued on the context stream.
 * Returns once enqueued (an agent outside the mesh found by a recompute is
 * reported by the next call that waits for the stream). */
int bfvm_step(bfvm_ctx* ctx, int64_t iterations);

/* Same, replayed from a HIP graph that captures `graph_steps` steps (the graph
 * is rebuilt when the positional data or the hooks change). */
int bfvm_step_graph(bfvm_ctx* ctx, int64_t iterations, int64_t graph_steps);

/* Waits for the context stream. */
int bfvm_sync(bfvm_ctx* ctx);

/* Field transfers in layout ((z*ny + y)*nx + x)*S + s; FP64 host arrays are
 * converted (data_manager::transfer_to_host/device, data_manager.cpp:35-93). */
int bfvm_download_field_f64(bfvm_ctx* ctx, double* host);
int bfvm_upload_field_f64(bfvm_ctx* ctx, const double* host);
int bfvm_download_field_f32(bfvm_ctx* ctx, float* host);
int bfvm_upload_field_f32(bfvm_ctx* ctx, const float* host);

/* Snapshots of the field for output while stepping goes on (replaces the
 * per-voxel get_substrate_density loop of vtk_serializer::serialize,
 * BFVM/src/vtk_serializer.cpp:54-79).  bfvm_snapshot_begin queues, after the
 * steps queued so far, a device-to-device copy of the FP32 field into a
 * snapshot buffer of the context (allocated on first use; the copy runs at
 * HBM speed on the context stream), then the copy of that snapshot to `host`
 * (nx*ny*nz*S floats, layout as above) on a separate copy stream, and returns
 * a ticket.  Steps queued afterwards run while the host copy is in flight
 * when `host` is pinned (bfvm_host_alloc); pageable memory is copied
 * synchronously.  A snapshot begun while the previous one is still being
 * copied makes the context stream wait for that copy.  bfvm_snapshot_wait
 * blocks until snapshot `ticket` is in `host`; it may be called from another
 * thread than the one stepping the context (its errors, and those of
 * bfvm_host_alloc/free, are reported by bfvm_last_error(NULL)). */
int bfvm_snapshot_begin(bfvm_ctx* ctx, float* host, int64_t* ticket);
int bfvm_snapshot_wait(bfvm_ctx* ctx, int64_t ticket);

/* Pinned host memory for bfvm_snapshot_begin. */
int bfvm_host_alloc(size_t bytes, void** out);
int bfvm_host_free(void* p);

/* Internalized substrates of every agent, [n*S] FP64. */
int bfvm_download_internalized(bfvm_ctx* ctx, double* host);

/* cell_solver::release_internalized_substrates (openmp_solver/src/cell_solver.cpp:351-374):
 * D += I*F/V at the agent's voxel, then I = 0.  No-op unless internalized
 * substrates are tracked.  Releases are logged and applied by the next call
 * that reads or writes the field or the internalized amounts (step, sync,
 * transfers, mass, device pointer/stream, set_agents): per voxel, in call
 * order, summed in FP64 and rounded to the FP32 field once, as the reference's
 * FP64 field accumulates them. */
int bfvm_release_internalized(bfvm_ctx* ctx, int64_t agent_index);

/* Total amount per substrate, sum over voxels of density * voxel volume,
 * reduced in FP64 on the device in a fixed order; out[S]. */
int bfvm_total_mass(bfvm_ctx* ctx, double* out);

/* bfvm_total_mass without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc).  The reduction's partial sums live in a static
 * 128 KiB device array of the library (16384 doubles; every context the
 * library accepts fits: it needs S * ceil(V / 65536) + S of them), one per GPU,
 * held until the process ends.  Calls on one GPU share that array: calls that
 * overlap on one GPU (contexts on different streams or threads) race and must
 * be serialized by the caller.  Pending releases of internalized substrates
 * (bfvm_release_internalized) are applied by an upload, so they make this call
 * fail with BFVM_ERR_STATE; a size beyond the array gives BFVM_ERR_UNSUPPORTED.
 * Same result as bfvm_total_mass, bit for bit. */
int bfvm_total_mass_static(bfvm_ctx* ctx, double* out);

/* Per-owner reduced sums after the last recompute, for parity checks:
 * owners[n_owners] agent indices (ascending), num/den/fac[n_owners*S] FP64,
 * *is_conflict.  Pass NULL arrays to query *n_owners only. */
int bfvm_cell_sums(bfvm_ctx* ctx, int64_t* n_owners, int64_t* owners, double* num, double* den, double* fac,
				   int32_t* is_conflict);

/* Device pointer of the FP32 field and the context stream (hipStream_t), for
 * zero-copy callers. */
int bfvm_field_device_ptr(bfvm_ctx* ctx, void** ptr);
int bfvm_stream(bfvm_ctx* ctx, void** stream);

/* Kernel timing on the context stream with HIP events: runs `steps` steps
 * (at most 64), queued back to back with an event between launches, and
 * reports the average duration in microseconds of each stage launch (with the
 * gap before it): out_us[0..5] = x-sweep, y-sweep, z-sweep, interior
 * Dirichlet, bulk, cells (0 when a stage does not run). */
int bfvm_time_stages(bfvm_ctx* ctx, int64_t steps, double* out_us);

/* Kernel timing of one sweep: `launches` back-to-back launches of the sweep
 * along `axis` (0 = x) between two HIP events on the context stream; *out_us
 * = the average per launch, without the gaps between separately timed
 * launches.  Measurement only: the field takes `launches` extra sweeps along
 * that axis (no counterpart in the reference). */
int bfvm_time_sweep(bfvm_ctx* ctx, int32_t axis, int64_t launches, double* out_us);

/* The sweeps a step launches (diagnostics; no counterpart in the reference).
 * kernels[axis] (3 entries; axes past dims are 0): 0 = lines solved inside
 * LDS tiles, 1 = one line per lane in VGPRs read from HBM (y/z columns,
 * strided x lines), 2 = x scan, 3 = x rows staged in LDS, 4 = pipelined
 * register lines, 5 = two lanes per y/z line, 6 = x chunked 16-byte
 * transposes, 7 = x chunked 4-byte transposes, 8 = x two lanes per line.
 * split (4 entries, may be NULL): the tiles of the first halves of the x and
 * y sweeps when the x sweep runs in two halves (bfvm_tuning.split_xy; 0 and
 * 0 when it does not), then the tile counts of the whole x and y sweeps. */
int bfvm_sweep_plan(const bfvm_ctx* ctx, int32_t* kernels, int32_t* split);

/* Kernel of the x sweep.  BFVM_X_AUTO picks per shape (measured on MI355X):
 * the chunked kernels for S = 4, 8 on rows of 96..200 voxels (16-byte chunks
 * up to 176 voxels, 4-byte chunks above; S = 4 also up to 32), for S = 3 and
 * (from 96 voxels) S = 5, 6, and for S = 2 on even rows up to 200 voxels; for
 * S = 2, 4 and 8 rows of 201..400 voxels the two-lane chunked kernel; the
 * scan kernel for S = 1 and other rows past 200 voxels where it fits (up to
 * 512 voxels, S*ceil(nx/64) <= 32); the strided lines otherwise.
 * When the chunked choice does not take the shape (fewer rows than a wave
 * holds, a 1-D domain with faces), the scan kernel runs where it fits.  An
 * explicitly chosen kernel that does not take the shape falls back to the
 * strided lines (rows up to 256 voxels) or LDS tiles. */
typedef enum bfvm_x_sweep
{
	BFVM_X_AUTO = 0,	  /* the library's default for the shape */
	BFVM_X_STRIDED = 1,	  /* one line per lane in VGPRs, read from HBM with stride S */
	BFVM_X_TRANSPOSE = 2, /* rows staged in LDS, one line per lane in VGPRs */
	BFVM_X_SCAN = 3,	  /* lanes along the row, lane-parallel affine scans */
	BFVM_X_PIPELINED = 4, /* rows prefetched into LDS while the previous tile is solved
						   * (FP64, S in {1, 2, 4, 8}, nx*S % 4 == 0) */
	BFVM_X_CHUNKED = 5,	  /* one line per lane in VGPRs, filled and emptied through an
						   * LDS transpose of 32-voxel chunks (FP64): 16-byte accesses
						   * for S in {4, 8}, 4-byte accesses otherwise (S <= 7); rows
						   * of up to 200 voxels (176 for S = 1, 192 for S = 5, 7) and
						   * at least one wave of rows (64/S, even for odd S).  Rows up
						   * to twice that take two lanes per line (4-byte chunks).
						   * Other shapes use the strided lines or LDS tiles */
	BFVM_X_CHUNKED4 = 6,  /* as BFVM_X_CHUNKED with 4-byte accesses for every S
						   * (1..8), loads streamed two chunks ahead of the
						   * elimination; rows past one lane's bucket take two
						   * lanes per line up to twice that */
	BFVM_X_TWO_LANE = 7	  /* two lanes per line (4-byte chunked halves, twisted
						   * elimination) for rows of 65 voxels up to twice the
						   * one-lane limit (FP64): twice the waves of the one-lane
						   * kernels, half the registers per line */
} bfvm_x_sweep;

/* Kernel selection and rounding.  persistent_wgs_per_cu: 0 = the default per
 * kernel, -1 = one workgroup per tile, > 0 = that many workgroups per CU, each
 * looping over tiles.  force_lds_sweeps = 1 solves every line inside an LDS
 * tile instead of in registers (the variant used for lines longer than 256).
 * x_sweep: a bfvm_x_sweep value.  pipeline = 1 solves y and z lines of
 * 192..256 voxels (FP64, nx*S % 4 == 0) with the pipelined register sweep, in
 * which the next tile of 64 lines streams into LDS while the current one is
 * solved.  round_nearest = 1 rounds FP64 results to the FP32 field to nearest
 * instead of stochastically (see bfvm_precision).  lines_per_block: lines per
 * workgroup of the one-lane register kernels (y/z column lines, strided and
 * chunked x lines): 64, 128 or 256 (one, two or four waves; 0 = 256).
 * split_xy (3-D domains whose x sweep runs register or scan lines and whose
 * y sweep runs column lines): 1 = launch the x sweep in two halves on two
 * streams, the y lines of the planes the first half covers starting as soon
 * as it is done (the x sweep's last partial round of waves then overlaps y
 * work instead of running alone); 2 = one launch per axis; 0 = the default:
 * halves when the x lanes (unless the x sweep is the scan kernel) and the y
 * lanes each need more than one round of the GPU's register file and the y
 * lines hold two waves per SIMD (buckets up to 208, or two lanes per line).
 * Neither setting changes the results.
 * bfvm_time_stages always times one launch per axis. */
typedef struct bfvm_tuning
{
	int32_t persistent_wgs_per_cu;
	int32_t force_lds_sweeps;
	int32_t x_sweep;
	int32_t pipeline;
	int32_t round_nearest;
	int32_t lines_per_block;
	int32_t split_xy;
	int32_t reserved[1];
} bfvm_tuning;
int bfvm_get_tuning(const bfvm_ctx* ctx, bfvm_tuning* out);
int bfvm_set_tuning(bfvm_ctx* ctx, const bfvm_tuning* t);

/*
 * Batches: several contexts (of any shapes, substrate counts and agent sets)
 * stepped together.  Each sweep is one persistent launch per group of domains
 * that share a kernel, over the concatenated work list of their line tiles;
 * the boundary, bulk and cell stages launch once per substrate count.  Every
 * domain follows the stage order of bfvm_step and gets the same results as
 * stepped alone.  The contexts are borrowed and stay usable on their own:
 * their streams are ordered before and after each batch step by events.  A
 * batch is not stepped, synced or queried after one of its contexts is
 * destroyed; waiting on its tickets and destroying it stay allowed (neither
 * touches the contexts).  No counterpart in the reference, whose solvers step
 * one microenvironment per call (openmp_solver.cpp:23-40); this is the
 * ensemble form of that call.
 */
typedef struct bfvm_batch bfvm_batch;
int bfvm_batch_create(bfvm_ctx* const* ctxs, int32_t count, bfvm_batch** out);
/* Waits for the batch's work in flight, then frees it (the contexts stay). */
void bfvm_batch_destroy(bfvm_batch* batch);
const char* bfvm_batch_last_error(const bfvm_batch* batch);

/* `iterations` steps of every domain; with graph_steps > 0 the loop replays a
 * HIP graph of graph_steps steps (rebuilt when a domain changes). */
int bfvm_batch_step(bfvm_batch* batch, int64_t iterations, int64_t graph_steps);
int bfvm_batch_sync(bfvm_batch* batch);

/* A ticket for the work queued on the batch so far (bfvm_batch_wait blocks
 * until it has run), so that a caller can queue the next steps before waiting
 * for the previous ones.  Up to 8 tickets are kept; older ones are refused. */
int bfvm_batch_mark(bfvm_batch* batch, int64_t* ticket);
int bfvm_batch_wait(bfvm_batch* batch, int64_t ticket);
/* Whether a ticket's work has run, without blocking: *done = 1 or 0. */
int bfvm_batch_query(bfvm_batch* batch, int64_t ticket, int32_t* done);

/* Launches per step: sweeps and the other stages (after any pending
 * recompute). */
int bfvm_batch_info(bfvm_batch* batch, int32_t* sweep_launches, int32_t* stage_launches);

/* Times `steps` steps of the batch under each candidate tuning (x kernel,
 * y/z pipelining, persistent workgroups per CU), applies the fastest to every
 * context and returns it in *chosen (may be NULL).  The candidates step the
 * domains' buffers; each domain's field, step count and internalized amounts
 * are copied aside first and restored afterwards, so the domains are left as
 * they were (a device copy of each field is held meanwhile). */
int bfvm_batch_autotune(bfvm_batch* batch, int64_t steps, bfvm_tuning* chosen);

#ifdef __cplusplus
}
#endif
