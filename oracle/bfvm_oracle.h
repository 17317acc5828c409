/*
 * bfvm_oracle.h -- CPU restatement of the reference BioFVM OpenMP solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * backend and the CPU baseline of the benchmark.  Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it.  Nothing in the
 * product (physicore_amd/) links, loads or calls it.
 *
 * What it restates (all paths relative to the reference checkout):
 *   step order        reactions-diffusion/biofvm/kernels/openmp_solver/src/openmp_solver.cpp:23-40
 *   Thomas factors    .../openmp_solver/src/diffusion_solver.cpp:29-86 (formulas: diffusion_solver.h:11-33)
 *   x/y/z sweeps      .../openmp_solver/src/diffusion_solver.cpp:116-202, 390-591, 593-794, 797-843
 *   Dirichlet         .../openmp_solver/src/dirichlet_solver.cpp:11-93
 *   bulk              .../openmp_solver/src/bulk_solver.cpp:12-49
 *   cells             .../openmp_solver/src/cell_solver.cpp:12-401
 *   voxel of a cell   reactions-diffusion/biofvm/src/mesh.cpp:37-61
 *   grid shape        reactions-diffusion/biofvm/src/mesh.cpp:8-28
 *
 * The field layout is ((z*ny + y)*nx + x)*S + s (substrate innermost, rows
 * unpadded).  The OpenMP backend pads each x*S row to the SIMD alignment; the
 * padding holds no values, so the arithmetic is the same.
 *
 * Pinning (see tests/test_oracle.py): the reference's own golden vectors for
 * the diffusion, bulk, cell and Dirichlet stages; the closed form of a uniform
 * field; a dense direct solve (scipy) of each implicit sweep on small grids;
 * brute-force cell sums; and the exact per-sweep mass identity of the Neumann
 * tridiagonal matrix.  The combined four-stage step has no golden vector in the
 * reference; its stage order is pinned only by reading openmp_solver.cpp:27-37.
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_domain
{
	int32_t dims;		/* 1, 2 or 3 */
	int32_t substrates; /* S */
	int64_t nx, ny, nz; /* grid shape, 1 for axes beyond dims */
	int64_t dx, dy, dz; /* voxel shape (integers, as cartesian_mesh::voxel_shape) */
	int64_t bbox_min[3];
	double dt;
	const double* diffusion; /* [S] */
	const double* decay;	 /* [S] */
	/* 1: reproduce the OpenMP backend's tile-relative y/z coefficient index
	 * (wrong coefficients when 48 % S != 0) and its n==1 factor; 0: per-substrate
	 * factors and b = 1/(1 + dt*decay/dims) for a single-voxel line, as the
	 * Thrust backend computes them (thrust_solver/src/diffusion_solver.cpp:89-100). */
	int32_t openmp_quirks;
} orc_domain;

typedef struct orc_dirichlet
{
	const double* min_values[3]; /* [S] or NULL (NULL = face off) */
	const uint8_t* min_conds[3];
	const double* max_values[3];
	const uint8_t* max_conds[3];
	int64_t interior_count;
	const int64_t* interior_voxels; /* [count*dims] */
	const double* interior_values;	/* [count*S] */
	const uint8_t* interior_conds;	/* [count*S] */
} orc_dirichlet;

/* Dense samples of a bulk_functor in field order [V*S]; all NULL = no bulk. */
typedef struct orc_bulk
{
	const double* supply;
	const double* uptake;
	const double* target;
} orc_bulk;

typedef struct orc_agents
{
	int64_t count;
	const double* positions; /* [count*dims] */
	const double* secretion; /* [count*S] */
	const double* saturation;
	const double* uptake;
	const double* net_export;
	double* internalized;			 /* [count*S], updated in place */
	const double* fraction_released; /* [count*S] */
	const double* volumes;			 /* [count] */
	const uint8_t* is_active;		 /* [count] */
} orc_agents;

/* Thomas factors of one axis (0,1,2) in FP64: b[n*S], e[(n-1)*S], c[S]. */
int orc_factors(const orc_domain* d, int axis, double* b, double* e, double* c);

/* Coefficient substrate used for column c (= x*S + s) of the y/z sweeps. */
int64_t orc_coef_substrate(const orc_domain* d, int64_t column);

/* Diffusion-decay only (the three sweeps), in place. */
int orc_diffuse_f64(const orc_domain* d, double* field, int64_t iterations);
int orc_diffuse_f32(const orc_domain* d, float* field, int64_t iterations);
/* A single sweep along one axis (0,1,2), in place. */
int orc_sweep_f64(const orc_domain* d, double* field, int axis);

int orc_dirichlet_f64(const orc_domain* d, double* field, const orc_dirichlet* dir);
int orc_bulk_f64(const orc_domain* d, double* field, const orc_bulk* bulk);

/* Full solver, mirroring openmp_solver: diffusion -> Dirichlet -> bulk -> cells.
 * All arrays are borrowed and read live on every step, as the OpenMP backend does. */
typedef struct orc_solver orc_solver;
orc_solver* orc_solver_create(const orc_domain* d, double* field, const orc_dirichlet* dir, const orc_bulk* bulk,
							  const orc_agents* agents, int compute_internalized);
orc_solver* orc_solver_create_f32(const orc_domain* d, float* field, const orc_dirichlet* dir, const orc_bulk* bulk,
								  const orc_agents* agents, int compute_internalized);
/* returns 0, or -1 if an active cell lies outside the grid */
int orc_solver_solve(orc_solver* s, int64_t iterations);
void orc_solver_recompute(orc_solver* s);
int orc_solver_release(orc_solver* s, int64_t agent_index);
/* Per-cell reduced sums after the last recompute ([count*S] each), the voxel
 * ballots ([V], -1 = none) and the conflict flag. */
int orc_solver_reduced(const orc_solver* s, double* num, double* den, double* fac, int64_t* ballots,
					   int32_t* is_conflict);
void orc_solver_destroy(orc_solver* s);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
