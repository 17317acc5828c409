// bfvm_oracle.cpp -- CPU restatement of the reference BioFVM OpenMP solver.
// TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline); the
// product never loads it.  Each function cites the reference lines it follows.
#include "bfvm_oracle.h"

#include <omp.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace {

constexpr int64_t tile_width = 48; // xs_tile_size_, openmp_solver/src/diffusion_solver.h:53

int64_t axis_n(const orc_domain& d, int axis) { return axis == 0 ? d.nx : (axis == 1 ? d.ny : d.nz); }
int64_t axis_shape(const orc_domain& d, int axis) { return axis == 0 ? d.dx : (axis == 1 ? d.dy : d.dz); }

// Coefficient substrate of column c (= x*S + s) in the y/z sweeps.
// openmp_solver/src/diffusion_solver.cpp:390-401 splits the merged x*S axis into
// blocks of S*copies (copies = ceil(48/S), :105) and those into tiles of 48, then
// indexes the diagonals with the tile-relative position (:427, :465, :515, :558).
int64_t coef_sub(const orc_domain& d, int64_t c)
{
	const int64_t S = d.substrates;
	if (!d.openmp_quirks)
		return c % S;
	const int64_t copies = (tile_width + S - 1) / S;
	const int64_t B = S * copies;
	const int64_t L = d.nx * S;
	const int64_t nb = L / B;
	const int64_t r = (c < nb * B) ? c % B : c - nb * B;
	return (r % tile_width) % S;
}

// diffusion_solver::precompute_values, openmp_solver/src/diffusion_solver.cpp:29-86
// (copies = 1; the y/z copies only replicate the same per-substrate values).
void factors(const orc_domain& d, int axis, double* b, double* e, double* c)
{
	const int64_t S = d.substrates;
	const int64_t n = axis_n(d, axis);
	const int64_t shape = axis_shape(d, axis);
	const double dims = (double)d.dims;
	const double h2 = (double)(shape * shape);
	const double dt = d.dt;

	for (int64_t s = 0; s < S; s++)
		c[s] = dt * d.diffusion[s] / h2;

	if (n == 1 && !d.openmp_quirks)
	{
		// thrust_solver/src/diffusion_solver.cpp:89-100: a single voxel has no
		// neighbours, so only the decay share remains on the diagonal.
		for (int64_t s = 0; s < S; s++)
			b[s] = 1 / (1 + d.decay[s] * dt / dims);
		return;
	}

	for (int64_t i = 0; i < n; i++)
		for (int64_t s = 0; s < S; s++)
		{
			double v = 1 + d.decay[s] * dt / dims + 2 * dt * d.diffusion[s] / h2;
			if (i == 0 || i == n - 1)
				v -= dt * d.diffusion[s] / h2;
			b[i * S + s] = v;
		}

	for (int64_t s = 0; s < S; s++)
		b[s] = 1 / b[s];

	for (int64_t i = 1; i < n; i++)
		for (int64_t s = 0; s < S; s++)
		{
			b[i * S + s] = 1 / (b[i * S + s] - c[s] * c[s] * b[(i - 1) * S + s]);
			e[(i - 1) * S + s] = c[s] * b[(i - 1) * S + s];
		}
}

template <typename real_t>
struct axis_factors
{
	int64_t n = 1;
	std::vector<real_t> b, e, c; // [n*S], [(n-1)*S], [S]
};

template <typename real_t>
axis_factors<real_t> make_factors(const orc_domain& d, int axis)
{
	const int64_t S = d.substrates;
	const int64_t n = axis_n(d, axis);
	std::vector<double> b(n * S), e(n > 1 ? (n - 1) * S : 1), c(S);
	factors(d, axis, b.data(), e.data(), c.data());
	axis_factors<real_t> f;
	f.n = n;
	f.b.assign(b.begin(), b.end());
	f.e.assign(e.begin(), e.end());
	f.c.assign(c.begin(), c.end());
	return f;
}

// Thomas recurrence along one line (formulas: openmp_solver/src/diffusion_solver.h:11-33).
//   forward   d_i += e_{i-1} * d_{i-1}
//   last      d_{n-1} *= b_{n-1}
//   backward  d_i = (d_i + c * d_{i+1}) * b_i
// x-sweep: solve_slice_x_2d_and_3d, openmp_solver/src/diffusion_solver.cpp:161-202
template <typename real_t>
void sweep_x(const orc_domain& d, real_t* f, const axis_factors<real_t>& F)
{
	const int64_t S = d.substrates, n = d.nx, rows = d.ny * d.nz;
	const real_t* b = F.b.data();
	const real_t* e = F.e.data();
	const real_t* c = F.c.data();
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < rows; r++)
	{
		real_t* row = f + r * n * S;
		for (int64_t i = 1; i < n; i++)
			for (int64_t s = 0; s < S; s++)
				row[i * S + s] = row[i * S + s] + e[(i - 1) * S + s] * row[(i - 1) * S + s];
		for (int64_t s = 0; s < S; s++)
			row[(n - 1) * S + s] = row[(n - 1) * S + s] * b[(n - 1) * S + s];
		for (int64_t i = n - 2; i >= 0; i--)
			for (int64_t s = 0; s < S; s++)
				row[i * S + s] = (row[i * S + s] + c[s] * row[(i + 1) * S + s]) * b[i * S + s];
	}
}

// y/z-sweeps: solve_slice_y_3d (diffusion_solver.cpp:390-591, parallel over z) and
// solve_slice_z_3d (:593-794, parallel over y); 2-D y: solve_slice_y_2d (:204-388).
// Columns are processed in tiles of 48 along the merged x*S axis with the line
// index marching inside the tile, as the reference does.  The per-column factor
// rows are expanded once per sweep so the inner loop stays vectorisable.
template <typename real_t>
void sweep_yz(const orc_domain& d, real_t* f, const axis_factors<real_t>& F, int axis)
{
	const int64_t S = d.substrates, L = d.nx * S;
	const int64_t n = F.n;
	const int64_t outer = (axis == 1) ? d.nz : d.ny;
	const int64_t outer_stride = (axis == 1) ? d.ny * L : L;
	const int64_t line_stride = (axis == 1) ? L : d.ny * L;

	std::vector<real_t> bx(n * L), ex(n > 1 ? (n - 1) * L : 1), cx(L);
	for (int64_t col = 0; col < L; col++)
	{
		const int64_t cs = coef_sub(d, col);
		cx[col] = F.c[cs];
		for (int64_t i = 0; i < n; i++)
			bx[i * L + col] = F.b[i * S + cs];
		for (int64_t i = 0; i + 1 < n; i++)
			ex[i * L + col] = F.e[i * S + cs];
	}

#pragma omp parallel for schedule(static)
	for (int64_t o = 0; o < outer; o++)
	{
		real_t* base = f + o * outer_stride;
		for (int64_t t0 = 0; t0 < L; t0 += tile_width)
		{
			const int64_t t1 = (t0 + tile_width < L) ? t0 + tile_width : L;
			for (int64_t i = 1; i < n; i++)
			{
				real_t* cur = base + i * line_stride;
				const real_t* prev = base + (i - 1) * line_stride;
				const real_t* er = ex.data() + (i - 1) * L;
				for (int64_t col = t0; col < t1; col++)
					cur[col] = cur[col] + er[col] * prev[col];
			}
			{
				real_t* last = base + (n - 1) * line_stride;
				const real_t* br = bx.data() + (n - 1) * L;
				for (int64_t col = t0; col < t1; col++)
					last[col] = last[col] * br[col];
			}
			for (int64_t i = n - 2; i >= 0; i--)
			{
				real_t* cur = base + i * line_stride;
				const real_t* next = base + (i + 1) * line_stride;
				const real_t* br = bx.data() + i * L;
				for (int64_t col = t0; col < t1; col++)
					cur[col] = (cur[col] + cx[col] * next[col]) * br[col];
			}
		}
	}
}

template <typename real_t>
struct diffusion
{
	orc_domain d;
	axis_factors<real_t> fx, fy, fz;

	explicit diffusion(const orc_domain& dom) : d(dom)
	{
		fx = make_factors<real_t>(d, 0);
		if (d.dims >= 2)
			fy = make_factors<real_t>(d, 1);
		if (d.dims >= 3)
			fz = make_factors<real_t>(d, 2);
	}

	// diffusion_solver::solve, openmp_solver/src/diffusion_solver.cpp:797-843:
	// every axis below dims is swept, x then y then z.
	void solve(real_t* f) const
	{
		sweep_x(d, f, fx);
		if (d.dims >= 2)
			sweep_yz(d, f, fy, 1);
		if (d.dims >= 3)
			sweep_yz(d, f, fz, 2);
	}
};

// dirichlet_solver::solve, openmp_solver/src/dirichlet_solver.cpp:44-93: faces
// x-min, x-max, then y, then z (later writes win), then the interior list.
template <typename real_t>
void dirichlet(const orc_domain& d, real_t* f, const orc_dirichlet* dir)
{
	if (!dir)
		return;
	const int64_t S = d.substrates, nx = d.nx, ny = d.ny, nz = d.nz;
	auto at = [&](int64_t x, int64_t y, int64_t z, int64_t s) -> real_t& {
		return f[((z * ny + y) * nx + x) * S + s];
	};
	for (int axis = 0; axis < d.dims; axis++)
	{
		for (int side = 0; side < 2; side++)
		{
			const double* vals = side == 0 ? dir->min_values[axis] : dir->max_values[axis];
			const uint8_t* conds = side == 0 ? dir->min_conds[axis] : dir->max_conds[axis];
			if (!vals || !conds)
				continue;
			const int64_t n = axis_n(d, axis);
			const int64_t fixed = side == 0 ? 0 : n - 1;
			for (int64_t z = 0; z < nz; z++)
				for (int64_t y = 0; y < ny; y++)
					for (int64_t x = 0; x < nx; x++)
					{
						const int64_t idx[3] = { x, y, z };
						if (idx[axis] != fixed)
							continue;
						for (int64_t s = 0; s < S; s++)
							if (conds[s])
								at(x, y, z, s) = (real_t)vals[s];
					}
		}
	}
	for (int64_t k = 0; k < dir->interior_count; k++)
	{
		const int64_t* v = dir->interior_voxels + k * d.dims;
		const int64_t x = v[0], y = d.dims >= 2 ? v[1] : 0, z = d.dims >= 3 ? v[2] : 0;
		for (int64_t s = 0; s < S; s++)
			if (dir->interior_conds[k * S + s])
				at(x, y, z, s) = (real_t)dir->interior_values[k * S + s];
	}
}

// bulk_solver::solve_single, openmp_solver/src/bulk_solver.cpp:12-41:
// D = (D + dt*S*T) / (1 + dt*(U+S)) for every (s,x,y,z).
template <typename real_t>
void bulk(const orc_domain& d, real_t* f, const orc_bulk* b)
{
	if (!b || !b->supply || !b->uptake || !b->target)
		return;
	const int64_t n = d.nx * d.ny * d.nz * d.substrates;
	const real_t dt = (real_t)d.dt;
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; i++)
	{
		const real_t S = (real_t)b->supply[i], U = (real_t)b->uptake[i], T = (real_t)b->target[i];
		f[i] = (f[i] + dt * S * T) / (1 + dt * (U + S));
	}
}

// cartesian_mesh::voxel_position, reactions-diffusion/biofvm/src/mesh.cpp:37-61:
// truncation of (pos - min) / voxel_shape in FP64.  A position on the max face
// gives index n (the reference only asserts); it is reported as an error here.
bool voxel_of(const orc_domain& d, const double* pos, int64_t* v)
{
	const int64_t shape[3] = { d.dx, d.dy, d.dz };
	const int64_t n[3] = { d.nx, d.ny, d.nz };
	int64_t idx[3] = { 0, 0, 0 };
	for (int a = 0; a < d.dims; a++)
	{
		const double q = (pos[a] - (double)d.bbox_min[a]) / (double)shape[a];
		if (!(q >= 0) || q >= (double)n[a])
			return false;
		idx[a] = (int64_t)(uint64_t)q;
	}
	// ballot layout: vectors<'x','y','z'>, cell_solver.cpp:336-338
	*v = idx[0] + n[0] * (idx[1] + n[1] * idx[2]);
	return true;
}

template <typename real_t>
struct solver
{
	orc_domain d;
	real_t* field;
	const orc_dirichlet* dir;
	const orc_bulk* blk;
	const orc_agents* ag;
	bool compute_internalized;
	diffusion<real_t> diff;

	bool recompute = true; // openmp_solver.h:14
	bool is_conflict = false;
	std::vector<double> num, den, fac;	  // per cell (cell_solver.h:47-49)
	std::vector<double> rnum, rden, rfac; // reduced, indexed by owner cell (:51-53)
	std::vector<int64_t> ballots;		  // per voxel (:57)
	std::vector<int64_t> cell_voxel;

	solver(const orc_domain& dom, real_t* f, const orc_dirichlet* di, const orc_bulk* b, const orc_agents* a,
		   bool internalized)
		: d(dom), field(f), dir(di), blk(b), ag(a), compute_internalized(internalized), diff(dom)
	{
		ballots.assign(d.nx * d.ny * d.nz, -1);
	}

	size_t V() const { return (size_t)(d.nx * d.ny * d.nz); }
	double voxel_volume() const { return (double)(d.dx * d.dy * d.dz); }

	// compute_intermediates + clear_ballots + ballot_and_sum,
	// cell_solver.cpp:22-129, run in cell-index order.  The first active cell of a
	// voxel is its owner; the owner's slot accumulates (den+1) for itself and den
	// for the others (:104-126), so the sums here are in the canonical order
	// owner-first then ascending cell index.
	bool do_recompute()
	{
		const int64_t n = ag ? ag->count : 0;
		const int64_t S = d.substrates;
		const double V = voxel_volume();
		const double dt = d.dt;
		num.assign(n * S, 0);
		den.assign(n * S, 0);
		fac.assign(n * S, 0);
		rnum.assign(n * S, 0);
		rden.assign(n * S, 0);
		rfac.assign(n * S, 0);
		cell_voxel.assign(n, -1);
		is_conflict = false;
		for (int64_t i = 0; i < n; i++)
		{
			if (!ag->is_active[i])
				continue;
			int64_t v;
			if (!voxel_of(d, ag->positions + i * d.dims, &v))
				return false;
			cell_voxel[i] = v;
			for (int64_t s = 0; s < S; s++)
			{
				const int64_t k = i * S + s;
				num[k] = ag->secretion[k] * ag->saturation[k] * dt * ag->volumes[i] / V;
				den[k] = (ag->uptake[k] + ag->secretion[k]) * dt * ag->volumes[i] / V;
				fac[k] = ag->net_export[k] * dt / V;
			}
			ballots[v] = -1;
		}
		for (int64_t i = 0; i < n; i++)
		{
			if (!ag->is_active[i])
				continue;
			const int64_t v = cell_voxel[i];
			if (ballots[v] < 0)
			{
				ballots[v] = i;
				for (int64_t s = 0; s < S; s++)
				{
					rnum[i * S + s] += num[i * S + s];
					rden[i * S + s] += den[i * S + s] + 1;
					rfac[i * S + s] += fac[i * S + s];
				}
			}
			else
			{
				is_conflict = true;
				const int64_t o = ballots[v];
				for (int64_t s = 0; s < S; s++)
				{
					rnum[o * S + s] += num[i * S + s];
					rden[o * S + s] += den[i * S + s];
					rfac[o * S + s] += fac[i * S + s];
				}
			}
		}
		return true;
	}

	// compute_result, cell_solver.cpp:187-244 (+ compute_fused :165-185,
	// compute_densities :146-163, compute_internalized :131-144).
	void apply()
	{
		const int64_t n = ag ? ag->count : 0;
		const int64_t S = d.substrates;
		const double V = voxel_volume();
		if (compute_internalized && !is_conflict)
		{
			for (int64_t i = 0; i < n; i++)
			{
				if (!ag->is_active[i] || cell_voxel[i] < 0)
					continue;
				real_t* D = field + cell_voxel[i] * S;
				for (int64_t s = 0; s < S; s++)
				{
					const double prev = (double)D[s];
					const double nv = (prev + rnum[i * S + s]) / rden[i * S + s] + rfac[i * S + s];
					D[s] = (real_t)nv;
					ag->internalized[i * S + s] += V * (prev - (double)D[s]);
				}
			}
			return;
		}
		for (int64_t i = 0; i < n; i++)
		{
			if (!ag->is_active[i] || cell_voxel[i] < 0)
				continue;
			if (ballots[cell_voxel[i]] != i)
				continue;
			real_t* D = field + cell_voxel[i] * S;
			for (int64_t s = 0; s < S; s++)
				D[s] = (real_t)(((double)D[s] + rnum[i * S + s]) / rden[i * S + s] + rfac[i * S + s]);
		}
		if (compute_internalized)
		{
			for (int64_t i = 0; i < n; i++)
			{
				if (!ag->is_active[i] || cell_voxel[i] < 0)
					continue;
				const real_t* D = field + cell_voxel[i] * S;
				for (int64_t s = 0; s < S; s++)
					ag->internalized[i * S + s] -=
						V * (num[i * S + s] - (double)D[s] * den[i * S + s] + fac[i * S + s]);
			}
		}
	}

	// openmp_solver::solve, openmp_solver.cpp:23-40
	int solve(int64_t iterations)
	{
		for (int64_t it = 0; it < iterations; it++)
		{
			diff.solve(field);
			dirichlet(d, field, dir);
			bulk(d, field, blk);
			if (ag && ag->count > 0)
			{
				// simulate_secretion_and_uptake, cell_solver.cpp:302-349
				if (recompute && !do_recompute())
					return -1;
				apply();
			}
		}
		recompute = false;
		return 0;
	}

	// release_internalized_substrates, cell_solver.cpp:351-374 and :274-289
	int release(int64_t index)
	{
		if (!compute_internalized)
			return 0;
		if (!ag || index < 0 || index >= ag->count)
			return -1;
		int64_t v;
		if (!voxel_of(d, ag->positions + index * d.dims, &v))
			return -1;
		const int64_t S = d.substrates;
		const double V = voxel_volume();
		real_t* D = field + v * S;
		for (int64_t s = 0; s < S; s++)
		{
			D[s] = (real_t)((double)D[s] + ag->internalized[index * S + s] * ag->fraction_released[index * S + s] / V);
			ag->internalized[index * S + s] = 0;
		}
		return 0;
	}
};

} // namespace

struct orc_solver
{
	std::unique_ptr<solver<double>> s64;
	std::unique_ptr<solver<float>> s32;
};

extern "C" {

int orc_factors(const orc_domain* d, int axis, double* b, double* e, double* c)
{
	if (!d || axis < 0 || axis > 2)
		return -1;
	factors(*d, axis, b, e, c);
	return 0;
}

int64_t orc_coef_substrate(const orc_domain* d, int64_t column) { return coef_sub(*d, column); }

int orc_diffuse_f64(const orc_domain* d, double* field, int64_t iterations)
{
	diffusion<double> diff(*d);
	for (int64_t i = 0; i < iterations; i++)
		diff.solve(field);
	return 0;
}

int orc_diffuse_f32(const orc_domain* d, float* field, int64_t iterations)
{
	diffusion<float> diff(*d);
	for (int64_t i = 0; i < iterations; i++)
		diff.solve(field);
	return 0;
}

int orc_sweep_f64(const orc_domain* d, double* field, int axis)
{
	if (axis == 0)
		sweep_x(*d, field, make_factors<double>(*d, 0));
	else
		sweep_yz(*d, field, make_factors<double>(*d, axis), axis);
	return 0;
}

int orc_dirichlet_f64(const orc_domain* d, double* field, const orc_dirichlet* dir)
{
	dirichlet(*d, field, dir);
	return 0;
}

int orc_bulk_f64(const orc_domain* d, double* field, const orc_bulk* b)
{
	bulk(*d, field, b);
	return 0;
}

orc_solver* orc_solver_create(const orc_domain* d, double* field, const orc_dirichlet* dir, const orc_bulk* bulk,
							  const orc_agents* agents, int compute_internalized)
{
	auto* s = new orc_solver;
	s->s64 = std::make_unique<solver<double>>(*d, field, dir, bulk, agents, compute_internalized != 0);
	return s;
}

orc_solver* orc_solver_create_f32(const orc_domain* d, float* field, const orc_dirichlet* dir, const orc_bulk* bulk,
								  const orc_agents* agents, int compute_internalized)
{
	auto* s = new orc_solver;
	s->s32 = std::make_unique<solver<float>>(*d, field, dir, bulk, agents, compute_internalized != 0);
	return s;
}

int orc_solver_solve(orc_solver* s, int64_t iterations)
{
	return s->s64 ? s->s64->solve(iterations) : s->s32->solve(iterations);
}

void orc_solver_recompute(orc_solver* s)
{
	if (s->s64)
		s->s64->recompute = true;
	else
		s->s32->recompute = true;
}

int orc_solver_release(orc_solver* s, int64_t index)
{
	return s->s64 ? s->s64->release(index) : s->s32->release(index);
}

int orc_solver_reduced(const orc_solver* s, double* num, double* den, double* fac, int64_t* ballots,
					   int32_t* is_conflict)
{
	auto copy = [&](const auto& sv) {
		const size_t n = sv.rnum.size();
		if (num)
			std::memcpy(num, sv.rnum.data(), n * sizeof(double));
		if (den)
			std::memcpy(den, sv.rden.data(), n * sizeof(double));
		if (fac)
			std::memcpy(fac, sv.rfac.data(), n * sizeof(double));
		if (ballots)
			std::memcpy(ballots, sv.ballots.data(), sv.ballots.size() * sizeof(int64_t));
		if (is_conflict)
			*is_conflict = sv.is_conflict ? 1 : 0;
	};
	if (s->s64)
		copy(*s->s64);
	else
		copy(*s->s32);
	return 0;
}

void orc_solver_destroy(orc_solver* s) { delete s; }

int orc_num_threads(void) { return omp_get_max_threads(); }

} // extern "C"
