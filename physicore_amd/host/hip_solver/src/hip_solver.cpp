#include <hip_solver/hip_solver.h>

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include <bfvm_hip.h>
#include <biofvm/solver_registry.h>
#include <hip_solver/register_solver.h>

using namespace physicore;
using namespace physicore::biofvm;
using namespace physicore::biofvm::kernels::hip_solver;

namespace {

int env_choice(const char* name, const char* a, int va, const char* b, int vb)
{
	const char* v = std::getenv(name);
	if (!v || !*v || std::strcmp(v, a) == 0)
		return va;
	if (std::strcmp(v, b) == 0)
		return vb;
	throw std::runtime_error(std::string("hip_solver: ") + name + " must be '" + a + "' or '" + b + "'");
}

} // namespace

void hip_solver::check(int rc, const char* what) const
{
	if (rc != BFVM_OK)
		throw std::runtime_error(std::string("hip_solver: ") + what + ": " + bfvm_last_error(ctx));
}

hip_solver::~hip_solver()
{
	if (ctx)
		bfvm_destroy(ctx);
}

// openmp_solver::initialize (kernels/openmp_solver/src/openmp_solver.cpp:8-21): runs once.
void hip_solver::initialize(microenvironment& m)
{
	if (initialized)
		return;
	nx = m.mesh.grid_shape[0];
	ny = m.mesh.grid_shape[1];
	nz = m.mesh.grid_shape[2];
	S = m.substrates_count;

	bfvm_domain_desc d {};
	d.dims = (int32_t)m.mesh.dims;
	d.substrates = (int32_t)S;
	for (int a = 0; a < 3; a++)
	{
		d.grid[a] = (int64_t)m.mesh.grid_shape[a];
		d.voxel[a] = m.mesh.voxel_shape[a] ? (int64_t)m.mesh.voxel_shape[a] : 1;
		d.bbox_min[a] = m.mesh.bounding_box_mins[a];
	}
	d.dt = m.diffusion_timestep;
	d.diffusion_coefficients = m.diffusion_coefficients.get();
	d.decay_rates = m.decay_rates.get();
	d.initial_conditions = m.initial_conditions.get();
	d.compute_internalized = m.compute_internalized_substrates ? 1 : 0;
	// PHYSICORE_HIP_PRECISION=fp32 selects FP32 recurrences, PHYSICORE_HIP_REDUCE=sorted the
	// bit-reproducible cell reduction, PHYSICORE_HIP_OPENMP_QUIRKS=1 the OpenMP factor indexing.
	d.precision = env_choice("PHYSICORE_HIP_PRECISION", "fp64", BFVM_SWEEP_FP64, "fp32", BFVM_SWEEP_FP32);
	d.reduce_mode = env_choice("PHYSICORE_HIP_REDUCE", "atomic", BFVM_REDUCE_ATOMIC, "sorted", BFVM_REDUCE_SORTED);
	d.openmp_quirks = env_choice("PHYSICORE_HIP_OPENMP_QUIRKS", "0", 0, "1", 1);

	bfvm_ctx* c = nullptr;
	if (bfvm_create(&d, &c) != BFVM_OK)
		throw std::runtime_error(std::string("hip_solver: ") + bfvm_last_error(nullptr));
	ctx = c;
	push_dirichlet(m);
	push_bulk(m);
	upload_agents = true;
	mirror_valid = false;
	initialized = true;
}

void hip_solver::push_dirichlet(microenvironment& m)
{
	std::vector<double> vals(6 * S, 0.0);
	std::vector<uint8_t> conds(6 * S, 0);
	for (index_t a = 0; a < m.mesh.dims && a < 3; a++)
	{
		for (int side = 0; side < 2; side++)
		{
			const real_t* v = side == 0 ? m.dirichlet_min_boundary_values[a].get() : m.dirichlet_max_boundary_values[a].get();
			const bool* c =
				side == 0 ? m.dirichlet_min_boundary_conditions[a].get() : m.dirichlet_max_boundary_conditions[a].get();
			if (!v || !c)
				continue;
			for (index_t s = 0; s < S; s++)
			{
				vals[(2 * a + side) * S + s] = v[s];
				conds[(2 * a + side) * S + s] = c[s] ? 1 : 0;
			}
		}
	}
	check(bfvm_set_dirichlet_faces(ctx, vals.data(), conds.data()), "Dirichlet faces");

	const index_t n = m.dirichlet_interior_voxels_count;
	std::vector<int64_t> vox(n * m.mesh.dims);
	std::vector<uint8_t> ic(n * S);
	for (index_t k = 0; k < n * m.mesh.dims; k++)
		vox[k] = (int64_t)m.dirichlet_interior_voxels[k];
	for (index_t k = 0; k < n * S; k++)
		ic[k] = m.dirichlet_interior_conditions[k] ? 1 : 0;
	check(bfvm_set_dirichlet_interior(ctx, (int64_t)n, vox.data(), n ? m.dirichlet_interior_values.get() : nullptr,
									  ic.data()),
		  "Dirichlet interior");
}

// The bulk functor is sampled once into a sparse list of the voxels where it
// supplies or takes up anything; elsewhere the update is the identity
// (kernels/openmp_solver/src/bulk_solver.cpp:29 with S = U = 0).
void hip_solver::push_bulk(microenvironment& m)
{
	if (!m.bulk_fnc)
	{
		check(bfvm_set_bulk(ctx, 0, nullptr, nullptr, nullptr, nullptr), "bulk");
		return;
	}
	bulk_functor& f = *m.bulk_fnc;
	std::vector<int64_t> vox;
	std::vector<double> sup, upt, tgt;
	std::vector<double> rs(S), ru(S), rt(S);
	for (index_t z = 0; z < nz; z++)
		for (index_t y = 0; y < ny; y++)
			for (index_t x = 0; x < nx; x++)
			{
				bool any = false;
				for (index_t s = 0; s < S; s++)
				{
					rs[s] = f.supply_rates(s, x, y, z);
					ru[s] = f.uptake_rates(s, x, y, z);
					rt[s] = f.supply_target_densities(s, x, y, z);
					any = any || rs[s] != 0 || ru[s] != 0;
				}
				if (!any)
					continue;
				vox.push_back((int64_t)(x + nx * (y + ny * z)));
				sup.insert(sup.end(), rs.begin(), rs.end());
				upt.insert(upt.end(), ru.begin(), ru.end());
				tgt.insert(tgt.end(), rt.begin(), rt.end());
			}
	check(bfvm_set_bulk(ctx, (int64_t)vox.size(), vox.data(), sup.data(), upt.data(), tgt.data()), "bulk");
}

// Copies the agent arrays to HBM (data_manager::transfer_to_device,
// kernels/thrust_solver/src/data_manager.cpp:61-93).  Unless the host copy is
// declared authoritative, internalized amounts accumulated on the device are
// pulled back first so a re-upload does not lose them.
void hip_solver::push_agents(microenvironment& m, bool host_authoritative)
{
	agent_data& a = retrieve_agent_data(*m.agents);
	const index_t n = a.base_data.agents_count;
	if (!host_authoritative && agents_on_device && device_agents == n && n > 0)
		check(bfvm_download_internalized(ctx, a.internalized_substrates.data()), "internalized download");
	check(bfvm_set_agents(ctx, (int64_t)n, n ? a.base_data.positions.data() : nullptr,
						  n ? a.secretion_rates.data() : nullptr, n ? a.saturation_densities.data() : nullptr,
						  n ? a.uptake_rates.data() : nullptr, n ? a.net_export_rates.data() : nullptr,
						  n ? a.internalized_substrates.data() : nullptr,
						  n ? a.fraction_released_at_death.data() : nullptr, n ? a.volumes.data() : nullptr,
						  n ? a.is_active.data() : nullptr),
		  "agent upload");
	agents_on_device = true;
	device_agents = n;
	upload_agents = false;
}

void hip_solver::sync_mirror() const
{
	if (!ctx)
		throw std::runtime_error("hip_solver: not initialized");
	if (mirror_valid)
		return;
	mirror.resize((std::size_t)nx * ny * nz * S);
	check(bfvm_download_field_f64(ctx, mirror.data()), "field download");
	mirror_valid = true;
}

void hip_solver::push_mirror()
{
	check(bfvm_upload_field_f64(ctx, mirror.data()), "field upload");
	mirror_dirty = false;
}

// openmp_solver::solve (kernels/openmp_solver/src/openmp_solver.cpp:23-40)
void hip_solver::solve(microenvironment& m, index_t iterations)
{
	initialize(m);
	if (mirror_dirty)
		push_mirror();
	if (upload_agents)
		push_agents(m, false);
	if (iterations == 0)
		return;
	const int64_t graph_steps = iterations >= 10 ? 10 : (int64_t)iterations;
	check(bfvm_step_graph(ctx, (int64_t)iterations, graph_steps), "step");
	mirror_valid = false;
}

bfvm_ctx* hip_solver::snapshot_context()
{
	if (ctx && mirror_dirty)
		push_mirror();
	return ctx;
}

real_t hip_solver::get_substrate_density(index_t s, index_t x, index_t y, index_t z) const
{
	sync_mirror();
	return mirror[offset(s, x, y, z)];
}

real_t& hip_solver::get_substrate_density(index_t s, index_t x, index_t y, index_t z)
{
	sync_mirror();
	mirror_dirty = true; // the caller may write through the reference
	return mirror[offset(s, x, y, z)];
}

void hip_solver::transfer_to_device(microenvironment& m)
{
	initialize(m);
	push_agents(m, true);
	if (mirror_dirty)
		push_mirror();
}

void hip_solver::transfer_to_host(microenvironment& m)
{
	if (!initialized)
		return;
	sync_mirror();
	agent_data& a = retrieve_agent_data(*m.agents);
	if (agents_on_device && device_agents == a.base_data.agents_count && device_agents > 0)
		check(bfvm_download_internalized(ctx, a.internalized_substrates.data()), "internalized download");
}

// the Thrust backend re-uploads the conditions (thrust_solver.cpp:61); the
// OpenMP one reads them live (openmp_solver.cpp:58-62)
void hip_solver::reinitialize_dirichlet(microenvironment& m)
{
	if (initialized)
		push_dirichlet(m);
}

// openmp_solver.cpp:64: the next solve re-reads positions and rates and rebuilds the sums
void hip_solver::recompute_positional_data(microenvironment& /*m*/) { upload_agents = true; }

void hip_solver::release_internalized_substrates(microenvironment& m, index_t index)
{
	initialize(m);
	if (!m.compute_internalized_substrates)
		return;
	if (upload_agents)
		push_agents(m, false);
	if (mirror_dirty)
		push_mirror();
	check(bfvm_release_internalized(ctx, (int64_t)index), "release");
	mirror_valid = false;
}

void physicore::biofvm::kernels::hip_solver::attach_to_registry()
{
	static const registry_adder<hip_solver> adder("hip_solver");
}
