// Test fixture: a microenvironment filled directly through its public data
// contract (reactions-diffusion/biofvm/include/biofvm/microenvironment.h:46-81),
// with the solver taken from the registry by name.  It stands in for the
// reference's builder in the ported tests; the backend itself never builds
// microenvironments.
#pragma once

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <biofvm/microenvironment.h>
#include <biofvm/solver_registry.h>

namespace fixture {

using namespace physicore;
using namespace physicore::biofvm;

struct substrate
{
	std::string name, units;
	real_t diffusion = 0, decay = 0, initial = 0;
};

// face Dirichlet values of one substrate, per axis: min faces, max faces
struct face_values
{
	index_t substrate;
	std::array<real_t, 3> min_values, max_values;
	std::array<bool, 3> min_on { true, true, true };
	std::array<bool, 3> max_on { true, true, true };
};

// one interior Dirichlet voxel, a value per substrate (all conditions on when `on` is empty)
struct interior_voxel
{
	std::array<index_t, 3> voxel;
	std::vector<real_t> values;
	std::vector<bool> on;
};

struct env_spec
{
	index_t dims = 3;
	std::array<sindex_t, 3> lo { 0, 0, 0 }, hi { 0, 0, 0 };
	std::array<index_t, 3> voxel { 1, 1, 1 };
	real_t dt = 0.01;
	std::string name, time_units, space_units;
	std::vector<substrate> substrates;
	std::vector<face_values> faces;
	std::vector<interior_voxel> interior;
	std::string solver = "hip_solver";
};

inline std::unique_ptr<microenvironment> make_env(const env_spec& sp)
{
	const index_t S = sp.substrates.size();
	auto m = std::make_unique<microenvironment>(cartesian_mesh(sp.dims, sp.lo, sp.hi, sp.voxel), S, sp.dt);
	m->name = sp.name;
	m->time_units = sp.time_units;
	m->space_units = sp.space_units;
	m->initial_conditions = std::make_unique<real_t[]>(S);
	m->diffusion_coefficients = std::make_unique<real_t[]>(S);
	m->decay_rates = std::make_unique<real_t[]>(S);
	for (index_t s = 0; s < S; s++)
	{
		const substrate& x = sp.substrates[s];
		m->substrates_names.push_back(x.name);
		m->substrates_units.push_back(x.units);
		m->initial_conditions[s] = x.initial;
		m->diffusion_coefficients[s] = x.diffusion;
		m->decay_rates[s] = x.decay;
	}

	const index_t n = sp.interior.size();
	m->dirichlet_interior_voxels_count = n;
	m->dirichlet_interior_voxels = std::make_unique<index_t[]>(n * sp.dims);
	m->dirichlet_interior_values = std::make_unique<real_t[]>(n * S);
	m->dirichlet_interior_conditions = std::make_unique<bool[]>(n * S);
	for (index_t i = 0; i < n; i++)
	{
		const interior_voxel& v = sp.interior[i];
		for (index_t a = 0; a < sp.dims; a++)
			m->dirichlet_interior_voxels[i * sp.dims + a] = v.voxel[a];
		for (index_t s = 0; s < S; s++)
		{
			m->dirichlet_interior_values[i * S + s] = s < v.values.size() ? v.values[s] : 0;
			m->dirichlet_interior_conditions[i * S + s] = v.on.empty() || (s < v.on.size() && v.on[s]);
		}
	}

	// a face's arrays exist when some substrate has a condition on it (null: no condition)
	for (index_t a = 0; a < sp.dims; a++)
		for (int side = 0; side < 2; side++)
		{
			auto& vals = side ? m->dirichlet_max_boundary_values[a] : m->dirichlet_min_boundary_values[a];
			auto& conds = side ? m->dirichlet_max_boundary_conditions[a] : m->dirichlet_min_boundary_conditions[a];
			for (const face_values& f : sp.faces)
			{
				if (!(side ? f.max_on[a] : f.min_on[a]))
					continue;
				if (!vals)
				{
					vals = std::make_unique<real_t[]>(S);
					conds = std::make_unique<bool[]>(S);
				}
			}
			if (!vals)
				continue;
			for (const face_values& f : sp.faces)
			{
				vals[f.substrate] = side ? f.max_values[a] : f.min_values[a];
				conds[f.substrate] = side ? f.max_on[a] : f.min_on[a];
			}
		}

	m->solver = solver_registry::instance().get(sp.solver);
	if (!m->solver)
		throw std::runtime_error("no solver registered as " + sp.solver);
	return m;
}

} // namespace fixture
