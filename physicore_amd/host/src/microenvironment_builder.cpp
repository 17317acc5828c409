#include <biofvm/microenvironment_builder.h>

#include <algorithm>
#include <stdexcept>

#include <biofvm/solver_registry.h>

using namespace physicore;
using namespace physicore::biofvm;

// reference: reactions-diffusion/biofvm/src/microenvironment_builder.cpp:14-212

void microenvironment_builder::set_name(std::string_view n) { name = n; }
void microenvironment_builder::set_time_units(std::string_view u) { time_units = u; }
void microenvironment_builder::set_space_units(std::string_view u) { space_units = u; }
void microenvironment_builder::set_time_step(real_t t) { timestep = t; }
void microenvironment_builder::set_simulation_time(real_t t) { simulation_time = t; }

void microenvironment_builder::resize(index_t dims, std::array<sindex_t, 3> mins, std::array<sindex_t, 3> maxs,
									  std::array<index_t, 3> voxel_shape)
{
	mesh = cartesian_mesh(dims, mins, maxs, voxel_shape);
}

void microenvironment_builder::add_density(const std::string& n, const std::string& units, real_t D, real_t decay,
										   real_t initial)
{
	substrates_names.push_back(n);
	substrates_units.push_back(units);
	diffusion_coefficients.push_back(D);
	decay_rates.push_back(decay);
	initial_conditions.push_back(initial);
	boundary_dirichlet_mins_values.push_back({ 0, 0, 0 });
	boundary_dirichlet_maxs_values.push_back({ 0, 0, 0 });
	boundary_dirichlet_mins_conditions.push_back({ false, false, false });
	boundary_dirichlet_maxs_conditions.push_back({ false, false, false });
}

std::size_t microenvironment_builder::get_density_index(const std::string& n) const
{
	auto it = std::find(substrates_names.begin(), substrates_names.end(), n);
	if (it == substrates_names.end())
		throw std::runtime_error("Density " + n + " not found");
	return (std::size_t)std::distance(substrates_names.begin(), it);
}

void microenvironment_builder::add_dirichlet_node(std::array<index_t, 3> voxel, std::vector<real_t> values,
												  std::vector<bool> conditions)
{
	if (values.size() != substrates_names.size())
		throw std::runtime_error("Dirichlet node values size does not match the number of densities");
	if (!conditions.empty() && conditions.size() != substrates_names.size())
		throw std::runtime_error("Dirichlet node conditions size does not match the number of densities");
	if (!mesh)
		throw std::runtime_error("Dirichlet node cannot be added without a mesh");
	for (index_t a = 0; a < mesh->dims; a++)
		dirichlet_voxels.push_back(voxel[a]);
	dirichlet_values.insert(dirichlet_values.end(), values.begin(), values.end());
	if (conditions.empty())
		conditions.assign(values.size(), true);
	dirichlet_conditions.insert(dirichlet_conditions.end(), conditions.begin(), conditions.end());
}

void microenvironment_builder::add_boundary_dirichlet_conditions(std::size_t k, std::array<real_t, 3> mins_values,
																 std::array<real_t, 3> maxs_values,
																 std::array<bool, 3> mins_conditions,
																 std::array<bool, 3> maxs_conditions)
{
	if (k >= substrates_names.size())
		throw std::runtime_error("Density index out of bounds");
	boundary_dirichlet_mins_values[k] = mins_values;
	boundary_dirichlet_maxs_values[k] = maxs_values;
	boundary_dirichlet_mins_conditions[k] = mins_conditions;
	boundary_dirichlet_maxs_conditions[k] = maxs_conditions;
}

void microenvironment_builder::set_bulk_functions(std::unique_ptr<bulk_functor> f) { bulk_fnc = std::move(f); }

void microenvironment_builder::do_compute_internalized_substrates() { compute_internalized_substrates = true; }

void microenvironment_builder::select_solver(const std::string& n) { solver_name = n; }

void microenvironment_builder::fill_dirichlet_vectors(microenvironment& m)
{
	// a face array exists only when some substrate has a condition on it
	// (reference microenvironment_builder.cpp:113-147)
	auto fill = [&](index_t d, const std::vector<std::array<real_t, 3>>& values,
					const std::vector<std::array<bool, 3>>& conds, std::array<std::unique_ptr<real_t[]>, 3>& out_v,
					std::array<std::unique_ptr<bool[]>, 3>& out_c) {
		bool any = false;
		for (index_t s = 0; s < m.substrates_count; s++)
			any = any || conds[s][d];
		if (!any)
			return;
		out_v[d] = std::make_unique<real_t[]>(m.substrates_count);
		out_c[d] = std::make_unique<bool[]>(m.substrates_count);
		for (index_t s = 0; s < m.substrates_count; s++)
		{
			out_v[d][s] = values[s][d];
			out_c[d][s] = conds[s][d];
		}
	};
	for (index_t d = 0; d < m.mesh.dims; d++)
	{
		fill(d, boundary_dirichlet_mins_values, boundary_dirichlet_mins_conditions, m.dirichlet_min_boundary_values,
			 m.dirichlet_min_boundary_conditions);
		fill(d, boundary_dirichlet_maxs_values, boundary_dirichlet_maxs_conditions, m.dirichlet_max_boundary_values,
			 m.dirichlet_max_boundary_conditions);
	}
}

std::unique_ptr<microenvironment> microenvironment_builder::build()
{
	if (!mesh)
		throw std::runtime_error("Microenvironment cannot be built without a mesh");
	if (substrates_names.empty())
		throw std::runtime_error("Microenvironment cannot be built wit no densities");

	auto m = std::make_unique<microenvironment>(*mesh, substrates_names.size(), timestep);
	m->name = std::move(name);
	m->time_units = std::move(time_units);
	m->space_units = std::move(space_units);
	m->simulation_time = simulation_time;
	m->substrates_names = std::move(substrates_names);
	m->substrates_units = std::move(substrates_units);

	const std::size_t S = m->substrates_count;
	m->initial_conditions = std::make_unique<real_t[]>(S);
	m->diffusion_coefficients = std::make_unique<real_t[]>(S);
	m->decay_rates = std::make_unique<real_t[]>(S);
	std::copy(initial_conditions.begin(), initial_conditions.end(), m->initial_conditions.get());
	std::copy(diffusion_coefficients.begin(), diffusion_coefficients.end(), m->diffusion_coefficients.get());
	std::copy(decay_rates.begin(), decay_rates.end(), m->decay_rates.get());

	m->dirichlet_interior_voxels_count = dirichlet_voxels.size() / m->mesh.dims;
	m->dirichlet_interior_voxels = std::make_unique<index_t[]>(dirichlet_voxels.size());
	std::copy(dirichlet_voxels.begin(), dirichlet_voxels.end(), m->dirichlet_interior_voxels.get());
	m->dirichlet_interior_values = std::make_unique<real_t[]>(dirichlet_values.size());
	std::copy(dirichlet_values.begin(), dirichlet_values.end(), m->dirichlet_interior_values.get());
	m->dirichlet_interior_conditions = std::make_unique<bool[]>(dirichlet_conditions.size());
	std::copy(dirichlet_conditions.begin(), dirichlet_conditions.end(), m->dirichlet_interior_conditions.get());

	fill_dirichlet_vectors(*m);
	m->bulk_fnc = std::move(bulk_fnc);
	m->compute_internalized_substrates = compute_internalized_substrates;

	auto s = solver_registry::instance().get(solver_name);
	if (!s)
		throw std::runtime_error("Can not find solver for microenvironment: " + solver_name);
	m->solver = std::move(s);
	return m;
}
