#include <biofvm/microenvironment.h>

#include <algorithm>
#include <stdexcept>

using namespace physicore;
using namespace physicore::biofvm;

microenvironment::microenvironment(const cartesian_mesh& mesh, index_t substrates_count, real_t timestep)
	: diffusion_timestep(timestep), mesh(mesh), substrates_count(substrates_count)
{
	auto base = std::make_unique<base_agent_data>(mesh.dims);
	auto data = std::make_unique<agent_data>(*base, substrates_count);
	agents = std::make_shared<agent_container>(std::move(base), std::move(data));
}

// reference: reactions-diffusion/biofvm/src/microenvironment.cpp:87
void microenvironment::run_single_timestep() { solver->solve(*this, 1); }

// reference: reactions-diffusion/biofvm/src/microenvironment.cpp:89-95 (the agents
// serializer is not part of this backend)
void microenvironment::serialize_state(real_t current_time)
{
	if (serializer)
		serializer->serialize(*this, current_time);
}

real_t microenvironment::get_substrate_density(index_t s, index_t x, index_t y, index_t z) const
{
	const biofvm::solver* sp = solver.get();
	return sp->get_substrate_density(s, x, y, z);
}

// reference: microenvironment.cpp:138-204 -- update an existing interior voxel or append one
// (values and conditions of the other substrates start at 0 / off)
void microenvironment::update_dirichlet_interior_voxel(std::array<index_t, 3> voxel, index_t substrate_idx,
													   real_t value, bool condition)
{
	if (substrate_idx >= substrates_count)
		throw std::runtime_error("Substrate index out of bounds");
	const index_t dims = mesh.dims;
	for (index_t i = 0; i < dirichlet_interior_voxels_count; ++i)
	{
		const index_t* v = &dirichlet_interior_voxels[i * dims];
		bool same = true;
		for (index_t a = 0; a < dims; a++)
			same = same && v[a] == voxel[a];
		if (!same)
			continue;
		dirichlet_interior_values[i * substrates_count + substrate_idx] = value;
		dirichlet_interior_conditions[i * substrates_count + substrate_idx] = condition;
		return;
	}
	const index_t n = dirichlet_interior_voxels_count + 1;
	auto nv = std::make_unique<index_t[]>(n * dims);
	auto nval = std::make_unique<real_t[]>(n * substrates_count);
	auto ncond = std::make_unique<bool[]>(n * substrates_count);
	const index_t old = dirichlet_interior_voxels_count;
	if (old)
	{
		std::copy(dirichlet_interior_voxels.get(), dirichlet_interior_voxels.get() + old * dims, nv.get());
		std::copy(dirichlet_interior_values.get(), dirichlet_interior_values.get() + old * substrates_count,
				  nval.get());
		std::copy(dirichlet_interior_conditions.get(), dirichlet_interior_conditions.get() + old * substrates_count,
				  ncond.get());
	}
	for (index_t a = 0; a < dims; a++)
		nv[old * dims + a] = voxel[a];
	for (index_t s = 0; s < substrates_count; ++s)
	{
		nval[old * substrates_count + s] = s == substrate_idx ? value : 0.0;
		ncond[old * substrates_count + s] = s == substrate_idx ? condition : false;
	}
	dirichlet_interior_voxels = std::move(nv);
	dirichlet_interior_values = std::move(nval);
	dirichlet_interior_conditions = std::move(ncond);
	dirichlet_interior_voxels_count = n;
}

namespace {
// reference: microenvironment.cpp:206-239
void update_boundary(const microenvironment& m, char dimension, index_t substrate_idx, real_t value, bool condition,
					 std::array<std::unique_ptr<real_t[]>, 3>& values, std::array<std::unique_ptr<bool[]>, 3>& conds)
{
	const auto dim = static_cast<index_t>(dimension - 'x');
	if (dim >= m.mesh.dims)
		throw std::runtime_error("Dimension index out of bounds");
	if (substrate_idx >= m.substrates_count)
		throw std::runtime_error("Substrate index out of bounds");
	if (!values[dim])
	{
		values[dim] = std::make_unique<real_t[]>(m.substrates_count);
		conds[dim] = std::make_unique<bool[]>(m.substrates_count);
		for (index_t s = 0; s < m.substrates_count; ++s)
		{
			values[dim][s] = 0.0;
			conds[dim][s] = false;
		}
	}
	values[dim][substrate_idx] = value;
	conds[dim][substrate_idx] = condition;
}
} // namespace

void microenvironment::update_dirichlet_boundary_min(char dimension, index_t substrate_idx, real_t value,
													 bool condition)
{
	update_boundary(*this, dimension, substrate_idx, value, condition, dirichlet_min_boundary_values,
					dirichlet_min_boundary_conditions);
}

void microenvironment::update_dirichlet_boundary_max(char dimension, index_t substrate_idx, real_t value,
													 bool condition)
{
	update_boundary(*this, dimension, substrate_idx, value, condition, dirichlet_max_boundary_values,
					dirichlet_max_boundary_conditions);
}

void microenvironment::update_dirichlet_conditions() { solver->reinitialize_dirichlet(*this); }
