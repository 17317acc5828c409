// The microenvironment a solver reads (reference: reactions-diffusion/biofvm/include/biofvm/microenvironment.h:20-82).
// The public fields are the data contract of every solver backend.  XML
// configuration is not part of this backend; VTK output is (vtk_serializer.h),
// attached by the caller.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include <common/timestep_executor.h>
#include <common/types.h>

#include "agent_container.h"
#include "bulk_functor.h"
#include "mesh.h"
#include "serializer.h"
#include "solver.h"

namespace physicore::biofvm {

class microenvironment : public timestep_executor
{
public:
	microenvironment(const cartesian_mesh& mesh, index_t substrates_count, real_t timestep);

	microenvironment(microenvironment&&) = delete;
	microenvironment(const microenvironment&) = delete;
	microenvironment& operator=(const microenvironment&) = delete;
	microenvironment& operator=(microenvironment&&) = delete;

	void run_single_timestep() override;
	void serialize_state(real_t current_time) override;

	real_t get_substrate_density(index_t s, index_t x, index_t y, index_t z) const;

	// Dirichlet condition modification methods
	void update_dirichlet_interior_voxel(std::array<index_t, 3> voxel, index_t substrate_idx, real_t value,
										 bool condition);
	void update_dirichlet_boundary_min(char dimension, index_t substrate_idx, real_t value, bool condition);
	void update_dirichlet_boundary_max(char dimension, index_t substrate_idx, real_t value, bool condition);
	void update_dirichlet_conditions();

	container_ptr agents;
	solver_ptr solver;
	serializer_ptr serializer; // called by serialize_state (reference: microenvironment.h:48)

	std::string name, time_units, space_units;
	real_t diffusion_timestep;
	real_t simulation_time = 0.0;
	cartesian_mesh mesh;

	index_t substrates_count;
	std::vector<std::string> substrates_names;
	std::vector<std::string> substrates_units;
	std::unique_ptr<real_t[]> initial_conditions;
	std::unique_ptr<real_t[]> diffusion_coefficients;
	std::unique_ptr<real_t[]> decay_rates;

	index_t dirichlet_interior_voxels_count = 0;
	std::unique_ptr<index_t[]> dirichlet_interior_voxels;  // [count * dims]
	std::unique_ptr<real_t[]> dirichlet_interior_values;   // [count * S]
	std::unique_ptr<bool[]> dirichlet_interior_conditions; // [count * S]

	std::array<std::unique_ptr<real_t[]>, 3> dirichlet_min_boundary_values = { nullptr, nullptr, nullptr };
	std::array<std::unique_ptr<real_t[]>, 3> dirichlet_max_boundary_values = { nullptr, nullptr, nullptr };
	std::array<std::unique_ptr<bool[]>, 3> dirichlet_min_boundary_conditions = { nullptr, nullptr, nullptr };
	std::array<std::unique_ptr<bool[]>, 3> dirichlet_max_boundary_conditions = { nullptr, nullptr, nullptr };

	std::unique_ptr<bulk_functor> bulk_fnc;

	bool compute_internalized_substrates = false;
};

} // namespace physicore::biofvm
