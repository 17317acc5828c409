// hip_solver: the MI355X backend behind biofvm::solver
// (reference interface: reactions-diffusion/biofvm/include/biofvm/solver.h:12-37; siblings
// openmp_solver, kernels/openmp_solver/src/openmp_solver.h:11-27, and thrust_solver,
// kernels/thrust_solver/src/thrust_solver.h:14-35).
//
// The field, the factors, the hooks and the agent arrays live in HBM behind the
// C ABI of libbfvm_hip.so (include/bfvm_hip.h).  The host keeps an FP64 mirror
// of the field for the real_t& accessor; it is refreshed lazily, and a mirror
// handed out through the non-const accessor is uploaded before the next solve.
// Internalized substrates stay on the device until transfer_to_host, as with
// the Thrust backend's data_manager (kernels/thrust_solver/src/data_manager.cpp:11-109).
//
// Includes only headers the reference tree has (common/, biofvm/ and
// include/bfvm_hip.h of this library), so the same sources build in the
// reference tree as a kernels/hip_solver backend.
#pragma once

#include <vector>

#include <biofvm/agent_container.h>
#include <biofvm/microenvironment.h>
#include <biofvm/solver.h>
#include <common/generic_agent_solver.h>

struct bfvm_ctx;

namespace physicore::biofvm::kernels::hip_solver {

class hip_solver : public solver, private generic_agent_solver<agent>
{
	bfvm_ctx* ctx = nullptr;
	bool initialized = false;
	bool agents_on_device = false;
	bool upload_agents = true; // re-read agent arrays from m before the next step
	index_t device_agents = 0;
	index_t nx = 1, ny = 1, nz = 1, S = 1;

	mutable std::vector<real_t> mirror; // FP64 copy of the field, [((z*ny+y)*nx+x)*S+s]
	mutable bool mirror_valid = false;	// equals the device field
	bool mirror_dirty = false;			// may have been written through the real_t& accessor

	void check(int rc, const char* what) const;
	void sync_mirror() const;
	void push_mirror();
	void push_agents(microenvironment& m, bool host_authoritative);
	void push_dirichlet(microenvironment& m);
	void push_bulk(microenvironment& m);
	std::size_t offset(index_t s, index_t x, index_t y, index_t z) const
	{
		return (((std::size_t)z * ny + y) * nx + x) * S + s;
	}

public:
	hip_solver() = default;
	hip_solver(const hip_solver&) = delete;
	hip_solver& operator=(const hip_solver&) = delete;
	~hip_solver() override;

	void initialize(microenvironment& m) override;
	void solve(microenvironment& m, index_t iterations) override;
	real_t get_substrate_density(index_t s, index_t x, index_t y, index_t z) const override;
	real_t& get_substrate_density(index_t s, index_t x, index_t y, index_t z) override;
	void transfer_to_device(microenvironment& m) override;
	void transfer_to_host(microenvironment& m) override;
	void reinitialize_dirichlet(microenvironment& m) override;
	void recompute_positional_data(microenvironment& m) override;

	// cell_solver::release_internalized_substrates (kernels/openmp_solver/src/cell_solver.cpp:351-374)
	void release_internalized_substrates(microenvironment& m, index_t index);

	// the underlying context, for callers that step or time it through the C ABI directly
	bfvm_ctx* context() const { return ctx; }

	// the context with writes made through the real_t& accessor uploaded, for
	// device snapshots of the field (null before initialize)
	bfvm_ctx* snapshot_context();
};

} // namespace physicore::biofvm::kernels::hip_solver
