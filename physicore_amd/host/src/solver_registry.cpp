#include <biofvm/solver_registry.h>

#include <hip_solver/register_solver.h>

using namespace physicore::biofvm;

// reference: reactions-diffusion/biofvm/src/solver_registry.cpp:13-52
bool solver_registry::register_factory(std::string solver_name, solver_factory_func_t&& f)
{
	auto [it, emplaced] = factory_registry.try_emplace(std::move(solver_name), std::move(f));
	(void)it;
	return emplaced;
}

solver_ptr solver_registry::get(const std::string& solver_name)
{
	auto it = factory_registry.find(solver_name);
	if (it == factory_registry.end())
		return nullptr;
	return it->second();
}

solver_registry& solver_registry::instance()
{
	static solver_registry r;
	return r;
}

namespace {
// built-in backends attach when the library loads, as the reference's
// attachment_point does for openmp_solver and the Thrust solvers
struct attachment_point
{
	attachment_point() { kernels::hip_solver::attach_to_registry(); }
};
const attachment_point ap;
} // namespace
