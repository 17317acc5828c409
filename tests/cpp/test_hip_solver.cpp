// The reference's solver tests, ported onto hip_solver through the public
// biofvm interface (microenvironment, solver, registry).  Microenvironments are
// filled through their data contract by env_fixture.h.
//
// Sources (reference checkout): reactions-diffusion/biofvm/kernels/openmp_solver/tests/
// test_diffusion_solver.cpp, test_dirichlet_solver.cpp, test_bulk_solver.cpp,
// test_cell_solver.cpp; reactions-diffusion/biofvm/tests/test_solver_registry.cpp,
// test_dirichlet_modification.cpp.  The reference drives each sub-solver
// directly; here every case goes through solve(), with D = decay = 0 where a
// test isolates the Dirichlet, bulk or cell stage (the sweeps are then the
// identity).  Values are read back after transfer_to_host, as the Thrust
// backend's tests do.  The field is FP32 on the device, so expected values are
// held to max(reference tolerance, 3e-7 relative).
//
// The VTK output tests port reactions-diffusion/biofvm/tests/test_vtk_serializer.cpp;
// the files are read back with the small reader below instead of the VTK library.
//
// Usage: test_hip_solver [--host-only] [filter-substring]
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iterator>
#include <map>
#include <memory>
#include <span>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <bfvm_hip.h>
#include <biofvm/microenvironment.h>
#include <biofvm/solver_registry.h>
#include <biofvm/vtk_serializer.h>
#include <hip_solver/hip_solver.h>

#include "env_fixture.h"
#include <common/generic_agent_solver.h>

using namespace physicore;
using namespace physicore::biofvm;
using physicore::biofvm::kernels::hip_solver::hip_solver;
using fixture::env_spec;
using fixture::make_env;

// ------------------------------------------------------------- harness

namespace {
// the agent arrays of a microenvironment, reached as a solver reaches them
struct agent_access : physicore::generic_agent_solver<agent>
{};
agent_data& agent_arrays_of(microenvironment& m) { return agent_access {}.retrieve_agent_data(*m.agents); }
} // namespace

namespace {
struct test_case
{
	const char* name;
	bool gpu;
	std::function<void()> fn;
};
std::vector<test_case>& registry_of_tests()
{
	static std::vector<test_case> t;
	return t;
}
struct adder
{
	adder(const char* n, bool gpu, std::function<void()> f) { registry_of_tests().push_back({ n, gpu, std::move(f) }); }
};
int failures = 0;
int checks = 0;
void fail(const char* file, int line, const std::string& what)
{
	failures++;
	std::printf("    FAIL %s:%d: %s\n", file, line, what.c_str());
}
} // namespace

#define TEST_GPU(suite, name)                                    \
	static void suite##_##name();                                \
	static adder add_##suite##_##name(#suite "." #name, true, suite##_##name); \
	static void suite##_##name()
#define TEST_HOST(suite, name)                                   \
	static void suite##_##name();                                \
	static adder add_##suite##_##name(#suite "." #name, false, suite##_##name); \
	static void suite##_##name()

#define EXPECT_NEAR(a, b, tol)                                                                              \
	do                                                                                                      \
	{                                                                                                       \
		checks++;                                                                                           \
		const double a_ = (double)(a), b_ = (double)(b), t_ = (double)(tol);                                \
		if (!(std::fabs(a_ - b_) <= t_))                                                                    \
			fail(__FILE__, __LINE__, #a " = " + std::to_string(a_) + ", expected " + std::to_string(b_) + \
										 " +- " + std::to_string(t_));                                     \
	} while (0)
// FP32 field: the reference tolerance or a few FP32 ulps of the value, whichever is larger
#define EXPECT_FIELD(a, b, tol) EXPECT_NEAR(a, b, std::max((double)(tol), 3e-7 * std::fabs((double)(b))))
#define EXPECT_EQ(a, b)                                                  \
	do                                                                   \
	{                                                                    \
		checks++;                                                        \
		if (!((a) == (b)))                                               \
			fail(__FILE__, __LINE__, std::string(#a " == " #b " failed")); \
	} while (0)
#define EXPECT_NE(a, b)                                                  \
	do                                                                   \
	{                                                                    \
		checks++;                                                        \
		if ((a) == (b))                                                  \
			fail(__FILE__, __LINE__, std::string(#a " != " #b " failed")); \
	} while (0)
#define EXPECT_TRUE(a) EXPECT_EQ((bool)(a), true)
#define EXPECT_FALSE(a) EXPECT_EQ((bool)(a), false)
#define EXPECT_THROW(stmt, exc)                                                     \
	do                                                                              \
	{                                                                               \
		checks++;                                                                   \
		bool t_ = false;                                                            \
		try                                                                         \
		{                                                                           \
			stmt;                                                                   \
		}                                                                           \
		catch (const exc&)                                                          \
		{                                                                           \
			t_ = true;                                                              \
		}                                                                           \
		if (!t_)                                                                    \
			fail(__FILE__, __LINE__, std::string(#stmt " did not throw " #exc));    \
	} while (0)

// ------------------------------------------------------------ fixtures

namespace {

std::unique_ptr<microenvironment> make_env(const cartesian_mesh& mesh, real_t dt, std::vector<real_t> D,
										   std::vector<real_t> lam, std::vector<real_t> init)
{
	const index_t S = D.size();
	auto m = std::make_unique<microenvironment>(mesh, S, dt);
	m->diffusion_coefficients = std::make_unique<real_t[]>(S);
	m->decay_rates = std::make_unique<real_t[]>(S);
	m->initial_conditions = std::make_unique<real_t[]>(S);
	for (index_t s = 0; s < S; s++)
	{
		m->diffusion_coefficients[s] = D[s];
		m->decay_rates[s] = lam[s];
		m->initial_conditions[s] = init[s];
	}
	m->solver = std::make_unique<hip_solver>();
	return m;
}

// test_diffusion_solver.cpp:12-34
std::unique_ptr<microenvironment> default_microenv(const cartesian_mesh& mesh)
{
	return make_env(mesh, 5, { 4, 2 }, { 5, 3 }, { 1, 1 });
}

// test_diffusion_solver.cpp:36-58
std::unique_ptr<microenvironment> biorobots_microenv(const cartesian_mesh& mesh)
{
	return make_env(mesh, 0.01, { 1000, 1000 }, { 0.1, 0.4 }, { 0, 0 });
}

// a stage in isolation: the sweeps are the identity with D = decay = 0
std::unique_ptr<microenvironment> still_microenv(const cartesian_mesh& mesh, real_t dt, std::vector<real_t> init)
{
	return make_env(mesh, dt, std::vector<real_t>(init.size(), 0.0), std::vector<real_t>(init.size(), 0.0), init);
}

real_t dens(microenvironment& m, index_t s, index_t x, index_t y = 0, index_t z = 0)
{
	return m.get_substrate_density(s, x, y, z);
}

} // namespace

// ----------------------------------------------------------- diffusion

TEST_GPU(DiffusionSolverTest, Uniform1D)
{
	auto m = default_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 80, 0, 0 }, { 20, 0, 0 }));
	m->solver->initialize(*m);
	m->run_single_timestep();
	for (index_t x = 0; x < m->mesh.grid_shape[0]; x++)
	{
		EXPECT_FIELD(dens(*m, 0, x), 0.03846154, 1e-6);
		EXPECT_FIELD(dens(*m, 1, x), 0.0625, 1e-6);
	}
}

TEST_GPU(DiffusionSolverTest, Uniform2D)
{
	auto m = default_microenv(cartesian_mesh(2, { 0, 0, 0 }, { 800, 800, 0 }, { 20, 20, 0 }));
	m->run_single_timestep();
	for (index_t y = 0; y < m->mesh.grid_shape[1]; y++)
		for (index_t x = 0; x < m->mesh.grid_shape[0]; x++)
		{
			EXPECT_FIELD(dens(*m, 0, x, y), 0.0054869675, 1e-6);
			EXPECT_FIELD(dens(*m, 1, x, y), 0.013840831, 1e-6);
		}
}

TEST_GPU(DiffusionSolverTest, Uniform3D)
{
	auto m = default_microenv(cartesian_mesh(3, { 0, 0, 0 }, { 200, 200, 200 }, { 20, 20, 20 }));
	m->run_single_timestep();
	for (index_t z = 0; z < 10; z++)
		for (index_t y = 0; y < 10; y++)
			for (index_t x = 0; x < 10; x++)
			{
				EXPECT_FIELD(dens(*m, 0, x, y, z), 0.0012299563, 1e-6);
				EXPECT_FIELD(dens(*m, 1, x, y, z), 0.0046296306, 1e-6);
			}
}

namespace {
// field = linear index s + x*S + y*S*nx + z*S*nx*ny, written through the real_t& accessor
void fill_linear(microenvironment& m)
{
	m.solver->initialize(m);
	const auto g = m.mesh.grid_shape;
	const index_t S = m.substrates_count;
	for (index_t z = 0; z < g[2]; z++)
		for (index_t y = 0; y < g[1]; y++)
			for (index_t x = 0; x < g[0]; x++)
				for (index_t s = 0; s < S; s++)
					m.solver->get_substrate_density(s, x, y, z) = (real_t)(s + S * (x + g[0] * (y + g[1] * z)));
}
void check_linear(microenvironment& m, const std::vector<double>& expected)
{
	const auto g = m.mesh.grid_shape;
	const index_t S = m.substrates_count;
	for (index_t z = 0; z < g[2]; z++)
		for (index_t y = 0; y < g[1]; y++)
			for (index_t x = 0; x < g[0]; x++)
				for (index_t s = 0; s < S; s++)
					EXPECT_FIELD(dens(m, s, x, y, z), expected[s + S * (x + g[0] * (y + g[1] * z))], 1e-6);
}
} // namespace

TEST_GPU(DiffusionSolverTest, Random1D)
{
	auto m = biorobots_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 60, 0, 0 }, { 20, 20, 20 }));
	fill_linear(*m);
	m->run_single_timestep();
	check_linear(*m, { 0.0486842592, 1.0444132121, 1.9980019980, 2.9880478088, 3.9473197368, 4.9316824055 });
}

TEST_GPU(DiffusionSolverTest, Random2D)
{
	auto m = biorobots_microenv(cartesian_mesh(2, { 0, 0, 0 }, { 60, 60, 0 }, { 20, 20, 20 }));
	fill_linear(*m);
	m->run_single_timestep();
	check_linear(*m, { 0.1948319355, 1.1899772978, 2.1441254507, 3.1335099015, 4.0934189658, 5.0770425053,
					   6.0427124809, 7.0205751090, 7.9920058, 8.9641077127, 9.9412995111, 10.9076403164,
					   11.8905930262, 12.8511729202, 13.8398865413, 14.7947055239, 15.7891800565, 16.7382381276 });
}

TEST_GPU(DiffusionSolverTest, Random3D)
{
	auto m = biorobots_microenv(cartesian_mesh(3, { 0, 0, 0 }, { 60, 60, 60 }, { 20, 20, 20 }));
	fill_linear(*m);
	m->run_single_timestep();
	check_linear(*m, { 0.6333066643,  1.6268066007,	 2.5825920996,	3.5703051208,  4.5318775349,  5.5138036408,
					   6.4811629703,  7.4573021609,	 8.4304484056,	9.4008006809,  10.3797338410, 11.3442992010,
					   12.3290192763, 13.2877977210, 14.2783047117, 15.2312962410, 16.2275901470, 17.1747947611,
					   18.1768755823, 19.1182932811, 20.1261610177, 21.0617918012, 22.0754464530, 23.0052903212,
					   24.0247318884, 24.9487888412, 25.9740173237, 26.8922873613, 27.9233027591, 28.8357858813,
					   29.8725881944, 30.7792844014, 31.8218736297, 32.7227829214, 33.7711590651, 34.6662814414,
					   35.7204445004, 36.6097799615, 37.6697299358, 38.5532784815, 39.6190153711, 40.4967770016,
					   41.5683008064, 42.4402755216, 43.5175862418, 44.3837740416, 45.4668716771, 46.3272725617,
					   47.4161571125, 48.2707710817, 49.3654425478, 50.2142696018, 51.3147279832, 52.1577681218 });
}

// ----------------------------------------------------------- Dirichlet

namespace {
// test_dirichlet_solver.cpp:39-60: only the first substrate is conditioned
void add_dirichlet_at(microenvironment& m, const std::vector<std::array<index_t, 3>>& idx,
					  const std::vector<real_t>& values)
{
	const index_t S = m.substrates_count, dims = m.mesh.dims;
	m.dirichlet_interior_voxels_count = idx.size();
	m.dirichlet_interior_voxels = std::make_unique<index_t[]>(idx.size() * dims);
	m.dirichlet_interior_values = std::make_unique<real_t[]>(idx.size() * S);
	m.dirichlet_interior_conditions = std::make_unique<bool[]>(idx.size() * S);
	for (index_t i = 0; i < idx.size(); i++)
	{
		for (index_t d = 0; d < dims; d++)
			m.dirichlet_interior_voxels[i * dims + d] = idx[i][d];
		m.dirichlet_interior_values[i * S] = values[i];
		m.dirichlet_interior_conditions[i * S] = true;
		for (index_t j = 1; j < S; j++)
			m.dirichlet_interior_conditions[i * S + j] = false;
	}
}
// test_dirichlet_solver.cpp:62-83
void add_boundary_dirichlet(microenvironment& m, index_t dim, bool min, real_t value)
{
	auto& values = min ? m.dirichlet_min_boundary_values[dim] : m.dirichlet_max_boundary_values[dim];
	auto& conds = min ? m.dirichlet_min_boundary_conditions[dim] : m.dirichlet_max_boundary_conditions[dim];
	if (!values)
	{
		values = std::make_unique<real_t[]>(m.substrates_count);
		conds = std::make_unique<bool[]>(m.substrates_count);
		for (index_t s = 0; s < m.substrates_count; s++)
		{
			values[s] = 42;
			conds[s] = false;
		}
	}
	values[0] = value;
	conds[0] = true;
}
} // namespace

TEST_GPU(DirichletSolverTest, Interior1D)
{
	auto m = still_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 100, 0, 0 }, { 20, 0, 0 }), 5, { 1, 1 });
	add_dirichlet_at(*m, { { 2, 0, 0 } }, { 41 });
	m->run_single_timestep();
	EXPECT_EQ(dens(*m, 0, 2), 41.0);
	EXPECT_EQ(dens(*m, 1, 2), 1.0);
}

TEST_GPU(DirichletSolverTest, Interior2D)
{
	auto m = still_microenv(cartesian_mesh(2, { 0, 0, 0 }, { 60, 60, 0 }, { 20, 20, 0 }), 5, { 1, 1 });
	add_dirichlet_at(*m, { { 1, 1, 0 }, { 2, 0, 0 } }, { 10, 11 });
	m->run_single_timestep();
	EXPECT_EQ(dens(*m, 0, 1, 1), 10.0);
	EXPECT_EQ(dens(*m, 0, 2, 0), 11.0);
	for (index_t y = 0; y < 3; y++)
		for (index_t x = 0; x < 3; x++)
			EXPECT_EQ(dens(*m, 1, x, y), 1.0);
}

TEST_GPU(DirichletSolverTest, Interior3D)
{
	auto m = still_microenv(cartesian_mesh(3, { 0, 0, 0 }, { 60, 60, 60 }, { 20, 20, 20 }), 5, { 1, 1 });
	add_dirichlet_at(*m, { { 1, 1, 1 }, { 1, 0, 2 }, { 0, 1, 2 } }, { 1000, 1001, 1002 });
	m->run_single_timestep();
	EXPECT_EQ(dens(*m, 0, 1, 1, 1), 1000.0);
	EXPECT_EQ(dens(*m, 0, 1, 0, 2), 1001.0);
	EXPECT_EQ(dens(*m, 0, 0, 1, 2), 1002.0);
}

namespace {
void boundary_case(index_t dims)
{
	auto m = still_microenv(cartesian_mesh(dims, { 0, 0, 0 }, { 100, 100, 100 }, { 20, 20, 20 }), 5, { 1, 1 });
	const real_t lo[3] = { 4, 6, 8 }, hi[3] = { 5, 7, 9 };
	for (index_t a = 0; a < dims; a++)
	{
		add_boundary_dirichlet(*m, a, true, lo[a]);
		add_boundary_dirichlet(*m, a, false, hi[a]);
	}
	m->run_single_timestep();
	const auto g = m->mesh.grid_shape;
	for (index_t z = 0; z < g[2]; z++)
		for (index_t y = 0; y < g[1]; y++)
			for (index_t x = 0; x < g[0]; x++)
			{
				// later faces win: x-min, x-max, y-min, y-max, z-min, z-max (test_dirichlet_solver.cpp:233-353)
				real_t want = 1;
				const index_t idx[3] = { x, y, z };
				for (index_t a = 0; a < dims; a++)
				{
					if (idx[a] == 0)
						want = lo[a];
					if (idx[a] == g[a] - 1)
						want = hi[a];
				}
				EXPECT_EQ(dens(*m, 0, x, y, z), want);
				EXPECT_EQ(dens(*m, 1, x, y, z), 1.0);
			}
}
} // namespace

TEST_GPU(DirichletSolverTest, Boundary1D) { boundary_case(1); }
TEST_GPU(DirichletSolverTest, Boundary2D) { boundary_case(2); }
TEST_GPU(DirichletSolverTest, Boundary3D) { boundary_case(3); }

// test_dirichlet_solver.cpp:355-505, through the builder and the registry
TEST_GPU(DirichletSolverTest, DynamicBoundaryModification)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 100.0, 0.01, 38.0 });
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 100, 100, 100 };
	sp.voxel = { 10, 10, 10 };
	sp.faces.push_back({ 0, { 160.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0 }, { true, false, false },
											  { false, false, false } });
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	env->solver->initialize(*env);
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	env->update_dirichlet_boundary_min('x', 0, 20.0, true);
	env->update_dirichlet_conditions();
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	EXPECT_EQ(env->get_substrate_density(0, 0, 1, 1), 20.0);
}

TEST_GPU(DirichletSolverTest, DynamicInteriorVoxelModification)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 100.0, 0.01, 38.0 });
	sp.substrates.push_back({ "Glucose", "mM", 50.0, 0.02, 5.0 });
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 100, 100, 100 };
	sp.voxel = { 10, 10, 10 };
	sp.interior.push_back({ { 5, 5, 5 }, { 160.0, 50.0 }, { true, true } });
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	env->solver->initialize(*env);
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 1, 0.0, true);
	env->update_dirichlet_conditions();
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	EXPECT_EQ(env->get_substrate_density(1, 5, 5, 5), 0.0);
}

TEST_GPU(DirichletSolverTest, DisableDirichletConditionsDynamically)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 100.0, 0.01, 38.0 });
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 100, 100, 100 };
	sp.voxel = { 10, 10, 10 };
	sp.faces.push_back({ 0, { 160.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0 }, { true, false, false },
											  { false, false, false } });
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	env->solver->initialize(*env);
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	env->update_dirichlet_boundary_min('x', 0, 160.0, false);
	env->update_dirichlet_conditions();
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	EXPECT_NE(env->get_substrate_density(1 - 1, 0, 1, 1), 160.0);
}

TEST_GPU(DirichletSolverTest, MultipleSubstrateModification)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 100.0, 0.01, 38.0 });
	sp.substrates.push_back({ "Glucose", "mM", 50.0, 0.02, 5.0 });
	sp.substrates.push_back({ "Lactate", "mM", 30.0, 0.015, 2.0 });
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 100, 100, 100 };
	sp.voxel = { 10, 10, 10 };
	sp.interior.push_back({ { 5, 5, 5 }, { 160.0, 50.0, 10.0 }, { true, true, true } });
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	env->solver->initialize(*env);
	EXPECT_EQ(env->dirichlet_interior_voxels_count, (index_t)1);
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 0, 80.0, true);
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 1, 100.0, true);
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 2, 0.0, true);
	env->update_dirichlet_interior_voxel({ 2, 2, 2 }, 0, 200.0, true);
	env->update_dirichlet_interior_voxel({ 2, 2, 2 }, 1, 60.0, true);
	env->update_dirichlet_interior_voxel({ 2, 2, 2 }, 2, 15.0, true);
	env->update_dirichlet_conditions();
	for (int i = 0; i < 5; ++i)
		env->run_single_timestep();
	EXPECT_EQ(env->get_substrate_density(0, 5, 5, 5), 80.0);
	EXPECT_EQ(env->get_substrate_density(1, 5, 5, 5), 100.0);
	EXPECT_EQ(env->get_substrate_density(2, 5, 5, 5), 0.0);
	EXPECT_EQ(env->get_substrate_density(0, 2, 2, 2), 200.0);
	EXPECT_EQ(env->get_substrate_density(1, 2, 2, 2), 60.0);
	EXPECT_EQ(env->get_substrate_density(2, 2, 2, 2), 15.0);
}

// ---------------------------------------------------------------- bulk

namespace {
// test_bulk_solver.cpp:39-61
struct test_functor : bulk_functor
{
	real_t supply_rates(index_t s, index_t x, index_t y, index_t z) override
	{
		return (x == 1 && y == 1 && z == 1 && s == 0) ? 5 : 0;
	}
	real_t supply_target_densities(index_t s, index_t x, index_t y, index_t z) override
	{
		return (x == 1 && y == 1 && z == 1 && s == 0) ? 6 : 0;
	}
	real_t uptake_rates(index_t s, index_t x, index_t y, index_t z) override
	{
		return (x == 1 && y == 1 && z == 1 && s == 0) ? 7 : 0;
	}
};
} // namespace

TEST_GPU(BulkSolverTest, SimulateBulkSource)
{
	auto m = still_microenv(cartesian_mesh(3, { 0, 0, 0 }, { 100, 100, 100 }, { 20, 20, 20 }), 0.01, { 10, 1 });
	m->bulk_fnc = std::make_unique<test_functor>();
	m->run_single_timestep();
	for (index_t z = 0; z < 5; z++)
		for (index_t y = 0; y < 5; y++)
			for (index_t x = 0; x < 5; x++)
				for (index_t s = 0; s < 2; s++)
				{
					if (x == 1 && y == 1 && z == 1 && s == 0)
						EXPECT_NEAR(dens(*m, s, x, y, z), 9.19643, 1e-4);
					else if (s == 1)
						EXPECT_EQ(dens(*m, s, x, y, z), 1.0);
					else
						EXPECT_EQ(dens(*m, s, x, y, z), 10.0);
				}
}

// --------------------------------------------------------------- cells

namespace {
// test_cell_solver.cpp:16-40 with D = decay = 0 (the cell stage alone)
std::unique_ptr<microenvironment> cell_microenv(const cartesian_mesh& mesh, bool internalized)
{
	auto m = still_microenv(mesh, 0.01, { 1, 1 });
	m->compute_internalized_substrates = internalized;
	return m;
}

// test_cell_solver.cpp:104-126
void set_default_agent_values(agent_interface* a, index_t off, index_t volume, std::array<real_t, 3> position,
							  index_t dims)
{
	a->secretion_rates()[0] = (real_t)(off + 100);
	a->secretion_rates()[1] = 0;
	a->uptake_rates()[0] = (real_t)(off + 200);
	a->uptake_rates()[1] = 0;
	a->saturation_densities()[0] = (real_t)(off + 300);
	a->saturation_densities()[1] = 0;
	a->net_export_rates()[0] = 0;
	a->net_export_rates()[1] = (real_t)(off + 400);
	a->fraction_released_at_death()[0] = 1;
	a->fraction_released_at_death()[1] = 1;
	a->volume() = (real_t)volume;
	for (index_t i = 0; i < dims; ++i)
		a->position()[i] = position[i];
}

hip_solver& hs(microenvironment& m) { return dynamic_cast<hip_solver&>(*m.solver); }

void simple_case(index_t dims, bool internalized, bool recompute)
{
	const std::array<sindex_t, 3> maxs = dims == 1 ? std::array<sindex_t, 3> { 60, 20, 20 }
										 : dims == 2 ? std::array<sindex_t, 3> { 60, 60, 20 }
													 : std::array<sindex_t, 3> { 60, 60, 60 };
	auto m = cell_microenv(cartesian_mesh(dims, { 0, 0, 0 }, maxs, { 20, 20, 20 }), internalized);
	auto* a1 = m->agents->create();
	auto* a2 = m->agents->create();
	auto* a3 = m->agents->create();
	set_default_agent_values(a1, 0, 1000, { 10, 10, 10 }, dims);
	set_default_agent_values(a2, 400, 1000, { 30, 30, 30 }, dims);
	set_default_agent_values(a3, 800, 1000, { 50, 50, 50 }, dims);
	auto at = [&](index_t k, index_t s) { return dens(*m, s, k, dims > 1 ? k : 0, dims > 2 ? k : 0); };

	m->run_single_timestep();
	m->solver->transfer_to_host(*m);
	if (internalized)
	{
		EXPECT_FIELD(a1->internalized_substrates()[0], -216000.0, 1e-6);
		EXPECT_FIELD(a1->internalized_substrates()[1], -4, 1e-6);
		EXPECT_FIELD(a2->internalized_substrates()[0], -1469052.6, 0.1);
		EXPECT_FIELD(a2->internalized_substrates()[1], -8, 1e-6);
		EXPECT_FIELD(a3->internalized_substrates()[0], -2927703.8, 0.1);
		EXPECT_FIELD(a3->internalized_substrates()[1], -12, 1e-6);
	}
	EXPECT_FIELD(at(0, 0), 28, 1e-6);
	EXPECT_FIELD(at(0, 1), 1.0005, 1e-6);
	EXPECT_FIELD(at(1, 0), 184.63158, 1e-5);
	EXPECT_FIELD(at(1, 1), 1.001, 1e-6);
	EXPECT_FIELD(at(2, 0), 366.963, 1e-3);
	EXPECT_FIELD(at(2, 1), 1.0015, 1e-6);

	if (recompute)
		m->solver->recompute_positional_data(*m);
	m->run_single_timestep();
	m->solver->transfer_to_host(*m);
	if (internalized)
	{
		EXPECT_FIELD(a1->internalized_substrates()[0], -373091, 0.1);
		EXPECT_FIELD(a1->internalized_substrates()[1], -8, 1e-6);
		EXPECT_FIELD(a2->internalized_substrates()[0], -2087601.1, 0.1);
		EXPECT_FIELD(a2->internalized_substrates()[1], -16, 1e-6);
		EXPECT_FIELD(a3->internalized_substrates()[0], -3795171.5, 0.1);
		EXPECT_FIELD(a3->internalized_substrates()[1], -24, 1e-6);
	}
	EXPECT_FIELD(at(0, 0), 47.636364, 1e-4);
	EXPECT_FIELD(at(0, 1), 1.001, 1e-6);
	EXPECT_FIELD(at(1, 0), 261.95, 1e-2);
	EXPECT_FIELD(at(1, 1), 1.002, 1e-6);
	EXPECT_FIELD(at(2, 0), 475.39642, 1e-4);
	EXPECT_FIELD(at(2, 1), 1.003, 1e-6);

	hs(*m).release_internalized_substrates(*m, 0);
	m->solver->transfer_to_host(*m);
	if (internalized)
	{
		// exactly 1 in the reference's FP64; within FP32 rounding here
		EXPECT_NEAR(at(0, 0), 1.0, 1e-6);
		EXPECT_NEAR(at(0, 1), 1.0, 1e-6);
		EXPECT_EQ(a1->internalized_substrates()[0], 0.0);
		EXPECT_EQ(a1->internalized_substrates()[1], 0.0);
	}
}

// brute-force expectations of test_cell_solver.cpp:42-102 (1-D meshes)
std::vector<real_t> expected_densities_1d(microenvironment& m)
{
	auto& data = agent_arrays_of(m);
	const index_t S = m.substrates_count, nx = m.mesh.grid_shape[0];
	const real_t V = (real_t)m.mesh.voxel_volume(), dt = m.diffusion_timestep;
	std::vector<real_t> out(nx * S);
	for (index_t s = 0; s < S; s++)
		for (index_t x = 0; x < nx; x++)
		{
			real_t num = 0, den = 0, fac = 0;
			for (index_t i = 0; i < data.agents_count; i++)
				if (m.mesh.voxel_position(std::span<const real_t>(&data.base_data.positions[i], 1))[0] == x)
				{
					num += data.secretion_rates[i * S + s] * data.saturation_densities[i * S + s] * dt *
						   data.volumes[i] / V;
					den += (data.secretion_rates[i * S + s] + data.uptake_rates[i * S + s]) * dt * data.volumes[i] / V;
					fac += data.net_export_rates[i * S + s] * dt / V;
				}
			out[x * S + s] = (dens(m, s, x) + num) / (1 + den) + fac;
		}
	return out;
}

void expected_internalized_1d(microenvironment& m, std::vector<real_t>& acc)
{
	auto& data = agent_arrays_of(m);
	const index_t S = m.substrates_count;
	const real_t V = (real_t)m.mesh.voxel_volume(), dt = m.diffusion_timestep;
	for (index_t i = 0; i < data.agents_count; i++)
		for (index_t s = 0; s < S; s++)
		{
			const real_t num = data.secretion_rates[i * S + s] * data.saturation_densities[i * S + s] * dt *
							   data.volumes[i];
			const real_t den =
				(data.secretion_rates[i * S + s] + data.uptake_rates[i * S + s]) * dt * data.volumes[i] / V;
			const real_t fac = data.net_export_rates[i * S + s] * dt;
			const index_t x = m.mesh.voxel_position(std::span<const real_t>(&data.base_data.positions[i], 1))[0];
			acc[i * S + s] -= (V * -den * dens(m, s, x) + num) + fac;
		}
}

void conflict_case(microenvironment& m, bool internalized, bool recompute)
{
	auto& data = agent_arrays_of(m);
	const index_t S = m.substrates_count, n = data.agents_count;
	std::vector<real_t> exp_int(n * S, 0);
	m.solver->initialize(m); // the expectations read the initial densities
	real_t peak = 1;
	const int steps = 2;
	for (int step = 0; step < steps; step++)
	{
		auto want = expected_densities_1d(m);
		if (step == 1 && recompute)
			m.solver->recompute_positional_data(m);
		m.run_single_timestep();
		m.solver->transfer_to_host(m);
		expected_internalized_1d(m, exp_int);
		// The expectation reads the densities as stored (FP32): each carries a
		// relative rounding of 2^-24, times V*den in the formula, so the bound
		// is relative (1e-6) where the reference's FP64 test uses 1e-6 absolute.
		if (internalized)
			for (index_t k = 0; k < n * S; k++)
				EXPECT_NEAR(data.internalized_substrates[k], exp_int[k], std::max(1e-6, 1e-6 * std::fabs(exp_int[k])));
		for (index_t x = 0; x < m.mesh.grid_shape[0]; x++)
			for (index_t s = 0; s < S; s++)
			{
				EXPECT_FIELD(dens(m, s, x), want[x * S + s], 1e-6);
				peak = std::max(peak, std::fabs(want[x * S + s]));
			}
	}
	for (index_t i = 0; i < n; i++)
		hs(m).release_internalized_substrates(m, i);
	m.solver->transfer_to_host(m);
	// Back to the initial density up to the FP32 storage of the intermediate
	// densities: one stochastic rounding (below one ulp, <= 2^-23 * peak) per
	// step (the internalized amounts are exact FP64 differences of the
	// unrounded densities).
	if (internalized)
		for (index_t x = 0; x < m.mesh.grid_shape[0]; x++)
			for (index_t s = 0; s < S; s++)
				EXPECT_NEAR(dens(m, s, x), 1.0, 1e-6 + steps * std::ldexp(peak, -23));
}
} // namespace

#define SIMPLE_CASES(D)                                                                       \
	TEST_GPU(CellSolverTest, Simple##D##D_internalized_recompute) { simple_case(D, true, true); } \
	TEST_GPU(CellSolverTest, Simple##D##D_internalized) { simple_case(D, true, false); }          \
	TEST_GPU(CellSolverTest, Simple##D##D_recompute) { simple_case(D, false, true); }             \
	TEST_GPU(CellSolverTest, Simple##D##D) { simple_case(D, false, false); }
SIMPLE_CASES(1)
SIMPLE_CASES(2)
SIMPLE_CASES(3)

namespace {
void conflict_small(bool internalized, bool recompute)
{
	auto m = cell_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 60, 20, 20 }, { 20, 20, 20 }), internalized);
	const index_t offs[6] = { 0, 600, 1100, 1600, 2100, 2600 }, vols[6] = { 500, 1000, 1500, 2000, 2500, 3000 };
	const real_t xs[6] = { 10, 30, 30, 50, 50, 50 };
	for (int i = 0; i < 6; i++)
		set_default_agent_values(m->agents->create(), offs[i], vols[i], { xs[i], 0, 0 }, 1);
	conflict_case(*m, internalized, recompute);
}
void conflict_big(bool internalized, bool recompute)
{
	cartesian_mesh mesh(1, { 0, 0, 0 }, { 2000, 20, 20 }, { 20, 20, 20 });
	auto m = cell_microenv(mesh, internalized);
	for (index_t i = 0; i < mesh.grid_shape[0]; i++)
		for (int j = 0; j < 50; j++)
			set_default_agent_values(m->agents->create(), 0, 500, mesh.voxel_center({ i, 0, 0 }), 1);
	conflict_case(*m, internalized, recompute);
}
} // namespace

TEST_GPU(CellSolverTest, Conflict_internalized_recompute) { conflict_small(true, true); }
TEST_GPU(CellSolverTest, Conflict_internalized) { conflict_small(true, false); }
TEST_GPU(CellSolverTest, Conflict_recompute) { conflict_small(false, true); }
TEST_GPU(CellSolverTest, Conflict) { conflict_small(false, false); }
TEST_GPU(CellSolverTest, ConflictBig_internalized_recompute) { conflict_big(true, true); }
TEST_GPU(CellSolverTest, ConflictBig_internalized) { conflict_big(true, false); }
TEST_GPU(CellSolverTest, ConflictBig) { conflict_big(false, false); }

TEST_GPU(CellSolverTest, InactiveAgentsSkipSecretion)
{
	auto m = cell_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 60, 20, 20 }, { 20, 20, 20 }), true);
	auto* a1 = m->agents->create();
	auto* a2 = m->agents->create();
	auto* a3 = m->agents->create();
	set_default_agent_values(a1, 0, 1000, { 10, 0, 0 }, 1);
	set_default_agent_values(a2, 400, 1000, { 30, 0, 0 }, 1);
	set_default_agent_values(a3, 800, 1000, { 50, 0, 0 }, 1);
	a2->is_active() = 0;
	m->run_single_timestep();
	m->solver->transfer_to_host(*m);
	EXPECT_FIELD(dens(*m, 0, 0), 28, 1e-6);
	EXPECT_FIELD(dens(*m, 1, 0), 1.0005, 1e-6);
	EXPECT_EQ(dens(*m, 0, 1), 1.0);
	EXPECT_EQ(dens(*m, 1, 1), 1.0);
	EXPECT_FIELD(dens(*m, 0, 2), 366.963, 1e-3);
	EXPECT_FIELD(dens(*m, 1, 2), 1.0015, 1e-6);
	EXPECT_FIELD(a1->internalized_substrates()[0], -216000.0, 1e-6);
	EXPECT_EQ(a2->internalized_substrates()[0], 0.0);
	EXPECT_EQ(a2->internalized_substrates()[1], 0.0);
	EXPECT_FIELD(a3->internalized_substrates()[0], -2927703.8, 0.1);
}

TEST_GPU(CellSolverTest, AllInactiveAgentsNoSecretion)
{
	auto m = cell_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 60, 20, 20 }, { 20, 20, 20 }), true);
	auto* a1 = m->agents->create();
	auto* a2 = m->agents->create();
	set_default_agent_values(a1, 0, 1000, { 10, 0, 0 }, 1);
	set_default_agent_values(a2, 400, 1000, { 30, 0, 0 }, 1);
	a1->is_active() = 0;
	a2->is_active() = 0;
	m->run_single_timestep();
	m->solver->transfer_to_host(*m);
	for (index_t x = 0; x < 3; x++)
		for (index_t s = 0; s < 2; s++)
			EXPECT_EQ(dens(*m, s, x), 1.0);
	EXPECT_EQ(a1->internalized_substrates()[0], 0.0);
	EXPECT_EQ(a2->internalized_substrates()[1], 0.0);
}

TEST_GPU(CellSolverTest, InactiveAgentReactivation)
{
	auto m = cell_microenv(cartesian_mesh(1, { 0, 0, 0 }, { 40, 20, 20 }, { 20, 20, 20 }), true);
	auto* a = m->agents->create();
	set_default_agent_values(a, 0, 1000, { 10, 0, 0 }, 1);
	a->is_active() = 1;
	m->run_single_timestep();
	EXPECT_FIELD(dens(*m, 0, 0), 28, 1e-6);
	EXPECT_FIELD(dens(*m, 1, 0), 1.0005, 1e-6);
	a->is_active() = 0;
	m->solver->recompute_positional_data(*m);
	m->run_single_timestep();
	EXPECT_FIELD(dens(*m, 0, 0), 28, 1e-6);
	EXPECT_FIELD(dens(*m, 1, 0), 1.0005, 1e-6);
	a->is_active() = 1;
	m->solver->recompute_positional_data(*m);
	m->run_single_timestep();
	EXPECT_FIELD(dens(*m, 0, 0), 47.636364, 1e-4);
	EXPECT_FIELD(dens(*m, 1, 0), 1.001, 1e-6);
}

// -------------------------------------------------- registry and host API

namespace {
class mock_solver : public solver
{
public:
	void initialize(microenvironment&) override {}
	void solve(microenvironment&, index_t) override {}
	real_t get_substrate_density(index_t, index_t, index_t, index_t) const override { return 0; }
	real_t& get_substrate_density(index_t, index_t, index_t, index_t) override
	{
		static real_t dummy = 0;
		return dummy;
	}
	void reinitialize_dirichlet(microenvironment&) override {}
	void recompute_positional_data(microenvironment&) override {}
};
} // namespace

// test_solver_registry.cpp:27-72
TEST_HOST(SolverRegistryTest, CheckPresentSolvers)
{
	EXPECT_NE(solver_registry::instance().get("hip_solver"), nullptr);
}

TEST_HOST(SolverRegistryTest, GetAndSet)
{
	solver_registry registry;
	EXPECT_EQ(registry.register_factory("solver_x", []() { return std::make_unique<mock_solver>(); }), true);
	EXPECT_NE(registry.get("solver_x"), nullptr);
	EXPECT_EQ(registry.register_factory("solver_x", []() { return std::make_unique<mock_solver>(); }), false);
	EXPECT_EQ(registry.get("solver_y"), nullptr);
}

TEST_HOST(SolverRegistryTest, RegistryAdder)
{
	const registry_adder<mock_solver> adder("solver_x");
	EXPECT_NE(solver_registry::instance().get("solver_x"), nullptr);
}

// test_dirichlet_modification.cpp: the host-side Dirichlet API
TEST_HOST(DirichletModification, SetInteriorVoxel)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 1.0, 0.01, 20.0 });
	sp.substrates.push_back({ "Glucose", "mM", 0.5, 0.02, 5.0 });
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 10, 10, 10 };
	sp.voxel = { 1, 1, 1 };
	sp.interior.push_back({ { 5, 5, 5 }, { 100.0, 50.0 }, { true, true } });
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	EXPECT_EQ(env->dirichlet_interior_voxels_count, (index_t)1);
	EXPECT_EQ(env->dirichlet_interior_values[0], 100.0);
	EXPECT_EQ(env->dirichlet_interior_values[1], 50.0);
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 0, 200.0, true);
	EXPECT_EQ(env->dirichlet_interior_values[0], 200.0);
	env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 1, 75.0, false);
	EXPECT_EQ(env->dirichlet_interior_values[1], 75.0);
	EXPECT_FALSE(env->dirichlet_interior_conditions[1]);
	EXPECT_THROW(env->update_dirichlet_interior_voxel({ 5, 5, 5 }, 2, 1.0, true), std::runtime_error);
	env->update_dirichlet_interior_voxel({ 1, 2, 3 }, 1, 9.0, true);
	EXPECT_EQ(env->dirichlet_interior_voxels_count, (index_t)2);
	EXPECT_EQ(env->dirichlet_interior_values[2], 0.0);
	EXPECT_FALSE(env->dirichlet_interior_conditions[2]);
	EXPECT_EQ(env->dirichlet_interior_values[3], 9.0);
}

TEST_HOST(DirichletModification, SetBoundary)
{
	env_spec sp;
	sp.substrates.push_back({ "O2", "mmHg", 1.0, 0.01, 20.0 });
	sp.substrates.push_back({ "Glucose", "mM", 0.5, 0.02, 5.0 });
	sp.dims = 2;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 10, 10, 0 };
	sp.voxel = { 1, 1, 1 };
	sp.solver = "hip_solver";
	auto env = make_env(sp);
	EXPECT_TRUE(env->dirichlet_min_boundary_values[0] == nullptr);
	env->update_dirichlet_boundary_min('x', 1, 5.0, true);
	EXPECT_EQ(env->dirichlet_min_boundary_values[0][1], 5.0);
	EXPECT_TRUE(env->dirichlet_min_boundary_conditions[0][1]);
	EXPECT_FALSE(env->dirichlet_min_boundary_conditions[0][0]);
	env->update_dirichlet_boundary_max('y', 0, 7.0, true);
	EXPECT_EQ(env->dirichlet_max_boundary_values[1][0], 7.0);
	EXPECT_THROW(env->update_dirichlet_boundary_min('z', 0, 1.0, true), std::runtime_error);
	EXPECT_THROW(env->update_dirichlet_boundary_max('x', 2, 1.0, true), std::runtime_error);
}

TEST_HOST(MeshTest, ShapeAndVoxelPosition)
{
	cartesian_mesh mesh(3, { -10, 0, 0 }, { 95, 40, 60 }, { 20, 20, 20 });
	EXPECT_EQ(mesh.grid_shape[0], (index_t)6);
	EXPECT_EQ(mesh.grid_shape[1], (index_t)2);
	EXPECT_EQ(mesh.grid_shape[2], (index_t)3);
	EXPECT_EQ(mesh.voxel_count(), (std::size_t)36);
	EXPECT_EQ(mesh.voxel_volume(), (index_t)8000);
	const real_t p[3] = { 9.999, 39.0, 0.0 };
	auto v = mesh.voxel_position(std::span<const real_t>(p, 3));
	EXPECT_EQ(v[0], (index_t)0);
	EXPECT_EQ(v[1], (index_t)1);
	EXPECT_EQ(v[2], (index_t)0);
	EXPECT_EQ(mesh.linearize(1, 1, 2), (std::size_t)(1 + 6 + 24));
}

TEST_HOST(AgentContainerTest, SwapRemove)
{
	microenvironment m(cartesian_mesh(3, { 0, 0, 0 }, { 60, 60, 60 }, { 20, 20, 20 }), 2, 0.01);
	auto* a = m.agents->create();
	auto* b = m.agents->create();
	auto* c = m.agents->create();
	a->volume() = 1;
	b->volume() = 2;
	c->volume() = 3;
	c->secretion_rates()[1] = 7;
	c->position()[2] = 5;
	m.agents->remove_at(0);
	EXPECT_EQ(m.agents->size(), (std::size_t)2);
	auto* moved = m.agents->get_agent_at(0);
	EXPECT_EQ(moved->volume(), 3.0);
	EXPECT_EQ(moved->secretion_rates()[1], 7.0);
	EXPECT_EQ(moved->position()[2], 5.0);
	EXPECT_EQ(m.agents->get_agent_at(1)->volume(), 2.0);
}

// ---------------------------------------------------------------- VTK output

namespace {
struct vti_file
{
	bool ok = false;
	std::array<long long, 6> extent {};
	std::array<double, 3> spacing {};
	std::vector<std::string> names, types;
	std::map<std::string, std::vector<double>> arrays;
};

std::string vti_attr(const std::string& tag, const std::string& key)
{
	const auto k = tag.find(" " + key + "=\"");
	if (k == std::string::npos)
		return {};
	const auto b = k + key.size() + 3;
	return tag.substr(b, tag.find('"', b) - b);
}

// ImageData with raw appended Float64 cell arrays and UInt64 block headers
vti_file read_vti(const std::filesystem::path& path)
{
	vti_file r;
	std::ifstream in(path, std::ios::binary);
	const std::string s((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	const auto img = s.find("<ImageData ");
	const auto app = s.find("<AppendedData encoding=\"raw\">");
	if (img == std::string::npos || app == std::string::npos || s.find("header_type=\"UInt64\"") == std::string::npos)
		return r;
	const std::string itag = s.substr(img, s.find('>', img) - img);
	std::istringstream ex(vti_attr(itag, "WholeExtent"));
	for (auto& e : r.extent)
		ex >> e;
	std::istringstream sp(vti_attr(itag, "Spacing"));
	for (auto& e : r.spacing)
		sp >> e;
	const auto base = s.find('_', app) + 1;
	for (auto pos = s.find("<DataArray ", img); pos != std::string::npos && pos < app;
		 pos = s.find("<DataArray ", pos + 1))
	{
		const std::string tag = s.substr(pos, s.find('>', pos) - pos);
		const std::string name = vti_attr(tag, "Name");
		r.names.push_back(name);
		r.types.push_back(vti_attr(tag, "type"));
		const std::size_t off = base + std::stoull(vti_attr(tag, "offset"));
		std::uint64_t nbytes = 0;
		if (off + 8 > s.size())
			return r;
		std::memcpy(&nbytes, s.data() + off, 8);
		if (off + 8 + nbytes > s.size() || nbytes % 8)
			return r;
		std::vector<double> v(nbytes / 8);
		std::memcpy(v.data(), s.data() + off + 8, nbytes);
		r.arrays[name] = std::move(v);
	}
	r.ok = true;
	return r;
}

std::filesystem::path fresh_dir(const char* name)
{
	auto d = std::filesystem::temp_directory_path() / name;
	std::filesystem::remove_all(d);
	std::filesystem::create_directories(d);
	return d;
}

std::string slurp(const std::filesystem::path& p)
{
	std::ifstream f(p);
	return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// create_test_microenvironment, test_vtk_serializer.cpp:44-70
std::unique_ptr<microenvironment> vtk_env(index_t dims = 3, std::array<index_t, 3> grid = { 3, 3, 3 },
										  std::array<index_t, 3> voxel = { 20, 20, 20 })
{
	env_spec sp;
	sp.name = "test_env";
	sp.time_units = "min";
	sp.space_units = "um";
	sp.dt = 0.01;
	sp.dims = dims;
	sp.lo = { 0, 0, 0 };
	sp.hi = { (sindex_t)(grid[0] * voxel[0]), (sindex_t)(grid[1] * voxel[1]), (sindex_t)(grid[2] * voxel[2]) };
	sp.voxel = voxel;
	sp.substrates.push_back({ "O2", "mmHg", 1.0, 0.01, 38.0 });
	sp.substrates.push_back({ "Glucose", "mM", 0.5, 0.02, 5.5 });
	sp.solver = "hip_solver";
	auto m = make_env(sp);
	m->solver->solve(*m, 1);
	return m;
}

// a solver whose densities are a function of (s, x, y, z): the synchronous read path
class pattern_solver : public solver
{
public:
	void initialize(microenvironment&) override {}
	void solve(microenvironment&, index_t) override {}
	real_t get_substrate_density(index_t s, index_t x, index_t y, index_t z) const override
	{
		return 1000.0 * s + 100.0 * z + 10.0 * y + x + 0.25;
	}
	real_t& get_substrate_density(index_t, index_t, index_t, index_t) override
	{
		static real_t dummy = 0;
		return dummy;
	}
	void reinitialize_dirichlet(microenvironment&) override {}
	void recompute_positional_data(microenvironment&) override {}
};
} // namespace

TEST_HOST(VtkSerializerTest, ReadPathOfOtherSolvers)
{
	const auto dir = fresh_dir("bfvm_vtk_host");
	microenvironment m(cartesian_mesh(3, { 40, 0, -20 }, { 120, 60, 40 }, { 20, 20, 20 }), 2, 0.01);
	m.substrates_names = { "A", "B" };
	m.solver = std::make_unique<pattern_solver>();
	{
		vtk_serializer ser(dir.string(), m);
		EXPECT_TRUE(std::filesystem::exists(dir / "vtk_microenvironment"));
		ser.serialize(m, 0.5);
		ser.serialize(m, 1.5);
	}
	const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000001.vti");
	EXPECT_TRUE(f.ok);
	const std::array<long long, 6> want_ext = { 2, 6, 0, 3, -1, 2 };
	for (int k = 0; k < 6; k++)
		EXPECT_EQ(f.extent[k], want_ext[k]);
	EXPECT_EQ(f.spacing[2], 20.0);
	EXPECT_EQ(f.names.size(), (std::size_t)2);
	EXPECT_EQ(f.types[0], std::string("Float64"));
	const auto& a = f.arrays.at("B");
	EXPECT_EQ(a.size(), m.mesh.voxel_count());
	for (index_t z = 0; z < 3; z++)
		for (index_t y = 0; y < 3; y++)
			for (index_t x = 0; x < 4; x++)
				EXPECT_EQ(a[m.mesh.linearize(x, y, z)], 1000.0 + 100.0 * z + 10.0 * y + x + 0.25);
	// vtk_serializer_base.cpp:22-37
	const std::string want_pvd = "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" "
								 "byte_order=\"LittleEndian\">\n  <Collection>\n"
								 "    <DataSet timestep=\"0.500000\" group=\"\" part=\"0\" "
								 "file=\"vtk_microenvironment/microenvironment_000000.vti\" />\n"
								 "    <DataSet timestep=\"1.500000\" group=\"\" part=\"0\" "
								 "file=\"vtk_microenvironment/microenvironment_000001.vti\" />\n"
								 "  </Collection>\n</VTKFile>";
	EXPECT_TRUE(slurp(dir / "microenvironment.pvd") == want_pvd);
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:74-160
TEST_GPU(VtkSerializerTest, ConstructorFilesAndPvd)
{
	const auto dir = fresh_dir("bfvm_vtk_files");
	auto m = vtk_env();
	vtk_serializer ser(dir.string(), *m);
	EXPECT_TRUE(std::filesystem::exists(dir / "vtk_microenvironment"));
	ser.serialize(*m, 0.1);
	ser.serialize(*m, 0.2);
	ser.serialize(*m, 0.0);
	ser.flush();
	for (const char* n : { "microenvironment_000000.vti", "microenvironment_000001.vti", "microenvironment_000002.vti" })
		EXPECT_TRUE(std::filesystem::exists(dir / "vtk_microenvironment" / n));
	const std::string pvd = slurp(dir / "microenvironment.pvd");
	for (const char* frag : { "<?xml version=\"1.0\"?>", "<VTKFile type=\"Collection\"", "<Collection>", "</Collection>",
							  "</VTKFile>", "timestep=\"0.1", "timestep=\"0.2", "microenvironment_000000.vti",
							  "microenvironment_000001.vti" })
		EXPECT_TRUE(pvd.find(frag) != std::string::npos);
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:162-205, :322-345
TEST_GPU(VtkSerializerTest, VtkFileStructure)
{
	const auto dir = fresh_dir("bfvm_vtk_structure");
	auto m = vtk_env();
	vtk_serializer ser(dir.string(), *m);
	ser.serialize(*m, 0.0);
	ser.flush();
	const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
	EXPECT_TRUE(f.ok);
	for (int a = 0; a < 3; a++)
	{
		EXPECT_EQ(f.extent[2 * a + 1] - f.extent[2 * a] + 1, 4LL); // points: grid_shape + 1
		EXPECT_EQ(f.spacing[a], 20.0);
	}
	EXPECT_EQ(f.arrays.count("O2"), (std::size_t)1);
	EXPECT_EQ(f.arrays.count("Glucose"), (std::size_t)1);
	EXPECT_EQ(f.arrays.at("O2").size(), (std::size_t)27);
	EXPECT_EQ(f.arrays.at("Glucose").size(), (std::size_t)27);
	for (const auto& t : f.types)
		EXPECT_EQ(t, std::string("Float64")); // real_t = double
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:207-227
TEST_GPU(VtkSerializerTest, HandleDifferentMeshDimensions)
{
	const auto dir = fresh_dir("bfvm_vtk_dims");
	{
		auto m1 = vtk_env(1, { 5, 1, 1 }, { 10, 1, 1 });
		vtk_serializer ser(dir.string(), *m1);
		ser.serialize(*m1, 0.0);
		ser.flush();
		const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
		EXPECT_TRUE(f.ok);
		EXPECT_EQ(f.extent[1], 5LL);
		EXPECT_EQ(f.extent[5], 0LL);
		EXPECT_EQ(f.spacing[2], 0.0);
		EXPECT_EQ(f.arrays.at("O2").size(), (std::size_t)5);
	}
	{
		auto m2 = vtk_env(2, { 4, 4, 1 }, { 15, 15, 1 });
		vtk_serializer ser(dir.string(), *m2);
		ser.serialize(*m2, 0.0);
		ser.flush();
		const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
		EXPECT_TRUE(f.ok);
		EXPECT_EQ(f.extent[3], 4LL);
		EXPECT_EQ(f.spacing[0], 15.0);
		EXPECT_EQ(f.arrays.at("Glucose").size(), (std::size_t)16);
	}
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:229-320
TEST_GPU(VtkSerializerTest, SingleAndManySubstrates)
{
	const auto dir = fresh_dir("bfvm_vtk_many");
	const std::vector<std::string> names = { "O2", "Glucose", "Lactate", "ATP", "CO2", "H2O" };
	env_spec sp;
	sp.name = "many_substrates_env";
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 40, 40, 40 };
	sp.voxel = { 20, 20, 20 };
	for (std::size_t i = 0; i < names.size(); ++i)
	{
		sp.substrates.push_back({ names[i], "unit", 1.0, 0.01, static_cast<real_t>(10 + i) });
		sp.faces.push_back({ 
			i, { (real_t)(15 + i), (real_t)(14 + i), (real_t)(13 + i) },
			{ (real_t)(20 + i), (real_t)(19 + i), (real_t)(18 + i) }, { true, true, true }, { true, true, true } });
	}
	sp.solver = "hip_solver";
	auto m = make_env(sp);
	m->solver->initialize(*m);
	vtk_serializer ser(dir.string(), *m);
	ser.serialize(*m, 0.0);
	ser.flush();
	const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
	EXPECT_TRUE(f.ok);
	EXPECT_EQ(f.names.size(), names.size());
	for (std::size_t i = 0; i < names.size(); ++i)
	{
		EXPECT_EQ(f.arrays.count(names[i]), (std::size_t)1);
		EXPECT_EQ(f.arrays.at(names[i])[0], (double)(10 + i)); // initial condition, before any step
	}
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:322-357
TEST_GPU(VtkSerializerTest, NonZeroBoundingBoxMins)
{
	const auto dir = fresh_dir("bfvm_vtk_offset");
	env_spec sp;
	sp.name = "offset_env";
	sp.dims = 3;
	sp.lo = { 100, 200, 300 };
	sp.hi = { 160, 260, 360 };
	sp.voxel = { 20, 20, 20 };
	sp.substrates.push_back({ "O2", "mmHg", 1.0, 0.01, 35.0 });
	sp.faces.push_back({ 0, { 38.0, 36.0, 34.0 }, { 42.0, 40.0, 38.0 }, { true, true, true },
											  { true, true, true } });
	sp.solver = "hip_solver";
	auto m = make_env(sp);
	m->solver->initialize(*m);
	vtk_serializer ser(dir.string(), *m);
	ser.serialize(*m, 0.0);
	ser.flush();
	const auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
	const std::array<long long, 6> want = { 5, 8, 10, 13, 15, 18 };
	for (int k = 0; k < 6; k++)
		EXPECT_EQ(f.extent[k], want[k]);
	std::filesystem::remove_all(dir);
}

// test_vtk_serializer.cpp:384-476
TEST_GPU(VtkSerializerTest, BoundaryConditionsEffect)
{
	const auto dir = fresh_dir("bfvm_vtk_boundary");
	env_spec sp;
	sp.name = "boundary_test_env";
	sp.dims = 3;
	sp.lo = { 0, 0, 0 };
	sp.hi = { 60, 60, 60 };
	sp.voxel = { 20, 20, 20 };
	sp.substrates.push_back({ "O2", "mmHg", 0, 0, 20.0 });
	sp.substrates.push_back({ "Glucose", "mM", 0.5, 0.02, 5.5 });
	sp.faces.push_back({ 0, { 100.0, 20.0, 20.0 }, { 5.0, 20.0, 20.0 }, { true, false, false },
											  { true, false, false } });
	sp.solver = "hip_solver";
	auto m = make_env(sp);
	m->solver->solve(*m, 1);
	vtk_serializer ser(dir.string(), *m);
	ser.serialize(*m, 0.0);
	ser.flush();
	auto f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
	EXPECT_TRUE(f.ok);
	EXPECT_EQ(f.arrays.at("O2").size(), m->mesh.voxel_count());
	for (index_t z = 0; z < 3; ++z)
		for (index_t y = 0; y < 3; ++y)
			for (index_t x = 0; x < 3; ++x)
			{
				const std::size_t v = m->mesh.linearize(x, y, z);
				EXPECT_EQ(f.arrays.at("O2")[v], x == 0 ? 100.0 : (x == 2 ? 5.0 : 20.0));
				EXPECT_EQ(f.arrays.at("Glucose")[v], m->get_substrate_density(1, x, y, z));
			}
	m->solver->solve(*m, 1);
	ser.serialize(*m, 0.1);
	ser.flush();
	f = read_vti(dir / "vtk_microenvironment" / "microenvironment_000001.vti");
	EXPECT_TRUE(f.ok);
	for (index_t z = 0; z < 3; ++z)
		for (index_t y = 0; y < 3; ++y)
			for (index_t x = 0; x < 3; ++x)
				EXPECT_EQ(f.arrays.at("Glucose")[m->mesh.linearize(x, y, z)], m->get_substrate_density(1, x, y, z));
	std::filesystem::remove_all(dir);
}

// no counterpart in the reference: snapshots taken while stepping goes on hold
// the field of the step they were taken at, and writes through the real_t&
// accessor reach them
TEST_GPU(VtkSerializerTest, SnapshotsWhileStepping)
{
	const auto dir = fresh_dir("bfvm_vtk_async");
	auto make = [] {
		env_spec sp;
		sp.name = "async_env";
		sp.dims = 3;
		sp.lo = { 0, 0, 0 };
		sp.hi = { 800, 640, 480 };
		sp.voxel = { 20, 20, 20 };
		sp.substrates.push_back({ "A", "u", 500.0, 0.1, 1.0 });
		sp.substrates.push_back({ "B", "u", 20.0, 0.5, 2.0 });
		sp.faces.push_back({ 0, { 10.0, 0, 0 }, { 0, 0, 0 }, { true, false, false },
												  { false, false, false } });
		sp.solver = "hip_solver";
		return make_env(sp);
	};
	auto m = make();
	auto twin = make();
	vtk_serializer ser(dir.string(), *m);
	m->solver->solve(*m, 3);
	m->solver->get_substrate_density(1, 5, 6, 7) = 42.0; // host write before the snapshot
	ser.serialize(*m, 0.03);
	m->solver->solve(*m, 7); // queued behind the snapshot, runs during its host copy
	ser.serialize(*m, 0.10);
	m->solver->solve(*m, 5);
	ser.flush();
	twin->solver->solve(*twin, 3);
	twin->solver->get_substrate_density(1, 5, 6, 7) = 42.0;
	auto f0 = read_vti(dir / "vtk_microenvironment" / "microenvironment_000000.vti");
	EXPECT_TRUE(f0.ok);
	std::size_t mismatches = 0;
	const auto& g = twin->mesh.grid_shape;
	for (index_t z = 0; z < g[2]; ++z)
		for (index_t y = 0; y < g[1]; ++y)
			for (index_t x = 0; x < g[0]; ++x)
				for (index_t s = 0; s < 2; ++s)
					mismatches += f0.arrays.at(s ? "B" : "A")[twin->mesh.linearize(x, y, z)] !=
								  twin->get_substrate_density(s, x, y, z);
	EXPECT_EQ(mismatches, (std::size_t)0);
	twin->solver->solve(*twin, 7);
	auto f1 = read_vti(dir / "vtk_microenvironment" / "microenvironment_000001.vti");
	mismatches = 0;
	for (index_t z = 0; z < g[2]; ++z)
		for (index_t y = 0; y < g[1]; ++y)
			for (index_t x = 0; x < g[0]; ++x)
				mismatches += f1.arrays.at("A")[twin->mesh.linearize(x, y, z)] != twin->get_substrate_density(0, x, y, z);
	EXPECT_EQ(mismatches, (std::size_t)0);
	std::filesystem::remove_all(dir);
}

// ---------------------------------------------------------------- main

int main(int argc, char** argv)
{
	bool host_only = false;
	const char* filter = nullptr;
	for (int i = 1; i < argc; i++)
	{
		if (std::strcmp(argv[i], "--host-only") == 0)
			host_only = true;
		else
			filter = argv[i];
	}
	if (!host_only && bfvm_device_count() < 1)
	{
		std::printf("no HIP device: run with --host-only\n");
		return 2;
	}
	int ran = 0, failed = 0;
	for (auto& t : registry_of_tests())
	{
		if ((host_only && t.gpu) || (filter && !std::strstr(t.name, filter)))
			continue;
		const int before = failures;
		std::printf("[ RUN  ] %s\n", t.name);
		std::fflush(stdout);
		try
		{
			t.fn();
		}
		catch (const std::exception& e)
		{
			fail(__FILE__, __LINE__, std::string("exception: ") + e.what());
		}
		ran++;
		const bool ok = failures == before;
		failed += ok ? 0 : 1;
		std::printf("[ %s ] %s\n", ok ? " OK " : "FAIL", t.name);
		std::fflush(stdout);
	}
	std::printf("%d tests, %d failed, %d checks\n", ran, failed, checks);
	return failed ? 1 : 0;
}
