// Drives the CPU oracle's C ABI through every solver stage on small grids, for
// a build under AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitizers.py): 1-, 2- and 3-D meshes, Dirichlet faces and
// interior voxels, a bulk, cells with shared voxels, inactive cells and
// internalized tracking, releases, recomputes, the FP32 solver and an
// out-of-grid cell.  Exits non-zero if a result is not finite or a call
// reports an unexpected status; the sanitizers report memory and UB errors.
#include <cmath>
#include <cstdio>
#include <vector>

#include "bfvm_oracle.h"

namespace {

int failures = 0;

void check(bool ok, const char* what)
{
	if (!ok)
	{
		std::printf("FAILED: %s\n", what);
		failures++;
	}
}

struct Case
{
	int dims, S;
	int64_t n[3];
	bool quirks;
};

double lcg(uint64_t& s)
{
	s = s * 6364136223846793005ull + 1442695040888963407ull;
	return (double)(s >> 11) / 9007199254740992.0;
}

void run_case(const Case& k, uint64_t seed)
{
	const int S = k.S;
	const int64_t V = k.n[0] * k.n[1] * k.n[2];
	std::vector<double> D(S), lam(S);
	for (int s = 0; s < S; s++)
	{
		D[s] = 10 + 1000 * lcg(seed);
		lam[s] = 0.01 + lcg(seed);
	}
	orc_domain d {};
	d.dims = k.dims;
	d.substrates = S;
	d.nx = k.n[0];
	d.ny = k.n[1];
	d.nz = k.n[2];
	d.dx = d.dy = d.dz = 20;
	d.dt = 0.01;
	d.diffusion = D.data();
	d.decay = lam.data();
	d.openmp_quirks = k.quirks ? 1 : 0;

	for (int a = 0; a < k.dims; a++)
	{
		std::vector<double> b(k.n[a] * S), e((k.n[a] > 1 ? k.n[a] - 1 : 1) * S), c(S);
		check(orc_factors(&d, a, b.data(), e.data(), c.data()) == 0, "factors");
	}
	for (int64_t col = 0; col < k.n[0] * S; col++)
	{
		const int64_t s = orc_coef_substrate(&d, col);
		check(s >= 0 && s < S, "coefficient substrate");
	}

	std::vector<double> field(V * S);
	for (auto& x : field)
		x = lcg(seed);
	std::vector<double> f2 = field;
	check(orc_diffuse_f64(&d, f2.data(), 2) == 0, "diffuse");
	for (int a = 0; a < k.dims; a++)
		check(orc_sweep_f64(&d, f2.data(), a) == 0, "sweep");
	std::vector<float> f32(field.begin(), field.end());
	check(orc_diffuse_f32(&d, f32.data(), 2) == 0, "diffuse f32");

	// Dirichlet: every face of the mesh, alternate conditions; two interior voxels
	std::vector<double> fv(6 * S);
	std::vector<uint8_t> fc(6 * S);
	for (int i = 0; i < 6 * S; i++)
	{
		fv[i] = lcg(seed);
		fc[i] = (uint8_t)(i % 2);
	}
	orc_dirichlet dir {};
	for (int a = 0; a < k.dims; a++)
	{
		dir.min_values[a] = fv.data() + (2 * a) * S;
		dir.min_conds[a] = fc.data() + (2 * a) * S;
		dir.max_values[a] = fv.data() + (2 * a + 1) * S;
		dir.max_conds[a] = fc.data() + (2 * a + 1) * S;
	}
	std::vector<int64_t> iv;
	for (int q = 0; q < 2; q++)
		for (int a = 0; a < k.dims; a++)
			iv.push_back((k.n[a] - 1) * q / 2);
	std::vector<double> ivals(2 * S, 0.5);
	std::vector<uint8_t> iconds(2 * S, 1);
	dir.interior_count = 2;
	dir.interior_voxels = iv.data();
	dir.interior_values = ivals.data();
	dir.interior_conds = iconds.data();
	check(orc_dirichlet_f64(&d, f2.data(), &dir) == 0, "dirichlet");

	std::vector<double> sup(V * S), upt(V * S), tgt(V * S);
	for (int64_t i = 0; i < V * S; i++)
	{
		sup[i] = (i % 3) ? lcg(seed) : 0.0;
		upt[i] = lcg(seed);
		tgt[i] = 2 * lcg(seed);
	}
	orc_bulk bulk { sup.data(), upt.data(), tgt.data() };
	check(orc_bulk_f64(&d, f2.data(), &bulk) == 0, "bulk");

	// cells: some share voxels (pairs at the same position), one inactive
	const int64_t n = 2 * V / 3 + 3;
	std::vector<double> pos(n * k.dims), sec(n * S), sat(n * S), cup(n * S), nex(n * S), frac(n * S), vol(n);
	std::vector<uint8_t> act(n, 1);
	const int64_t ext[3] = { k.n[0] * 20, k.n[1] * 20, k.n[2] * 20 };
	for (int64_t i = 0; i < n; i++)
	{
		for (int a = 0; a < k.dims; a++)
			pos[i * k.dims + a] = (i % 2 && i > 0) ? pos[(i - 1) * k.dims + a] : lcg(seed) * ext[a] * 0.999;
		vol[i] = 100 + 1000 * lcg(seed);
		for (int s = 0; s < S; s++)
		{
			sec[i * S + s] = lcg(seed);
			sat[i * S + s] = 1;
			cup[i * S + s] = lcg(seed);
			nex[i * S + s] = 0.1 * lcg(seed);
			frac[i * S + s] = lcg(seed);
		}
	}
	act[n / 2] = 0;
	for (int tracking = 0; tracking < 2; tracking++)
	{
		std::vector<double> internalized(n * S, 1.0);
		orc_agents ag { n, pos.data(), sec.data(), sat.data(), cup.data(), nex.data(), internalized.data(),
						frac.data(), vol.data(), act.data() };
		std::vector<double> f = field;
		orc_solver* s = orc_solver_create(&d, f.data(), &dir, &bulk, &ag, tracking);
		check(s != nullptr, "solver");
		if (!s)
			continue;
		check(orc_solver_solve(s, 3) == 0, "solve");
		std::vector<double> num(n * S), den(n * S), fac(n * S);
		std::vector<int64_t> ballots(V);
		int32_t conflict = 0;
		check(orc_solver_reduced(s, num.data(), den.data(), fac.data(), ballots.data(), &conflict) == 0, "reduced");
		check(orc_solver_release(s, 0) == 0, "release");
		pos[0] = ext[0] * 0.5; // moves: the next solve after a recompute rebins it
		orc_solver_recompute(s);
		check(orc_solver_solve(s, 2) == 0, "solve after recompute");
		for (double x : f)
			check(std::isfinite(x), "finite field");
		for (double x : internalized)
			check(std::isfinite(x), "finite internalized");
		// an active cell outside the mesh is reported
		const double keep = pos[k.dims];
		pos[k.dims] = ext[0] * 2.0;
		orc_solver_recompute(s);
		check(orc_solver_solve(s, 1) == -1, "out-of-grid cell reported");
		pos[k.dims] = keep;
		orc_solver_destroy(s);
	}
	std::vector<float> g32(field.begin(), field.end());
	orc_solver* s32 = orc_solver_create_f32(&d, g32.data(), &dir, &bulk, nullptr, 0);
	check(s32 != nullptr, "f32 solver");
	if (s32)
	{
		check(orc_solver_solve(s32, 2) == 0, "f32 solve");
		orc_solver_destroy(s32);
	}
}

} // namespace

int main()
{
	const Case cases[] = {
		{ 1, 3, { 17, 1, 1 }, false }, { 2, 2, { 9, 7, 1 }, false }, { 3, 5, { 6, 5, 4 }, true },
		{ 3, 1, { 1, 8, 3 }, false },  { 3, 7, { 50, 3, 2 }, true },
	};
	uint64_t seed = 12345;
	for (const Case& k : cases)
		run_case(k, seed++);
	std::printf("oracle driver: %d failed\n", failures);
	return failures ? 1 : 0;
}
