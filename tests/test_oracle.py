"""Pins the CPU oracle: the reference's golden vectors, closed forms, a dense
direct solve of each implicit sweep, brute-force cell sums and the exact
per-sweep mass identity.  CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as orc
import scenarios as sc

G = sc.GOLDEN


# ------------------------------------------------------------------ diffusion

@pytest.mark.parametrize("name,dims", [("Uniform1D", 1), ("Uniform2D", 2), ("Uniform3D", 3)])
def test_uniform_golden_and_closed_form(name, dims):
    t = G["diffusion"][name]
    env = G["diffusion"][t["env"]]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    f = sc.constant_field(p, env["initial"])
    p.diffuse(f)
    for s, want in enumerate(t["expected_per_substrate"]):
        assert np.abs(f[..., s] - want).max() < t["tol"]
        # a uniform field only decays: rho / (1 + dt*lambda/dims)^dims
        closed = env["initial"][s] / (1 + env["dt"] * env["decay"][s] / dims) ** dims
        np.testing.assert_allclose(f[..., s], closed, rtol=1e-12)


@pytest.mark.parametrize("name,dims", [("Random1D", 1), ("Random2D", 2), ("Random3D", 3)])
def test_random_golden(name, dims):
    t = G["diffusion"][name]
    env = G["diffusion"][t["env"]]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    f = sc.linear_index_field(p)
    p.diffuse(f)
    np.testing.assert_allclose(f.ravel(), t["expected"], atol=t["tol"], rtol=0)


def tridiag(n, c, k):
    """Neumann LOD matrix of one axis: diagonal 1+k+2c (1+k+c at the ends), -c off it."""
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 1 + k + 2 * c - (c if i in (0, n - 1) else 0)
        if i > 0:
            A[i, i - 1] = -c
        if i < n - 1:
            A[i, i + 1] = -c
    if n == 1:
        A[0, 0] = 1 + k
    return A


def dense_sweep(p, field, axis):
    """Independent restatement: solve every line of one axis with a banded LU."""
    out = field.copy()
    n = (p.nx, p.ny, p.nz)[axis]
    ax = {0: 2, 1: 1, 2: 0}[axis]
    moved = np.moveaxis(out, ax, 0)
    shape = (p.voxel[0], p.voxel[1], p.voxel[2])[axis]
    for s in range(p.S):
        c = p.dt * p.diffusion[s] / shape ** 2
        k = p.dt * p.decay[s] / p.dims
        A = tridiag(n, c, k)
        ab = np.zeros((3, n))
        ab[1] = np.diag(A)
        if n > 1:
            ab[0, 1:] = np.diag(A, 1)
            ab[2, :-1] = np.diag(A, -1)
        rhs = moved[..., s].reshape(n, -1)
        moved[..., s] = sla.solve_banded((1, 1), ab, rhs).reshape(moved[..., s].shape)
    return out


@pytest.mark.parametrize("grid,S", [((5, 4, 3), 3), ((7, 1, 1), 2), ((1, 6, 4), 1), ((9, 8, 7), 4)])
def test_sweeps_match_banded_solve(grid, S):
    rng = np.random.default_rng(1)
    D = 10 ** rng.uniform(0, 5, S)
    lam = rng.uniform(0.01, 1, S)
    dims = 3
    p = orc.Problem(dims, grid, (20, 20, 20), 0.01, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    for axis in range(3):
        f = f0.copy()
        p.sweep(f, axis)
        np.testing.assert_allclose(f, dense_sweep(p, f0, axis), rtol=1e-11, atol=1e-13)
    # the full step is x, then y, then z
    want = dense_sweep(p, dense_sweep(p, dense_sweep(p, f0, 0), 1), 2)
    got = p.diffuse(f0.copy())
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-13)


def test_dense_lod_kronecker_tiny():
    """Brute force: the LOD step equals three Kronecker-structured dense solves."""
    rng = np.random.default_rng(2)
    nx, ny, nz, S = 3, 4, 2, 2
    D, lam = [300.0, 7.0], [0.2, 0.9]
    p = orc.Problem(3, (nx, ny, nz), (10, 20, 30), 0.05, D, lam)
    f0 = rng.uniform(0, 1, p.shape)
    got = p.diffuse(f0.copy())
    for s in range(S):
        k = p.dt * lam[s] / 3
        Ax = tridiag(nx, p.dt * D[s] / 100, k)
        Ay = tridiag(ny, p.dt * D[s] / 400, k)
        Az = tridiag(nz, p.dt * D[s] / 900, k)
        u = f0[..., s].reshape(-1)  # index (z*ny + y)*nx + x
        Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
        u = np.linalg.solve(np.kron(Iz, np.kron(Iy, Ax)), u)
        u = np.linalg.solve(np.kron(Iz, np.kron(Ay, Ix)), u)
        u = np.linalg.solve(np.kron(Az, np.kron(Iy, Ix)), u)
        np.testing.assert_allclose(got[..., s].reshape(-1), u, rtol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sweep_mass_identity(axis):
    """Columns of the Neumann matrix sum to 1+dt*lambda/dims, so each sweep divides
    every line's mass by exactly that."""
    rng = np.random.default_rng(3)
    S = 3
    p = orc.Problem(3, (11, 9, 13), (20, 20, 20), 0.01, [1e5, 30.0, 1.0], [0.5, 0.01, 1.0])
    f = rng.uniform(0, 1, p.shape)
    m0 = f.sum(axis=(0, 1, 2))
    p.sweep(f, axis)
    m1 = f.sum(axis=(0, 1, 2))
    np.testing.assert_allclose(m1, m0 / (1 + p.dt * p.decay / 3), rtol=1e-12)
    del S


def test_factors_formula_and_single_voxel_line():
    p = orc.Problem(3, (5, 1, 3), (20, 20, 20), 0.01, [1000.0, 5.0], [0.3, 0.7])
    b, e, c = p.factors(0)
    np.testing.assert_allclose(c, 0.01 * np.array([1000.0, 5.0]) / 400)
    for s in range(2):
        A = tridiag(5, c[s], 0.01 * p.decay[s] / 3)
        # b' is the reciprocal of the LU pivots; e_i = c*b'_i
        piv = np.zeros(5)
        piv[0] = A[0, 0]
        for i in range(1, 5):
            piv[i] = A[i, i] - c[s] ** 2 / piv[i - 1]
        np.testing.assert_allclose(b[:, s], 1 / piv, rtol=1e-14)
        np.testing.assert_allclose(e[:, s], c[s] / piv[:-1], rtol=1e-14)
    # single-voxel y line: Thrust-style 1/(1+dt*lambda/dims) by default ...
    by, _, _ = p.factors(1)
    np.testing.assert_allclose(by[0], 1 / (1 + 0.01 * p.decay / 3), rtol=1e-15)
    # ... and the OpenMP formula (one end subtraction) in quirk mode
    q = orc.Problem(3, (5, 1, 3), (20, 20, 20), 0.01, [1000.0, 5.0], [0.3, 0.7], openmp_quirks=True)
    qy, _, qc = q.factors(1)
    np.testing.assert_allclose(qy[0], 1 / (1 + 0.01 * q.decay / 3 + qc), rtol=1e-15)


def _split_static(length, size):
    """noarr::into_blocks_static(size) on a dimension of `length`: the body holds
    length // size full blocks of `size`, the border one block of the
    length % size trailing elements.  Returns (is_border, blocks, block_len)."""
    return [(0, length // size, size), (1, 1, length % size)]


def _walk_openmp_yz_columns(L, S, tile=48):
    """The loop nest of solve_slice_y_3d (openmp_solver/src/diffusion_solver.cpp:390-591;
    solve_slice_z_3d, :593-794, nests its loops the same way) over the merged x*S
    axis of length L: merge_blocks<'x','s','c'>, then into_blocks_static<'c','b','x','s'>
    (S * copies, copies = ceil(48/S), :105, :832), then into_blocks_static<'s','p','S','s'>(48)
    (:399-401).  Visits body blocks (b=0: x over blocks; p=0: S over tiles, s over
    the tile; p=1: s over the tile border), then the border block (b=1, x=0) the
    same way, and yields (column, s): `s` is the index every diagonal access of
    the loop body uses, diag_l | get_at<'i','c'>(b, i, s) (:427, :465, :515, :558),
    with c_[s] for the off-diagonal."""
    copies = -(-tile // S)
    B = S * copies
    nb = L // B
    for b, x_len, blk in _split_static(L, B):
        for x in range(x_len):
            for p, t_len, s_len in _split_static(blk, tile):
                for t in range(t_len):
                    for s in range(s_len):
                        r = (blk // tile) * tile + s if p else t * tile + s
                        yield (nb * B if b else x * B) + r, s


@pytest.mark.parametrize("S", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("nx", [1, 7, 10, 23, 48, 49, 50, 97])
def test_openmp_tile_quirk_map(S, nx):
    """The oracle's OpenMP-compatible coefficient map against a walk of the
    reference's blocked loop nest: every column of the row is visited once, and
    the diagonal entry it reads (entry k of the copies*S diagonal columns is
    substrate k % S: get_diagonal_layout_c, diffusion_solver.cpp:19-22, filled
    per copy at :34-83) is the substrate the oracle's map gives.  Rows cover one
    or several blocks, with and without a border block.  With the map off the
    oracle uses the column's own substrate."""
    L = nx * S
    q = orc.Problem(3, (nx, 2, 2), (20, 20, 20), 0.01, [1.0] * S, [0.1] * S, openmp_quirks=True)
    p = orc.Problem(3, (nx, 2, 2), (20, 20, 20), 0.01, [1.0] * S, [0.1] * S)
    seen = {}
    for col, s in _walk_openmp_yz_columns(L, S):
        assert col not in seen
        seen[col] = s % S
    assert sorted(seen) == list(range(L))
    for col in range(L):
        assert q.coef_substrate(col) == seen[col], (col, q.coef_substrate(col), seen[col])
        assert p.coef_substrate(col) == col % S
    wrong = [c for c in range(L) if seen[c] != c % S]
    if 48 % S == 0:
        assert wrong == []
    if S == 5 and nx >= 10:
        assert wrong[:2] == [48, 49]
    if S == 7 and nx >= 7:
        assert wrong[0] == 48


def test_openmp_tile_quirk_sweep():
    """In quirk mode column c is solved with the matrix of substrate coef(c)."""
    rng = np.random.default_rng(5)
    S = 5
    D = [1e4, 10.0, 300.0, 2.0, 5e4]
    lam = [0.1, 0.9, 0.5, 0.02, 0.3]
    q = orc.Problem(3, (12, 6, 5), (20, 20, 20), 0.01, D, lam, openmp_quirks=True)
    f0 = rng.uniform(0, 1, q.shape)
    for axis in (1, 2):
        f = f0.copy()
        q.sweep(f, axis)
        n = (q.ny, q.nz)[axis - 1]
        ax = {1: 1, 2: 0}[axis]
        src = np.moveaxis(f0, ax, 0).reshape(n, -1, q.nx * S)
        out = np.moveaxis(f, ax, 0).reshape(n, -1, q.nx * S)
        for col in range(q.nx * S):
            cs = q.coef_substrate(col)
            A = tridiag(n, q.dt * D[cs] / 400, q.dt * lam[cs] / 3)
            np.testing.assert_allclose(out[:, :, col], np.linalg.solve(A, src[:, :, col]), rtol=1e-11)


# ----------------------------------------------------------------------- bulk

def test_bulk_golden():
    t = G["bulk"]["SimulateBulkSource"]
    p = sc.problem_from_env(t, 3, t["bbox_max"], t["voxel"])
    f = sc.constant_field(p, t["initial"])
    sup, upt, tgt = (np.zeros(p.shape) for _ in range(3))
    x, y, z = t["point"]
    s0 = t["point_substrate"]
    sup[z, y, x, s0], tgt[z, y, x, s0], upt[z, y, x, s0] = t["supply"], t["target"], t["uptake"]
    orc.bulk(p, f, sup, upt, tgt)
    assert abs(f[z, y, x, s0] - t["expected_point"]) < t["tol_point"]
    mask = np.ones(p.shape, dtype=bool)
    mask[z, y, x, s0] = False
    ref = sc.constant_field(p, t["initial"])
    assert np.array_equal(f[mask], ref[mask])
    # closed form of the implicit update
    want = (t["initial"][0] + p.dt * 5 * 6) / (1 + p.dt * (7 + 5))
    assert f[z, y, x, s0] == pytest.approx(want, rel=1e-15)


# ---------------------------------------------------------------------- cells

def _cells_step(p, ag, internalized):
    f = sc.constant_field(p, G["cells"]["default_microenv"]["initial"])
    s = orc.Solver(p, f, agents=ag, compute_internalized=internalized)
    return f, s


def cells_only(p, f, ag, internalized):
    """A solver whose diffusion is the identity (D = lambda = 0) so a step is the
    cell stage alone, as the reference cell tests call it directly."""
    q = orc.Problem(p.dims, (p.nx, p.ny, p.nz), p.voxel, p.dt, [0.0] * p.S, [0.0] * p.S)
    return orc.Solver(q, f, agents=ag, compute_internalized=internalized)


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("internalized", [True, False])
@pytest.mark.parametrize("recompute", [True, False])
def test_cells_simple_golden(dims, internalized, recompute):
    g = G["cells"]["Simple"]
    p, ag = sc.simple_cells(dims)
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, internalized)
    s.solve(1)
    diag = [(0, 0, 0), (1, 1, 1), (2, 2, 2)]

    def at(k, sub):
        x, y, z = diag[k]
        return f[z if dims > 2 else 0, y if dims > 1 else 0, x, sub]

    for k in range(3):
        for sub in range(2):
            want, tol = g["step1_density"][2 * k + sub]
            assert abs(at(k, sub) - want) < tol
            if internalized:
                want, tol = g["step1_internalized"][2 * k + sub]
                assert abs(ag.internalized[k, sub] - want) < tol
    if recompute:
        s.recompute()
    s.solve(1)
    for k in range(3):
        for sub in range(2):
            want, tol = g["step2_density"][2 * k + sub]
            assert abs(at(k, sub) - want) < tol
            if internalized:
                want, tol = g["step2_internalized"][2 * k + sub]
                assert abs(ag.internalized[k, sub] - want) < tol
    s.release(0)
    if internalized:
        assert at(0, 0) == 1.0 and at(0, 1) == 1.0
        assert ag.internalized[0, 0] == 0 and ag.internalized[0, 1] == 0


@pytest.mark.parametrize("builder", [sc.conflict_cells, sc.conflict_big_cells])
@pytest.mark.parametrize("internalized", [True, False])
@pytest.mark.parametrize("recompute", [True, False])
def test_cells_conflict_brute_force(builder, internalized, recompute):
    p, ag = builder()
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, internalized)
    exp_int = np.zeros((ag.n, 2))
    for step in range(2):
        want = sc.brute_force_densities_1d(p, f, ag)
        if step == 1 and recompute:
            s.recompute()
        s.solve(1)
        np.testing.assert_allclose(f.reshape(p.nx, 2), want, atol=1e-6, rtol=0)
        exp_int = sc.brute_force_internalized_1d(p, f, ag, exp_int)
        if internalized:
            np.testing.assert_allclose(ag.internalized, exp_int, atol=1e-6, rtol=0)
    for i in range(ag.n):
        s.release(i)
    if internalized:
        np.testing.assert_allclose(f, 1.0, atol=1e-6)


def test_cells_inactive_skip():
    p, ag = sc.simple_cells(1)
    ag.is_active[1] = 0
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, True)
    s.solve(1)
    assert abs(f[0, 0, 0, 0] - 28) < 1e-6 and abs(f[0, 0, 0, 1] - 1.0005) < 1e-6
    assert f[0, 0, 1, 0] == 1 and f[0, 0, 1, 1] == 1
    assert abs(f[0, 0, 2, 0] - 366.963) < 1e-3 and abs(f[0, 0, 2, 1] - 1.0015) < 1e-6
    assert ag.internalized[1, 0] == 0 and ag.internalized[1, 1] == 0
    assert abs(ag.internalized[0, 0] + 216000) < 1e-6
    assert abs(ag.internalized[2, 0] + 2927703.8) < 0.1


def test_cells_all_inactive():
    p, ag = sc.simple_cells(1)
    ag = orc.AgentSpec(2, 2, 1)
    sc.default_agent(ag, 0, 0, 1000, [10, 0, 0])
    sc.default_agent(ag, 1, 400, 1000, [30, 0, 0])
    ag.is_active[:] = 0
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, True)
    s.solve(1)
    assert np.all(f == 1) and np.all(ag.internalized == 0)


def test_cells_reactivation():
    g = G["cells"]["InactiveAgentReactivation"]
    p, _ = sc.simple_cells(1)
    p = sc.problem_from_env(G["cells"]["default_microenv"], 1, [40, 20, 20], [20, 20, 20])
    ag = orc.AgentSpec(1, 2, 1)
    sc.default_agent(ag, 0, 0, 1000, [10, 0, 0])
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, True)
    s.solve(1)
    np.testing.assert_allclose(f[0, 0, 0], g["after_first"], atol=1e-6)
    ag.is_active[0] = 0
    s.recompute()
    s.solve(1)
    np.testing.assert_allclose(f[0, 0, 0], g["after_deactivated"], atol=1e-6)
    ag.is_active[0] = 1
    s.recompute()
    s.solve(1)
    (w0, t0), (w1, t1) = g["after_reactivated"]
    assert abs(f[0, 0, 0, 0] - w0) < t0 and abs(f[0, 0, 0, 1] - w1) < t1


def test_cells_reduced_sums_canonical_order():
    """The owner is the lowest active index; sums run owner-first, then ascending."""
    p, ag = sc.conflict_cells()
    f = sc.constant_field(p, [1, 1])
    s = cells_only(p, f, ag, False)
    s.solve(1)
    num, den, fac, ballots, conflict = s.reduced()
    assert conflict
    assert list(ballots[:3]) == [0, 1, 3]
    V = 8000.0
    n = ag.secretion * ag.saturation * p.dt * ag.volumes[:, None] / V
    d = (ag.uptake + ag.secretion) * p.dt * ag.volumes[:, None] / V
    assert np.array_equal(num[3], (n[3] + n[4]) + n[5])
    assert np.array_equal(den[3], ((d[3] + 1) + d[4]) + d[5])


# ------------------------------------------------------------------ Dirichlet

@pytest.mark.parametrize("name,dims", [("Interior1D", 1), ("Interior2D", 2), ("Interior3D", 3)])
def test_dirichlet_interior(name, dims):
    t = G["dirichlet"][name]
    env = G["dirichlet"]["default_microenv"]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    f = sc.constant_field(p, env["initial"])
    spec = orc.DirichletSpec(2, dims)
    vox = [v[:dims] for v in t["voxels"]]
    vals = [[v, 0] for v in t["values"]]
    conds = [[1, 0] for _ in t["values"]]
    spec.set_interior(vox, vals, conds)
    orc.dirichlet(p, f, spec)
    for v, val in zip(t["voxels"], t["values"]):
        x, y, z = v
        assert f[z, y, x, 0] == val
    assert np.all(f[..., 1] == 1)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_dirichlet_boundary_precedence(dims):
    t = G["dirichlet"]["Boundary"]
    env = G["dirichlet"]["default_microenv"]
    p = sc.problem_from_env(env, dims, t["bbox_max"], t["voxel"])
    f = sc.constant_field(p, env["initial"])
    spec = orc.DirichletSpec(2, dims)
    for a, ax in enumerate("xyz"[:dims]):
        lo, hi = t["face_values"][ax]
        spec.set_face(a, "min", [lo, 42], [1, 0])
        spec.set_face(a, "max", [hi, 42], [1, 0])
    orc.dirichlet(p, f, spec)
    n = [p.nx, p.ny, p.nz]
    for z in range(p.nz):
        for y in range(p.ny):
            for x in range(p.nx):
                want = 1.0
                idx = [x, y, z]
                for a, ax in enumerate("xyz"[:dims]):
                    lo, hi = t["face_values"][ax]
                    if idx[a] == 0:
                        want = lo
                    if idx[a] == n[a] - 1:
                        want = hi
                assert f[z, y, x, 0] == want
                assert f[z, y, x, 1] == 1.0


def _builder_env(densities, face_xmin=None, nodes=()):
    """microenvironment_builder: 3-D box [0,100]^3 at dx=10, dt default 0.01
    (microenvironment_builder.h:27), densities (D, decay, initial)."""
    S = len(densities)
    p = orc.Problem(3, (10, 10, 10), (10, 10, 10), 0.01, [d[0] for d in densities],
                    [d[1] for d in densities])
    f = sc.constant_field(p, [d[2] for d in densities])
    spec = orc.DirichletSpec(S, 3)
    if face_xmin is not None:
        spec.set_face(0, "min", face_xmin[0], face_xmin[1])
    if nodes:
        spec.set_interior([n[0] for n in nodes], [n[1] for n in nodes], [n[2] for n in nodes])
    return p, f, spec


def test_dirichlet_dynamic_boundary():
    p, f, spec = _builder_env([(100.0, 0.01, 38.0)], face_xmin=([160.0], [1]))
    s = orc.Solver(p, f, dirichlet=spec)
    s.solve(5)
    spec.min_values[0][0] = 20.0  # update_dirichlet_boundary_min('x', 0, 20, true)
    s.solve(5)
    assert f[1, 1, 0, 0] == 20.0


def test_dirichlet_dynamic_disable():
    p, f, spec = _builder_env([(100.0, 0.01, 38.0)], face_xmin=([160.0], [1]))
    s = orc.Solver(p, f, dirichlet=spec)
    s.solve(5)
    spec.min_conds[0][0] = 0
    s.solve(5)
    assert f[1, 1, 0, 0] != 160.0


def test_dirichlet_dynamic_interior_multi():
    dens = [(100.0, 0.01, 38.0), (50.0, 0.02, 5.0), (30.0, 0.015, 2.0)]
    # the interior list grows from one voxel to two, as update_dirichlet_interior_voxel
    # appends (microenvironment.cpp:165-203)
    p, f, spec = _builder_env(dens, nodes=[((5, 5, 5), (160.0, 50.0, 10.0), (1, 1, 1))])
    s = orc.Solver(p, f, dirichlet=spec)
    s.solve(5)
    spec2 = orc.DirichletSpec(3, 3)
    spec2.set_interior([(5, 5, 5), (2, 2, 2)], [(80.0, 100.0, 0.0), (200.0, 60.0, 15.0)],
                       [(1, 1, 1), (1, 1, 1)])
    s2 = orc.Solver(p, f, dirichlet=spec2)
    s2.solve(5)
    for sub, vox, val in G["dirichlet"]["MultipleSubstrateModification"]["expected"]:
        x, y, z = vox
        assert f[z, y, x, sub] == val


# --------------------------------------------------------------- full step

def test_full_step_stage_order_config1_mass():
    """Config-1 style step: diffusion, then Dirichlet faces, then cells.  The FP32
    method stays within 1e-5 of FP64 over 100 steps."""
    p, f, dirs, _, ag = sc.config1()
    f32 = f.astype(np.float32)
    ag32 = orc.AgentSpec(ag.n, 1, 3)
    for k in ("positions", "secretion", "saturation", "uptake", "net_export", "volumes"):
        setattr(ag32, k, getattr(ag, k).copy())
    s64 = orc.Solver(p, f, dirichlet=dirs, agents=ag)
    s32 = orc.Solver(p, f32, dirichlet=dirs, agents=ag32)
    s64.solve(100)
    s32.solve(100)
    occupied = np.zeros(p.shape[:3], dtype=bool)
    vox = (ag.positions // 20).astype(int)
    occupied[vox[:, 2], vox[:, 1], vox[:, 0]] = True
    assert np.all(f[:, :, 0][~occupied[:, :, 0]] == 38.0)
    assert np.all(f[-1][~occupied[-1]] == 38.0)
    err = np.abs(f32 - f).max() / np.abs(f).max()
    assert err < 1e-5


def _bulk_and_cell_voxel_case():
    """A 1-D row of 3 voxels with D = decay = 0 (the sweeps are the identity),
    bulk supply/uptake at voxel 1 and one cell in voxel 1."""
    dt, V = 0.01, 20.0 * 20 * 20
    p = orc.Problem(1, (3, 1, 1), (20, 20, 20), dt, [0.0], [0.0])
    f = np.full(p.shape, 5.0)
    sup, upt, tgt = (np.zeros(p.shape) for _ in range(3))
    sup[0, 0, 1], upt[0, 0, 1], tgt[0, 0, 1] = 3.0, 2.0, 7.0
    ag = orc.AgentSpec(1, 1, 1)
    ag.positions[0] = [30.0]
    ag.secretion[0], ag.saturation[0], ag.uptake[0], ag.net_export[0] = 40.0, 9.0, 25.0, 600.0
    ag.volumes[0] = 2000.0
    bulk = lambda d: (d + dt * 3.0 * 7.0) / (1 + dt * (2.0 + 3.0))  # bulk_solver.cpp:29
    num = 40.0 * 9.0 * dt * 2000.0 / V                                   # cell_solver.cpp:51-77
    den = (40.0 + 25.0) * dt * 2000.0 / V
    fac = 600.0 * dt / V
    cell = lambda d: (d + num) / (1 + den) + fac                         # cell_solver.cpp:146-163
    return p, f, (sup, upt, tgt), ag, bulk, cell


def test_bulk_then_cells_closed_form():
    """openmp_solver.cpp:27-37 runs bulk before cells.  For a voxel with both, the
    step gives cells(bulk(D)), and differs from bulk(cells(D)) by more than the
    FP64 rounding; voxels with neither keep their value."""
    p, f, bulk_arrays, ag, bulk, cell = _bulk_and_cell_voxel_case()
    orc.Solver(p, f, bulk=bulk_arrays, agents=ag).solve(1)
    want = cell(bulk(5.0))
    other = bulk(cell(5.0))
    assert abs(want - other) > 1e-3 * abs(want)
    assert abs(f[0, 0, 1, 0] - want) <= 1e-14 * abs(want)
    assert f[0, 0, 0, 0] == 5.0 and f[0, 0, 2, 0] == 5.0
