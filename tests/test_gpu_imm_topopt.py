"""GPU parity of the B-rep immersed path: ray-parity classification and
volume fractions, the legacy cut-cell face stabilisation, the immersed solve
and the immersed SIMP loop (rust/src/immersed.rs, rust/src/topopt_immersed.rs)
against the oracle and the reference's tilted-box fixtures."""
import numpy as np
import pytest

import oracle_py as O
from conftest import load_fixture

pytestmark = pytest.mark.gpu

icc = pytest.importorskip("immersed_catmull_clark_amd")


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = icc.Context(0)
    yield c
    c.close()


def _tilted(name="immersed_tilted_box.json"):
    d = load_fixture(name)
    m = icc.TriangleMesh3D(np.asarray(d["nodes"], dtype=np.float64).reshape(-1, 3),
                           np.asarray(d["elements"], dtype=np.int64).reshape(-1, 3))
    return d, m


def _hull(seed=3, n=40):
    """Closed convex triangle surface (a randomly perturbed ellipsoid's hull)."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= np.array([1.3, 0.8, 0.6]) * (1.0 + 0.1 * rng.random((n, 1)))
    return icc.TriangleMesh3D(p, ConvexHull(p).simplices.astype(np.int64))


def test_padded_bounds_and_classification_fixture(ctx):
    """immersed_vs_matlab.rs:20-40: grid bounds, status exact, weights < 1e-7;
    bit-exact against the oracle's ray caster."""
    d, m = _tilted()
    b = icc.padded_bounds(m, 0.05)
    assert np.array_equal(b.ravel(), O.padded_bounds(m.vertices, 0.05))
    assert np.all(np.abs(b.ravel() - np.asarray(d["grid_bounds"]).ravel()) < 1e-14)
    w, st = icc.element_weights(ctx, m, b, d["grid_res"])
    assert np.array_equal(st, np.asarray(d["status"], dtype=np.int8))
    assert np.max(np.abs(w - np.asarray(d["weights"]))) < 1e-7
    ost, ow = O.classify_mesh(m.vertices, m.triangles, b, d["grid_res"])
    assert np.array_equal(st, ost) and np.array_equal(w, ow)


@pytest.mark.parametrize("case", ["tilted", "hull"])
@pytest.mark.parametrize("sub", [(4, 4, 4), (3, 2, 5)])
def test_classification_matches_oracle_bitwise(ctx, case, sub):
    m = _tilted()[1] if case == "tilted" else _hull()
    res = (23, 17, 13)
    b = icc.padded_bounds(m, 0.05)
    w, st = icc.element_weights(ctx, m, b, res, sub)
    ost, ow = O.classify_mesh(m.vertices, m.triangles, b, res, sub)
    assert np.array_equal(st, ost)
    assert np.array_equal(w, ow)
    assert (st == -1).sum() > 0 and (st == 1).sum() > 0 and (st == 0).sum() > 0


def _fixture_operator(ctx, d, m, free):
    res = d["grid_res"]
    g = icc.Grid(ctx, res, np.asarray(d["grid_bounds"]).ravel(), young=d["young"], poisson=d["poisson"])
    w, st = icc.element_weights(ctx, m, icc.padded_bounds(m, 0.05), res)
    op = icc.Operator(g, scale=d["emin"] + (1.0 - d["emin"]) * w, free=free)
    nf = op.set_legacy_stabilization(st, d["gamma_gp"])
    assert nf > 0
    return g, op


def test_legacy_stabilization_fixture_probe(ctx):
    """immersed_vs_matlab.rs:45-55: (K + K_gp) v on the probe vector < 1e-12."""
    d, m = _tilted()
    g, op = _fixture_operator(ctx, d, m, None)
    v = O.probe_vectors(g.ndof, 1)[0]
    assert O.rel_err(op.apply(v), d["k_probe"]) < 1e-12
    # |K_gp|_F of the assembled stabilisation (immersed_vs_matlab.rs:41-44)
    kf = d["kgp_frobenius"]
    assert abs(op.face_frobenius() - kf) <= 1e-12 * kf


@pytest.mark.parametrize("case", ["tilted", "hull"])
def test_legacy_operator_matches_oracle(ctx, case):
    """M (K_s + S) M x with the matrix-free face springs vs the oracle's
    assembled stiffness + legacy stabilisation (CSR)."""
    if case == "tilted":
        m, res = _tilted()[1], [9, 7, 6]
    else:
        m, res = _hull(), [12, 8, 7]
    pb = O.Immersed(m.vertices, m.triangles, res, stab_kind=1, gamma=0.05, emin=1e-4)
    g = icc.Grid(ctx, res, pb.bounds)
    op = icc.Operator(g, scale=1e-4 + (1.0 - 1e-4) * pb.weights, free=pb.free)
    op.set_legacy_stabilization(pb.status, 0.05)
    fr = pb.free.astype(np.float64)
    rng = np.random.default_rng(5)
    for _ in range(2):
        x = rng.standard_normal(pb.ndof)
        assert O.rel_err(op.apply(x), fr * pb.matvec(fr * x)) < 1e-13
    # diagonal: K's plus the springs' (the shared-point block is zero)
    e = np.zeros(pb.ndof)
    dref = np.zeros(pb.ndof)
    for i in range(0, pb.ndof, 37):
        e[:] = 0.0
        e[i] = 1.0
        dref[i] = pb.matvec(e)[i]
    sel = np.arange(0, pb.ndof, 37)
    assert np.max(np.abs(op.diag()[sel] - dref[sel])) < 1e-13 * np.max(np.abs(dref))


def test_immersed_mesh_solve_fixture(ctx):
    """immersed_vs_matlab.rs:56-73: PCG(1e-12) on the masked operator vs the
    MATLAB direct solution (< 1e-6) and compliance (< 1e-9)."""
    d, m = _tilted()
    free = (np.asarray(d["free_mask"]) != 0).astype(np.uint8)
    g, op = _fixture_operator(ctx, d, m, free)
    load = np.asarray(d["load"])
    b = np.where(free != 0, load, 0.0)
    diag = np.maximum(op.diag(), 1e-6)
    u, it, res, _ = op.pcg(b, diag, 1e-12, 20 * g.ndof)
    u = np.where(free != 0, u, 0.0)
    assert O.rel_err(u, d["u_direct"]) < 1e-6
    c = float(load @ u)
    assert abs(c - d["compliance"]) / d["compliance"] < 1e-9


def _topopt_cfg(d, max_iter=None, res=None, solver=0):
    return icc.ImmersedTopOptConfig(grid_res=tuple(res or d["grid_res"]), volfrac=d["volfrac"], penal=d["penal"],
                                    rmin=d["rmin"], max_iter=int(max_iter or d["max_iter"]),
                                    stabilization=icc.Stabilization.legacy(d["gamma_gp"]), pcg_tol=1e-10,
                                    solver=solver)


def test_imm_topopt_fixture(ctx):
    """imm_topopt_vs_matlab.rs:18-49: compliance per iteration < 1e-4 relative,
    final densities < 1e-3 (MATLAB ran FP32 PCG)."""
    d, m = _tilted("imm_topopt_tilted_box.json")
    opt = icc.ImmersedTopOpt(ctx, m, _topopt_cfg(d))
    seen = []
    h = opt.run(lambda it, c, v, ch, n: seen.append(it))
    assert seen == list(range(1, d["max_iter"] + 1))
    cref = np.asarray(d["compliance"])
    assert np.all(np.abs(np.asarray(h.compliance) - cref) / cref < 1e-4), (h.compliance, cref)
    assert np.max(np.abs(h.density - np.asarray(d["xPhys"]))) < 1e-3


def test_imm_topopt_problem_matches_oracle(ctx):
    """Design domain, CAD volume, clamp and load vs the oracle's restatement."""
    d, m = _tilted("imm_topopt_tilted_box.json")
    res = (10, 8, 6)
    opt = icc.ImmersedTopOpt(ctx, m, _topopt_cfg(d, res=res))
    pb = O.Immersed(m.vertices, m.triangles, res, stab_kind=1, gamma=d["gamma_gp"], clamp=0)
    assert np.array_equal(opt.grid_bounds.ravel(), pb.bounds)
    assert np.array_equal(opt.status, pb.status) and np.array_equal(opt.weights, pb.weights)
    assert np.array_equal(opt.active, pb.weights >= 0.05)
    assert opt.cad_volume == float(np.sum(pb.weights[pb.weights >= 0.05]))
    assert np.array_equal(opt.free, pb.free.astype(bool))
    assert np.array_equal(opt.force, pb.force)


@pytest.mark.parametrize("res", [(3, 3, 3), (10, 8, 6)])
def test_imm_topopt_matches_oracle(ctx, res):
    """The whole loop vs the oracle's restatement of ImmersedTopOpt::run at
    pcg_tol 1e-10: compliance and volume to 1e-8, densities and change to 1e-6,
    PCG iteration counts within 2 + 6 % (the summation order differs and
    these small cut-cell systems are ill-conditioned, so round-off moves the
    stopping iteration by a few)."""
    d, m = _tilted("imm_topopt_tilted_box.json")
    iters = 6
    opt = icc.ImmersedTopOpt(ctx, m, _topopt_cfg(d, max_iter=iters, res=res))
    h = opt.run()
    comp, vol, dens, change, its = O.immersed_topopt_run(m.vertices, m.triangles, res, iters, d["volfrac"], d["penal"],
                                                         d["rmin"], d["gamma_gp"], pcg_tol=1e-10, full=True)
    assert np.all(np.abs(np.asarray(h.compliance) - comp) / comp < 1e-8), (h.compliance, comp)
    assert np.all(np.abs(np.asarray(h.volume) - vol) < 1e-9)
    assert np.max(np.abs(h.density - dens)) < 1e-6
    assert np.max(np.abs(np.asarray(h.change) - change)) < 1e-6
    assert np.all(np.abs(np.asarray(h.pcg_iterations) - its) <= 2 + 0.06 * its), (h.pcg_iterations, its)


@pytest.mark.parametrize("solver", [1, 2])
def test_imm_topopt_mixed_solver(ctx, solver):
    """solver 1 (FP32 PCG with FP64 reliable updates; 2: split-FP16 inner
    contraction) reaches the same true residual, so the loop tracks the FP64
    one closely."""
    d, m = _tilted("imm_topopt_tilted_box.json")
    res = (10, 8, 6)
    h64 = icc.ImmersedTopOpt(ctx, m, _topopt_cfg(d, max_iter=4, res=res)).run()
    h32 = icc.ImmersedTopOpt(ctx, m, _topopt_cfg(d, max_iter=4, res=res, solver=solver)).run()
    c64, c32 = np.asarray(h64.compliance), np.asarray(h32.compliance)
    assert np.all(np.abs(c32 - c64) / c64 < 1e-6), (c32, c64)
    assert np.max(np.abs(h32.density - h64.density)) < 1e-4


@pytest.mark.parametrize("case", ["all_cut", "tilted"])
def test_ghost_penalty_operator_matches_oracle(ctx, case):
    """The matrix-free ghost penalty (scale 0 isolates it) vs the oracle's
    assembled ghost_penalty matrix (immersed.rs:234-370): unmasked and masked
    products and the diagonal."""
    if case == "all_cut":
        bounds, res = np.array([0.0, 1.3, -0.2, 0.7, 0.1, 0.6]), (6, 5, 4)
        status = -np.ones(int(np.prod(res)), dtype=np.int8)
    else:
        _, m = _tilted()
        res = (11, 9, 7)
        bounds = icc.padded_bounds(m, 0.05).ravel()
        status = icc.element_weights(ctx, m, bounds, res)[1]
    gamma, young = 1e-2, 2.5
    g = icc.Grid(ctx, res, bounds, young=young)
    op = icc.Operator(g, scale=np.zeros(g.nel_tot))
    assert op.set_ghost_penalty(status, gamma) > 0
    rng = np.random.default_rng(2)
    x = rng.standard_normal(g.ndof)
    want, fro = O.ghost_penalty_matvec(bounds, res, 2, status, gamma, young, x)
    assert fro > 0.0
    assert O.rel_err(op.apply(x), want) < 1e-13
    assert abs(op.face_frobenius() - fro) <= 1e-12 * fro
    fr = (rng.random(g.ndof) > 0.2).astype(np.uint8)
    op.set_free(fr)
    want_m, _ = O.ghost_penalty_matvec(bounds, res, 2, status, gamma, young, fr * x)
    assert O.rel_err(op.apply(x), fr * want_m) < 1e-13
    sel = np.arange(0, g.ndof, 29)
    d = op.diag()
    for i in sel:
        e = np.zeros(g.ndof)
        e[i] = 1.0
        gi, _ = O.ghost_penalty_matvec(bounds, res, 2, status, gamma, young, e)
        assert abs(d[i] - gi[i]) <= 1e-13 * fro


def test_imm_topopt_ghost_penalty_matches_oracle(ctx):
    """ImmersedTopOptConfig::default's stabilisation (ghost penalty, gamma
    1e-2) through the whole loop vs the oracle (stab_kind 2)."""
    d, m = _tilted("imm_topopt_tilted_box.json")
    res, iters = (10, 8, 6), 5
    cfg = icc.ImmersedTopOptConfig(grid_res=res, max_iter=iters, pcg_tol=1e-10)
    assert cfg.stabilization.kind == icc.Stabilization().kind and cfg.stabilization.gamma == 1e-2
    opt = icc.ImmersedTopOpt(ctx, m, cfg)
    h = opt.run()
    comp, vol, dens, change, its = O.immersed_topopt_run(m.vertices, m.triangles, res, iters, cfg.volfrac, cfg.penal,
                                                         cfg.rmin, 1e-2, emin=cfg.emin, pcg_tol=1e-10, stab_kind=2,
                                                         full=True)
    assert np.all(np.abs(np.asarray(h.compliance) - comp) / comp < 1e-8), (h.compliance, comp)
    assert np.all(np.abs(np.asarray(h.volume) - vol) < 1e-9)
    assert np.max(np.abs(h.density - dens)) < 1e-6
    assert np.all(np.abs(np.asarray(h.pcg_iterations) - its) <= 2 + 0.06 * its), (h.pcg_iterations, its)


def test_mesh_and_topopt_input_errors(ctx):
    """Empty meshes and out-of-range triangle indices are rejected before any
    device work; a B-rep that leaves no design cell fails at run time."""
    d, m = _tilted()
    bad = icc.TriangleMesh3D(m.vertices, np.array([[0, 1, 99]]))
    with pytest.raises(icc.IccError) as ei:
        icc.element_weights(ctx, bad, icc.padded_bounds(m, 0.05), (4, 4, 4))
    assert ei.value.code == icc.ICC_EINVAL
    with pytest.raises(icc.IccError):
        icc.ImmersedTopOpt(ctx, bad, icc.ImmersedTopOptConfig(grid_res=(4, 4, 4), max_iter=1))
    # two tiny boxes at opposite corners of their bounding box: no cell
    # reaches w >= 0.05
    a = icc.TriangleMesh3D.new_box((0.0, 0.0, 0.0), (0.01, 0.01, 0.01))
    b = icc.TriangleMesh3D.new_box((0.99, 0.99, 0.99), (1.0, 1.0, 1.0))
    two = icc.TriangleMesh3D(np.vstack([a.vertices, b.vertices]), np.vstack([a.triangles, b.triangles + 8]))
    opt = icc.ImmersedTopOpt(ctx, two, icc.ImmersedTopOptConfig(grid_res=(4, 4, 3), max_iter=1))
    assert opt.n_active == 0
    with pytest.raises(icc.IccError) as ei:
        opt.run()
    assert ei.value.code == icc.ICC_EINVAL


def test_face_terms_without_cut_cells(ctx):
    """A status with no CUT cell stabilises no face: the operator is the plain
    one, for both kinds (to round-off: the apply's atomic scatter order
    varies between launches)."""
    g = icc.Grid(ctx, (5, 4, 3), [0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    op = icc.Operator(g, scale=np.ones(g.nel_tot))
    ref = icc.Operator(g, scale=np.ones(g.nel_tot))
    status = np.ones(g.nel_tot, dtype=np.int8)
    assert op.set_legacy_stabilization(status, 0.05) == 0
    assert op.set_ghost_penalty(status, 1e-2) == 0
    x = np.random.default_rng(0).standard_normal(g.ndof)
    assert O.rel_err(op.apply(x), ref.apply(x)) < 1e-14


@pytest.mark.parametrize("nel", [(1, 1, 1), (2, 1, 3), (3, 2, 2), (4, 5, 1)])
def test_ghost_penalty_small_grids_match_oracle(ctx, nel):
    """Grids of one to five cells per direction, where the jump tables see
    the repeated end knots, against the oracle's assembled matrix."""
    bounds = np.array([0.0, 1.1, 0.0, 0.7, 0.0, 0.9])
    rng = np.random.default_rng(sum(nel))
    status = rng.choice(np.array([-1, 1, 0], dtype=np.int8), size=int(np.prod(nel)), p=[0.6, 0.3, 0.1])
    g = icc.Grid(ctx, nel, bounds, young=1.7)
    op = icc.Operator(g, scale=np.zeros(g.nel_tot))
    op.set_ghost_penalty(status, 0.3)
    x = rng.standard_normal(g.ndof)
    want, fro = O.ghost_penalty_matvec(bounds, nel, 2, status, 0.3, 1.7, x)
    got = op.apply(x)
    if fro == 0.0:
        assert np.max(np.abs(got)) == 0.0
    else:
        assert O.rel_err(got, want) < 1e-13


def test_build_immersed_problem_fixture(ctx):
    """build_immersed_problem / solve_immersed_problem through the B-rep
    handle (immersed_vs_matlab.rs:20-73): status exact, weights < 1e-7, load
    < 1e-15, free mask exact, K v probe < 1e-12 (unmasked: all free), PCG(1e-12)
    vs the MATLAB direct solution < 1e-6, compliance < 1e-9."""
    d, m = _tilted()
    opts = icc.ImmersedOptions(grid_res=tuple(d["grid_res"]), young=d["young"], poisson=d["poisson"],
                               stabilization=icc.Stabilization.legacy(d["gamma_gp"]), emin=d["emin"])
    pb = icc.build_immersed_problem(ctx, m, opts)
    assert np.all(np.abs(pb.grid_bounds.ravel() - np.asarray(d["grid_bounds"]).ravel()) < 1e-14)
    assert np.array_equal(pb.status, np.asarray(d["status"], dtype=np.int8))
    assert np.max(np.abs(pb.weights - np.asarray(d["weights"]))) < 1e-7
    assert O.rel_err(pb.force, d["load"]) < 1e-15
    assert np.array_equal(pb.free, np.asarray(d["free_mask"]) != 0)
    sol = icc.solve_immersed_problem(pb, 1e-12, 20 * pb.ndof)
    assert sol.status == icc.ICC_OK
    assert O.rel_err(sol.u, d["u_direct"]) < 1e-6
    assert abs(sol.compliance - d["compliance"]) / d["compliance"] < 1e-9
    free_all = icc.ImmersedOptions(grid_res=tuple(d["grid_res"]), stabilization=icc.Stabilization.legacy(d["gamma_gp"]),
                                   emin=d["emin"], clamped_face=None)
    pb2 = icc.build_immersed_problem(ctx, m, free_all)
    v = O.probe_vectors(pb2.ndof, 1)[0]
    assert O.rel_err(pb2.apply(v), d["k_probe"]) < 1e-12


@pytest.mark.parametrize("stab,clamp", [("ghost", "bottom"), ("legacy", "left"), ("none", "bottom")])
def test_build_immersed_problem_matches_oracle(ctx, stab, clamp):
    """Other stabilisations and clamped faces against the oracle's
    build_immersed_problem / solve_immersed_problem: solution to 1e-8,
    compliance to 1e-9, same free mask and load."""
    m = _hull()
    res = (12, 9, 7)
    kind = {"ghost": (icc.Stabilization.ghost_penalty(1e-2), 2), "legacy": (icc.Stabilization.legacy(0.05), 1),
            "none": (icc.Stabilization.none(), 0)}[stab]
    ckind = {"bottom": 0, "top": 1, "left": 2}[clamp]
    opts = icc.ImmersedOptions(grid_res=res, stabilization=kind[0], clamped_face=clamp, pcg_tol=1e-12,
                               pcg_maxit=200000)
    pb = icc.build_immersed_problem(ctx, m, opts)
    ref = O.Immersed(m.vertices, m.triangles, res, stab_kind=kind[1], gamma=kind[0].gamma, emin=1e-4, clamp=ckind)
    assert np.array_equal(pb.free, ref.free.astype(bool))
    assert O.rel_err(pb.force, ref.force) < 1e-15
    sol = pb.solve()
    u, c, it, res_ = ref.solve(1e-12, 200000)
    assert O.rel_err(sol.u, u) < 1e-8, (O.rel_err(sol.u, u), sol.iterations, it)
    assert abs(sol.compliance - c) / abs(c) < 1e-9
    for solver in (1, 2):
        mixed = pb.solve(tol=1e-10, solver=solver)
        assert mixed.true_residual <= 1e-10 and O.rel_err(mixed.u, u) < 1e-6, solver


def test_build_immersed_problem_custom_force_top_clamp(ctx):
    """A caller's force with the top face clamped (the default load would
    sit on fixed DOFs): the free mask matches the oracle's top clamp, the
    solution satisfies M (K + S) M u = M f to the tolerance, u is zero on the
    clamp, and the compliance is F.u."""
    m = _hull()
    res = (10, 8, 6)
    opts = icc.ImmersedOptions(grid_res=res, clamped_face="top", pcg_tol=1e-10, pcg_maxit=100000)
    probe = icc.build_immersed_problem(ctx, m, opts)
    f = np.random.default_rng(12).standard_normal(probe.ndof) * probe.free
    opts.force = f
    pb = icc.build_immersed_problem(ctx, m, opts)
    ref = O.Immersed(m.vertices, m.triangles, res, stab_kind=2, gamma=1e-2, emin=1e-4, clamp=1)
    assert np.array_equal(pb.free, ref.free.astype(bool))
    assert np.array_equal(pb.force, f)
    sol = pb.solve()
    assert sol.status == icc.ICC_OK
    r = np.where(pb.free, f - pb.apply(sol.u), 0.0)
    assert np.linalg.norm(r) / np.linalg.norm(f) < 1e-9
    assert np.all(sol.u[~pb.free] == 0.0)
    assert abs(sol.compliance - float(f @ sol.u)) <= 1e-12 * abs(sol.compliance)

