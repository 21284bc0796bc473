"""GPU parity of the surface boundary conditions on triangle surfaces
(rust/src/immersed_bc.rs) and ImmersedProblem::add_surface_terms
(rust/src/immersed.rs:541-548) against the oracle's restatement, which
tests/test_oracle_surface.py pins (Dunavant exactness, rigid-translation
and load totals, the reference's Timoshenko cantilever)."""
import numpy as np
import pytest

import oracle_py as O

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


def _tilted_box():
    m = icc.TriangleMesh3D.new_box([0.1, -0.2, 0.05], [1.3, 0.4, 0.55])
    rot = np.array([[0.95, -0.25, 0.08], [0.25, 0.96, 0.04], [-0.08, 0.0, 1.0]])
    return icc.TriangleMesh3D(m.vertices @ rot.T, m.triangles)


def _conditions():
    """(kind, triangles, keyword arguments of the oracle's Surface.add, product condition)."""
    D, T, P, R = O.Surface.DIRICHLET, O.Surface.TRACTION, O.Surface.PRESSURE, O.Surface.ROBIN
    return [
        (D, [8, 9], dict(comps=(1, 1, 1), values=(0.0, 0.0, 0.0)), icc.Dirichlet((True, True, True), (0, 0, 0))),
        (D, [0, 1], dict(comps=(0, 1, 1), values=(0.0, 0.25, -0.5)), icc.Dirichlet((False, True, True), (0, 0.25, -0.5))),
        (T, [10, 11], dict(values=(0.1, -2.0, 0.3)), icc.NeumannTraction((0.1, -2.0, 0.3))),
        (P, [2, 3, 4], dict(pressure=1.5), icc.NeumannPressure(1.5)),
        (R, [6, 7], dict(kn=3.0, kt=0.5), icc.RobinFoundation(3.0, 0.5)),
    ]


def test_scalar_basis_at_matches_oracle_bitwise(ctx):
    bounds = [[0.0, 1.3], [-0.2, 0.7], [0.1, 0.6]]
    g = icc.Grid(ctx, (7, 5, 4), np.asarray(bounds).ravel())
    sp = O.Space(bounds, [7, 5, 4], 2)
    rng = np.random.default_rng(7)
    pts = [np.array([rng.uniform(lo - 0.1, hi + 0.1) for lo, hi in bounds]) for _ in range(200)]
    pts += [np.array([bounds[0][0] + k * 1.3 / 7, 0.3, 0.35]) for k in range(8)]  # on the knots
    for p in pts:
        idx, val = icc.scalar_basis_at(g, p)
        io, vo = O.scalar_basis_at(sp, p)
        assert np.array_equal(idx, io) and np.array_equal(val, vo), p


def test_surface_terms_match_oracle(ctx):
    """All four condition kinds accumulated into one SurfaceTerms: the
    per-element blocks formed on the GPU give the oracle's triplets (same
    order, same values bit for bit) and its load to round-off."""
    mesh = _tilted_box()
    opts = icc.ImmersedOptions(grid_res=(9, 7, 6), clamped_face=None)
    pb = icc.build_immersed_problem(ctx, mesh, opts)
    ref = O.Immersed(mesh.vertices, mesh.triangles, (9, 7, 6), stab_kind=2, gamma=1e-2, emin=1e-4, clamp=-1)
    beta = 1e3 * 1.0 / pb.h_min
    st = icc.SurfaceTerms(pb)
    so = O.Surface(ref, immersed=True)
    for kind, tris, kw, bc in _conditions():
        st.add(mesh, tris, bc, beta=beta)
        so.add(kind, mesh.vertices, mesh.triangles, tris, beta=beta, **kw)
    r, c, v, f = st.terms(pb.ndof)
    ro, co, vo, fo = so.terms()
    nt, nb, npts = st.sizes()
    assert nt == len(vo) and npts > 0 and nb > 0
    assert np.array_equal(r, ro) and np.array_equal(c, co)
    assert np.array_equal(v, vo)
    # the load is summed per element block here and point by point in the oracle
    assert O.rel_err(f, fo) < 1e-14


def test_dirichlet_penalty_fn_matches_oracle(ctx):
    mesh = _tilted_box()
    pb = icc.build_immersed_problem(ctx, mesh, icc.ImmersedOptions(grid_res=(8, 6, 5), clamped_face=None))
    ref = O.Immersed(mesh.vertices, mesh.triangles, (8, 6, 5), stab_kind=2, gamma=1e-2, clamp=-1)
    g0 = np.array([0.3, -0.2, 0.1])
    G = np.array([[1.0, 0.5, 0.0], [0.0, 2.0, -1.0], [0.25, 0.0, 0.5]])
    beta = 40.0
    st = icc.dirichlet_penalty_fn(pb, mesh, np.arange(12), beta, lambda x: g0 + G @ x)
    so = O.Surface(ref, immersed=True)
    so.dirichlet_affine(mesh.vertices, mesh.triangles, np.arange(12), beta, g0, G)
    r, c, v, f = st.terms(pb.ndof)
    ro, co, vo, fo = so.terms()
    assert np.array_equal(r, ro) and np.array_equal(c, co) and np.array_equal(v, vo)
    assert O.rel_err(f, fo) < 1e-14


@pytest.mark.parametrize("stab", ["ghost", "legacy"])
def test_problem_with_surface_terms_matches_oracle(ctx, stab):
    """add_surface_terms on the B-rep problem (no clamped face; penalty clamp,
    traction, pressure and Robin terms): the operator M (K + S + K_s) M x
    against the oracle's assembled matrix, the load, and the FP64 and mixed
    solves against the oracle's PCG."""
    mesh = _tilted_box()
    res = (10, 7, 6)
    kind = (icc.Stabilization.ghost_penalty(1e-2), 2) if stab == "ghost" else (icc.Stabilization.legacy(0.05), 1)
    pb = icc.build_immersed_problem(ctx, mesh, icc.ImmersedOptions(grid_res=res, stabilization=kind[0],
                                                                   clamped_face=None, pcg_tol=1e-12,
                                                                   pcg_maxit=400000))
    ref = O.Immersed(mesh.vertices, mesh.triangles, res, stab_kind=kind[1], gamma=kind[0].gamma, clamp=-1)
    pb.set_force(np.zeros(pb.ndof))
    ref.zero_force()
    beta = 1e3 / pb.h_min
    st = icc.SurfaceTerms(pb)
    so = O.Surface(ref, immersed=True)
    for k, tris, kw, bc in _conditions():
        st.add(mesh, tris, bc, beta=beta)
        so.add(k, mesh.vertices, mesh.triangles, tris, beta=beta, **kw)
    pb.add_surface_terms(st)
    ref.add_surface(so)
    assert O.rel_err(pb.force, ref.force) < 1e-14
    x = np.random.default_rng(4).standard_normal(pb.ndof)
    assert O.rel_err(pb.apply(x), ref.matvec(x)) < 1e-13
    sol = pb.solve()
    u, comp, it, res_ = ref.solve(1e-12, 400000)
    assert sol.status == icc.ICC_OK
    assert O.rel_err(sol.u, u) < 1e-8, (O.rel_err(sol.u, u), sol.iterations, it)
    assert abs(sol.compliance - comp) / abs(comp) < 1e-9
    mixed = pb.solve(tol=1e-10, solver=1)
    assert mixed.true_residual <= 1e-10 and O.rel_err(mixed.u, u) < 1e-6


def test_immersed_cantilever_tip_deflection(ctx):
    """immersed_cad_bc.rs:115-151 on the GPU: a 4 x 0.5 x 0.5 box on a
    44 x 6 x 6 grid, penalty clamp (factor 1e3) on x = 0, tip traction
    -1e-3 / (b h), ghost penalty 1e-2, PCG to 1e-10: the tip deflection is
    within 5 % of Timoshenko beam theory and within 1e-7 of the oracle's."""
    (l, b, h) = (4.0, 0.5, 0.5)
    p_total = 1e-3
    mesh = icc.TriangleMesh3D.new_box([0, 0, 0], [l, b, h])
    pb = icc.build_immersed_problem(ctx, mesh, icc.ImmersedOptions(grid_res=(44, 6, 6), clamped_face=None))
    pb.set_force(np.zeros(pb.ndof))
    st = icc.surface_boundary_terms(pb, mesh, [([8, 9], icc.Dirichlet()),
                                               ([10, 11], icc.NeumannTraction((0.0, -p_total / (b * h), 0.0)))],
                                    1e3, 1.0, pb.h_min)
    pb.add_surface_terms(st)
    sol = pb.solve(tol=1e-10, maxit=20 * pb.ndof)
    assert sol.status == icc.ICC_OK and sol.residual < 1e-9
    x, w, _ = icc.surface_quadrature(mesh, [10, 11], 0.05)
    nsc = pb.ndof // 3
    delta = -sum(wq * icc.eval_vector_at(pb, sol.u, xq, nsc)[1] for xq, wq in zip(x, w)) / w.sum()
    e, nu = 1.0, 0.3
    g = e / (2 * (1 + nu))
    timo = p_total * l ** 3 / (3 * e * b * h ** 3 / 12) + p_total * l / (5 / 6 * g * b * h)
    assert abs(delta - timo) / timo < 0.05, (delta, timo)
    ref = O.Immersed(mesh.vertices, mesh.triangles, (44, 6, 6), stab_kind=2, gamma=1e-2, emin=1e-4, clamp=-1)
    ref.zero_force()
    so = O.Surface(ref, immersed=True)
    so.add(O.Surface.DIRICHLET, mesh.vertices, mesh.triangles, [8, 9], beta=1e3 / pb.h_min)
    so.add(O.Surface.TRACTION, mesh.vertices, mesh.triangles, [10, 11], values=(0.0, -p_total / (b * h), 0.0))
    ref.add_surface(so)
    u, comp, it, res_ = ref.solve(1e-10, 20 * ref.ndof)
    d_ref = -sum(wq * ref.eval_vector_at(u, xq)[1] for xq, wq in zip(x, w)) / w.sum()
    assert abs(delta - d_ref) / abs(d_ref) < 1e-7


def test_surface_input_errors(ctx):
    mesh = _tilted_box()
    pb = icc.build_immersed_problem(ctx, mesh, icc.ImmersedOptions(grid_res=(6, 5, 4), clamped_face=None))
    st = icc.SurfaceTerms(pb)
    with pytest.raises(icc.IccError):
        st.add(mesh, [12], icc.NeumannTraction((1, 0, 0)))  # triangle index out of range
    with pytest.raises(TypeError):
        st.add(mesh, [0], object())
    assert st.sizes() == (0, 0, 0)
    other = icc.Grid(ctx, (3, 3, 3), np.array([0, 1, 0, 1, 0, 1.0]))
    so = icc.SurfaceTerms(other)
    so.add(mesh, [0], icc.NeumannTraction((1, 0, 0)))
    with pytest.raises(icc.IccError):
        pb.add_surface_terms(so)
