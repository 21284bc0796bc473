"""GPU parity of the HIP hot path against the CPU oracle, through the C ABI.

Tolerances: FP64 paths 1e-13 relative (operator) / 1e-8 (solves); FP32
contraction 2e-6 relative on the operator.
"""
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


def _oracle_space(bounds, nel, young=1.0, poisson=0.3):
    sp = O.Space(bounds, nel, 2)
    lam, mu = O.lame(young, poisson)
    types, tid = sp.elem_mats(lam, mu)
    return sp, types, tid


GRIDS = [
    ([[0.5, 2.5], [-1, 0.2], [0.1, 0.4]], [4, 3, 3]),   # iga_kernel.json case 5
    ([[0, 1.2], [0, 0.6], [0, 0.3]], [8, 4, 2]),        # topopt fixture grid
    ([[0, 2.0], [0, 1.0], [0, 1.0]], [16, 8, 8]),       # config 0 grid
    ([[0, 1.0], [0, 1.0], [0, 1.0]], [1, 1, 1]),
    ([[0, 1.0], [0, 0.5], [0, 2.0]], [2, 1, 5]),
    ([[0, 3.0], [0, 1.0], [0, 1.0]], [13, 7, 3]),
    ([[0, 2.0], [0, 1.0], [0, 1.0]], [32, 3, 3]),      # the reference's key splits a type at nel = 32
]


def test_mfma_layout(ctx):
    """D[i][j] = i + 16 (j % 4) by construction; register r of lane l must hold
    row i = 4 r + l/16 (f64) or 4 (l/16) + r (f32) of column l % 16."""
    f64, f32 = ctx.mfma_layout()
    lane = np.arange(64)[:, None]
    r = np.arange(4)[None, :]
    col = lane % 16
    assert np.array_equal(f64, (4 * r + lane // 16 + 16 * (col % 4)).astype(np.float64)), f64
    assert np.array_equal(f32, (4 * (lane // 16) + r + 16 * (col % 4)).astype(np.float64)), f32


@pytest.mark.parametrize("bounds,nel", GRIDS)
def test_grid_matches_oracle(ctx, bounds, nel):
    g = icc.Grid(ctx, nel, bounds)
    sp, types, tid = _oracle_space(bounds, nel)
    assert g.ndof == sp.ndof and g.nel_tot == sp.nel
    assert np.array_equal(g.connectivity(), sp.connectivity())
    # the reference's numbering, also where its rounded key splits a type
    # (nel = 32 on a length-2 side: 36 types)
    gt = g.type_id()
    assert g.ntypes == sp.ntypes and np.array_equal(gt, tid)
    # every element's matrix agrees with the reference's matrix of that element
    mats = [g.element_matrix(t) for t in range(g.ntypes)]
    for e in range(g.nel_tot):
        K, R = mats[gt[e]], types[tid[e]]
        assert np.max(np.abs(K - R)) <= 1e-13 * np.abs(R).max()


def test_apply_fixture_probes(ctx):
    d = load_fixture("iga_kernel.json")
    case = d["cases"][4]
    assert case["degree"] == 2 and len(case["nsub"]) == 3
    g = icc.Grid(ctx, case["nsub"], case["bounds"], young=1.0, poisson=0.3)
    op = icc.Operator(g)
    probes = np.asarray(case["probes"])
    for j, v in enumerate(O.probe_vectors(g.ndof)):
        y = op.apply(v)
        assert O.rel_err(y, probes[:, j]) < 1e-13


@pytest.mark.parametrize("bounds,nel", GRIDS)
def test_apply_scaled_masked_matches_oracle(ctx, bounds, nel):
    rng = np.random.default_rng(sum(nel))
    g = icc.Grid(ctx, nel, bounds)
    sp, _, _ = _oracle_space(bounds, nel)
    s = rng.uniform(0.01, 1.0, g.nel_tot) ** 3
    free = (rng.uniform(size=g.ndof) > 0.2).astype(np.uint8)
    op = icc.Operator(g, scale=s, free=free)
    sp.assemble(s)
    x = rng.standard_normal(g.ndof)
    ref = free * sp.matvec(free * x)
    y = op.apply(x)
    assert O.rel_err(y, ref) < 1e-13
    y32 = op.apply(x, prec=icc.ICC_F32)
    assert O.rel_err(y32, ref) < 2e-6
    # Jacobi diagonal of the unmasked operator
    assert O.rel_err(op.diag(), sp.diag()) < 1e-14


def test_energies_and_sensitivities(ctx):
    bounds, nel = [[0, 2.0], [0, 1.0], [0, 1.0]], [16, 8, 8]
    g = icc.Grid(ctx, nel, bounds)
    sp, _, _ = _oracle_space(bounds, nel)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(g.ndof)
    x = rng.uniform(0.01, 1.0, g.nel_tot)
    op = icc.Operator(g)
    ee = op.energies(u)
    ref = sp.energies(u)
    assert O.rel_err(ee, ref) < 1e-13
    dc = op.sens(u, x, 3.0, 1e-9)
    dref = -3.0 * (1.0 - 1e-9) * x ** 2 * ref
    assert O.rel_err(dc, dref) < 1e-13
    w = rng.uniform(0, 1, g.nel_tot)
    dcw = op.sens(u, x, 3.0, 1e-4, w=w)
    assert O.rel_err(dcw, -w * (3.0 * (1 - 1e-4) * x ** 2 * ref)) < 1e-13


@pytest.mark.parametrize("nel,rmin", [([8, 4, 2], 1.5), ([16, 8, 8], 1.5), ([13, 7, 3], 2.3)])
def test_filters_match_oracle(ctx, nel, rmin):
    bounds = [[0, 1.2], [0, 0.6], [0, 0.3]]
    g = icc.Grid(ctx, nel, bounds)
    sp = O.Space(bounds, nel, 2)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.001, 1.0, g.nel_tot)
    dc = -rng.uniform(0.0, 5.0, g.nel_tot)
    f = icc.Filter(g, rmin)
    assert O.rel_err(f.sens(x, dc), sp.filter_apply(rmin, 0, x, dc)) < 1e-15
    assert O.rel_err(f.fwd(x), sp.filter_apply(rmin, 1, None, x)) < 1e-15
    assert O.rel_err(f.bwd(dc), sp.filter_apply(rmin, 2, None, dc)) < 1e-15


@pytest.mark.parametrize("nel,rmin,frac", [([8, 4, 2], 1.5, 1.0), ([13, 7, 3], 2.3, 0.6), ([16, 12, 8], 1.5, 0.3)])
def test_design_filter_matches_oracle(ctx, nel, rmin, frac):
    """Design-domain filter (topopt_immersed.rs:117-151) on a random active set,
    every mode against the oracle's rows."""
    bounds = [[0, 1.2], [0, 0.6], [0, 0.3]]
    g = icc.Grid(ctx, nel, bounds)
    sp = O.Space(bounds, nel, 2)
    rng = np.random.default_rng(11)
    act = (rng.uniform(size=g.nel_tot) < frac).astype(np.uint8)
    x = rng.uniform(0.001, 1.0, g.nel_tot)
    dc = -rng.uniform(0.0, 5.0, g.nel_tot)
    f = icc.Filter(g, rmin, active=act)
    assert O.rel_err(f.sens(x, dc), sp.design_filter_apply(rmin, 0, act, x, dc)) < 1e-15
    assert O.rel_err(f.fwd(x), sp.design_filter_apply(rmin, 1, act, None, x)) < 1e-15
    assert O.rel_err(f.bwd(dc), sp.design_filter_apply(rmin, 2, act, None, dc)) < 1e-15
    f.close()


def test_design_filter_full_box_equals_structured(ctx):
    """All cells active on cubic elements: the design-domain filter equals the
    structured one (weights differ by the factor h only)."""
    nel = [16, 8, 8]
    g = icc.Grid(ctx, nel, [[0, 2.0], [0, 1.0], [0, 1.0]])
    rng = np.random.default_rng(12)
    x = rng.uniform(0.001, 1.0, g.nel_tot)
    v = rng.standard_normal(g.nel_tot)
    a = icc.Filter(g, 1.5, active=np.ones(g.nel_tot, dtype=np.uint8))
    b = icc.Filter(g, 1.5)
    assert O.rel_err(a.sens(x, v), b.sens(x, v)) < 1e-14
    assert O.rel_err(a.fwd(v), b.fwd(v)) < 1e-14
    assert O.rel_err(a.bwd(v), b.bwd(v)) < 1e-14
    with pytest.raises(ValueError):
        icc.Filter(g, 1.5, active=np.ones(5, dtype=np.uint8))


def _cantilever_problem(nel, seed, size=(2.0, 1.0, 1.0), rho=(0.01, 1.0), emin=1e-9):
    """Clamp x = 0, consistent downward unit traction on x = L; SIMP scale."""
    bounds = [[0, size[0]], [0, size[1]], [0, size[2]]]
    sp, _, _ = _oracle_space(bounds, nel)
    rng = np.random.default_rng(seed)
    x = rng.uniform(rho[0], rho[1], sp.nel)
    s = emin + (1 - emin) * x ** 3
    ncp = sp.ndof_dir
    i = np.arange(sp.ndof_sc) % ncp[0]
    free = np.ones(sp.ndof, dtype=np.uint8)
    for c in range(3):
        free[c * sp.ndof_sc:(c + 1) * sp.ndof_sc][i == 0] = 0
    f = sp.traction_xmax([0.0, -1.0, 0.0]) * free
    return bounds, sp, s, free, f


def test_pcg_matches_oracle(ctx):
    nel = [8, 4, 4]
    bounds, sp, s, free, f = _cantilever_problem(nel, 11, rho=(0.3, 1.0), emin=1e-3)
    sp.assemble(s)
    diag = np.maximum(sp.diag(), 1e-12)
    x_ref, it_ref, res_ref = sp.pcg(f, diag, free, 1e-10, 20 * sp.ndof)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    d = op.diag()
    assert O.rel_err(d, sp.diag()) < 1e-14
    x, it, res, rc = op.pcg(f, np.maximum(d, 1e-12), 1e-10, 20 * g.ndof)
    assert rc == icc.ICC_OK
    assert res < 1e-10
    assert abs(it - it_ref) <= max(2, it_ref // 50), (it, it_ref)
    assert O.rel_err(x, x_ref) < 1e-8


def test_pcg_reports_iteration_limit(ctx):
    nel = [8, 4, 4]
    bounds, sp, s, free, f = _cantilever_problem(nel, 12)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    x, it, res, rc = op.pcg(f, np.maximum(op.diag(), 1e-12), 1e-14, 5)
    assert rc == icc.ICC_ENOCONV and it == 5 and res > 1e-14


@pytest.mark.parametrize("inner", ["f64", "f32"])
def test_iterative_refinement_true_residual(ctx, inner):
    nel = [16, 8, 8]
    bounds, sp, s, free, f = _cantilever_problem(nel, 13)
    sp.assemble(s)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    prec = icc.ICC_F64 if inner == "f64" else icc.ICC_F32
    u, st, rc = op.solve_ir(f, tol=1e-8, inner=prec, inner_tol=1e-3, max_outer=60, max_inner=20000)
    assert rc == icc.ICC_OK, st
    # the true residual, evaluated by the oracle's assembled FP64 operator
    r = free * (f - sp.matvec(free * u))
    assert np.linalg.norm(r) / np.linalg.norm(free * f) <= 1.01e-8
    assert st["true_res"] <= 1e-8


@pytest.mark.parametrize("delta", [0.1, 1e-3])
def test_mixed_reliable_updates_true_residual(ctx, delta):
    """FP32 recurrence with FP64 residual replacement: the true residual of the
    returned u, recomputed by the oracle's assembled FP64 operator, meets 1e-8,
    and u agrees with the FP64 PCG solution to the conditioning of the solve."""
    nel = [16, 8, 8]
    bounds, sp, s, free, f = _cantilever_problem(nel, 14)
    sp.assemble(s)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    u, st, rc = op.solve_mixed(f, tol=1e-8, delta=delta)
    assert rc == icc.ICC_OK, st
    r = free * (f - sp.matvec(free * u))
    assert np.linalg.norm(r) / np.linalg.norm(free * f) <= 1.01e-8
    assert st["true_res"] <= 1e-8 and st["updates"] >= 1
    u64, it, res, rc64 = op.pcg(f, np.maximum(op.diag(), 1e-12), 1e-10, 20 * op.ndof)
    assert rc64 == icc.ICC_OK
    assert O.rel_err(u, u64) < 1e-5
    # a converged start returns at once with no FP32 iterations
    u2, st2, rc2 = op.solve_mixed(f, tol=1e-6, delta=delta, u0=u)
    assert rc2 == icc.ICC_OK and st2["inner_iters"] == 0 and st2["updates"] == 0


def test_mixed_iteration_limit(ctx):
    nel = [8, 4, 4]
    bounds, sp, s, free, f = _cantilever_problem(nel, 15)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    u, st, rc = op.solve_mixed(f, tol=1e-14, delta=0.5, maxit=7)
    assert rc == icc.ICC_ENOCONV and st["inner_iters"] == 7


def test_apply_properties_config1_size(ctx):
    """Size-independent properties at the 96x48x48 grid: linearity, symmetry,
    rigid-body null space, FP32 vs FP64."""
    nel = [96, 48, 48]
    bounds = [[0, 2.0], [0, 1.0], [0, 1.0]]
    g = icc.Grid(ctx, nel, bounds)
    assert g.ndof == 735000 and g.ntypes == 27
    rng = np.random.default_rng(3)
    s = 1e-9 + (1 - 1e-9) * rng.uniform(0.01, 1.0, g.nel_tot) ** 3
    op = icc.Operator(g, scale=s)
    x = rng.standard_normal(g.ndof)
    y = rng.standard_normal(g.ndof)
    ax, ay = op.apply(x), op.apply(y)
    axy = op.apply(2.0 * x - 0.5 * y)
    assert O.rel_err(axy, 2.0 * ax - 0.5 * ay) < 1e-12
    assert abs(x @ ay - y @ ax) < 1e-11 * abs(x @ ay)
    ax32 = op.apply(x, prec=icc.ICC_F32)
    assert O.rel_err(ax32, ax) < 2e-6
    # rigid translation in y: K t = 0
    t = np.zeros(g.ndof)
    t[g.ndof_sc:2 * g.ndof_sc] = 1.0
    assert np.linalg.norm(op.apply(t)) < 1e-12 * np.linalg.norm(ax) * np.linalg.norm(t) / np.linalg.norm(x)


@pytest.mark.parametrize("nel", [[4, 3, 2], [8, 5, 3], [12, 7, 6]])
def test_wq_diag_matches_oracle(ctx, nel):
    """Diagonal of the WQ stiffness (wq.rs:329-514, interior layout, element
    densities) vs the oracle's assembled WQ matrix; at uniform density it is
    the Gauss diagonal (wq_vs_matlab.rs:35-51)."""
    bounds = [[0, 1.2], [0, 0.6], [0, 0.3]]
    g = icc.Grid(ctx, nel, bounds)
    sp, _, _ = _oracle_space(bounds, nel)
    x = np.random.default_rng(9).uniform(0.1, 1.0, g.nel_tot)
    sp.assemble_wq(1.0, 0.3, 1, x, 3.0, 1e-3, layout=0)
    assert O.rel_err(g.wq_diag(x, 3.0, 1e-3), sp.diag()) < 1e-12
    sp.assemble(None)
    assert O.rel_err(g.wq_diag(None), sp.diag()) < 1e-12


def _cylinders(seed, n=3):
    rng = np.random.default_rng(seed)
    return np.array([[rng.uniform(0.3, 1.7), rng.uniform(0.25, 0.75), rng.uniform(0.08, 0.15)] for _ in range(n)])


@pytest.mark.parametrize("res,padded", [([32, 16, 16], False), ([24, 14, 9], True), ([128, 64, 64], False)])
def test_classification_bit_exact(ctx, res, padded):
    """Cell status (INSIDE/OUTSIDE/CUT) and 4^3 sub-cell volume fractions
    (immersed.rs:88-155) with the box-minus-cylinders predicate: bit-exact
    with the oracle; the active list (w > 0) follows."""
    lo, hi = [0.0, 0.0, 0.0], [2.0, 1.0, 1.0]
    gb = [[-0.1, 2.1], [-0.05, 1.05], [-0.05, 1.05]] if padded else [[0.0, 2.0], [0.0, 1.0], [0.0, 1.0]]
    cyl = _cylinders(sum(res))
    st_ref, w_ref = O.classify_cylinders(lo, hi, cyl, gb, res)
    st, w = icc.classify_cylinders(ctx, lo, hi, cyl, gb, res)
    assert np.array_equal(st, st_ref)
    assert np.array_equal(w, w_ref)
    assert np.array_equal(np.flatnonzero(w > 0), np.flatnonzero(w_ref > 0))
    # INSIDE cells weigh 1, OUTSIDE 0; a CUT cell may sample 0..sub^3 points inside
    assert set(np.unique(st)) <= {-1, 0, 1}
    assert np.all(w[st == 1] == 1.0) and np.all(w[st == 0] == 0.0)
    assert np.all((w >= 0.0) & (w <= 1.0))


def test_tiling_variants_and_autotune(ctx):
    """Every tiling variant and workgroup count gives the same masked operator
    (the per-tile mask flags follow the re-tiling); the autotuner keeps one."""
    nel = [20, 10, 6]
    bounds, sp, s, free, f = _cantilever_problem(nel, 17)
    sp.assemble(s)
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    x = np.random.default_rng(5).standard_normal(g.ndof)
    ref = free * sp.matvec(free * x)
    for v in range(12):
        for wg in (1, 2, 4):
            g.set_tiling(v, wg)
            assert g.tiling[:2] == (v, wg)
            assert O.rel_err(op.apply(x), ref) < 1e-13, (v, wg)
            assert O.rel_err(op.apply(x, prec=icc.ICC_F32), ref) < 2e-6, (v, wg)
    v, wg, us = op.autotune(icc.ICC_F64, reps=3)
    assert g.tiling[:2] == (v, wg) and us > 0
    assert O.rel_err(op.apply(x), ref) < 1e-13


def test_balanced_partition_splits_tiles_and_keeps_the_operator(ctx):
    """On the 96x48x48 grid the work-balanced partition (variants 0..5) splits
    tiles; every partition and workgroup count gives the same masked operator
    (FP64 to round-off of the scatter order; FP32 and BF16 within their
    rounding; the clamped face puts fixed DOFs into split tiles too) and the
    same PCG iterates (the per-workgroup p.Ap partials)."""
    nel = (96, 48, 48)
    g = icc.Grid(ctx, nel, [[0, 2.0], [0, 1.0], [0, 1.0]])
    rng = np.random.default_rng(21)
    s = 1e-9 + (1 - 1e-9) * rng.uniform(0.01, 1.0, g.nel_tot) ** 3
    free = np.ones(g.ndof, dtype=np.uint8)
    ix = np.arange(g.ndof_sc) % g.ncp[0]
    for c in range(3):
        free[c * g.ndof_sc:(c + 1) * g.ndof_sc][ix == 0] = 0
    op = icc.Operator(g, scale=s, free=free)
    x = rng.standard_normal(g.ndof)
    b = free * rng.standard_normal(g.ndof)
    g.set_tiling(6, 2)
    plain_tiles = g.tiling[2]
    ref = op.apply(x)
    ref32 = op.apply(x, prec=icc.ICC_F32)
    refb = op.apply(x, prec=icc.ICC_BF16)
    diag = np.maximum(op.diag(), 1e-12)
    xp, _, _, _ = op.pcg(b, diag, 0.0, 30)
    assert np.array_equal(ref * (1 - free), np.zeros(g.ndof))
    for v, wg in ((0, 2), (1, 2), (4, 2), (0, 1), (8, 4)):
        g.set_tiling(v, wg)
        if v == 0:
            assert g.tiling[2] > plain_tiles  # tiles were split
        assert O.rel_err(op.apply(x), ref) < 1e-14, (v, wg)
        assert O.rel_err(op.apply(x, prec=icc.ICC_F32), ref32) < 1e-6, (v, wg)
        assert O.rel_err(op.apply(x, prec=icc.ICC_BF16), refb) < 1e-5, (v, wg)
        xv, _, _, _ = op.pcg(b, diag, 0.0, 30)
        assert O.rel_err(xv, xp) < 1e-10, (v, wg)


def _random_spd_addon(n, rng, rows_by_len):
    """Symmetric, diagonally dominant sparse add-on whose rows have the given
    numbers of off-diagonal entries (before symmetrisation), as CSR arrays."""
    import scipy.sparse as sps
    r, c, v = [], [], []
    for ln, rows in rows_by_len:
        for i in rows:
            cols = rng.choice(n, ln, replace=False)
            r += [i] * ln
            c += list(cols)
            v += list(rng.uniform(-1.0, 1.0, ln))
    t = sps.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    s = (t + t.T).tocsr()
    s = s + sps.diags(np.asarray(abs(s).sum(axis=1)).ravel() + 1e-3 * (np.asarray(abs(s).sum(axis=1)).ravel() > 0))
    s = s.tocsr()
    s.sort_indices()
    return s


def test_sparse_addon_long_rows_apply_and_pcg(ctx):
    # the CSR add-on (y = S x beside the element term) with rows of 0 to a few
    # hundred entries: the wave-per-row product takes 64 entries per pass, so
    # rows of 64, 65, 130 and more exercise its remainder passes; applies in
    # FP64 and FP32 against scipy, and a PCG solve of M (K + S) M u = f whose
    # true residual is recomputed with the oracle's K and scipy's S
    nel = [8, 4, 4]
    bounds, sp, s, free, f = _cantilever_problem(nel, 17, rho=(0.3, 1.0), emin=1e-3)
    n = sp.ndof
    rng = np.random.default_rng(29)
    pick = rng.permutation(n)
    S = _random_spd_addon(n, rng, [(10, pick[:40]), (63, pick[40:50]), (64, pick[50:55]), (65, pick[55:60]),
                                   (130, pick[60:64]), (300, pick[64:66])])
    assert S.getnnz(axis=1).max() > 300
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    ops = icc.Operator(g, scale=s, free=free)
    ops.set_sparse(S.indptr, S.indices, S.data)
    m = free.astype(np.float64)
    for trial in range(2):
        x = rng.standard_normal(n)
        ref = m * (S @ (m * x))
        y = ops.apply(x) - op.apply(x)
        assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
        y32 = ops.apply(x, prec=icc.ICC_F32) - op.apply(x, prec=icc.ICC_F32)
        assert np.linalg.norm(y32 - ref) <= 2e-6 * np.linalg.norm(ops.apply(x))
    d = np.maximum(ops.diag(), 1e-12)
    u, it, res, rc = ops.pcg(f, d, 1e-10, 20 * n)
    assert rc == icc.ICC_OK and res < 1e-10
    sp.assemble(s)
    r = m * (f - sp.matvec(m * u) - S @ (m * u))
    assert np.linalg.norm(r) / np.linalg.norm(f) < 1e-9
    op.close()
    ops.close()
    g.close()
