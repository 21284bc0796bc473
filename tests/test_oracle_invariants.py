"""Invariants that pin the oracle beyond the fixtures (the mathematics fixes
them): rigid-body null space and symmetry of K (iga_kernel.rs:50-92), the
structured filter rows equal the O(nel^2) scan of topopt_iga.rs:109-125, the
consistent traction integrates to t * area, the density filter is adjoint to
its backward (structured and design-domain variants), the design-domain
filter reduces to the structured one on a full box of cubic cells, the WQ operator equals Gauss at uniform density
(wq_vs_matlab.rs:35-51)."""
import numpy as np

import oracle_py as O


def test_rigid_body_modes_and_symmetry():
    sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], [3, 2, 2], 2)
    lam, mu = O.lame(1.0, 0.3)
    types, _ = sp.elem_mats(lam, mu)
    for K in types:
        assert np.max(np.abs(K - K.T)) < 1e-13 * np.abs(K).max()
    sp.assemble(None)
    g = sp.greville()
    fro = sp.frobenius()
    modes = []
    for c in range(3):
        m = np.zeros(sp.ndof)
        m[c * sp.ndof_sc:(c + 1) * sp.ndof_sc] = 1.0
        modes.append(m)
    for a, b in [(0, 1), (1, 2), (0, 2)]:
        m = np.zeros(sp.ndof)
        m[a * sp.ndof_sc:(a + 1) * sp.ndof_sc] = -g[:, b]
        m[b * sp.ndof_sc:(b + 1) * sp.ndof_sc] = g[:, a]
        modes.append(m)
    for m in modes:
        assert np.linalg.norm(sp.matvec(m)) / (fro * np.linalg.norm(m)) < 1e-13


def test_assembly_paths_agree():
    sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], [4, 3, 3], 2)
    lam, mu = O.lame(1.0, 0.3)
    sp.elem_mats(lam, mu)
    rng = np.random.default_rng(0)
    s = rng.uniform(0.01, 1.0, sp.nel)
    v = rng.standard_normal(sp.ndof)
    sp.assemble(s)
    y1 = sp.matvec(v)
    sp.assemble(s, pattern=True)
    y2 = sp.matvec(v)
    assert O.rel_err(y1, y2) < 1e-14


def test_structured_filter_equals_bruteforce():
    for nel, rmin in [([8, 4, 2], 1.5), ([7, 5, 3], 2.3), ([6, 6, 6], 1.0)]:
        sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], nel, 2)
        assert sp.filter_rows_match(rmin)


def test_density_filter_adjoint():
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], [8, 4, 4], 2)
    rng = np.random.default_rng(1)
    x = rng.uniform(size=sp.nel)
    g = rng.standard_normal(sp.nel)
    fx = sp.filter_apply(1.5, 1, None, x)
    bg = sp.filter_apply(1.5, 2, None, g)
    assert abs(fx @ g - x @ bg) < 1e-12 * abs(fx @ g)


def test_design_filter_all_active_equals_structured():
    """With cubic elements and every cell active, the design-domain weights
    r - |d h| are h times the structured weights rmin - |d| (the boundary
    offsets weigh zero in both), so every normalised mode agrees."""
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], [8, 4, 4], 2)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.001, 1.0, sp.nel)
    v = rng.standard_normal(sp.nel)
    act = np.ones(sp.nel, dtype=np.uint8)
    for rmin in (1.0, 1.5, 2.3):
        for mode in (0, 1, 2):
            a = sp.design_filter_apply(rmin, mode, act, x, v)
            b = sp.filter_apply(rmin, mode, x if mode == 0 else None, v)
            assert O.rel_err(a, b) < 1e-14, (rmin, mode)


def test_design_filter_support_and_adjoint():
    """Inactive rows are empty, inactive cells never contribute, and the
    backward mode is the adjoint of the forward mode."""
    sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], [9, 6, 4], 2)
    rng = np.random.default_rng(4)
    act = (rng.uniform(size=sp.nel) < 0.7).astype(np.uint8)
    x = rng.uniform(0.001, 1.0, sp.nel)
    v = rng.standard_normal(sp.nel)
    g = rng.standard_normal(sp.nel)
    rmin = 1.7
    outs = [sp.design_filter_apply(rmin, m, act, x, v) for m in (0, 1, 2)]
    for o in outs:
        assert np.all(o[act == 0] == 0.0)
    v2 = v.copy()
    v2[act == 0] = 1e6
    x2 = x.copy()
    x2[act == 0] = 0.5
    for m, o in zip((0, 1, 2), outs):
        assert np.array_equal(sp.design_filter_apply(rmin, m, act, x2, v2)[act == 1], o[act == 1])
    # adjoint on the active subspace (inactive entries of g are ignored)
    va = v * act
    fx = sp.design_filter_apply(rmin, 1, act, None, va)
    bg = sp.design_filter_apply(rmin, 2, act, None, g * act)
    assert abs(fx @ (g * act) - va @ bg) < 1e-12 * abs(fx @ (g * act))
    # the forward filter preserves constants on each active cell with a neighbour set
    one = sp.design_filter_apply(rmin, 1, act, None, act.astype(np.float64))
    assert np.max(np.abs(one[act == 1] - 1.0)) < 1e-14


def test_traction_total_force():
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], [6, 3, 4], 2)
    f = sp.traction_xmax([0.0, -1.0, 0.0])
    nsc = sp.ndof_sc
    assert abs(f[nsc:2 * nsc].sum() - (-1.0 * 1.0 * 1.0)) < 1e-13
    assert abs(f[:nsc].sum()) < 1e-15 and abs(f[2 * nsc:].sum()) < 1e-15


def test_wq_equals_gauss_uniform_density():
    sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], [4, 3, 2], 2)
    lam, mu = O.lame(1.0, 0.3)
    sp.elem_mats(lam, mu)
    v = O.probe_vectors(sp.ndof, 1)[0]
    sp.assemble(None)
    yg = sp.matvec(v)
    dg = sp.diag()
    for layout in (0, 1):
        sp.assemble_wq(1.0, 0.3, 0, None, 3.0, 1e-3, layout=layout)
        assert O.rel_err(sp.matvec(v), yg) < 1e-12
        assert O.rel_err(sp.diag(), dg) < 1e-12


def test_sensitivity_identity():
    """dc = -p x^(p-1)(1-emin) u_e^T K_e u_e (testIgaKernel.m:130-142) equals the
    quadrature form fast_sensitivities (topopt2d.rs:13-48)."""
    sp = O.Space([[0, 1.2], [0, 0.6], [0, 0.3]], [4, 3, 2], 2)
    lam, mu = O.lame(1.0, 0.3)
    sp.elem_mats(lam, mu)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(sp.ndof)
    x = rng.uniform(0.1, 1.0, sp.nel)
    ee = sp.energies(u)
    dc = -3.0 * x ** 2 * (1 - 1e-3) * ee
    assert O.rel_err(sp.fast_sens(u, x, False, 3.0, 1e-3, 1.0, 0.3), dc) < 1e-11


def test_structured_csr_equals_assembly_pattern():
    """The CPU-baseline CSR builder equals AssemblyPattern (sparse.rs:128-175)
    bit for bit, pattern and values."""
    for nel in ([4, 3, 3], [7, 5, 2], [1, 2, 3]):
        sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], nel, 2)
        lam, mu = O.lame(1.0, 0.3)
        sp.elem_mats(lam, mu)
        s = np.random.default_rng(3).uniform(0.01, 1.0, sp.nel) ** 3
        assert sp.assemble_structured(s) > 0
        assert sp.compare_with_pattern(s) == 0


def test_sdf_immersed_problem_invariants():
    """The analytic immersed problem of the oracle: traction total = t * face
    area, penalty normalisation beta * area on a rigid translation (K 1 = 0),
    symmetry, support mask, and mode 0 = reference scale on every cell."""
    cyl = O.seeded_cylinders(5)
    lo, hi = [0.0, 0.0, 0.0], [2.0, 1.0, 1.0]
    pb = O.SdfImmersed(lo, hi, cyl, [16, 8, 8], mode=1)
    nsc = pb.ndof // 3
    assert abs(pb.force[nsc:2 * nsc].sum() + 1.0) < 1e-13 and abs(pb.force[:nsc].sum()) == 0.0
    assert abs(pb.beta - 1e3 * 1.0 / 0.125) < 1e-9
    # weights: a cylinder of radius r removes pi r^2 per unit length from the box volume
    vol = pb.weights.sum() * (2.0 / 16) * (1.0 / 8) * (1.0 / 8)
    exact = 2.0 - np.pi * (cyl[:, 2] ** 2).sum()
    assert abs(vol - exact) < 0.02 * exact
    assert np.array_equal(pb.active, (pb.weights >= 0.05).astype(np.uint8))
    assert np.array_equal(pb.scale, np.where(pb.weights >= 0.05, pb.weights, 0.0))
    # rigid translation in y: only the penalty acts, total = beta * area of the x = 0 face
    t = np.zeros(pb.ndof)
    t[nsc:2 * nsc] = 1.0
    y = pb.matvec(t * pb.free)
    assert abs(y[nsc:2 * nsc].sum() - pb.beta * 1.0) < 1e-9 * pb.beta
    assert np.abs(y[:nsc]).max() < 1e-9 * pb.beta
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(pb.ndof), rng.standard_normal(pb.ndof)
    assert abs(a @ pb.matvec(b) - b @ pb.matvec(a)) < 1e-11 * abs(a @ pb.matvec(b))
    # DOFs outside every active cell's support are fixed: with a hole large
    # enough to leave unsupported control points, the mask is the union of
    # the supports of the active cells
    big = O.SdfImmersed(lo, hi, [[1.0, 0.5, 0.36]], [16, 8, 8], mode=1)
    nel = (16, 8, 8)
    act = big.active.reshape(nel[2], nel[1], nel[0])
    sup = np.zeros((nel[2] + 2, nel[1] + 2, nel[0] + 2), dtype=bool)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                sup[dz:dz + nel[2], dy:dy + nel[1], dx:dx + nel[0]] |= act.astype(bool)
    assert np.array_equal(big.free, np.tile(sup.ravel(), 3).astype(np.uint8))
    assert 0 < big.free.sum() < big.ndof
    pb0 = O.SdfImmersed(lo, hi, cyl, [16, 8, 8], mode=0, emin=1e-4)
    assert np.array_equal(pb0.scale, 1e-4 + (1 - 1e-4) * pb0.weights) and pb0.free.all()


def _open_knots(nel, p):
    return np.concatenate([np.zeros(p), np.arange(nel + 1) / nel, np.ones(p)])


def _blossom_coeffs(nel, p, m):
    """B-spline coefficients of xi^m (m <= p) on the open uniform knots:
    Marsden's identity, c_i = e_m(t_(i+1), ..., t_(i+p)) / binom(p, m)."""
    from itertools import combinations
    from math import comb
    t = _open_knots(nel, p)
    out = []
    for i in range(nel + p):
        args = t[i + 1:i + p + 1]
        out.append(sum(np.prod(c) for c in combinations(args, m)) / comb(p, m) if m > 0 else 1.0)
    return np.asarray(out)


def test_ghost_penalty_annihilates_polynomials_up_to_degree_p():
    """immersed_cad_bc.rs:63-93: j(u, v) vanishes on global polynomials of
    degree <= p (fields built here from Marsden's identity, not by the
    reference's interpolation) and penalises a generic field."""
    bounds = [0.0, 1.3, -0.2, 0.7, 0.1, 0.6]
    res = (6, 5, 4)
    for p in (2, 3):
        nd = [r + p for r in res]
        nsc = int(np.prod(nd))
        status = -np.ones(int(np.prod(res)), dtype=np.int8)
        rnd = ((np.arange(3 * nsc) * 7919) % 101) / 101.0
        y, fro = O.ghost_penalty_matvec(bounds, res, p, status, 1.0, 1.0, rnd)
        assert fro > 0.0 and np.linalg.norm(y) > 1e-3 * fro * np.linalg.norm(rnd)
        for d in range(3):
            for m in range(p + 1):
                cd = _blossom_coeffs(res[d], p, m)
                u = np.zeros(3 * nsc)
                idx = np.arange(nsc)
                s = [idx % nd[0], (idx // nd[0]) % nd[1], idx // (nd[0] * nd[1])]
                u[((d + 1) % 3) * nsc + idx] = cd[s[d]]
                y, _ = O.ghost_penalty_matvec(bounds, res, p, status, 1.0, 1.0, u)
                assert np.linalg.norm(y) < 1e-10 * fro * np.linalg.norm(u), (p, d, m)


def test_ghost_penalty_single_interior_face_closed_form():
    """One x-face deep inside a uniform p = 2 grid: the block is
    coef (j j^T) (x) M_y (x) M_z per component, with the jump of the second
    derivative j = (-1, 3, -3, 1) / h_x^2, the uniform quadratic element mass
    M = h/120 [[6,13,1],[13,54,13],[1,13,6]] and coef = gamma E h_mean^3."""
    bounds = [0.0, 1.2, 0.0, 0.8, 0.0, 1.0]
    res = (8, 8, 8)
    h = np.array([1.2, 0.8, 1.0]) / 8
    gamma, young = 0.37, 2.5
    nel = 512
    status = np.zeros(nel, dtype=np.int8)
    lin = lambda i, j, k: i + 8 * (j + 8 * k)
    status[lin(3, 4, 4)] = -1
    status[lin(4, 4, 4)] = 1
    nd = 10
    nsc = nd ** 3
    jv = np.array([-1.0, 3.0, -3.0, 1.0]) / h[0] ** 2
    M = lambda hh: hh / 120.0 * np.array([[6.0, 13, 1], [13, 54, 13], [1, 13, 6]])
    coef = gamma * young * np.mean(h) ** 3
    # local DOFs: x index 3..6, y index 4..6, z index 4..6
    loc = [(t, a, b) for b in range(3) for a in range(3) for t in range(4)]
    gl = [(3 + t) + nd * ((4 + a) + nd * (4 + b)) for (t, a, b) in loc]
    blk = np.array([[coef * jv[t1] * jv[t2] * M(h[1])[a1, a2] * M(h[2])[b1, b2] for (t2, a2, b2) in loc]
                    for (t1, a1, b1) in loc])
    for c in (0, 2):
        for col in (0, 13, 35):
            e = np.zeros(3 * nsc)
            e[c * nsc + gl[col]] = 1.0
            y, _ = O.ghost_penalty_matvec(bounds, res, 2, status, gamma, young, e)
            want = np.zeros(3 * nsc)
            want[c * nsc + np.asarray(gl)] = blk[:, col]
            assert np.max(np.abs(y - want)) < 1e-13 * np.max(np.abs(blk)), (c, col)
