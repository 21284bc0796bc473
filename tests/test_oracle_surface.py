"""Pins the oracle's restatement of the surface boundary terms
(rust/src/immersed_bc.rs: scalar_basis_at, surface_quadrature,
SurfaceAccumulator, surface_boundary_terms, dirichlet_penalty_fn) and of
ImmersedProblem::add_surface_terms (rust/src/immersed.rs:541-548).

What fixes each function (none of it is a reference vector; the reference
holds none for these functions):
* the Dunavant rule: exact integrals of every monomial of degree <= 4 on a
  triangle, a!b!/(a+b+2)!, and inexact at degree 6;
* the splitting: areas add up, every sub-edge is <= max_edge, 6 * 4^k points;
* scalar_basis_at: partition of unity, nonnegativity, reproduction of x, y,
  z through the Greville abscissae (exact for B-splines), and the support =
  the element's connectivity;
* the accumulator: u^T K_s u = beta * area for a rigid translation (the
  invariant of immersed_cad_bc.rs), the 3x3 structure of the Robin term, a
  dense brute-force assembly of beta * sum_q w_q N_a N_b;
* traction and pressure loads: the totals t * A and -p * A * n
  (immersed_cad_bc.rs:95-112);
* dirichlet_penalty_fn: sum_a f_a = beta * int g for an affine g;
* the immersed cantilever of immersed_cad_bc.rs:115-151 (penalty clamp,
  tip traction): tip deflection within 5 % of Timoshenko beam theory.
"""
import math

import numpy as np
import pytest

import oracle_py as O

BOX_TRIS = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7], [2, 7, 6],
                     [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.int64)


def box_nodes(mn, mx):
    """TriangleMesh3D::new_box vertices (cut_cell.rs:27-53)."""
    a, b = mn, mx
    return np.array([[a[0], a[1], a[2]], [b[0], a[1], a[2]], [b[0], b[1], a[2]], [a[0], b[1], a[2]],
                     [a[0], a[1], b[2]], [b[0], a[1], b[2]], [b[0], b[1], b[2]], [a[0], b[1], b[2]]], dtype=np.float64)


def test_dunavant_rule_exact_to_degree_four():
    nodes = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0]])
    x, w, n = O.surface_quadrature(nodes, np.array([[0, 1, 2]]), [0], 10.0)
    assert len(w) == 6
    np.testing.assert_allclose(n, np.tile([0.0, 0.0, 1.0], (6, 1)))
    for a in range(5):
        for b in range(5 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = float(np.sum(w * x[:, 0] ** a * x[:, 1] ** b))
            assert abs(got - exact) < 2e-14, (a, b, got, exact)
    # degree 6 is not integrated exactly (a rule of higher degree would pass this)
    assert abs(float(np.sum(w * x[:, 0] ** 6)) - math.factorial(6) / math.factorial(8)) > 1e-6


def test_splitting_preserves_area_and_bounds_edges():
    rng = np.random.default_rng(3)
    nodes = rng.uniform(-1, 1, (3, 3))
    e = [np.linalg.norm(nodes[i] - nodes[(i + 1) % 3]) for i in range(3)]
    area = 0.5 * np.linalg.norm(np.cross(nodes[1] - nodes[0], nodes[2] - nodes[0]))
    for max_edge in (10.0, max(e) * 0.9, max(e) * 0.3, max(e) * 0.1):
        x, w, n = O.surface_quadrature(nodes, np.array([[0, 1, 2]]), [0], max_edge)
        k = 0
        while max(e) / 2 ** k > max_edge:
            k += 1
        assert len(w) == 6 * 4 ** k
        assert abs(w.sum() - area) < 1e-13 * area
        # every point lies in the triangle's plane
        nrm = np.cross(nodes[1] - nodes[0], nodes[2] - nodes[0])
        assert np.max(np.abs((x - nodes[0]) @ nrm)) < 1e-13


def test_scalar_basis_at_partition_support_and_reproduction():
    bounds = [[0.0, 1.3], [-0.2, 0.7], [0.1, 0.6]]
    nsub = [6, 5, 4]
    sp = O.Space(bounds, nsub, 2)
    conn = sp.connectivity().reshape(sp.nel, 81)[:, :27]
    rng = np.random.default_rng(5)
    gre = np.asarray(sp.greville()).reshape(-1, 3)  # [ndof_sc, 3] Greville abscissae (physical)
    for _ in range(40):
        p = np.array([rng.uniform(lo, hi) for lo, hi in bounds])
        idx, val = O.scalar_basis_at(sp, p)
        assert len(idx) == 27
        assert abs(val.sum() - 1.0) < 1e-14
        assert np.all(val >= 0.0)
        e = [min(int((p[d] - bounds[d][0]) / (bounds[d][1] - bounds[d][0]) * nsub[d]), nsub[d] - 1) for d in range(3)]
        el = e[0] + nsub[0] * (e[1] + nsub[1] * e[2])
        assert sorted(idx) == sorted(conn[el]), (p, el)
        # x_d = sum_a G_d(a) N_a(p): linear reproduction through the Greville points
        for d in range(3):
            assert abs(float(val @ gre[idx, d]) - p[d]) < 1e-13
    # points outside the grid are clamped to the boundary
    idx_o, val_o = O.scalar_basis_at(sp, [-5.0, 0.25, 0.35])
    idx_b, val_b = O.scalar_basis_at(sp, [0.0, 0.25, 0.35])
    assert np.array_equal(idx_o, idx_b) and np.array_equal(val_o, val_b)


def _csr_dense(r, c, v, n):
    K = np.zeros((n, n))
    np.add.at(K, (r, c), v)
    return K


def test_dirichlet_block_equals_dense_assembly_and_rigid_translation():
    mn, mx = [0.1, 0.05, 0.2], [0.9, 0.55, 0.6]
    nodes = box_nodes(mn, mx)
    sp = O.Space([[0.0, 1.0], [0.0, 0.6], [0.1, 0.7]], [5, 4, 3], 2)
    beta = 37.5
    sel = [0, 1, 8, 9]  # -z and -x faces
    surf = O.Surface(sp)
    surf.add(O.Surface.DIRICHLET, nodes, BOX_TRIS, sel, beta=beta, comps=(1, 1, 1), values=(0.5, -1.0, 2.0))
    r, c, v, f = surf.terms()
    n = sp.ndof
    K = _csr_dense(r, c, v, n)
    # brute force: beta * sum_q w_q N_a N_b per component, from the oracle's own points
    hmin = min(1.0 / 5, 0.6 / 4, 0.6 / 3)
    x, w, _ = O.surface_quadrature(nodes, BOX_TRIS, sel, 0.25 * hmin)
    Kb = np.zeros((n, n))
    fb = np.zeros(n)
    for q in range(len(w)):
        idx, val = O.scalar_basis_at(sp, x[q])
        for comp, g in enumerate((0.5, -1.0, 2.0)):
            ii = comp * sp.ndof_sc + idx
            Kb[np.ix_(ii, ii)] += beta * w[q] * np.outer(val, val)
            fb[ii] += beta * w[q] * val * g
    assert np.max(np.abs(K - Kb)) < 1e-12 * np.max(np.abs(Kb))
    assert np.max(np.abs(f - fb)) < 1e-12 * np.max(np.abs(fb))
    # (k N_a) N_b and (k N_b) N_a round differently: symmetric to round-off only, as in the reference
    assert np.max(np.abs(K - K.T)) < 1e-15 * np.max(np.abs(K))
    area = 0.8 * 0.5 + 0.5 * 0.4  # -z face (0.8 x 0.5) and -x face (0.5 x 0.4)
    for comp in range(3):
        u = np.zeros(n)
        u[comp * sp.ndof_sc:(comp + 1) * sp.ndof_sc] = 1.0
        assert abs(u @ K @ u - beta * area) < 1e-12 * beta * area
    # the load is beta * int g_c over the faces
    for comp, g in enumerate((0.5, -1.0, 2.0)):
        assert abs(f[comp * sp.ndof_sc:(comp + 1) * sp.ndof_sc].sum() - beta * g * area) < 1e-12 * beta * area


def test_traction_pressure_robin_totals():
    (l, b, h) = (4.0, 0.5, 0.5)
    nodes = box_nodes([0, 0, 0], [l, b, h])
    # immersed_cad_bc.rs:95-112: grid [16, 3, 3] on the padded bounds
    imm = O.Immersed(nodes, BOX_TRIS, (16, 3, 3), stab_kind=2, gamma=1e-2, emin=1e-4, clamp=-1)
    tip = [10, 11]
    surf = O.Surface(imm, immersed=True)
    surf.add(O.Surface.TRACTION, nodes, BOX_TRIS, tip, values=(0.0, -2.0, 0.0))
    r, c, v, f = surf.terms()
    nsc = imm.ndof // 3
    assert len(v) == 0
    assert abs(f[nsc:2 * nsc].sum() - (-2.0 * b * h)) < 1e-12
    assert abs(f[:nsc].sum()) < 1e-12 and abs(f[2 * nsc:].sum()) < 1e-12
    x, w, _ = O.surface_quadrature(nodes, BOX_TRIS, tip, 0.05)
    assert abs(w.sum() - b * h) < 1e-12
    # pressure on the closed surface integrates to zero; on the tip face to -p A n
    s2 = O.Surface(imm, immersed=True)
    s2.add(O.Surface.PRESSURE, nodes, BOX_TRIS, np.arange(12), pressure=3.0)
    f2 = s2.terms()[3]
    for comp in range(3):
        assert abs(f2[comp * nsc:(comp + 1) * nsc].sum()) < 1e-12
    s3 = O.Surface(imm, immersed=True)
    s3.add(O.Surface.PRESSURE, nodes, BOX_TRIS, tip, pressure=3.0)
    f3 = s3.terms()[3]
    assert abs(f3[:nsc].sum() + 3.0 * b * h) < 1e-12
    # Robin: kn n n^T + kt (I - n n^T); on the x = 0 face, n = -e_x
    s4 = O.Surface(imm, immersed=True)
    s4.add(O.Surface.ROBIN, nodes, BOX_TRIS, [8, 9], kn=5.0, kt=2.0)
    r4, c4, v4, f4 = s4.terms()
    K4 = _csr_dense(r4, c4, v4, imm.ndof)
    assert np.max(np.abs(f4)) == 0.0
    for comp, k in enumerate((5.0, 2.0, 2.0)):
        u = np.zeros(imm.ndof)
        u[comp * nsc:(comp + 1) * nsc] = 1.0
        assert abs(u @ K4 @ u - k * b * h) < 1e-12 * k * b * h
    # no coupling between components for an axis-aligned normal
    assert np.all(np.asarray(r4) // nsc == np.asarray(c4) // nsc)


def test_dirichlet_penalty_fn_affine_load():
    nodes = box_nodes([0.2, 0.1, 0.15], [0.8, 0.7, 0.45])
    imm = O.Immersed(nodes, BOX_TRIS, (6, 6, 4), stab_kind=0, clamp=-1)
    g0 = np.array([0.3, -0.2, 0.1])
    G = np.array([[1.0, 0.5, 0.0], [0.0, 2.0, -1.0], [0.25, 0.0, 0.5]])
    beta = 11.0
    surf = O.Surface(imm, immersed=True)
    surf.dirichlet_affine(nodes, BOX_TRIS, np.arange(12), beta, g0, G)
    r, c, v, f = surf.terms()
    nsc = imm.ndof // 3
    # beta * int_S g dA: int_S dA = total area, int_S x dA = area * centroid for each face
    lo, hi = np.array([0.2, 0.1, 0.15]), np.array([0.8, 0.7, 0.45])
    ext = hi - lo
    total_area, first = 0.0, np.zeros(3)
    for d in range(3):
        a = ext[(d + 1) % 3] * ext[(d + 2) % 3]
        for side in (lo[d], hi[d]):
            cen = 0.5 * (lo + hi)
            cen[d] = side
            total_area += a
            first += a * cen
    for comp in range(3):
        exact = beta * (g0[comp] * total_area + G[comp] @ first)
        assert abs(f[comp * nsc:(comp + 1) * nsc].sum() - exact) < 1e-11 * abs(exact)
    K = _csr_dense(r, c, v, imm.ndof)
    u = np.zeros(imm.ndof)
    u[:nsc] = 1.0
    assert abs(u @ K @ u - beta * total_area) < 1e-12 * beta * total_area


@pytest.mark.parametrize("stab", [2])
def test_immersed_cantilever_tip_deflection_near_timoshenko(stab):
    """immersed_cad_bc.rs:115-151 on the oracle: penalty clamp (factor 1e3)
    on x = 0, traction on x = L, ghost penalty 1e-2, grid 44 x 6 x 6."""
    (l, b, h) = (4.0, 0.5, 0.5)
    p_total = 1e-3
    nodes = box_nodes([0, 0, 0], [l, b, h])
    imm = O.Immersed(nodes, BOX_TRIS, (44, 6, 6), stab_kind=stab, gamma=1e-2, emin=1e-4, clamp=-1)
    imm.zero_force()
    hmin = min((imm.bounds[2 * d + 1] - imm.bounds[2 * d]) / n for d, n in enumerate((44, 6, 6)))
    surf = O.Surface(imm, immersed=True)
    surf.add(O.Surface.DIRICHLET, nodes, BOX_TRIS, [8, 9], beta=1e3 * 1.0 / hmin, comps=(1, 1, 1))
    surf.add(O.Surface.TRACTION, nodes, BOX_TRIS, [10, 11], values=(0.0, -p_total / (b * h), 0.0))
    imm.add_surface(surf)
    u, comp, it, res = imm.solve(1e-10, 20 * imm.ndof)
    assert res < 1e-9
    x, w, _ = O.surface_quadrature(nodes, BOX_TRIS, [10, 11], 0.05)
    s = sum(wq * imm.eval_vector_at(u, xq)[1] for xq, wq in zip(x, w))
    delta = -s / w.sum()
    e, nu = 1.0, 0.3
    g = e / (2 * (1 + nu))
    inertia = b * h ** 3 / 12
    timo = p_total * l ** 3 / (3 * e * inertia) + p_total * l / (5 / 6 * g * b * h)
    assert abs(delta - timo) / timo < 0.05, (delta, timo)
