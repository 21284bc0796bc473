"""The opt-in deterministic mode (icc_op_set_deterministic): per-element
outputs and a fixed-order gather in place of the atomic tile scatter, so that
applies, PCG runs and mixed solves repeat bit for bit, as the reference's
row-parallel CSR product does (rust/src/sparse.rs:60-68).  Parity with the
oracle's assembled operator to the usual tolerances; repeatability exact."""
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


GRIDS = [
    ([[0.5, 2.5], [-1, 0.2], [0.1, 0.4]], [4, 3, 3]),
    ([[0, 2.0], [0, 1.0], [0, 1.0]], [16, 8, 8]),
    ([[0, 3.0], [0, 1.0], [0, 1.0]], [13, 7, 3]),
    ([[0, 1.0], [0, 1.0], [0, 1.0]], [1, 1, 1]),
]


def _cantilever(ctx, nel, seed, det=True):
    bounds = [[0, 2.0], [0, 1.0], [0, 1.0]]
    rng = np.random.default_rng(seed)
    sp = O.Space(bounds, nel, 2)
    sp.elem_mats(*O.lame(1.0, 0.3))
    s = 1e-9 + (1 - 1e-9) * rng.uniform(0.01, 1.0, sp.nel) ** 3
    i = np.arange(sp.ndof_sc) % sp.ndof_dir[0]
    free = np.ones(sp.ndof, dtype=np.uint8)
    for c in range(3):
        free[c * sp.ndof_sc:(c + 1) * sp.ndof_sc][i == 0] = 0
    f = sp.traction_xmax([0.0, -1.0, 0.0]) * free
    g = icc.Grid(ctx, nel, bounds)
    op = icc.Operator(g, scale=s, free=free)
    if det:
        op.set_deterministic(True)
    return sp, s, free, f, g, op


@pytest.mark.parametrize("bounds,nel", GRIDS)
def test_det_apply_matches_oracle_and_repeats(ctx, bounds, nel):
    rng = np.random.default_rng(sum(nel) + 7)
    g = icc.Grid(ctx, nel, bounds)
    sp = O.Space(bounds, nel, 2)
    sp.elem_mats(*O.lame(1.0, 0.3))
    s = rng.uniform(0.01, 1.0, g.nel_tot) ** 3
    free = (rng.uniform(size=g.ndof) > 0.2).astype(np.uint8)
    op = icc.Operator(g, scale=s, free=free)
    sp.assemble(s)
    x = rng.standard_normal(g.ndof)
    ref = free * sp.matvec(free * x)
    y_atomic = op.apply(x)
    op.set_deterministic(True)
    y1 = op.apply(x)
    y2 = op.apply(x)
    assert np.array_equal(y1, y2)
    assert O.rel_err(y1, ref) < 1e-13 and O.rel_err(y1, y_atomic) < 1e-13
    z1 = op.apply(x, prec=icc.ICC_F32)
    z2 = op.apply(x, prec=icc.ICC_F32)
    assert np.array_equal(z1, z2) and O.rel_err(z1, ref) < 2e-6
    # with a sparse add-on: CSR then the gather, still repeatable
    n = g.ndof
    rows = rng.integers(0, n, 200)
    cols = rng.integers(0, n, 200)
    vals = rng.uniform(0.5, 1.5, 200)
    op.add_triplets(rows, cols, vals)
    S = np.zeros((n, n))
    np.add.at(S, (rows, cols), vals)
    ref2 = ref + free * (S @ (free * x))
    w1, w2 = op.apply(x), op.apply(x)
    assert np.array_equal(w1, w2) and O.rel_err(w1, ref2) < 1e-13
    op.set_deterministic(False)
    assert O.rel_err(op.apply(x), ref2) < 1e-13


def test_det_pcg_repeats_bitwise(ctx):
    sp, s, free, f, g, op = _cantilever(ctx, [12, 6, 6], 3)
    d = np.maximum(op.diag(), 1e-12)
    # 30 iterations, twice: the same iterate and residual bit for bit
    x1, it1, r1, rc1 = op.pcg(f, d, 1e-14, 30)
    x2, it2, r2, rc2 = op.pcg(f, d, 1e-14, 30)
    assert it1 == it2 == 30 and rc1 == rc2 == icc.ICC_ENOCONV
    assert np.array_equal(x1, x2) and r1 == r2
    # full solves: equal iteration counts and solutions, and the oracle's
    # PCG (the same algorithm, sums in another order) within one iteration
    x1, it1, r1, rc1 = op.pcg(f, d, 1e-10, 20 * g.ndof)
    x2, it2, r2, rc2 = op.pcg(f, d, 1e-10, 20 * g.ndof)
    assert rc1 == rc2 == icc.ICC_OK and it1 == it2 and np.array_equal(x1, x2)
    sp.assemble(s)
    xo, ito, reso = sp.pcg(f, np.maximum(sp.diag(), 1e-12), free, 1e-10, 20 * sp.ndof)
    assert abs(it1 - ito) <= 1 and O.rel_err(x1, xo) < 1e-8
    # the allocation-free entry point runs the same iterations
    import torch
    fd = torch.from_numpy(f).cuda()
    dd = torch.from_numpy(d).cuda()
    xd = torch.zeros(g.ndof, dtype=torch.float64, device="cuda")
    it = np.zeros(1, dtype=np.int64)
    res = np.zeros(1)
    rc = icc.lib().icc_pcg_dev_noalloc(op.h, fd.data_ptr(), dd.data_ptr(), 1e-10, 20 * g.ndof, xd.data_ptr(),
                                      it.ctypes.data, res.ctypes.data)
    torch.cuda.synchronize()
    assert rc == icc.ICC_OK and int(it[0]) == it1 and np.array_equal(xd.cpu().numpy(), x1)


def test_det_mixed_solve_repeats_bitwise(ctx):
    sp, s, free, f, g, op = _cantilever(ctx, [16, 8, 8], 4)
    u1, st1, rc1 = op.solve_mixed(f, tol=1e-8, delta=0.1)
    u2, st2, rc2 = op.solve_mixed(f, tol=1e-8, delta=0.1)
    assert rc1 == rc2 == icc.ICC_OK
    assert st1 == st2 and np.array_equal(u1, u2)
    sp.assemble(s)
    r = free * (f - sp.matvec(free * u1))
    assert np.linalg.norm(r) / np.linalg.norm(f) <= 1.01e-8
    with pytest.raises(icc.IccError):
        op.solve_mixed(f, tol=1e-8, inner=icc.ICC_F16X)


def test_det_refuses_face_terms(ctx):
    g = icc.Grid(ctx, (6, 5, 4), [[0, 1.2], [0, 1.0], [0, 0.8]])
    op = icc.Operator(g)
    status = np.where(np.arange(g.nel_tot) % 3 == 0, -1, 1).astype(np.int8)
    op.set_deterministic(True)
    with pytest.raises(icc.IccError):
        op.set_legacy_stabilization(status, 0.05)
    op.set_deterministic(False)
    assert op.set_legacy_stabilization(status, 0.05) > 0
    with pytest.raises(icc.IccError):
        op.set_deterministic(True)
