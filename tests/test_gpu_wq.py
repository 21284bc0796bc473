"""GPU parity of the weighted-quadrature stiffness formation (wq_stiffness /
wq_stiffness_layout, rust/src/wq.rs:324-514, interior and Calabro layouts,
p = 2) against the reference's 3D fixture probe and the oracle's restatement."""
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


def test_wq_form_fixture_probe(ctx):
    """wq_vs_matlab.rs:90-92: 3D element-density WQ stiffness (4x3x2 on
    [0,1.2]x[0,0.6]x[0,0.3], penal 3, emin 1e-3), probe < 1e-12."""
    d = load_fixture("wq_fastformation.json")
    g = icc.Grid(ctx, (4, 3, 2), [0.0, 1.2, 0.0, 0.6, 0.0, 0.3])
    m = icc.WqMatrix(g).form(icc.ICC_DENSITY_ELEMENT, np.asarray(d["xe3"]), 3.0, 1e-3)
    v = O.probe_vectors(g.ndof, 1)[0]
    assert O.rel_err(m.apply(v), d["Kf3_probe"]) < 1e-12


def _oracle_wq(bounds, nel, kind, x, penal=3.0, emin=1e-3, layout=0):
    sp = O.Space([[bounds[0], bounds[1]], [bounds[2], bounds[3]], [bounds[4], bounds[5]]], list(nel), 2)
    sp.assemble_wq(1.0, 0.3, kind, x, penal, emin, layout=layout)
    return sp


@pytest.mark.parametrize("layout", ["interior", "calabro"])
@pytest.mark.parametrize("nel", [(5, 4, 3), (9, 6, 7)])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_wq_form_matches_oracle(ctx, nel, kind, layout):
    """Products, diagonal, Frobenius norm and symmetry of the GPU-formed WQ
    matrix vs the oracle's assembled CSR for no density, element densities
    and spline (control-point) densities, in both point layouts (the Calabro
    layout puts WQ points on the knots, where an element density is the
    right-hand element's)."""
    bounds = [0.0, 1.3, -0.2, 0.7, 0.1, 0.6]
    rng = np.random.default_rng(31 + kind)
    g = icc.Grid(ctx, nel, bounds)
    x = None
    if kind == 1:
        x = rng.uniform(0.05, 1.0, g.nel_tot)
    elif kind == 2:
        x = rng.uniform(0.05, 1.0, g.ndof_sc)
    m = icc.WqMatrix(g, layout).form(kind, x, 3.0, 1e-3)
    sp = _oracle_wq(bounds, nel, kind, np.zeros(1) if x is None else x, layout=icc.WqMatrix.LAYOUTS[layout])
    for _ in range(2):
        v = rng.standard_normal(g.ndof)
        assert O.rel_err(m.apply(v), sp.matvec(v)) < 1e-13
    ip, ix, dd = m.csr()
    n = g.ndof
    assert ip[-1] == len(ix) == len(dd)
    rows = np.repeat(np.arange(n), np.diff(ip))
    diag = np.zeros(n)
    sel = rows == ix
    diag[rows[sel]] = dd[sel]
    assert O.rel_err(diag, sp.diag()) < 1e-13
    fro = float(np.sqrt(np.sum(dd * dd)))
    assert abs(fro - sp.frobenius()) <= 1e-13 * sp.frobenius()
    # symmetric by construction: K[r, c] == K[c, r]
    key = rows * n + ix
    tkey = ix * n + rows
    order = np.argsort(key)
    pos = np.searchsorted(key[order], tkey)
    assert np.all(key[order][pos] == tkey)
    assert np.max(np.abs(dd[order][pos] - dd)) <= 1e-15 * np.max(np.abs(dd))


@pytest.mark.parametrize("layout", ["interior", "calabro"])
def test_wq_form_uniform_equals_gauss(ctx, layout):
    """At uniform density the WQ rules are exact on the product of the test
    function with the trial space (wq_vs_matlab.rs:35-51): the WQ matrix
    equals the Gauss stiffness, in either layout."""
    nel, bounds = (6, 5, 4), [0.0, 1.3, -0.2, 0.7, 0.1, 0.6]
    g = icc.Grid(ctx, nel, bounds)
    m = icc.WqMatrix(g, layout).form(icc.ICC_DENSITY_NONE)
    op = icc.Operator(g, scale=np.ones(g.nel_tot))
    v = np.random.default_rng(4).standard_normal(g.ndof)
    assert O.rel_err(m.apply(v), op.apply(v)) < 1e-12


def test_wq_reform_on_device(ctx):
    """Re-forming for new densities from a device array gives the host-path
    matrix."""
    import torch
    nel, bounds = (6, 4, 5), [0.0, 1.0, 0.0, 0.7, 0.0, 0.9]
    g = icc.Grid(ctx, nel, bounds)
    rng = np.random.default_rng(8)
    x1, x2 = rng.uniform(0.1, 1.0, g.nel_tot), rng.uniform(0.1, 1.0, g.nel_tot)
    m = icc.WqMatrix(g).form(icc.ICC_DENSITY_ELEMENT, x1)
    m.form(icc.ICC_DENSITY_ELEMENT, torch.from_numpy(x2).cuda())
    ctx.sync()
    ref = icc.WqMatrix(g).form(icc.ICC_DENSITY_ELEMENT, x2)
    v = rng.standard_normal(g.ndof)
    assert O.rel_err(m.apply(v), ref.apply(v)) < 1e-15


@pytest.mark.parametrize("nel,layout", [((1, 1, 1), "interior"), ((1, 2, 3), "interior"), ((2, 2, 1), "interior"),
                                        ((2, 2, 2), "calabro"), ((2, 3, 4), "calabro")])
def test_wq_form_tiny_grids_match_oracle(ctx, nel, layout):
    """One- and two-element directions (every WQ rule touches the end knots;
    in the Calabro layout two elements share their end-element points)."""
    bounds = [0.0, 0.8, 0.0, 1.1, 0.0, 0.6]
    rng = np.random.default_rng(7)
    g = icc.Grid(ctx, nel, bounds)
    x = rng.uniform(0.1, 1.0, g.nel_tot)
    m = icc.WqMatrix(g, layout).form(icc.ICC_DENSITY_ELEMENT, x, 3.0, 1e-3)
    sp = _oracle_wq(bounds, nel, 1, x, layout=icc.WqMatrix.LAYOUTS[layout])
    v = rng.standard_normal(g.ndof)
    assert O.rel_err(m.apply(v), sp.matvec(v)) < 1e-13


def test_wq_errors(ctx):
    g = icc.Grid(ctx, (3, 3, 3), [0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    m = icc.WqMatrix(g)
    with pytest.raises(icc.IccError):
        m.apply(np.zeros(g.ndof))  # not formed yet
    with pytest.raises(icc.IccError):
        m.form(icc.ICC_DENSITY_ELEMENT, None)  # density field missing
    with pytest.raises(icc.IccError):
        m.form(7, np.ones(g.nel_tot))
    with pytest.raises(ValueError):
        icc.WqMatrix(g, "gauss")
    # one element in a direction: the Calabro rule of its middle function is singular
    g1 = icc.Grid(ctx, (3, 1, 3), [0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    with pytest.raises(icc.IccError, match="Calabro"):
        icc.WqMatrix(g1, "calabro")
