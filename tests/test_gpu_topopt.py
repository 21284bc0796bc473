"""GPU parity of the CantileverTopOpt driver (rust/src/topopt_iga.rs) and of
the adjoint sensitivities against finite differences."""
import math

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


def test_topopt_fixture(ctx):
    """topopt_vs_matlab.rs:15-40: 6 compliances < 1e-7 rel, final densities
    < 1e-6 vs the MATLAB direct-solve run, pcg_tol 1e-12."""
    d = load_fixture("topopt3d_8x4x2.json")
    cfg = icc.CantileverTopOptConfig(nel=tuple(d["nel"]), volfrac=d["volfrac"], penal=d["penal"], rmin=d["rmin"],
                                     max_iter=d["max_iter"], pcg_tol=1e-12)
    t = icc.CantileverTopOpt(ctx, cfg)
    seen = []
    h = t.run(report=lambda it, c, v, ch: seen.append(it))
    cref = np.asarray(d["compliance"])
    assert len(h.compliance) == len(cref) and seen == list(range(1, len(cref) + 1))
    assert np.all(np.abs(np.asarray(h.compliance) - cref) / cref < 1e-7), (h.compliance, cref)
    assert np.max(np.abs(h.density - np.asarray(d["xPhys"]))) < 1e-6


def test_topopt_matches_oracle_run(ctx):
    """Same loop as the oracle restatement, default 16x8x4 grid, 8 iterations."""
    comp, change, its, dens = O.topopt_run([16, 8, 4], max_iter=8, pcg_tol=1e-12)
    t = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(max_iter=8, pcg_tol=1e-12))
    h = t.run()
    assert np.all(np.abs(np.asarray(h.compliance) - comp) / comp < 1e-8)
    assert np.max(np.abs(h.density - dens)) < 1e-7


def test_solve_state_matches_oracle(ctx):
    cfg = icc.CantileverTopOptConfig(nel=(12, 6, 4), pcg_tol=1e-11, emin=1e-3)
    t = icc.CantileverTopOpt(ctx, cfg)
    ref = O.TopOpt([12, 6, 4], pcg_tol=1e-11, emin=1e-3)
    assert np.array_equal(t.free, ref.free)
    assert np.array_equal(t.force, ref.force)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.1, 1.0, t.nel)
    u0 = rng.standard_normal(t.ndof) * 0.01 * ref.free
    u, it, res = t.solve_state(x, u0)
    ur, itr, resr = ref.solve_state(x, u0)
    assert O.rel_err(u, ur) < 1e-8
    assert res < 1e-11 * 1.0001 or abs(res - resr) < 1e-12


@pytest.mark.parametrize("filt", [0, 1])
def test_adjoint_vs_finite_differences(ctx, filt):
    """dc/dx from the hand-written sensitivity kernel (+ density filter
    backward) vs central differences of c(x), h = 1e-4, each solve to 1e-13.
    With c ~ 137 the difference quotient at h = 1e-6 carries round-off of a
    few 1e-6 relative even with exact direct solves; at h = 1e-4 the
    truncation and round-off errors are both below 1e-7."""
    cfg = icc.CantileverTopOptConfig(nel=(12, 6, 6), size=(2.0, 1.0, 1.0), emin=1e-9, rmin=1.5, pcg_tol=1e-13,
                                     load=1, filter=filt)
    t = icc.CantileverTopOpt(ctx, cfg)
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], [12, 6, 6], 2)
    rng = np.random.default_rng(21)
    x = rng.uniform(0.2, 1.0, t.nel)
    c, g, u = t.compliance(x)

    def c_hat(xd, u0):
        # 2 F.u - u^T K(x) u: equals c at the solution, error quadratic in the
        # solve residual (F.u alone is first order in it); summed exactly
        # (fsum) so round-off does not dominate the h = 1e-6 difference
        _, _, uu = t.compliance(xd, u0=u0, grad=False)
        xt = sp.filter_apply(cfg.rmin, 1, None, xd) if filt == 1 else xd
        s = cfg.emin + (1 - cfg.emin) * xt ** cfg.penal
        terms = np.concatenate([2.0 * t.force * uu, -(s * t.element_energies(uu))])
        return math.fsum(terms)

    assert abs(c_hat(x, u) - c) < 1e-10 * c
    hstep = 1e-4
    for e in rng.choice(t.nel, 8, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += hstep
        xm[e] -= hstep
        fd = (c_hat(xp, u) - c_hat(xm, u)) / (2 * hstep)
        assert abs(g[e] - fd) <= 1e-6 * abs(fd), (e, g[e], fd)


def test_traction_load_total(ctx):
    cfg = icc.CantileverTopOptConfig(nel=(10, 5, 4), size=(2.0, 1.0, 0.5), load=1)
    t = icc.CantileverTopOpt(ctx, cfg)
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 0.5]], [10, 5, 4], 2)
    ref = sp.traction_xmax([0.0, -1.0, 0.0])
    assert O.rel_err(t.force, ref) < 1e-14
    nsc = sp.ndof_sc
    assert abs(t.force[nsc:2 * nsc].sum() + 0.5) < 1e-13
