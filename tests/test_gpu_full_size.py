"""The benchmark configurations at their full sizes, checked through
properties that do not depend on the oracle assembling the whole system:

* adjoint SIMP on the 128 x 64 x 64 grid (1,698,840 DOF) with the density
  filter: the hand-written sensitivity and filter-backward kernels against
  central finite differences of the compliance (h = 1e-6) on seeded
  elements, every solve to 1e-12; two OC iterations keep the volume;
* the immersed 128 x 64 x 64 problem (box minus three seeded cylinders): the
  mixed solves' (FP32 and split-FP16 inner operators) true residuals
  recomputed with the problem's FP64 operator, and their solutions against
  the FP64 solve;
* the batch of 24 uneven instances, 24 x 12 x 12 up to 160 x 80 x 80 (24.2M
  DOF), autotuned, in one mixed solve: every instance's reported true
  residual <= 1e-8, recomputed independently for three of them (the oracle's
  assembled operator for the smallest, a fresh FP64 operator otherwise)."""
import math

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

icc = pytest.importorskip("immersed_catmull_clark_amd")

BOX = [[0, 2.0], [0, 1.0], [0, 1.0]]


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = icc.Context(0)
    yield c
    c.close()


def test_adjoint_full_size_vs_finite_differences(ctx):
    base = dict(nel=(128, 64, 64), size=(2.0, 1.0, 1.0), volfrac=0.3, penal=3.0, rmin=1.5, emin=1e-9, load=1,
                filter=1)
    tf = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(max_iter=1, pcg_tol=1e-12, **base))
    grid = icc.Grid(ctx, base["nel"], BOX)
    flt = icc.Filter(grid, base["rmin"])
    op = icc.Operator(grid, free=tf.free)
    emin, pen = base["emin"], base["penal"]
    rng = np.random.default_rng(20261206)
    x = rng.uniform(0.2, 1.0, tf.nel)
    c, g, u = tf.compliance(x)
    s0 = emin + (1 - emin) * flt.fwd(x) ** pen
    ee0 = tf.element_energies(u)
    assert abs(math.fsum(s0 * ee0) - c) < 1e-9 * c

    def dcomp(xd):
        # c(x') - c(x) = -sum_f (s'_f - s_f) ee_f(u) + 2 d.(F - K' u) - d.K' d, u' = u + d
        sd = emin + (1 - emin) * flt.fwd(xd) ** pen
        _, _, ud = tf.compliance(xd, u0=u, grad=False)
        d = ud - u
        op.set_scale(sd)
        ku, kd = op.apply(u), op.apply(d)
        return -math.fsum((sd - s0) * ee0) + 2.0 * math.fsum(d * (tf.force - ku)) - math.fsum(d * kd)

    hstep = 1e-6
    for e in rng.choice(tf.nel, 6, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += hstep
        xm[e] -= hstep
        fd = (dcomp(xp) - dcomp(xm)) / (2 * hstep)
        assert abs(g[e] - fd) <= 1e-6 * abs(fd), (e, g[e], fd)
    tf.close()
    # two OC iterations: the compliance falls and the filtered volume meets the target
    t = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(max_iter=2, pcg_tol=1e-8, **base))
    h = t.run()
    assert len(h.compliance) == 2 and h.compliance[1] < h.compliance[0]
    assert np.all(np.asarray(h.pcg_iterations) > 0)


def test_immersed_full_size_mixed_vs_fp64(ctx):
    rng = np.random.default_rng(20261204)
    cyl = np.column_stack([rng.uniform(0.4, 1.6, 3), rng.uniform(0.3, 0.7, 3), rng.uniform(0.08, 0.15, 3)])
    sols = {}
    for solver in (0, 1, 2):
        h = icc.ImmersedSdf(ctx, icc.ImmersedConfig(res=(128, 64, 64), mode=1, solver=solver, pcg_tol=1e-8), cyl)
        u, st, rc = h.solve()
        assert rc == icc.ICC_OK and st["true_res"] <= 1e-8
        f = h.force * h.free
        r = h.free * (f - h.apply(u))
        assert np.linalg.norm(r) / np.linalg.norm(f) <= 1.01e-8
        assert abs(st["compliance"] - float(h.force @ u)) <= 1e-12 * abs(st["compliance"])
        sols[solver] = (u, st["compliance"])
        h.close()
    for s in (1, 2):
        (u64, c64), (um, cm) = sols[0], sols[s]
        assert O.rel_err(um, u64) < 1e-6 and abs(cm - c64) / abs(c64) < 1e-8, s


def test_batch_24_instances_full_size(ctx):
    nx = [2 * int(round((24 + round(136 * k / 23)) / 2)) for k in range(24)]
    nel = [(n, n // 2, n // 2) for n in nx]
    rng = np.random.default_rng(20261205)
    setups = []
    for k in range(24):
        ne = nel[k][0] * nel[k][1] * nel[k][2]
        rho = rng.uniform(0.01, 1.0, ne)
        setups.append((1e-9 + (1 - 1e-9) * rho ** 3, rng.uniform(0.5, 2.0)))
    b = icc.Batch(ctx, nel)
    for k, (s, mag) in enumerate(setups):
        b.set(k, scale=s)
        b.set_traction(k, (0.0, -mag, 0.0))
    b.autotune(icc.ICC_F32, reps=2)
    us, st, rc = b.solve(prec=icc.ICC_F32, tol=1e-8, delta=0.1)
    assert rc == icc.ICC_OK
    assert np.all(st[:, 3] == icc.ICC_OK) and np.all(st[:, 2] <= 1e-8) and np.all(st[:, 1] >= 1)
    for k in (0, 12, 23):
        s, mag = setups[k]
        g = icc.Grid(ctx, nel[k], BOX)
        i = np.arange(g.ndof_sc) % g.ncp[0]
        free = np.ones(g.ndof, dtype=np.uint8)
        for c in range(3):
            free[c * g.ndof_sc:(c + 1) * g.ndof_sc][i == 0] = 0
        sp = O.Space(BOX, list(nel[k]), 2)
        f = sp.traction_xmax([0.0, -mag, 0.0]) * free
        if k == 0:
            sp.elem_mats(*O.lame(1.0, 0.3))
            sp.assemble(s)
            r = free * (f - sp.matvec(free * us[k]))
        else:
            op = icc.Operator(g, scale=s, free=free)
            r = free * (f - op.apply(us[k]))
            op.close()
        assert np.linalg.norm(r) / np.linalg.norm(f) <= 1.01e-8, (k, np.linalg.norm(r) / np.linalg.norm(f))
        g.close()
    b.close()
