"""Batched solves of structured instances of uneven sizes (icc_batch_*): the
tiles of all unfinished instances run as one apply launch per PCG iteration,
the vector updates as one launch per stage over the instances' segments, and
each instance stops on its own.  Every instance's solution is checked with
the oracle's assembled operator, and against the instance solved alone."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

icc = pytest.importorskip("immersed_catmull_clark_amd")

# eight uneven instances: sizes from 96 to 8,192 elements, one of them
# finishing far earlier than the others (the plan is rebuilt without it)
NELS = [(8, 4, 4), (12, 6, 6), (16, 8, 8), (10, 4, 6), (32, 16, 16), (6, 4, 4), (20, 10, 10), (24, 12, 12)]
BOX = [[0, 2.0], [0, 1.0], [0, 1.0]]


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = icc.Context(0)
    yield c
    c.close()


def _instances(seed=31):
    rng = np.random.default_rng(seed)
    out = []
    for nel in NELS:
        sp = O.Space(BOX, list(nel), 2)
        sp.elem_mats(*O.lame(1.0, 0.3))
        rho = rng.uniform(0.01, 1.0, sp.nel)
        s = 1e-9 + (1 - 1e-9) * rho ** 3
        mag = rng.uniform(0.5, 2.0)
        i = np.arange(sp.ndof_sc) % sp.ndof_dir[0]
        free = np.ones(sp.ndof, dtype=np.uint8)
        for c in range(3):
            free[c * sp.ndof_sc:(c + 1) * sp.ndof_sc][i == 0] = 0
        f = sp.traction_xmax([0.0, -mag, 0.0]) * free
        out.append((sp, s, free, f))
    return out


def _oracle_residual(sp, s, free, f, u):
    sp.assemble(s)
    r = free * (f - sp.matvec(free * u))
    return np.linalg.norm(r) / np.linalg.norm(f)


def test_batch_fp64_matches_oracle_and_single_solves(ctx):
    inst = _instances()
    b = icc.Batch(ctx, NELS)
    for k, (sp, s, free, f) in enumerate(inst):
        b.set(k, scale=s, free=free, force=f)
    b.autotune(icc.ICC_F64, reps=2)
    us, st, rc = b.solve(prec=icc.ICC_F64, tol=1e-9)
    assert rc == icc.ICC_OK, st
    for k, (sp, s, free, f) in enumerate(inst):
        res = _oracle_residual(sp, s, free, f, us[k])
        assert res < 2e-9, (k, res)
        assert abs(st[k, 2] - res) < 1e-12
        g = icc.Grid(ctx, NELS[k], BOX)
        op = icc.Operator(g, scale=s, free=free)
        x, it, r1, rc1 = op.pcg(f, np.maximum(op.diag(), 1e-12), 1e-9, 20 * g.ndof)
        assert rc1 == icc.ICC_OK
        # the same algorithm in another launch layout: both scatter with
        # global atomics, whose order varies from run to run, and the
        # iterations amplify that round-off (the stopping test may move by one)
        assert abs(int(st[k, 0]) - it) <= 1, (k, st[k, 0], it)
        assert O.rel_err(us[k], x) < 1e-9


def test_batch_mixed_true_residuals(ctx):
    inst = _instances(32)
    b = icc.Batch(ctx, NELS)
    for k, (sp, s, free, f) in enumerate(inst):
        b.set(k, scale=s, free=free, force=f)
    b.autotune(icc.ICC_F32, reps=2)
    us, st, rc = b.solve(prec=icc.ICC_F32, tol=1e-8, delta=0.1)
    assert rc == icc.ICC_OK, st
    for k, (sp, s, free, f) in enumerate(inst):
        assert st[k, 1] >= 1 and st[k, 3] == icc.ICC_OK and st[k, 2] <= 1e-8
        res = _oracle_residual(sp, s, free, f, us[k])
        assert res <= 1.01e-8, (k, res)
        # the same instance alone: about the same number of FP32 iterations
        # (the FP32 recurrence turns the run-to-run order of the atomic
        # scatter into iteration counts that vary between runs: 255 and 290
        # were seen for the first instance)
        g = icc.Grid(ctx, NELS[k], BOX)
        op = icc.Operator(g, scale=s, free=free)
        u1, st1, rc1 = op.solve_mixed(f, tol=1e-8, delta=0.1)
        assert rc1 == icc.ICC_OK
        assert abs(int(st[k, 0]) - st1["inner_iters"]) <= max(8, 0.25 * st1["inner_iters"]), (k, st[k, 0], st1)


def test_batch_single_instance_and_errors(ctx):
    inst = _instances(33)[:1]
    b = icc.Batch(ctx, NELS[:1])
    sp, s, free, f = inst[0]
    b.set(0, scale=s, free=free, force=f)
    us, st, rc = b.solve(prec=icc.ICC_F64, tol=1e-10)
    assert rc == icc.ICC_OK and _oracle_residual(sp, s, free, f, us[0]) < 2e-10
    with pytest.raises(icc.IccError):
        b.solve(prec=icc.ICC_F16X, tol=1e-8)
    # an iteration limit ends with ENOCONV and the instance's status
    us, st, rc = b.solve(prec=icc.ICC_F64, tol=1e-12, maxit=5)
    assert rc == icc.ICC_ENOCONV and st[0, 3] == icc.ICC_ENOCONV and st[0, 0] == 5
