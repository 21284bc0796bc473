"""icc_pcg_dev_noalloc: the PCG entry point that allocates nothing, against
icc_pcg_dev on the same device inputs, two calls in a row (the second on a
smaller operator, so the shared workspace holds the first call's contents),
with the device's free memory unchanged across the second call (the first
one also lets the runtime make its own one-time allocations, such as
scratch for kernels that spill); and the size limit of the fixed workspace.

Bar: both run the same FP64 kernels; the apply's FP64 atomics round in a
run-dependent order, so iteration counts agree to +-2 and x to 1e-8
relative (the PCG tolerance is 1e-10)."""
import ctypes

import numpy as np
import pytest

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


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _problem(ctx, nel, seed):
    import torch
    rng = np.random.default_rng(seed)
    g = icc.Grid(ctx, nel, [0.0, 2.0, 0.0, 1.0, 0.0, 1.0])
    s = 1e-9 + (1 - 1e-9) * rng.uniform(0.1, 1.0, g.nel_tot) ** 3
    free = np.ones(g.ndof, dtype=np.uint8)
    ix = np.arange(g.ndof) % g.ndof_sc % (nel[0] + 2)
    free[ix == 0] = 0
    op = icc.Operator(g, scale=s, free=free)
    f = free * rng.standard_normal(g.ndof)
    d = np.maximum(op.diag(), 1e-12)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    return g, op, dev(f), dev(d)


def _solve(name, ctx, op, b, d, tol, maxit, x=None):
    import torch
    from immersed_catmull_clark_amd._abi import lib
    x = torch.zeros_like(b) if x is None else x
    it = np.zeros(1, dtype=np.int64)
    res = np.zeros(1)
    torch.cuda.synchronize()
    rc = getattr(lib(), name)(op.h, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(d.data_ptr()), float(tol),
                              int(maxit), ctypes.c_void_p(x.data_ptr()), it.ctypes.data, res.ctypes.data)
    ctx.sync()
    return x.cpu().numpy(), int(it[0]), float(res[0]), rc


def test_noalloc_pcg_matches_pcg_dev_twice(ctx):
    tol, maxit = 1e-10, 20000
    for k, nel in enumerate([(48, 24, 24), (24, 12, 12)]):
        import torch
        g, op, b, d = _problem(ctx, nel, 3 + k)
        x0 = torch.zeros_like(b)
        torch.cuda.synchronize()
        free0 = _free_bytes()
        x_na, it_na, res_na, rc_na = _solve("icc_pcg_dev_noalloc", ctx, op, b, d, tol, maxit, x0)
        if k == 1:
            # a fresh operator: icc_pcg_dev would allocate its workspace here
            # (13 vectors, 1.6 MB at this size)
            assert _free_bytes() == free0, (free0, _free_bytes())
        x_ref, it_ref, res_ref, rc_ref = _solve("icc_pcg_dev", ctx, op, b, d, tol, maxit)
        print(nel, "noalloc", it_na, res_na, "pcg_dev", it_ref, res_ref)
        assert rc_na == icc.ICC_OK and rc_ref == icc.ICC_OK, (rc_na, rc_ref)
        assert abs(it_na - it_ref) <= 2, (it_na, it_ref)
        assert res_na < tol and res_ref < tol
        assert np.linalg.norm(x_na - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


def test_noalloc_pcg_rejects_oversize(ctx):
    from immersed_catmull_clark_amd._abi import lib
    g, op, b, d = _problem(ctx, (72, 72, 70), 9)
    assert g.ndof > (1 << 20)
    x, it, res, rc = _solve("icc_pcg_dev_noalloc", ctx, op, b, d, 1e-8, 10)
    assert rc == icc.ICC_EINVAL
    assert b"exceed" in lib().icc_last_error()
