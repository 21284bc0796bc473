"""Batched structured instances of uneven sizes over libicc's graph-batched
solver (icc_batch_*): the PCG iterations of every unfinished instance are
captured into one HIP graph, and each instance stops on its own."""
import ctypes as C

import numpy as np

from ._abi import ICC_BREAKDOWN_CODES, ICC_F32, ICC_OK, check, lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Batch:
    def __init__(self, ctx, nel, size=None, young=1.0, poisson=0.3):
        self.ctx = ctx
        ne = np.ascontiguousarray(np.asarray(nel, dtype=np.int64).reshape(-1, 3))
        sz = None if size is None else np.ascontiguousarray(np.asarray(size, dtype=np.float64).reshape(-1, 3))
        self.n = len(ne)
        h = C.c_void_p()
        check(lib().icc_batch_create(ctx.h, self.n, _p(ne), _p(sz), float(young), float(poisson), C.byref(h)))
        self.h = h
        self.ndof = np.zeros(self.n, dtype=np.int64)
        self.nel = np.zeros(self.n, dtype=np.int64)
        check(lib().icc_batch_sizes(self.h, _p(self.ndof), _p(self.nel)))
        self.offsets = np.concatenate([[0], np.cumsum(self.ndof)])

    def set(self, k, scale=None, free=None, force=None):
        s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        fr = None if free is None else np.ascontiguousarray(free, dtype=np.uint8)
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        check(lib().icc_batch_set(self.h, int(k), _p(s), _p(fr), _p(f)))

    def set_traction(self, k, t):
        tt = np.ascontiguousarray(t, dtype=np.float64)
        check(lib().icc_batch_set_traction(self.h, int(k), _p(tt)))

    def autotune(self, prec=ICC_F32, reps=3):
        """Tile shapes per instance, by timing each instance's apply."""
        check(lib().icc_batch_autotune(self.h, int(prec), int(reps)))

    def solve(self, prec=ICC_F32, tol=1e-8, delta=0.1, maxit=0):
        """Returns (list of u per instance, stats [n, 4]: iterations, reliable
        updates, true relative residual, status; return code)."""
        u = np.zeros(int(self.offsets[-1]))
        st = np.zeros((self.n, 4))
        rc = lib().icc_batch_solve(self.h, int(prec), float(tol), float(delta), int(maxit), _p(u), _p(st))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        return [u[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n)], st, rc

    def close(self):
        if getattr(self, "h", None):
            lib().icc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
