"""Immersed solve with an analytic geometry (a box minus cylinders parallel to
z) over libicc's C++ driver (icc_immersed_*): the build_immersed_problem /
solve_immersed_problem pair of the reference crate (rust/src/immersed.rs:466-570)
on a p = 2 background grid aligned to the box.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._abi import ICC_BREAKDOWN_CODES, ICC_OK, ImmersedConfigC, check, lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class ImmersedConfig:
    """Fields of icc_immersed_config; defaults of icc_immersed_config_default.
    mode 0: scale emin + (1 - emin) w on every cell (immersed.rs:484); mode 1:
    scale w on cells with w >= active_min, other cells dropped.  solver 0: FP64
    Jacobi PCG (|r|/|b| < pcg_tol); 1: FP32 PCG with FP64 reliable updates to a
    true relative residual <= pcg_tol; 2: the same with the split-FP16 inner
    contraction."""
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (2.0, 1.0, 1.0)
    res: tuple = (32, 16, 16)
    sub: tuple = (4, 4, 4)
    young: float = 1.0
    poisson: float = 0.3
    emin: float = 1e-4
    active_min: float = 0.05
    penalty_factor: float = 1e3
    traction: tuple = (0.0, -1.0, 0.0)
    pcg_tol: float = 1e-8
    pcg_maxit: int = 0
    mode: int = 1
    solver: int = 0

    def to_c(self):
        c = ImmersedConfigC()
        lib().icc_immersed_config_default(C.byref(c))
        for k in ("box_lo", "box_hi", "traction"):
            getattr(c, k)[:] = [float(v) for v in getattr(self, k)]
        for k in ("res", "sub"):
            getattr(c, k)[:] = [int(v) for v in getattr(self, k)]
        for k in ("young", "poisson", "emin", "active_min", "penalty_factor", "pcg_tol"):
            setattr(c, k, float(getattr(self, k)))
        c.pcg_maxit = int(self.pcg_maxit)
        c.mode = int(self.mode)
        c.solver = int(self.solver)
        return c


class ImmersedSdf:
    """Immersed elasticity problem on the box minus the cylinders `cyl`
    (rows cx, cy, r); classification on the GPU, solve through the MFMA apply
    plus the sparse penalty face term."""

    def __init__(self, ctx, cfg, cyl):
        self.ctx = ctx
        self.cfg = cfg
        self._c = cfg.to_c()
        cy = np.ascontiguousarray(np.asarray(cyl, dtype=np.float64).reshape(-1, 3))
        h = C.c_void_p()
        check(lib().icc_immersed_create(ctx.h, C.byref(self._c), _p(cy), len(cy), C.byref(h)))
        self.h = h
        sz = np.zeros(5, dtype=np.int64)
        beta = C.c_double()
        check(lib().icc_immersed_sizes(self.h, _p(sz), C.byref(beta)))
        self.nel, self.ndof, self.n_active, self.n_free, self.nnz_penalty = (int(v) for v in sz)
        self.beta = beta.value
        self.status = np.zeros(self.nel, dtype=np.int8)
        self.weights = np.zeros(self.nel)
        self.active = np.zeros(self.nel, dtype=np.uint8)
        self.scale = np.zeros(self.nel)
        self.free = np.zeros(self.ndof, dtype=np.uint8)
        self.force = np.zeros(self.ndof)
        check(lib().icc_immersed_problem(self.h, _p(self.status), _p(self.weights), _p(self.active), _p(self.scale),
                                         _p(self.free), _p(self.force)))

    def apply(self, x):
        y = np.zeros(self.ndof)
        check(lib().icc_immersed_apply(self.h, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y)))
        return y

    def solve(self):
        """Returns (u, stats dict, status code)."""
        u = np.zeros(self.ndof)
        st = np.zeros(8)
        rc = lib().icc_immersed_solve(self.h, _p(u), _p(st))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        return u, dict(iterations=int(st[0]), second=float(st[1]), true_res=float(st[2]), compliance=float(st[3]),
                       fp64_applies=int(st[4])), rc

    def close(self):
        if getattr(self, "h", None):
            lib().icc_immersed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
