"""Structured SIMP cantilever, mirroring the reference crate's
`CantileverTopOpt` / `CantileverTopOptConfig` (rust/src/topopt_iga.rs) over
libicc's C++ driver (icc_topopt_*).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import REPORT_FN, TopOptConfigC, check, lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class CantileverTopOptConfig:
    """Defaults of CantileverTopOptConfig::default (topopt_iga.rs:35-51).
    Extensions: load 0 = reference nodal load, 1 = unit downward traction on
    x = L; filter 0 = sensitivity filter, 1 = density filter (adjoint);
    solver 0 = FP64 PCG, 1 = FP32 PCG with FP64 reliable updates, 2 = the same with the
    split-FP16 inner contraction."""
    nel: tuple = (16, 8, 4)
    size: tuple = (1.2, 0.6, 0.3)
    degree: int = 2
    volfrac: float = 0.3
    penal: float = 3.0
    rmin: float = 1.5
    max_iter: int = 40
    young: float = 1.0
    poisson: float = 0.3
    emin: float = 1e-3
    move_limit: float = 0.2
    pcg_tol: float = 1e-10
    load: int = 0
    filter: int = 0
    solver: int = 0

    def to_c(self):
        c = TopOptConfigC()
        lib().icc_topopt_config_default(C.byref(c))
        for d in range(3):
            c.nel[d] = int(self.nel[d])
            c.size[d] = float(self.size[d])
        for k in ("volfrac", "penal", "rmin", "young", "poisson", "emin", "move_limit", "pcg_tol"):
            setattr(c, k, float(getattr(self, k)))
        c.max_iter = int(self.max_iter)
        c.load, c.filter, c.solver = int(self.load), int(self.filter), int(self.solver)
        return c


@dataclass
class TopOptHistory:
    """TopOptHistory (topopt_iga.rs:54-59)."""
    compliance: list = field(default_factory=list)
    change: list = field(default_factory=list)
    pcg_iterations: list = field(default_factory=list)
    density: np.ndarray = None


class CantileverTopOpt:
    def __init__(self, ctx, cfg=None):
        if cfg is None:
            cfg = CantileverTopOptConfig()
        if cfg.degree != 2:
            raise ValueError("only degree 2 is implemented")
        self.ctx = ctx
        self.cfg = cfg
        self._c = cfg.to_c()
        h = C.c_void_p()
        check(lib().icc_topopt_create(ctx.h, C.byref(self._c), C.byref(h)))
        self.h = h
        s = np.zeros(2, dtype=np.int64)
        check(lib().icc_topopt_sizes(h, _p(s)))
        self.nel, self.ndof = int(s[0]), int(s[1])
        self.force = np.zeros(self.ndof)
        self.free = np.zeros(self.ndof, dtype=np.uint8)
        check(lib().icc_topopt_force(h, _p(self.force), _p(self.free)))

    def solve_state(self, x, u0):
        """(u, iters, res) as solve_state (topopt_iga.rs:137-162)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        u = np.zeros(self.ndof)
        it = np.zeros(1, dtype=np.int64)
        res = np.zeros(1)
        check(lib().icc_topopt_solve_state(self.h, _p(x), _p(u0), _p(u), _p(it), _p(res)))
        return u, int(it[0]), float(res[0])

    def element_energies(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        ee = np.zeros(self.nel)
        check(lib().icc_topopt_element_energies(self.h, _p(u), _p(ee)))
        return ee

    def compliance(self, x, u0=None, grad=True):
        """Compliance of design x and dc/dx through the configured filter.
        Returns (c, dcdx, u)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        u = np.zeros(self.ndof) if u0 is None else np.ascontiguousarray(u0, dtype=np.float64).copy()
        c = np.zeros(1)
        g = np.zeros(self.nel) if grad else None
        check(lib().icc_topopt_compliance(self.h, _p(x), _p(u), _p(c), _p(g)))
        return float(c[0]), g, u

    def run(self, report=None):
        mi = int(self.cfg.max_iter)
        comp = np.zeros(mi)
        ch = np.zeros(mi)
        its = np.zeros(mi, dtype=np.int64)
        dens = np.zeros(self.nel)
        n = np.zeros(1, dtype=np.int64)
        cb = REPORT_FN(lambda it, c, v, chg, _u: report(int(it), c, v, chg)) if report else None
        check(lib().icc_topopt_run(self.h, cb, None, _p(comp), _p(ch), _p(its), _p(dens), _p(n)))
        k = int(n[0])
        return TopOptHistory(list(comp[:k]), list(ch[:k]), [int(v) for v in its[:k]], dens)

    def close(self):
        if self.h:
            lib().icc_topopt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
