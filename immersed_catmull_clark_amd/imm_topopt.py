"""Immersed geometry from a triangle B-rep and immersed SIMP topology
optimisation, mirroring the reference crate's `TriangleMesh3D`
(rust/src/cut_cell.rs:21-53), `padded_bounds` / `element_weights` with a
`ParityRayCaster` (rust/src/immersed.rs:29-155, 450-463) and
`ImmersedTopOpt` / `ImmersedTopOptConfig` (rust/src/topopt_immersed.rs) over
libicc's C++ driver (icc_imm_topopt_*).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import (ICC_BREAKDOWN_CODES, ICC_CLAMP_BOTTOM, ICC_CLAMP_LEFT, ICC_CLAMP_NONE, ICC_CLAMP_TOP, ICC_OK,
                   ICC_STAB_GHOST, ICC_STAB_LEGACY, ICC_STAB_NONE, IMM_REPORT_FN, ImmersedMeshConfigC,
                   ImmTopOptConfigC, check, lib)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class TriangleMesh3D:
    """Closed triangle surface: vertices [nv, 3], triangles [nt, 3] (indices)."""
    vertices: np.ndarray
    triangles: np.ndarray

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(np.asarray(self.vertices, dtype=np.float64).reshape(-1, 3))
        self.triangles = np.ascontiguousarray(np.asarray(self.triangles, dtype=np.int64).reshape(-1, 3))

    @staticmethod
    def new_box(mn, mx):
        """Axis-aligned box with outward-oriented triangles (cut_cell.rs:27-53)."""
        (a, b, c), (d, e, f) = mn, mx
        v = [(a, b, c), (d, b, c), (d, e, c), (a, e, c), (a, b, f), (d, b, f), (d, e, f), (a, e, f)]
        t = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7), (2, 7, 6), (0, 4, 7),
             (0, 7, 3), (1, 2, 6), (1, 6, 5)]
        return TriangleMesh3D(np.asarray(v, dtype=np.float64), np.asarray(t, dtype=np.int64))


def padded_bounds(mesh, padding):
    """[[x0, x1], [y0, y1], [z0, z1]] (immersed.rs:450-463)."""
    out = np.zeros(6)
    check(lib().icc_mesh_padded_bounds(_p(mesh.vertices), len(mesh.vertices), float(padding), _p(out)))
    return out.reshape(3, 2)


def element_weights(ctx, mesh, bounds, res, sub=(4, 4, 4)):
    """(weights [nel], status [nel]) of the background cells by ray parity,
    computed on the GPU (immersed.rs:121-155)."""
    gb = np.ascontiguousarray(np.asarray(bounds, dtype=np.float64).reshape(6))
    r = np.asarray(res, dtype=np.int64)
    s = np.asarray(sub, dtype=np.int64)
    nel = int(np.prod(r))
    status = np.zeros(nel, dtype=np.int8)
    w = np.zeros(nel)
    check(lib().icc_classify_mesh(ctx.h, _p(mesh.vertices), len(mesh.vertices), _p(mesh.triangles),
                                  len(mesh.triangles), _p(gb), _p(r), _p(s), _p(status), _p(w)))
    return w, status


@dataclass
class Stabilization:
    """Stabilization::{None, Legacy{gamma}, GhostPenalty{gamma}} (immersed.rs)."""
    kind: int = ICC_STAB_GHOST
    gamma: float = 1e-2

    @staticmethod
    def none():
        return Stabilization(ICC_STAB_NONE, 0.0)

    @staticmethod
    def legacy(gamma):
        return Stabilization(ICC_STAB_LEGACY, float(gamma))

    @staticmethod
    def ghost_penalty(gamma):
        return Stabilization(ICC_STAB_GHOST, float(gamma))


_CLAMP = {None: ICC_CLAMP_NONE, "bottom": ICC_CLAMP_BOTTOM, "top": ICC_CLAMP_TOP, "left": ICC_CLAMP_LEFT}


@dataclass
class ImmersedOptions:
    """Fields and defaults of ImmersedOptions (immersed.rs:392-429); degree
    must be 2.  clamped_face: "bottom", "top", "left" or None; force: None
    for the default top-centre load.  Extension: solver 0 = FP64 Jacobi
    PCG, 1 = FP32 PCG with FP64 reliable updates, 2 = the same with the
    split-FP16 inner contraction."""
    grid_res: tuple = (24, 16, 12)
    padding: float = 0.05
    degree: int = 2
    young: float = 1.0
    poisson: float = 0.3
    stabilization: Stabilization = field(default_factory=Stabilization)
    emin: float = 1e-4
    subcell_res: tuple = (4, 4, 4)
    clamped_face: object = "bottom"
    force: np.ndarray = None
    pcg_tol: float = 1e-8
    pcg_maxit: int = 5000
    solver: int = 0

    def to_c(self):
        if self.degree != 2:
            raise ValueError("only degree 2 is implemented")
        c = ImmersedMeshConfigC()
        lib().icc_immersed_mesh_config_default(C.byref(c))
        c.grid_res[:] = [int(v) for v in self.grid_res]
        c.subcell_res[:] = [int(v) for v in self.subcell_res]
        for k in ("padding", "young", "poisson", "emin", "pcg_tol"):
            setattr(c, k, float(getattr(self, k)))
        c.stabilization = int(self.stabilization.kind)
        c.gamma = float(self.stabilization.gamma)
        c.clamped_face = _CLAMP[self.clamped_face]
        c.pcg_maxit = int(self.pcg_maxit)
        c.solver = int(self.solver)
        return c


@dataclass
class ImmersedSolution:
    """ImmersedSolution (immersed.rs:541-547) plus the true relative residual."""
    u: np.ndarray
    compliance: float
    iterations: int
    residual: float
    true_residual: float
    status: int = ICC_OK


class ImmersedProblem:
    """build_immersed_problem(mesh, opts) on the GPU (immersed.rs:466-538)."""

    def __init__(self, ctx, mesh, opts=None):
        if opts is None:
            opts = ImmersedOptions()
        self.ctx, self.mesh, self.opts = ctx, mesh, opts
        self._c = opts.to_c()
        f = None if opts.force is None else np.ascontiguousarray(opts.force, dtype=np.float64)
        h = C.c_void_p()
        check(lib().icc_immersed_mesh_create(ctx.h, C.byref(self._c), _p(mesh.vertices), len(mesh.vertices),
                                             _p(mesh.triangles), len(mesh.triangles), _p(f), C.byref(h)))
        self.h = h
        sz = np.zeros(4, dtype=np.int64)
        gb = np.zeros(6)
        check(lib().icc_immersed_mesh_sizes(h, _p(sz), _p(gb)))
        self.nel, self.ndof, self.n_faces, self.n_free = (int(v) for v in sz)
        self.grid_bounds = gb.reshape(3, 2)
        self.status = np.zeros(self.nel, dtype=np.int8)
        self.weights = np.zeros(self.nel)
        self.force = np.zeros(self.ndof)
        fr = np.zeros(self.ndof, dtype=np.uint8)
        check(lib().icc_immersed_mesh_problem(h, _p(self.status), _p(self.weights), _p(self.force), _p(fr)))
        self.free = fr.astype(bool)

    def grid_handle(self):
        g = C.c_void_p()
        check(lib().icc_immersed_mesh_grid(self.h, C.byref(g)))
        return g

    @property
    def h_min(self):
        """Smallest background element size (SpaceBox::element_size)."""
        gb = self.grid_bounds
        return min((gb[d, 1] - gb[d, 0]) / self.opts.grid_res[d] for d in range(3))

    def set_force(self, force):
        """Replaces the problem's force (e.g. zeroed before surface loads)."""
        f = np.ascontiguousarray(force, dtype=np.float64).reshape(self.ndof)
        check(lib().icc_immersed_mesh_set_force(self.h, _p(f)))
        self.force = f.copy()

    def add_surface_terms(self, st):
        """ImmersedProblem::add_surface_terms (immersed.rs:541-548)."""
        check(lib().icc_immersed_mesh_add_surface(self.h, st.h))
        fr = np.zeros(self.ndof, dtype=np.uint8)
        check(lib().icc_immersed_mesh_problem(self.h, None, None, _p(self.force), _p(fr)))

    def apply(self, x):
        """M (K + S) M x."""
        y = np.zeros(self.ndof)
        check(lib().icc_immersed_mesh_apply(self.h, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y)))
        return y

    def solve(self, tol=None, maxit=None, solver=None):
        """solve_immersed_problem(pb, tol, maxit) (immersed.rs:551-570)."""
        o = self.opts
        check(lib().icc_immersed_mesh_set_solver(self.h, float(o.pcg_tol if tol is None else tol),
                                                 int(o.pcg_maxit if maxit is None else maxit),
                                                 int(o.solver if solver is None else solver)))
        u = np.zeros(self.ndof)
        st = np.zeros(8)
        rc = lib().icc_immersed_mesh_solve(self.h, _p(u), _p(st))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        return ImmersedSolution(u, float(st[3]), int(st[0]), float(st[1]), float(st[2]), rc)

    def close(self):
        if getattr(self, "h", None):
            lib().icc_immersed_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_immersed_problem(ctx, mesh, opts=None):
    return ImmersedProblem(ctx, mesh, opts)


def solve_immersed_problem(pb, tol, maxit):
    return pb.solve(tol, maxit)


@dataclass
class ImmersedTopOptConfig:
    """Fields and defaults of ImmersedTopOptConfig (topopt_immersed.rs:21-52),
    ghost penalty gamma 1e-2 by default.  Extension: solver 0 = FP64 Jacobi
    PCG; 1 = FP32 PCG with FP64 reliable updates; 2 = the same with the
    split-FP16 inner contraction."""
    grid_res: tuple = (28, 20, 16)
    volfrac: float = 0.35
    penal: float = 3.0
    rmin: float = 1.8
    max_iter: int = 30
    young: float = 1.0
    poisson: float = 0.3
    emin: float = 1e-4
    stabilization: Stabilization = field(default_factory=Stabilization)
    pcg_tol: float = 1e-8
    pcg_maxit: int = 20000
    solver: int = 0

    def to_c(self):
        c = ImmTopOptConfigC()
        lib().icc_imm_topopt_config_default(C.byref(c))
        c.grid_res[:] = [int(v) for v in self.grid_res]
        for k in ("volfrac", "penal", "rmin", "young", "poisson", "emin", "pcg_tol"):
            setattr(c, k, float(getattr(self, k)))
        c.max_iter = int(self.max_iter)
        c.pcg_maxit = int(self.pcg_maxit)
        c.stabilization = int(self.stabilization.kind)
        c.gamma = float(self.stabilization.gamma)
        c.solver = int(self.solver)
        return c


@dataclass
class ImmersedTopOptHistory:
    """ImmersedTopOptHistory (topopt_immersed.rs:69-75)."""
    compliance: list = field(default_factory=list)
    volume: list = field(default_factory=list)
    change: list = field(default_factory=list)
    pcg_iterations: list = field(default_factory=list)
    density: np.ndarray = None


class ImmersedTopOpt:
    """ImmersedTopOpt::new(brep, cfg) / run(report) on the GPU."""

    def __init__(self, ctx, brep, cfg=None):
        if cfg is None:
            cfg = ImmersedTopOptConfig()
        self.ctx = ctx
        self.cfg = cfg
        self.brep = brep
        self._c = cfg.to_c()
        h = C.c_void_p()
        check(lib().icc_imm_topopt_create(ctx.h, C.byref(self._c), _p(brep.vertices), len(brep.vertices),
                                          _p(brep.triangles), len(brep.triangles), C.byref(h)))
        self.h = h
        sz = np.zeros(5, dtype=np.int64)
        cad = C.c_double()
        gb = np.zeros(6)
        check(lib().icc_imm_topopt_sizes(h, _p(sz), C.byref(cad), _p(gb)))
        self.nel, self.ndof, self.n_active, self.n_faces, self.n_free = (int(v) for v in sz)
        self.cad_volume = cad.value
        self.grid_bounds = gb.reshape(3, 2)
        self.status = np.zeros(self.nel, dtype=np.int8)
        self.weights = np.zeros(self.nel)
        act = np.zeros(self.nel, dtype=np.uint8)
        self.force = np.zeros(self.ndof)
        fr = np.zeros(self.ndof, dtype=np.uint8)
        check(lib().icc_imm_topopt_problem(h, _p(self.status), _p(self.weights), _p(act), _p(self.force), _p(fr)))
        self.active = act.astype(bool)
        self.free = fr.astype(bool)

    def run(self, report=None):
        """Runs cfg.max_iter iterations; report(iter, c, vol, change, pcg_its)
        is called after each (topopt_immersed.rs:155-247)."""
        m = int(self.cfg.max_iter)
        comp, vol, ch = np.zeros(m), np.zeros(m), np.zeros(m)
        its = np.zeros(m, dtype=np.int64)
        dens = np.zeros(self.nel)
        cb = None
        if report is not None:
            cb = IMM_REPORT_FN(lambda it, c, v, d, n, _u: report(int(it), c, v, d, int(n)))
        check(lib().icc_imm_topopt_run(self.h, cb, None, _p(comp), _p(vol), _p(ch), _p(its), _p(dens)))
        return ImmersedTopOptHistory(list(comp), list(vol), list(ch), [int(v) for v in its], dens)

    def state(self):
        u = np.zeros(self.ndof)
        check(lib().icc_imm_topopt_state(self.h, _p(u)))
        return u

    def close(self):
        if getattr(self, "h", None):
            lib().icc_imm_topopt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
