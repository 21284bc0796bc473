"""Surface boundary conditions on triangle surfaces, mirroring the reference
crate's rust/src/immersed_bc.rs (`surface_quadrature`, `scalar_basis_at`,
`SurfaceTerms`, `surface_boundary_terms`, `dirichlet_penalty_fn`) and
`eval_vector_at` (rust/src/iga/fields.rs:118-146), over libicc (the
per-element surface blocks are formed on the GPU, icc_surface_*).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._abi import check, lib

ICC_BC_DIRICHLET, ICC_BC_TRACTION, ICC_BC_PRESSURE, ICC_BC_ROBIN = 0, 1, 2, 3
POINT_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _grid_handle(obj):
    """The icc_grid of a Grid or an ImmersedProblem."""
    if hasattr(obj, "grid_handle"):
        return obj.grid_handle()
    return obj.h


def _sel(sel):
    return np.ascontiguousarray(np.asarray(sel, dtype=np.int64).reshape(-1))


def surface_quadrature(mesh, tris, max_edge):
    """(x [n, 3], w [n], normal [n, 3]) on the triangles `tris` of `mesh`
    (immersed_bc.rs:109-148)."""
    sel = _sel(tris)
    n = C.c_int64()
    check(lib().icc_surface_quadrature(_p(mesh.vertices), len(mesh.vertices), _p(mesh.triangles), len(mesh.triangles),
                                       _p(sel), len(sel), float(max_edge), 0, None, C.byref(n)))
    out = np.zeros((n.value, 7))
    check(lib().icc_surface_quadrature(_p(mesh.vertices), len(mesh.vertices), _p(mesh.triangles), len(mesh.triangles),
                                       _p(sel), len(sel), float(max_edge), n.value, _p(out), C.byref(n)))
    return out[:, :3].copy(), out[:, 3].copy(), out[:, 4:].copy()


def scalar_basis_at(space, x):
    """(global indices [27], values [27]) of the nonzero scalar basis
    functions at x (immersed_bc.rs:25-50); space: a Grid or an ImmersedProblem."""
    idx = np.zeros(27, dtype=np.int64)
    val = np.zeros(27)
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))
    check(lib().icc_scalar_basis_at(_grid_handle(space), _p(xx), _p(idx), _p(val)))
    return idx, val


def eval_vector_at(space, u, x, ndof_sc):
    """u(x) of a component-blocked vector field (fields.rs:118-146)."""
    idx, val = scalar_basis_at(space, x)
    out = np.zeros(3)
    for a in range(27):
        for c in range(3):
            out[c] += val[a] * u[c * ndof_sc + idx[a]]
    return out


@dataclass
class Dirichlet:
    """BoundaryConditionType::Dirichlet: penalty on the flagged components."""
    components: tuple = (True, True, True)
    values: tuple = (0.0, 0.0, 0.0)


@dataclass
class NeumannTraction:
    traction: tuple = (0.0, 0.0, 0.0)


@dataclass
class NeumannPressure:
    pressure: float = 0.0


@dataclass
class RobinFoundation:
    normal_stiffness: float = 0.0
    tangential_stiffness: float = 0.0


class SurfaceTerms:
    """SurfaceTerms (immersed_bc.rs:152-160) on a Grid or an ImmersedProblem:
    stiffness triplets and a load; conditions accumulate in call order."""

    def __init__(self, space):
        self._space = space  # keeps the grid alive
        h = C.c_void_p()
        check(lib().icc_surface_create(_grid_handle(space), C.byref(h)))
        self.h = h
        self._keep = []

    def add(self, mesh, tris, bc, beta=0.0, max_edge=0.0):
        """One labelled condition of surface_boundary_terms on the triangles
        `tris`; beta is the Dirichlet penalty (factor * E / h_min);
        max_edge <= 0 takes 0.25 h_min like the reference."""
        par = np.zeros(8)
        if isinstance(bc, Dirichlet):
            kind = ICC_BC_DIRICHLET
            par[:3] = [1.0 if c else 0.0 for c in bc.components]
            par[3:6] = bc.values
            par[6] = beta
        elif isinstance(bc, NeumannTraction):
            kind = ICC_BC_TRACTION
            par[:3] = bc.traction
        elif isinstance(bc, NeumannPressure):
            kind = ICC_BC_PRESSURE
            par[0] = bc.pressure
        elif isinstance(bc, RobinFoundation):
            kind = ICC_BC_ROBIN
            par[:2] = [bc.normal_stiffness, bc.tangential_stiffness]
        else:
            raise TypeError(f"unknown boundary condition {bc!r}")
        sel = _sel(tris)
        check(lib().icc_surface_add(self.h, kind, _p(par), _p(mesh.vertices), len(mesh.vertices),
                                    _p(mesh.triangles), len(mesh.triangles), _p(sel), len(sel), float(max_edge)))
        return self

    def add_dirichlet_fn(self, mesh, tris, beta, g, max_edge=0.0):
        """dirichlet_penalty_fn's term (immersed_bc.rs:161-182): beta int (u - g(x)).v."""
        def cb(x, out, _user):
            r = g(np.array([x[0], x[1], x[2]]))
            out[0], out[1], out[2] = float(r[0]), float(r[1]), float(r[2])

        fn = POINT_FN(cb)
        self._keep.append(fn)
        sel = _sel(tris)
        check(lib().icc_surface_add_dirichlet_fn(self.h, float(beta), C.cast(fn, C.c_void_p), None,
                                                 _p(mesh.vertices), len(mesh.vertices), _p(mesh.triangles),
                                                 len(mesh.triangles), _p(sel), len(sel), float(max_edge)))
        return self

    def sizes(self):
        """(triplets, element blocks, quadrature points)."""
        s = np.zeros(3, dtype=np.int64)
        check(lib().icc_surface_sizes(self.h, _p(s)))
        return tuple(int(v) for v in s)

    def terms(self, ndof):
        """(rows, cols, vals, force [ndof])."""
        n = self.sizes()[0]
        r = np.zeros(n, dtype=np.int64)
        c = np.zeros(n, dtype=np.int64)
        v = np.zeros(n)
        f = np.zeros(ndof)
        check(lib().icc_surface_terms(self.h, _p(r), _p(c), _p(v), _p(f)))
        return r, c, v, f

    def close(self):
        if getattr(self, "h", None):
            lib().icc_surface_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def surface_boundary_terms(space, mesh, conditions, penalty_factor, young, h_min):
    """surface_boundary_terms (immersed_bc.rs:186-247) over an explicit list
    of (triangle indices, condition) pairs; beta = penalty_factor * young / h_min."""
    st = SurfaceTerms(space)
    beta = penalty_factor * young / h_min
    for tris, bc in conditions:
        st.add(mesh, tris, bc, beta=beta)
    return st


def dirichlet_penalty_fn(space, mesh, tris, beta, g):
    """dirichlet_penalty_fn (immersed_bc.rs:161-182)."""
    return SurfaceTerms(space).add_dirichlet_fn(mesh, tris, beta, g)
