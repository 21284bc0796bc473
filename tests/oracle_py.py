"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, the smoke check and the bench's
cpu_baseline leg only, to check the product against.  The product never
imports it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None

i64 = C.c_int64
dbl = C.c_double
vp = C.c_void_p


def _p(a):
    return a.ctypes.data_as(vp) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_space_new.restype = vp
        L.orc_space_new.argtypes = [i64, vp, vp, i64]
        L.orc_space_free.argtypes = [vp]
        L.orc_space_sizes.argtypes = [vp, vp]
        L.orc_space_connectivity.argtypes = [vp, vp]
        L.orc_space_elem_mats.restype = i64
        L.orc_space_elem_mats.argtypes = [vp, dbl, dbl]
        L.orc_space_types.argtypes = [vp, vp, vp]
        L.orc_space_load_fixture.argtypes = [vp, vp]
        L.orc_space_traction_xmax.argtypes = [vp, vp, vp]
        L.orc_space_greville.argtypes = [vp, vp]
        L.orc_space_assemble.restype = i64
        L.orc_space_assemble.argtypes = [vp, vp]
        L.orc_space_assemble_pattern.restype = i64
        L.orc_space_assemble_pattern.argtypes = [vp, vp]
        L.orc_csr_matvec.argtypes = [vp, vp, vp]
        L.orc_csr_diag.argtypes = [vp, vp]
        L.orc_csr_frobenius.restype = dbl
        L.orc_csr_frobenius.argtypes = [vp]
        L.orc_csr_pcg.argtypes = [vp, vp, vp, vp, dbl, i64, vp, vp, vp]
        L.orc_space_energies.argtypes = [vp, vp, vp]
        L.orc_fast_sens.argtypes = [vp, vp, vp, i64, dbl, dbl, dbl, dbl, vp]
        L.orc_filter_apply.argtypes = [vp, dbl, i64, vp, vp, vp]
        L.orc_design_filter_apply.restype = None
        L.orc_design_filter_apply.argtypes = [vp, dbl, i64, vp, vp, vp, vp]
        L.orc_filter_rows_match.restype = i64
        L.orc_filter_rows_match.argtypes = [vp, dbl]
        L.orc_topopt_run.restype = i64
        L.orc_topopt_run.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_topopt_new.restype = vp
        L.orc_topopt_new.argtypes = [vp, vp, vp]
        L.orc_topopt_free.argtypes = [vp]
        L.orc_sdf_new.restype = vp
        L.orc_sdf_new.argtypes = [vp, vp, vp, i64, vp, vp, vp, C.c_int]
        L.orc_sdf_free.argtypes = [vp]
        L.orc_sdf_sizes.restype = dbl
        L.orc_sdf_sizes.argtypes = [vp, vp]
        L.orc_sdf_get.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_sdf_matvec.argtypes = [vp, vp, vp]
        L.orc_sdf_solve.argtypes = [vp, dbl, i64, vp, vp]
        L.orc_topopt_force_free.argtypes = [vp, vp, vp]
        L.orc_topopt_solve_state.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_immersed_new.restype = vp
        L.orc_immersed_new.argtypes = [vp, i64, vp, i64, vp, vp]
        L.orc_immersed_free.argtypes = [vp]
        L.orc_immersed_ndof.restype = i64
        L.orc_immersed_ndof.argtypes = [vp]
        L.orc_immersed_info.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_immersed_matvec.argtypes = [vp, vp, vp]
        L.orc_immersed_solve.argtypes = [vp, dbl, i64, vp, vp, vp, vp]
        L.orc_immersed_topopt_run.restype = i64
        L.orc_immersed_topopt_run.argtypes = [vp, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.orc_classify_cylinders.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, vp]
        L.orc_classify_mesh.argtypes = [vp, i64, vp, i64, vp, vp, vp, vp, vp]
        L.orc_padded_bounds.argtypes = [vp, i64, dbl, vp]
        L.orc_ghost_penalty_matvec.restype = dbl
        L.orc_ghost_penalty_matvec.argtypes = [vp, vp, i64, vp, dbl, dbl, vp, vp]
        L.orc_set_threads.argtypes = [i64]
        L.orc_wq_rule.restype = i64
        L.orc_wq_rule.argtypes = [i64, i64, i64, i64, i64, vp]
        L.orc_space_assemble_wq.restype = i64
        L.orc_space_assemble_wq.argtypes = [vp, dbl, dbl, i64, vp, dbl, dbl, i64]
        L.orc_space_assemble_structured.restype = i64
        L.orc_space_assemble_structured.argtypes = [vp, vp]
        L.orc_compare_with_pattern.restype = i64
        L.orc_compare_with_pattern.argtypes = [vp, vp]
        L.orc_num_threads.restype = i64
        L.orc_num_threads.argtypes = []
        L.orc_surface_quadrature.restype = i64
        L.orc_surface_quadrature.argtypes = [vp, i64, vp, i64, vp, i64, dbl, vp]
        L.orc_scalar_basis_at.restype = i64
        L.orc_scalar_basis_at.argtypes = [vp, i64, vp, vp, vp]
        L.orc_surface_new.restype = vp
        L.orc_surface_new.argtypes = [vp, i64]
        L.orc_surface_free.argtypes = [vp]
        L.orc_surface_add.argtypes = [vp, i64, vp, i64, vp, i64, vp, i64, vp, dbl]
        L.orc_surface_dirichlet_affine.argtypes = [vp, vp, i64, vp, i64, vp, i64, dbl, vp, vp]
        L.orc_surface_get.restype = i64
        L.orc_surface_get.argtypes = [vp, vp, vp, vp, vp]
        L.orc_immersed_add_surface.argtypes = [vp, vp]
        L.orc_immersed_zero_force.argtypes = [vp]
        L.orc_immersed_eval_vector_at.argtypes = [vp, vp, vp, vp]
        _lib = L
    return _lib


def lame(young, poisson):
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    mu = young / (2.0 * (1.0 + poisson))
    return lam, mu


class Space:
    """Oracle SpaceBox (rust/src/iga/space.rs) with element matrices and CSR."""

    def __init__(self, bounds, nsub, degree):
        b = np.ascontiguousarray(np.asarray(bounds, dtype=np.float64).ravel())
        n = np.ascontiguousarray(np.asarray(nsub, dtype=np.int64))
        self.h = lib().orc_space_new(len(nsub), _p(b), _p(n), int(degree))
        s = np.zeros(12, dtype=np.int64)
        lib().orc_space_sizes(self.h, _p(s))
        self.dim, self.nel, self.ndof_sc, self.ndof, self.nsh, self.nsh_sc = (int(v) for v in s[:6])
        self.ndof_dir = [int(v) for v in s[6:6 + self.dim]]
        self.nel_dir = [int(v) for v in s[9:9 + self.dim]]
        self.ntypes = None

    def __del__(self):
        try:
            lib().orc_space_free(self.h)
        except Exception:
            pass

    def connectivity(self):
        out = np.zeros(self.nel * self.nsh, dtype=np.int64)
        lib().orc_space_connectivity(self.h, _p(out))
        return out

    def elem_mats(self, lam, mu):
        self.ntypes = int(lib().orc_space_elem_mats(self.h, lam, mu))
        types = np.zeros((self.ntypes, self.nsh, self.nsh))
        tid = np.zeros(self.nel, dtype=np.int64)
        lib().orc_space_types(self.h, _p(types), _p(tid))
        return types, tid

    def load_fixture(self):
        out = np.zeros(self.ndof)
        lib().orc_space_load_fixture(self.h, _p(out))
        return out

    def traction_xmax(self, t):
        t = np.ascontiguousarray(np.asarray(t, dtype=np.float64))
        out = np.zeros(self.ndof)
        lib().orc_space_traction_xmax(self.h, _p(t), _p(out))
        return out

    def greville(self):
        out = np.zeros((self.ndof_sc, 3))
        lib().orc_space_greville(self.h, _p(out))
        return out

    def assemble(self, scale=None, pattern=False):
        s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        if pattern:
            return int(lib().orc_space_assemble_pattern(self.h, _p(s)))
        return int(lib().orc_space_assemble(self.h, _p(s)))

    def assemble_structured(self, scale):
        return int(lib().orc_space_assemble_structured(self.h, _p(np.ascontiguousarray(scale, dtype=np.float64))))

    def compare_with_pattern(self, scale):
        return int(lib().orc_compare_with_pattern(self.h, _p(np.ascontiguousarray(scale, dtype=np.float64))))

    def assemble_wq(self, young, poisson, density_kind, x, penal, emin, layout=0):
        xx = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        return int(lib().orc_space_assemble_wq(self.h, young, poisson, density_kind, _p(xx), penal, emin, layout))

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        lib().orc_csr_matvec(self.h, _p(x), _p(y))
        return y

    def diag(self):
        d = np.zeros(self.ndof)
        lib().orc_csr_diag(self.h, _p(d))
        return d

    def frobenius(self):
        return float(lib().orc_csr_frobenius(self.h))

    def pcg(self, b, diag, free, tol, maxit):
        b = np.ascontiguousarray(b, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        fr = np.ascontiguousarray(free, dtype=np.uint8)
        x = np.zeros(self.ndof)
        it = np.zeros(1, dtype=np.int64)
        res = np.zeros(1)
        lib().orc_csr_pcg(self.h, _p(b), _p(diag), _p(fr), tol, int(maxit), _p(x), _p(it), _p(res))
        return x, int(it[0]), float(res[0])

    def energies(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(self.nel)
        lib().orc_space_energies(self.h, _p(u), _p(out))
        return out

    def fast_sens(self, u, x, spline, penal, emin, young, poisson):
        u = np.ascontiguousarray(u, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.ndof_sc if spline else self.nel)
        lib().orc_fast_sens(self.h, _p(u), _p(x), int(spline), penal, emin, young, poisson, _p(out))
        return out

    def filter_apply(self, rmin, mode, x, v):
        x = np.ascontiguousarray(x if x is not None else np.zeros(self.nel), dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.nel)
        lib().orc_filter_apply(self.h, rmin, mode, _p(x), _p(v), _p(out))
        return out

    def design_filter_apply(self, rmin, mode, active, x, v):
        """Design-domain filter (topopt_immersed.rs:117-151); modes as filter_apply."""
        act = np.ascontiguousarray(np.asarray(active).reshape(-1) != 0, dtype=np.uint8)
        x = np.ascontiguousarray(x, dtype=np.float64) if x is not None else np.zeros(act.size)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(act.size)
        lib().orc_design_filter_apply(self.h, rmin, mode, _p(act), _p(x), _p(v), _p(out))
        return out

    def filter_rows_match(self, rmin):
        return bool(lib().orc_filter_rows_match(self.h, rmin))


def topopt_params(volfrac=0.3, penal=3.0, rmin=1.5, max_iter=40, young=1.0, poisson=0.3, emin=1e-3,
                  move=0.2, pcg_tol=1e-10):
    return np.array([volfrac, penal, rmin, max_iter, young, poisson, emin, move, pcg_tol], dtype=np.float64)


def topopt_run(nel, size=(1.2, 0.6, 0.3), **kw):
    par = topopt_params(**kw)
    nel_a = np.asarray(nel, dtype=np.int64)
    size_a = np.asarray(size, dtype=np.float64)
    mi = int(par[3])
    comp = np.zeros(mi)
    ch = np.zeros(mi)
    its = np.zeros(mi, dtype=np.int64)
    dens = np.zeros(int(np.prod(nel)))
    n = lib().orc_topopt_run(_p(nel_a), _p(size_a), _p(par), _p(comp), _p(ch), _p(its), _p(dens))
    return comp[:n], ch[:n], its[:n], dens


class TopOpt:
    def __init__(self, nel, size=(1.2, 0.6, 0.3), **kw):
        self.par = topopt_params(**kw)
        self.nel_a = np.asarray(nel, dtype=np.int64)
        self.size_a = np.asarray(size, dtype=np.float64)
        self.h = lib().orc_topopt_new(_p(self.nel_a), _p(self.size_a), _p(self.par))
        self.space = Space([[0, size[0]], [0, size[1]], [0, size[2]]], nel, 2)
        self.ndof = self.space.ndof
        self.nel = self.space.nel
        self.force = np.zeros(self.ndof)
        self.free = np.zeros(self.ndof, dtype=np.uint8)
        lib().orc_topopt_force_free(self.h, _p(self.force), _p(self.free))

    def __del__(self):
        try:
            lib().orc_topopt_free(self.h)
        except Exception:
            pass

    def solve_state(self, x, u0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        u = np.zeros(self.ndof)
        it = np.zeros(1, dtype=np.int64)
        res = np.zeros(1)
        lib().orc_topopt_solve_state(self.h, _p(x), _p(u0), _p(u), _p(it), _p(res))
        return u, int(it[0]), float(res[0])


class Immersed:
    def __init__(self, nodes, tris, res, degree=2, young=1.0, poisson=0.3, stab_kind=1, gamma=0.05, emin=1e-4,
                 clamp=0, sub=(4, 4, 4), padding=0.05):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        tris = np.ascontiguousarray(tris, dtype=np.int64)
        ipar = np.array([res[0], res[1], res[2], degree, stab_kind, clamp, sub[0], sub[1], sub[2]], dtype=np.int64)
        par = np.array([young, poisson, gamma, emin, padding], dtype=np.float64)
        self.h = lib().orc_immersed_new(_p(nodes), len(nodes), _p(tris), len(tris), _p(ipar), _p(par))
        self.ndof = int(lib().orc_immersed_ndof(self.h))
        nel = int(np.prod(res))
        self.bounds = np.zeros(6)
        fro = np.zeros(1)
        self.status = np.zeros(nel, dtype=np.int8)
        self.weights = np.zeros(nel)
        self.force = np.zeros(self.ndof)
        self.free = np.zeros(self.ndof, dtype=np.uint8)
        lib().orc_immersed_info(self.h, _p(self.bounds), _p(fro), _p(self.status), _p(self.weights), _p(self.force),
                                _p(self.free))
        self.stab_frobenius = float(fro[0])

    def __del__(self):
        try:
            lib().orc_immersed_free(self.h)
        except Exception:
            pass

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        lib().orc_immersed_matvec(self.h, _p(x), _p(y))
        return y

    def solve(self, tol, maxit):
        u = np.zeros(self.ndof)
        c = np.zeros(1)
        it = np.zeros(1, dtype=np.int64)
        res = np.zeros(1)
        lib().orc_immersed_solve(self.h, tol, int(maxit), _p(u), _p(c), _p(it), _p(res))
        return u, float(c[0]), int(it[0]), float(res[0])

    def zero_force(self):
        lib().orc_immersed_zero_force(self.h)
        self.force[:] = 0.0

    def add_surface(self, surf):
        """ImmersedProblem::add_surface_terms (immersed.rs:541-548)."""
        lib().orc_immersed_add_surface(self.h, surf.h)
        self.force += surf.terms()[3]

    def eval_vector_at(self, u, x):
        u = np.ascontiguousarray(u, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(3)
        lib().orc_immersed_eval_vector_at(self.h, _p(u), _p(x), _p(out))
        return out


def _mesh_args(nodes, tris, sel):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int64)
    sel = np.ascontiguousarray(sel, dtype=np.int64)
    return nodes, tris, sel


def surface_quadrature(nodes, tris, sel, max_edge):
    """(x [n,3], w [n], normal [n,3]) of surface_quadrature (immersed_bc.rs:109-148)."""
    nodes, tris, sel = _mesh_args(nodes, tris, sel)
    n = lib().orc_surface_quadrature(_p(nodes), len(nodes), _p(tris), len(tris), _p(sel), len(sel), max_edge, None)
    out = np.zeros((n, 7))
    lib().orc_surface_quadrature(_p(nodes), len(nodes), _p(tris), len(tris), _p(sel), len(sel), max_edge, _p(out))
    return out[:, :3].copy(), out[:, 3].copy(), out[:, 4:].copy()


def scalar_basis_at(space, x, immersed=False):
    """(indices, values) of scalar_basis_at (immersed_bc.rs:25-50) on an oracle Space or Immersed."""
    idx = np.zeros(64, dtype=np.int64)
    val = np.zeros(64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = lib().orc_scalar_basis_at(space.h, 1 if immersed else 0, _p(x), _p(idx), _p(val))
    return idx[:n].copy(), val[:n].copy()


class Surface:
    """Surface terms (immersed_bc.rs SurfaceTerms) on an oracle Space or Immersed problem."""

    DIRICHLET, TRACTION, PRESSURE, ROBIN = 0, 1, 2, 3

    def __init__(self, space, immersed=False):
        self._owner = space
        self.h = lib().orc_surface_new(space.h, 1 if immersed else 0)

    def __del__(self):
        try:
            lib().orc_surface_free(self.h)
        except Exception:
            pass

    def add(self, kind, nodes, tris, sel, beta=0.0, comps=(1, 1, 1), values=(0, 0, 0), pressure=0.0, kn=0.0, kt=0.0):
        """One labelled condition of surface_boundary_terms (immersed_bc.rs:186-247)."""
        nodes, tris, sel = _mesh_args(nodes, tris, sel)
        par = np.array(list(comps) + list(values) + [pressure, kn, kt], dtype=np.float64)
        lib().orc_surface_add(self.h, kind, _p(nodes), len(nodes), _p(tris), len(tris), _p(sel), len(sel), _p(par),
                              beta)

    def dirichlet_affine(self, nodes, tris, sel, beta, g0, G):
        """dirichlet_penalty_fn (immersed_bc.rs:161-182) with g(x) = g0 + G x."""
        nodes, tris, sel = _mesh_args(nodes, tris, sel)
        g0 = np.ascontiguousarray(g0, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64).reshape(9)
        lib().orc_surface_dirichlet_affine(self.h, _p(nodes), len(nodes), _p(tris), len(tris), _p(sel), len(sel),
                                           beta, _p(g0), _p(G))

    def terms(self):
        """(rows, cols, vals, force) after into_triplets."""
        n = lib().orc_surface_get(self.h, None, None, None, None)
        r = np.zeros(n, dtype=np.int64)
        c = np.zeros(n, dtype=np.int64)
        v = np.zeros(n)
        f = np.zeros(self._owner.ndof)
        lib().orc_surface_get(self.h, _p(r), _p(c), _p(v), _p(f))
        return r, c, v, f


def immersed_topopt_run(nodes, tris, res, max_iter, volfrac, penal, rmin, gamma, young=1.0, poisson=0.3, emin=1e-4,
                        pcg_tol=1e-10, stab_kind=1, full=False):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int64)
    ipar = np.array([res[0], res[1], res[2], max_iter, stab_kind], dtype=np.int64)
    par = np.array([volfrac, penal, rmin, young, poisson, emin, gamma, pcg_tol], dtype=np.float64)
    comp = np.zeros(max_iter)
    vol = np.zeros(max_iter)
    dens = np.zeros(int(np.prod(res)))
    change = np.zeros(max_iter)
    its = np.zeros(max_iter, dtype=np.int64)
    n = lib().orc_immersed_topopt_run(_p(nodes), len(nodes), _p(tris), len(tris), _p(ipar), _p(par), _p(comp),
                                      _p(vol), _p(dens), _p(change), _p(its))
    if full:
        return comp[:n], vol[:n], dens, change[:n], its[:n]
    return comp[:n], vol[:n], dens


def classify_mesh(nodes, tris, grid_bounds, res, sub=(4, 4, 4)):
    """(status, weights) by ray parity (immersed.rs:29-155)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int64)
    gb = np.ascontiguousarray(np.asarray(grid_bounds, dtype=np.float64).ravel())
    r = np.asarray(res, dtype=np.int64)
    s = np.asarray(sub, dtype=np.int64)
    nel = int(np.prod(res))
    status = np.zeros(nel, dtype=np.int8)
    w = np.zeros(nel)
    lib().orc_classify_mesh(_p(nodes), len(nodes), _p(tris), len(tris), _p(gb), _p(r), _p(s), _p(status), _p(w))
    return status, w


def ghost_penalty_matvec(bounds, res, degree, status, gamma, young, x):
    """(G x, |G|_F) for the ghost-penalty matrix G (immersed.rs:234-370)."""
    b = np.ascontiguousarray(np.asarray(bounds, dtype=np.float64).ravel())
    r = np.asarray(res, dtype=np.int64)
    st = np.ascontiguousarray(status, dtype=np.int8)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    fro = lib().orc_ghost_penalty_matvec(_p(b), _p(r), int(degree), _p(st), float(gamma), float(young), _p(x), _p(y))
    return y, fro


def padded_bounds(nodes, padding):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    out = np.zeros(6)
    lib().orc_padded_bounds(_p(nodes), len(nodes), float(padding), _p(out))
    return out


def classify_cylinders(box_lo, box_hi, cyl, grid_bounds, res, sub=(4, 4, 4)):
    box_lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    box_hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    cyl = np.ascontiguousarray(np.asarray(cyl, dtype=np.float64).ravel())
    gb = np.ascontiguousarray(np.asarray(grid_bounds, dtype=np.float64).ravel())
    r = np.asarray(res, dtype=np.int64)
    s = np.asarray(sub, dtype=np.int64)
    nel = int(np.prod(res))
    status = np.zeros(nel, dtype=np.int8)
    w = np.zeros(nel)
    lib().orc_classify_cylinders(_p(box_lo), _p(box_hi), _p(cyl), len(cyl) // 3, _p(gb), _p(r), _p(s), _p(status),
                                 _p(w))
    return status, w


class SdfImmersed:
    """Oracle restatement of the immersed problem for a box minus z-cylinders
    (build_immersed_problem / solve_immersed_problem, immersed.rs:466-570)."""

    def __init__(self, box_lo, box_hi, cyl, res, sub=(4, 4, 4), young=1.0, poisson=0.3, emin=1e-4, active_min=0.05,
                 penalty_factor=1e3, traction=(0.0, -1.0, 0.0), mode=1):
        lo = np.ascontiguousarray(box_lo, dtype=np.float64)
        hi = np.ascontiguousarray(box_hi, dtype=np.float64)
        cy = np.ascontiguousarray(np.asarray(cyl, dtype=np.float64).ravel())
        r = np.ascontiguousarray(res, dtype=np.int64)
        sb = np.ascontiguousarray(sub, dtype=np.int64)
        par = np.ascontiguousarray([young, poisson, emin, active_min, penalty_factor] + list(traction), dtype=np.float64)
        self.h = lib().orc_sdf_new(_p(lo), _p(hi), _p(cy), len(cy) // 3, _p(r), _p(sb), _p(par), int(mode))
        sz = np.zeros(3, dtype=np.int64)
        self.beta = float(lib().orc_sdf_sizes(self.h, _p(sz)))
        self.nel, self.ndof, self.nnz = (int(v) for v in sz)
        self.status = np.zeros(self.nel, dtype=np.int8)
        self.weights = np.zeros(self.nel)
        self.active = np.zeros(self.nel, dtype=np.uint8)
        self.scale = np.zeros(self.nel)
        self.free = np.zeros(self.ndof, dtype=np.uint8)
        self.force = np.zeros(self.ndof)
        lib().orc_sdf_get(self.h, _p(self.status), _p(self.weights), _p(self.active), _p(self.scale), _p(self.free),
                          _p(self.force))

    def __del__(self):
        try:
            lib().orc_sdf_free(self.h)
        except Exception:
            pass

    def matvec(self, x):
        y = np.zeros(self.ndof)
        lib().orc_sdf_matvec(self.h, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y))
        return y

    def solve(self, tol, maxit):
        u = np.zeros(self.ndof)
        st = np.zeros(3)
        lib().orc_sdf_solve(self.h, tol, int(maxit), _p(u), _p(st))
        return u, int(st[0]), float(st[1]), float(st[2])


def seeded_cylinders(seed, n=3, r=(0.08, 0.15), cx=(0.4, 1.6), cy=(0.3, 0.7)):
    """Cylinder holes (cx, cy, r) parallel to z, axes away from the box faces."""
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(*cx, n), rng.uniform(*cy, n), rng.uniform(*r, n)])


def wq_rule(nel, p, layout, which, i=0):
    n = int(lib().orc_wq_rule(nel, p, layout, which, i, None))
    out = np.zeros(n)
    lib().orc_wq_rule(nel, p, layout, which, i, _p(out))
    return out


def probe_vectors(n, count=3):
    """v_j(i) = sin(0.37 i j), i = 1..n, j = 1..count (iga_vs_matlab.rs:52-60)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    return [np.sin(0.37 * i * j) for j in range(1, count + 1)]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
