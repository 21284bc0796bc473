"""Python host API over libicc: contexts, grids, operators, filters.

numpy arrays go through the host-pointer entry points; torch CUDA tensors go
through the *_dev entry points on the context's HIP stream (call
Context.sync() before reading them from torch's stream).
"""
import ctypes as C

import numpy as np

from ._abi import (ICC_BREAKDOWN_CODES, ICC_F32, ICC_F64, ICC_OK, check, lib)


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())  # torch tensor


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _is_dev(a):
    return a is not None and not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)


class Context:
    """One device + one HIP stream (icc_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().icc_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        check(lib().icc_ctx_sync(self.h))

    @property
    def stream(self):
        return lib().icc_ctx_stream(self.h)

    def close(self):
        if self.h:
            lib().icc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mfma_layout(self):
        """(f64, f32) accumulator contents [lane, reg] of the layout probe."""
        out = np.zeros(512)
        check(lib().icc_debug_mfma_layout(self.h, _p(out)))
        return out[:256].reshape(64, 4), out[256:].reshape(64, 4)

    def mfma_layout_bf16(self):
        """Accumulator contents [lane, reg] of the BF16 16x16x32 layout probe."""
        out = np.zeros(256)
        check(lib().icc_debug_mfma_layout_bf16(self.h, _p(out)))
        return out.reshape(64, 4)

    def mfma_layout_f16(self):
        """(16x16x32, 16x16x16) accumulator contents [lane, reg] of the FP16
        layout probe."""
        out = np.zeros(512)
        check(lib().icc_debug_mfma_layout_f16(self.h, _p(out)))
        return out[:256].reshape(64, 4), out[256:].reshape(64, 4)


def element_types(nel, bounds):
    """(ntypes, type_id) of elasticity_element_matrices (rust/src/iga/
    elasticity.rs:30-98) for a p = 2 grid, on the host without a device:
    the numbering Grid.type_id() reports."""
    n = np.ascontiguousarray(np.asarray(nel, dtype=np.int64).reshape(3))
    b = _f64(np.asarray(bounds, dtype=np.float64).ravel())
    nt = np.zeros(1, dtype=np.int64)
    tid = np.zeros(int(np.prod(n)), dtype=np.int64)
    check(lib().icc_element_types(_p(n), _p(b), _p(nt), _p(tid)))
    return int(nt[0]), tid


class Grid:
    """p = 2 B-spline space on a box with its exact element matrices
    (SpaceBox::new + elasticity_element_matrices, rust/src/iga)."""

    def __init__(self, ctx, nel, bounds, young=1.0, poisson=0.3, degree=2):
        self.ctx = ctx
        n = np.asarray(nel, dtype=np.int64)
        b = _f64(np.asarray(bounds, dtype=np.float64).ravel())
        h = C.c_void_p()
        check(lib().icc_grid_create(ctx.h, _p(n), _p(b), int(degree), float(young), float(poisson), C.byref(h)))
        self.h = h
        s = np.zeros(11, dtype=np.int64)
        check(lib().icc_grid_sizes(h, _p(s)))
        self.nel_tot, self.ndof_sc, self.ndof, self.ntypes, self.nsh = (int(v) for v in s[:5])
        self.ncp = [int(v) for v in s[5:8]]
        self.nel = [int(v) for v in s[8:11]]
        self.young, self.poisson = float(young), float(poisson)
        self.bounds = b.reshape(3, 2)

    def set_split_tiling(self, max_groups):
        """Element groups per tile of the split-FP16 apply (1..16)."""
        check(lib().icc_grid_set_split_tiling(self.h, int(max_groups)))

    def connectivity(self):
        out = np.zeros(self.nel_tot * 81, dtype=np.int64)
        check(lib().icc_grid_connectivity(self.h, _p(out)))
        return out

    def type_id(self):
        out = np.zeros(self.nel_tot, dtype=np.int64)
        check(lib().icc_grid_type_id(self.h, _p(out)))
        return out

    def element_matrix(self, t):
        out = np.zeros((81, 81))
        check(lib().icc_grid_element_matrix(self.h, int(t), _p(out)))
        return out

    def set_tiling(self, variant, wg_per_cu=2):
        """Tile shape order variant % 6; variants 0..5 balance the work per
        persistent workgroup (splitting tiles), 6..11 give every workgroup an
        equal number of whole tiles.  wg_per_cu: 1, 2 or 4."""
        check(lib().icc_grid_set_tiling(self.h, int(variant), int(wg_per_cu)))

    @property
    def tiling(self):
        """(variant, workgroups per CU, tile count)."""
        out = np.zeros(3, dtype=np.int64)
        check(lib().icc_grid_tiling(self.h, _p(out)))
        return tuple(int(v) for v in out)

    def wq_diag(self, x=None, penal=3.0, emin=1e-3):
        d = np.zeros(self.ndof)
        xx = None if x is None else _f64(x)
        check(lib().icc_wq_diag(self.h, _p(xx), float(penal), float(emin), _p(d)))
        return d

    def close(self):
        if self.h:
            lib().icc_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """y = M sum_e s_e C_e^T K_t(e) C_e (M x): the matvec closure of the
    reference's PCG (topopt_iga.rs:151-159), matrix-free on the GPU."""

    def __init__(self, grid, scale=None, free=None):
        self.grid = grid
        h = C.c_void_p()
        check(lib().icc_op_create(grid.h, C.byref(h)))
        self.h = h
        if scale is not None:
            self.set_scale(scale)
        if free is not None:
            self.set_free(free)

    @property
    def ndof(self):
        return self.grid.ndof

    def set_scale(self, s):
        if _is_dev(s):
            check(lib().icc_op_set_scale_dev(self.h, _p(s)))
        else:
            check(lib().icc_op_set_scale(self.h, _p(_f64(s))))

    def set_free(self, free):
        fr = None if free is None else np.ascontiguousarray(free, dtype=np.uint8)
        check(lib().icc_op_set_free_mask(self.h, _p(fr)))

    def set_sparse(self, indptr, indices, data):
        ip = np.ascontiguousarray(indptr, dtype=np.int64)
        ix = np.ascontiguousarray(indices, dtype=np.int64)
        dd = _f64(data)
        check(lib().icc_op_set_sparse(self.h, len(ip) - 1, _p(ip), _p(ix), _p(dd)))

    def set_deterministic(self, on=True):
        """Bitwise-repeatable FP64 / FP32 applies and solves (per-element
        outputs and a fixed-order gather instead of the atomic scatter)."""
        check(lib().icc_op_set_deterministic(self.h, 1 if on else 0))

    def add_triplets(self, rows, cols, vals):
        """S = S + T (CsrMatrix::plus_triplets, sparse.rs:98-113)."""
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        v = _f64(vals)
        check(lib().icc_op_add_triplets(self.h, len(v), _p(r), _p(c), _p(v)))

    def set_legacy_stabilization(self, status, gamma):
        """Legacy cut-cell stabilisation (immersed.rs:168-222) as a matrix-free
        face term; status None removes it.  Returns the number of faces."""
        nf = C.c_int64(0)
        st = None if status is None else np.ascontiguousarray(status, dtype=np.int8)
        check(lib().icc_op_set_legacy_stabilization(self.h, _p(st), float(gamma), C.byref(nf)))
        return int(nf.value)

    def set_ghost_penalty(self, status, gamma, young=None):
        """Consistent ghost penalty (immersed.rs:234-370) as a matrix-free
        face term; status None removes it.  Returns the number of faces."""
        nf = C.c_int64(0)
        st = None if status is None else np.ascontiguousarray(status, dtype=np.int8)
        yg = self.grid.young if young is None else float(young)
        check(lib().icc_op_set_ghost_penalty(self.h, _p(st), float(gamma), yg, C.byref(nf)))
        return int(nf.value)

    def face_frobenius(self):
        """Frobenius norm of the assembled face stabilisation term."""
        f = C.c_double(0.0)
        check(lib().icc_op_face_frobenius(self.h, C.byref(f)))
        return f.value

    def apply(self, x, prec=ICC_F64, out=None):
        if _is_dev(x):
            check(lib().icc_apply_dev(self.h, prec, _p(x), _p(out)))
            return out
        y = np.zeros(self.ndof)
        check(lib().icc_apply(self.h, prec, _p(_f64(x)), _p(y)))
        return y

    def apply_accumulate(self, x, out, prec=ICC_F64):
        """out += A x on device tensors (tile kernel only, no sparse term)."""
        check(lib().icc_apply_accumulate_dev(self.h, prec, _p(x), _p(out)))
        return out

    def diag(self, out=None):
        if out is not None and _is_dev(out):
            check(lib().icc_diag_dev(self.h, _p(out)))
            return out
        d = np.zeros(self.ndof)
        check(lib().icc_diag(self.h, _p(d)))
        return d

    def pcg(self, b, diag, tol, maxit):
        """PcgSolver::solve (solver.rs:14-76) with this operator as the closure.
        Returns (x, iters, res, status)."""
        it = np.zeros(1, dtype=np.int64)
        res = np.zeros(1)
        x = np.zeros(self.ndof)
        rc = lib().icc_pcg(self.h, _p(_f64(b)), _p(_f64(diag)), float(tol), int(maxit), _p(x), _p(it), _p(res))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        return x, int(it[0]), float(res[0]), rc

    def solve_ir(self, f, tol=1e-8, inner=ICC_F32, inner_tol=1e-3, max_outer=50, max_inner=100000, u0=None):
        """Mixed-precision iterative refinement (FP64 outer residual).
        Returns (u, stats dict, status)."""
        u = np.zeros(self.ndof) if u0 is None else _f64(u0).copy()
        stats = np.zeros(8)
        rc = lib().icc_solve_ir(self.h, int(inner), _p(_f64(f)), float(tol), float(inner_tol), int(max_outer),
                                int(max_inner), _p(u), _p(stats))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        st = dict(outer=int(stats[0]), inner_iters=int(stats[1]), true_res=float(stats[2]),
                  fp64_applies=int(stats[3]), inner_applies=int(stats[4]))
        return u, st, rc

    def autotune(self, prec=ICC_F64, reps=5):
        """Picks the grid's tiling by timing this operator's apply; returns
        (variant, workgroups per CU, microseconds per apply)."""
        v = np.zeros(1, dtype=np.int64)
        wg = np.zeros(1, dtype=np.int64)
        us = np.zeros(1)
        check(lib().icc_grid_autotune(self.h, int(prec), int(reps), _p(v), _p(wg), _p(us)))
        return int(v[0]), int(wg[0]), float(us[0])

    def solve_mixed(self, f, tol=1e-8, delta=0.1, maxit=1000000, u0=None, inner=ICC_F32):
        """PCG with FP32 vectors and the inner contraction in `inner` (ICC_F32,
        ICC_F16X or ICC_BF16), with reliable updates against the FP64
        operator: the inner residual is replaced by the FP64 true residual
        each time it drops by `delta`.  Returns (u, stats dict, status)."""
        u = np.zeros(self.ndof) if u0 is None else _f64(u0).copy()
        stats = np.zeros(8)
        rc = lib().icc_solve_mixed_prec(self.h, int(inner), _p(_f64(f)), float(tol), float(delta), int(maxit), _p(u),
                                        _p(stats))
        check(rc, allowed=(ICC_OK,) + ICC_BREAKDOWN_CODES)
        st = dict(updates=int(stats[0]), inner_iters=int(stats[1]), true_res=float(stats[2]),
                  fp64_applies=int(stats[3]), inner_applies=int(stats[4]))
        return u, st, rc

    def energies(self, u):
        ee = np.zeros(self.grid.nel_tot)
        check(lib().icc_energies(self.h, _p(_f64(u)), _p(ee)))
        return ee

    def sens(self, u, x, penal, emin, w=None):
        dc = np.zeros(self.grid.nel_tot)
        ww = None if w is None else _f64(w)
        check(lib().icc_sens(self.h, _p(_f64(u)), _p(_f64(x)), float(penal), float(emin), _p(ww), _p(dc)))
        return dc

    def close(self):
        if self.h:
            lib().icc_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Filter:
    """Filter on element centres.

    Without ``active``: the structured filter of TopOpt3D (topopt_iga.rs:100-126),
    radius ``rmin`` in elements.  With ``active`` (one flag per element): the
    design-domain filter of ImmersedTopOpt::new (topopt_immersed.rs:117-151),
    radius ``rmin * mean(h)``, only active neighbours of active cells; inactive
    rows are empty.
    """

    def __init__(self, grid, rmin, active=None):
        self.grid = grid
        h = C.c_void_p()
        if active is None:
            check(lib().icc_filter_create(grid.h, float(rmin), C.byref(h)))
        else:
            act = np.ascontiguousarray(np.asarray(active).reshape(-1) != 0, dtype=np.uint8)
            if act.size != grid.nel_tot:
                raise ValueError(f"active has {act.size} entries, the grid has {grid.nel_tot} elements")
            check(lib().icc_filter_create_active(grid.h, float(rmin), _p(act), C.byref(h)))
        self.h = h

    def sens(self, x, dc):
        out = np.zeros(self.grid.nel_tot)
        check(lib().icc_filter_sens(self.h, _p(_f64(x)), _p(_f64(dc)), _p(out)))
        return out

    def fwd(self, x):
        out = np.zeros(self.grid.nel_tot)
        check(lib().icc_filter_fwd(self.h, _p(_f64(x)), _p(out)))
        return out

    def bwd(self, g):
        out = np.zeros(self.grid.nel_tot)
        check(lib().icc_filter_bwd(self.h, _p(_f64(g)), _p(out)))
        return out

    def close(self):
        if self.h:
            lib().icc_filter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def classify_cylinders(ctx, box_lo, box_hi, cyl, grid_bounds, res, sub=(4, 4, 4)):
    """classify_background_cells + element_weights (immersed.rs:88-155) with
    the box-minus-cylinders SDF predicate.  Returns (status int8, weights)."""
    nel = int(np.prod(res))
    status = np.zeros(nel, dtype=np.int8)
    w = np.zeros(nel)
    cyl = _f64(np.asarray(cyl, dtype=np.float64).ravel())
    check(lib().icc_classify_cylinders(ctx.h, _p(_f64(box_lo)), _p(_f64(box_hi)), _p(cyl), len(cyl) // 3,
                                       _p(_f64(np.asarray(grid_bounds).ravel())),
                                       _p(np.asarray(res, dtype=np.int64)), _p(np.asarray(sub, dtype=np.int64)),
                                       _p(status), _p(w)))
    return status, w


class WqMatrix:
    """Weighted-quadrature elasticity stiffness of a grid, formed on the GPU
    (wq_stiffness / wq_stiffness_layout, rust/src/wq.rs:324-514):
    form(density, x, penal, emin) with density ICC_DENSITY_NONE,
    ICC_DENSITY_ELEMENT (x[nel]) or ICC_DENSITY_SPLINE (x[ndof_sc]).
    layout: "interior" (WqLayout::Interior, the default of wq_stiffness) or
    "calabro" (WqLayout::Calabro)."""

    LAYOUTS = {"interior": 0, "calabro": 1}

    def __init__(self, grid, layout="interior"):
        if layout not in self.LAYOUTS:
            raise ValueError(f"unknown WQ layout {layout!r}; expected one of {sorted(self.LAYOUTS)}")
        self.grid = grid
        self.layout = layout
        h = C.c_void_p()
        check(lib().icc_wq_create_layout(grid.h, self.LAYOUTS[layout], C.byref(h)))
        self.h = h

    def form(self, density=0, x=None, penal=3.0, emin=1e-3):
        if x is not None and _is_dev(x):
            check(lib().icc_wq_form_dev(self.h, int(density), _p(x), float(penal), float(emin)))
            return self
        xx = None if x is None else _f64(x)
        check(lib().icc_wq_form(self.h, int(density), _p(xx), float(penal), float(emin)))
        return self

    def apply(self, x, out=None):
        if _is_dev(x):
            check(lib().icc_wq_apply_dev(self.h, _p(x), _p(out)))
            return out
        y = np.zeros(self.grid.ndof)
        check(lib().icc_wq_apply(self.h, _p(_f64(x)), _p(y)))
        return y

    def csr(self):
        """(indptr, indices, data) of the formed matrix, rows/columns
        component-blocked, columns ascending."""
        nnz = C.c_int64(0)
        check(lib().icc_wq_csr(self.h, C.byref(nnz), None, None, None))
        ip = np.zeros(self.grid.ndof + 1, dtype=np.int64)
        ix = np.zeros(nnz.value, dtype=np.int64)
        dd = np.zeros(nnz.value)
        check(lib().icc_wq_csr(self.h, C.byref(nnz), _p(ip), _p(ix), _p(dd)))
        return ip, ix, dd

    def close(self):
        if getattr(self, "h", None):
            lib().icc_wq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

