"""ctypes binding of libicc.so (the C ABI declared in include/icc.h).

The library is required: importing the package on a machine without the
built library raises, and every entry point runs the HIP path; there is no
CPU fallback.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libicc.so")

ICC_OK, ICC_EINVAL, ICC_ENOMEM, ICC_EHIP, ICC_EBREAKDOWN, ICC_ENOCONV, ICC_ENOTSUP = range(7)
ICC_F64, ICC_F32, ICC_BF16, ICC_F16X = 0, 1, 2, 3

i64 = C.c_int64
dbl = C.c_double
vp = C.c_void_p
cint = C.c_int
vpp = C.POINTER(C.c_void_p)

# (name, restype, argtypes)
SIGNATURES = [
    ("icc_last_error", C.c_char_p, []),
    ("icc_version", C.c_char_p, []),
    ("icc_ctx_create", cint, [cint, vpp]),
    ("icc_ctx_destroy", cint, [vp]),
    ("icc_ctx_sync", cint, [vp]),
    ("icc_ctx_stream", vp, [vp]),
    ("icc_grid_create", cint, [vp, vp, vp, i64, dbl, dbl, vpp]),
    ("icc_grid_destroy", cint, [vp]),
    ("icc_grid_sizes", cint, [vp, vp]),
    ("icc_grid_connectivity", cint, [vp, vp]),
    ("icc_grid_type_id", cint, [vp, vp]),
    ("icc_element_types", cint, [vp, vp, vp, vp]),
    ("icc_grid_element_matrix", cint, [vp, i64, vp]),
    ("icc_op_create", cint, [vp, vpp]),
    ("icc_op_destroy", cint, [vp]),
    ("icc_op_set_scale", cint, [vp, vp]),
    ("icc_op_set_scale_dev", cint, [vp, vp]),
    ("icc_op_set_free_mask", cint, [vp, vp]),
    ("icc_op_set_sparse", cint, [vp, i64, vp, vp, vp]),
    ("icc_apply", cint, [vp, cint, vp, vp]),
    ("icc_apply_dev", cint, [vp, cint, vp, vp]),
    ("icc_apply_accumulate_dev", cint, [vp, cint, vp, vp]),
    ("icc_diag", cint, [vp, vp]),
    ("icc_diag_dev", cint, [vp, vp]),
    ("icc_pcg", cint, [vp, vp, vp, dbl, i64, vp, vp, vp]),
    ("icc_pcg_dev", cint, [vp, vp, vp, dbl, i64, vp, vp, vp]),
    ("icc_pcg_dev_noalloc", cint, [vp, vp, vp, dbl, i64, vp, vp, vp]),
    ("icc_solve_ir_dev", cint, [vp, cint, vp, dbl, dbl, i64, i64, vp, vp]),
    ("icc_solve_ir", cint, [vp, cint, vp, dbl, dbl, i64, i64, vp, vp]),
    ("icc_solve_mixed_dev", cint, [vp, vp, dbl, dbl, i64, vp, vp]),
    ("icc_solve_mixed", cint, [vp, vp, dbl, dbl, i64, vp, vp]),
    ("icc_solve_mixed_prec_dev", cint, [vp, cint, vp, dbl, dbl, i64, vp, vp]),
    ("icc_solve_mixed_prec", cint, [vp, cint, vp, dbl, dbl, i64, vp, vp]),
    ("icc_energies", cint, [vp, vp, vp]),
    ("icc_energies_dev", cint, [vp, vp, vp]),
    ("icc_sens_dev", cint, [vp, vp, vp, dbl, dbl, vp, vp]),
    ("icc_sens", cint, [vp, vp, vp, dbl, dbl, vp, vp]),
    ("icc_filter_create", cint, [vp, dbl, vpp]),
    ("icc_filter_destroy", cint, [vp]),
    ("icc_filter_create_active", cint, [vp, dbl, vp, vpp]),
    ("icc_filter_sens_dev", cint, [vp, vp, vp, vp]),
    ("icc_filter_fwd_dev", cint, [vp, vp, vp]),
    ("icc_filter_bwd_dev", cint, [vp, vp, vp]),
    ("icc_filter_sens", cint, [vp, vp, vp, vp]),
    ("icc_filter_fwd", cint, [vp, vp, vp]),
    ("icc_filter_bwd", cint, [vp, vp, vp]),
    ("icc_wq_diag", cint, [vp, vp, dbl, dbl, vp]),
    ("icc_wq_diag_dev", cint, [vp, vp, dbl, dbl, vp]),
    ("icc_classify_cylinders", cint, [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp]),
    ("icc_debug_mfma_layout", cint, [vp, vp]),
    ("icc_topopt_config_default", None, [vp]),
    ("icc_topopt_create", cint, [vp, vp, vpp]),
    ("icc_topopt_destroy", cint, [vp]),
    ("icc_topopt_sizes", cint, [vp, vp]),
    ("icc_topopt_force", cint, [vp, vp, vp]),
    ("icc_topopt_solve_state", cint, [vp, vp, vp, vp, vp, vp]),
    ("icc_topopt_element_energies", cint, [vp, vp, vp]),
    ("icc_topopt_compliance", cint, [vp, vp, vp, vp, vp]),
    ("icc_topopt_run", cint, [vp, vp, vp, vp, vp, vp, vp, vp]),
    ("icc_grid_set_tiling", cint, [vp, cint, cint]),
    ("icc_grid_tiling", cint, [vp, vp]),
    ("icc_grid_autotune", cint, [vp, cint, cint, vp, vp, vp]),
    ("icc_batch_create", cint, [vp, i64, vp, vp, dbl, dbl, vpp]),
    ("icc_batch_destroy", cint, [vp]),
    ("icc_batch_sizes", cint, [vp, vp, vp]),
    ("icc_batch_set", cint, [vp, i64, vp, vp, vp]),
    ("icc_batch_set_traction", cint, [vp, i64, vp]),
    ("icc_batch_autotune", cint, [vp, cint, cint]),
    ("icc_batch_solve", cint, [vp, cint, dbl, dbl, i64, vp, vp]),
    ("icc_immersed_config_default", None, [vp]),
    ("icc_immersed_create", cint, [vp, vp, vp, i64, vpp]),
    ("icc_immersed_destroy", cint, [vp]),
    ("icc_immersed_sizes", cint, [vp, vp, vp]),
    ("icc_immersed_problem", cint, [vp, vp, vp, vp, vp, vp, vp]),
    ("icc_immersed_apply", cint, [vp, vp, vp]),
    ("icc_immersed_solve", cint, [vp, vp, vp]),
    ("icc_immersed_solve_dev", cint, [vp, vp, vp]),
    ("icc_op_set_legacy_stabilization", cint, [vp, vp, dbl, vp]),
    ("icc_op_set_ghost_penalty", cint, [vp, vp, dbl, dbl, vp]),
    ("icc_op_face_frobenius", cint, [vp, vp]),
    ("icc_classify_mesh", cint, [vp, vp, i64, vp, i64, vp, vp, vp, vp, vp]),
    ("icc_mesh_padded_bounds", cint, [vp, i64, dbl, vp]),
    ("icc_imm_topopt_config_default", None, [vp]),
    ("icc_imm_topopt_create", cint, [vp, vp, vp, i64, vp, i64, vpp]),
    ("icc_imm_topopt_destroy", cint, [vp]),
    ("icc_imm_topopt_sizes", cint, [vp, vp, vp, vp]),
    ("icc_imm_topopt_problem", cint, [vp, vp, vp, vp, vp, vp]),
    ("icc_imm_topopt_run", cint, [vp, vp, vp, vp, vp, vp, vp, vp]),
    ("icc_imm_topopt_state", cint, [vp, vp]),
    ("icc_debug_mfma_layout_bf16", cint, [vp, vp]),
    ("icc_debug_mfma_layout_f16", cint, [vp, vp]),
    ("icc_immersed_mesh_config_default", None, [vp]),
    ("icc_immersed_mesh_create", cint, [vp, vp, vp, i64, vp, i64, vp, vpp]),
    ("icc_immersed_mesh_destroy", cint, [vp]),
    ("icc_immersed_mesh_sizes", cint, [vp, vp, vp]),
    ("icc_immersed_mesh_problem", cint, [vp, vp, vp, vp, vp]),
    ("icc_immersed_mesh_set_solver", cint, [vp, dbl, i64, cint]),
    ("icc_immersed_mesh_apply", cint, [vp, vp, vp]),
    ("icc_immersed_mesh_solve", cint, [vp, vp, vp]),
    ("icc_immersed_mesh_set_force", cint, [vp, vp]),
    ("icc_immersed_mesh_grid", cint, [vp, vpp]),
    ("icc_immersed_mesh_add_surface", cint, [vp, vp]),
    ("icc_op_add_triplets", cint, [vp, i64, vp, vp, vp]),
    ("icc_op_set_deterministic", cint, [vp, cint]),
    ("icc_grid_set_split_tiling", cint, [vp, cint]),
    ("icc_surface_quadrature", cint, [vp, i64, vp, i64, vp, i64, dbl, i64, vp, vp]),
    ("icc_scalar_basis_at", cint, [vp, vp, vp, vp]),
    ("icc_surface_create", cint, [vp, vpp]),
    ("icc_surface_destroy", cint, [vp]),
    ("icc_surface_add", cint, [vp, cint, vp, vp, i64, vp, i64, vp, i64, dbl]),
    ("icc_surface_add_dirichlet_fn", cint, [vp, dbl, vp, vp, vp, i64, vp, i64, vp, i64, dbl]),
    ("icc_surface_sizes", cint, [vp, vp]),
    ("icc_surface_terms", cint, [vp, vp, vp, vp, vp]),
    ("icc_wq_create", cint, [vp, vpp]),
    ("icc_wq_create_layout", cint, [vp, cint, vpp]),
    ("icc_debug_quantize_kmat", dbl, [vp, cint, vp]),
    ("icc_debug_wq_rule", i64, [i64, cint, cint, i64, vp]),
    ("icc_wq_destroy", cint, [vp]),
    ("icc_wq_form", cint, [vp, cint, vp, dbl, dbl]),
    ("icc_wq_form_dev", cint, [vp, cint, vp, dbl, dbl]),
    ("icc_wq_apply", cint, [vp, vp, vp]),
    ("icc_wq_apply_dev", cint, [vp, vp, vp]),
    ("icc_wq_csr", cint, [vp, vp, vp, vp, vp]),
]

REPORT_FN = C.CFUNCTYPE(None, i64, dbl, dbl, dbl, vp)
IMM_REPORT_FN = C.CFUNCTYPE(None, i64, dbl, dbl, dbl, i64, vp)
ICC_STAB_NONE, ICC_STAB_LEGACY, ICC_STAB_GHOST = range(3)
ICC_DENSITY_NONE, ICC_DENSITY_ELEMENT, ICC_DENSITY_SPLINE = range(3)
ICC_CLAMP_NONE, ICC_CLAMP_BOTTOM, ICC_CLAMP_TOP, ICC_CLAMP_LEFT = -1, 0, 1, 2


class TopOptConfigC(C.Structure):
    """icc_topopt_config (include/icc.h)."""
    _fields_ = [("nel", i64 * 3), ("size", dbl * 3), ("volfrac", dbl), ("penal", dbl), ("rmin", dbl),
                ("max_iter", i64), ("young", dbl), ("poisson", dbl), ("emin", dbl), ("move_limit", dbl),
                ("pcg_tol", dbl), ("load", C.c_int32), ("filter", C.c_int32), ("solver", C.c_int32),
                ("pad", C.c_int32)]


class ImmersedConfigC(C.Structure):
    """icc_immersed_config (include/icc.h)."""
    _fields_ = [("box_lo", dbl * 3), ("box_hi", dbl * 3), ("res", i64 * 3), ("sub", i64 * 3), ("young", dbl),
                ("poisson", dbl), ("emin", dbl), ("active_min", dbl), ("penalty_factor", dbl),
                ("traction", dbl * 3), ("pcg_tol", dbl), ("pcg_maxit", i64), ("mode", C.c_int32),
                ("solver", C.c_int32)]


class ImmersedMeshConfigC(C.Structure):
    """icc_immersed_mesh_config (include/icc.h)."""
    _fields_ = [("grid_res", i64 * 3), ("padding", dbl), ("young", dbl), ("poisson", dbl),
                ("stabilization", C.c_int32), ("clamped_face", C.c_int32), ("gamma", dbl), ("emin", dbl),
                ("subcell_res", i64 * 3), ("pcg_tol", dbl), ("pcg_maxit", i64), ("solver", C.c_int32),
                ("pad", C.c_int32)]


class ImmTopOptConfigC(C.Structure):
    """icc_imm_topopt_config (include/icc.h)."""
    _fields_ = [("grid_res", i64 * 3), ("volfrac", dbl), ("penal", dbl), ("rmin", dbl), ("max_iter", i64),
                ("young", dbl), ("poisson", dbl), ("emin", dbl), ("stabilization", C.c_int32),
                ("solver", C.c_int32), ("gamma", dbl), ("pcg_tol", dbl), ("pcg_maxit", i64)]


_lib = None


class IccError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libicc error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m immersed_catmull_clark_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, allowed=(ICC_OK,)):
    if rc not in allowed:
        msg = lib().icc_last_error()
        raise IccError(rc, msg.decode() if msg else "")
    return rc

ICC_BREAKDOWN_CODES = (ICC_EBREAKDOWN, ICC_ENOCONV)
