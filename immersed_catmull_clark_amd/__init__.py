"""MI355X-native matrix-free p = 2 immersed-IGA elasticity hot path.

Host API over libicc.so (HIP kernels for gfx950 behind a C ABI, include/icc.h).
"""
from ._abi import (ICC_BF16, ICC_F16X, ICC_F32, ICC_F64, ICC_OK, ICC_EINVAL, ICC_ENOMEM, ICC_EHIP, ICC_EBREAKDOWN,  # noqa: F401
                   ICC_ENOCONV, ICC_ENOTSUP, ICC_DENSITY_NONE, ICC_DENSITY_ELEMENT, ICC_DENSITY_SPLINE, IccError,
                   LIB_PATH, lib)
from .core import Context, Filter, Grid, Operator, WqMatrix, classify_cylinders, element_types  # noqa: F401
from .topopt import CantileverTopOpt, CantileverTopOptConfig, TopOptHistory  # noqa: F401
from .immersed import ImmersedConfig, ImmersedSdf  # noqa: F401
from .batch import Batch  # noqa: F401
from .imm_topopt import (ImmersedOptions, ImmersedProblem, ImmersedSolution, ImmersedTopOpt,  # noqa: F401
                         ImmersedTopOptConfig, ImmersedTopOptHistory, Stabilization, TriangleMesh3D,
                         build_immersed_problem, element_weights, padded_bounds, solve_immersed_problem)
from .surface import (Dirichlet, NeumannPressure, NeumannTraction, RobinFoundation, SurfaceTerms,  # noqa: F401
                      dirichlet_penalty_fn, eval_vector_at, scalar_basis_at, surface_boundary_terms,
                      surface_quadrature)

__all__ = ["Context", "Grid", "Operator", "Filter", "classify_cylinders", "ICC_F64", "ICC_F32", "ICC_BF16", "ICC_F16X", "IccError"]
