"""MI355X (gfx950) implementation of g2o's per-iteration Levenberg-Marquardt
linear-system path (BlockSolver<BlockSolverTraits<6,3>> with Schur, block-Jacobi
PCG; BlockSolver<BlockSolverTraits<9,3>> for BAL problems; 6x6 pose graphs) behind a C ABI (include/g2o_hip.h), with a Python mirror of the
SparseOptimizer / Solver API."""
from ._lib import lib, LIB_PATH, FP32, FP64, SCHUR_EXPLICIT, SCHUR_IMPLICIT, KERNEL_NAMES, LEVENBERG, GAUSS_NEWTON, DOGLEG  # noqa: F401,E501
from .solver import BALAdjuster, BundleAdjuster, Context, G2oHipError, PoseGraphOptimizer, linear_solve_pcg, default_options  # noqa: F401
from .solver import full_block_rows, linear_solve_pcg_device  # noqa: F401
