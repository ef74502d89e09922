"""sfdhip -- MI355X split-step path of SuperfluidDynamics.jl behind the
reference's Grid / Field / NumModel API (host side of libsfdhip.so)."""
from ._lib import SfdError, ShapeError, ArgError, build, load, EXPORTS, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    HipArray, ComplexField, RealField, Grid, Field, norm, normalize, normalize_,
    PotentialZero, PotentialQuadratic, PotentialArray, GrossPitaevskiiParameters,
    NumModelSplit1, NumModelSplit2, NumModelSplitImag, timeStep, energy, energy_all, solve, steps_with_energy,
    solveLapRot, solveNL, BatchImagSolver, adjoint_gradient, FidelityLoss, fidelity_loss, project, inner,
    mul_x, mul_y, mul_z, ldiv_x, ldiv_y, ldiv_z, grid_x, grid_y, grid_z, linrange_coords, fftfreq,
    Plan, GradientField, computeDerivatives,
)
from .io import WriterSave, write, read, CheckpointError  # noqa: F401

__all__ = [n for n in dir() if not n.startswith("_")]
