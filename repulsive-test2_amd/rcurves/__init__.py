"""Host-side mirror of the Repulsive Curves src/core API over librcurves.so."""
from . import _lib, core, curves, dispatch  # noqa: F401
from .core import (  # noqa: F401
    DEFAULTS, PenaltyConfig, SobolevStepOptions, TangentPointEnergy, assemble_a, barycenter_block,
    barycenter_target, calculate_disjoint_pairs, calculate_energy, edge_lengths, edge_lengths_block,
    gradient_analytical, gradient_finite_diff, penalty_energy_and_gradient, point_block, sobolev_step,
    sobolev_step_set, solve_constrained_gradient_set_frozen, step, StepOptions, total_length,
    total_length_block,
)
