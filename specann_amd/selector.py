"""The solvers of this package by the `solver_name` a run gives in configs/ (utils/selector.py:4-16)."""
import importlib

SOLVERS = {
    'spectral_annealing': ('specann_amd.spectral_annealing', 'SpectralAnnealing'),   # SpecAnn
    'spec_adj': ('specann_amd.endpoints', 'Alpha0GPU'),                             # SpecAdj (alpha = 0)
    'spec_signed_lap': ('specann_amd.endpoints', 'Alpha1GPU'),                      # SpecSignedLap (alpha = 1)
}


def get_solver_class(solver_name: str):
    """Import and return the class of a solver name."""
    if solver_name not in SOLVERS:
        raise ValueError(f"Unknown solver {solver_name!r}; choose one of {sorted(SOLVERS)}")
    module, cls = SOLVERS[solver_name]
    return getattr(importlib.import_module(module), cls)
