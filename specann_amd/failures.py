"""Exceptions raised by the solvers and the out-of-memory test (utils/failures.py:1-25).

Same contract as the reference: NumericalFailure when no alpha produced a solution, and
is_out_of_memory walks the cause chain and matches the usual allocator messages, so a run that
runs out of device or host memory is tagged "out of memory" and the CLI exits with status 2.
"""
from __future__ import annotations

from ._lib import DeviceOutOfMemory, EigensolverBreakdown  # noqa: F401  (re-exported)


class NumericalFailure(RuntimeError):
    """The solver ran but produced no valid solution (e.g. the eigensolver failed at every alpha)."""


_OUT_OF_MEMORY_TEXT = ('hip out of memory', 'hiperroroutofmemory', 'cuda out of memory', 'out of memory on device',
                       'outofmemoryerror', 'out of memory', 'unable to allocate', 'cannot allocate memory',
                       'bad_alloc', 'out of space', 'memoryerror')


def is_out_of_memory(exc: BaseException) -> bool:
    """True if the exception or one in its cause chain means the host or the GPU ran out of memory."""
    seen = set()
    while exc is not None and id(exc) not in seen and len(seen) < 10:
        seen.add(id(exc))
        if isinstance(exc, MemoryError):
            return True
        if any(s in f"{type(exc).__name__}: {exc}".lower() for s in _OUT_OF_MEMORY_TEXT):
            return True
        exc = exc.__cause__ if exc.__cause__ is not None else exc.__context__
    return False


def is_numerical(exc: BaseException) -> bool:
    """A numerical failure (the alpha is skipped), not a bug or a broken environment
    (models/spectral_annealing.py:74-77)."""
    import numpy as np
    return (isinstance(exc, (EigensolverBreakdown, ArithmeticError, np.linalg.LinAlgError))
            or type(exc).__name__ in ('LinAlgError',))
