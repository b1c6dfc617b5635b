"""Precision policy.

The reference toggles JAX's process-wide x64 flag (``config.py:29-62``).
This engine has no global numeric mode: each simulator picks complex64 or
complex128 explicitly, and every reduction (norms, expectations, energies,
gradients) accumulates in FP64 whatever the state precision.  The functions
below keep the reference's names; they record the caller's preference, which
``circuit_to_energy_fn`` uses as its default state precision.
"""
import functools

_explicit = False
_float64 = True
_compensated = True


def set_precision(float64: bool = True) -> None:
    """Choose the default state precision of energy functions (c128 if True)."""
    global _explicit, _float64
    _float64 = bool(float64)
    _explicit = True


def ensure_x64() -> None:
    """Default to complex128 unless set_precision() chose otherwise."""
    global _float64
    if not _explicit:
        _float64 = True


def default_float64() -> bool:
    return _float64


def set_compensation(enable: bool = True) -> None:
    """Default gate arithmetic of complex64 simulators: compensated matrices
    (each FP32 matrix applied as hi + lo, lo part first; on by default) or
    plain FP32-rounded matrices (the reference's complex64 arithmetic,
    ``circuits/compiler.py:172-179``).  Has no effect on complex128."""
    global _compensated
    _compensated = bool(enable)


def default_compensation() -> bool:
    return _compensated


def with_x64(fn):
    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        ensure_x64()
        return fn(*args, **kwargs)
    return wrapper
