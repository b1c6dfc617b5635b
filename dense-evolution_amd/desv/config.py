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


def with_x64(fn):
    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        ensure_x64()
        return fn(*args, **kwargs)
    return wrapper
