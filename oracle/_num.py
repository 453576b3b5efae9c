"""Integer powers in the multiplication order of JAX's binary exponentiation.

``x ** n`` with a Python integer ``n`` lowers to ``lax.integer_pow`` and ``jnp.power`` with an
integer array exponent to the masked square-and-multiply loop of ``jax.numpy.power``; both
multiply the powers x^(2^i) of the set bits in increasing order of i.  Using the same order
here keeps the FP64 oracle and the FP64 kernel on the same rounding path (e.g. x^3 = x*(x*x),
x^4 = (x*x)*(x*x)).
"""
from __future__ import annotations

import numpy as np


def ipow(x, n):
    """x ** n for a non-negative integer scalar or integer array n (at most 6 bits)."""
    x = np.asarray(x, dtype=np.float64)
    n = np.asarray(n, dtype=np.int64)
    acc = np.where((x == 0.0) & (n != 0), 0.0, 1.0)
    base = x
    for _ in range(6):
        acc = np.where((n & 1) != 0, acc * base, acc)
        base = base * base
        n = n >> 1
    return acc
