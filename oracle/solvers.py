"""Linear solvers (FP64).

* ``tridia``  -- Thomas algorithm in the order of ``python-port/ibis/physics/solvers.py:17-83``
  (F77 ``tridia``, ``fortran/ibis_2.6b4/math.f:189``).  Pinned by the 3x3 known answer
  x = [1, 1, 1] (``tests/test_solvers.py:23-36``) and by ``scipy.linalg.solve_banded`` on
  random diagonally dominant systems (``tests/test_solvers.py:38-61``).
* ``gauss_nopivot`` -- no-pivot Gaussian elimination in F77 ``linsolve`` loop order
  (``python-port/ibis/physics/canopy_exchange.py:26-61``).  Pinned by ``numpy.linalg.solve``
  on well-conditioned systems and by a 2x2 closed form.
"""
from __future__ import annotations

import numpy as np


def tridia(a, b, c, r):
    """Solve tridiagonal systems row-wise; inputs (N, K), a[:, 0] and c[:, -1] unused."""
    n, k = b.shape
    gam = np.empty((n, k))
    x = np.empty((n, k))
    alpha = 1.0 / b[:, 0]
    gam[:, 0] = c[:, 0] * alpha
    x[:, 0] = r[:, 0] * alpha
    for j in range(1, k - 1):
        alpha = 1.0 / (b[:, j] - a[:, j] * gam[:, j - 1])
        gam[:, j] = c[:, j] * alpha
        x[:, j] = (r[:, j] - a[:, j] * x[:, j - 1]) * alpha
    j = k - 1
    x[:, j] = (r[:, j] - a[:, j] * x[:, j - 1]) / (b[:, j] - a[:, j] * gam[:, j - 1])
    for j in range(k - 2, -1, -1):
        x[:, j] = x[:, j] - gam[:, j] * x[:, j + 1]
    return x


def gauss_nopivot(A, rhs):
    """Solve (N, n, n) systems by forward elimination without pivoting, then back-substitution.

    Row updates cover the whole row, as the reference does; entries that are zero stay
    exactly zero, so the sparsity of the canopy system does not change the rounding.
    """
    A = np.array(A, dtype=np.float64, copy=True)
    b = np.array(rhs, dtype=np.float64, copy=True)
    n = A.shape[1]
    for p in range(n - 1):
        piv = A[:, p, p]
        for i in range(p + 1, n):
            f = A[:, i, p] / piv
            A[:, i, :] = A[:, i, :] - f[:, None] * A[:, p, :]
            b[:, i] = b[:, i] - f * b[:, p]
    x = np.zeros_like(b)
    for p in range(n - 1, -1, -1):
        s = np.zeros(A.shape[0])
        for j in range(p + 1, n):
            s = s + A[:, p, j] * x[:, j]
        x[:, p] = (b[:, p] - s) / A[:, p, p]
    return x
