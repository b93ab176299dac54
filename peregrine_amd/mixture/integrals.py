"""Collision integrals of the Stockmayer potential at any reduced temperature and dipole strength.

Follows ``peregrine/simulator/mixture/transport/integrals.py:1-60``: each delta* column of the
Monchick-Mason tables (data/integrals.yaml) is a natural cubic spline in ln T* (the rows above
T* = 0), and between columns the value is the quadratic through the three nearest columns in
delta*, kept inside the table.
"""

import numpy as np

from peregrine_amd.mixture.species import database


class _Table:
    """One tabulated integral: knots ln T*, values per delta* column, and each column's second
    derivatives at the knots."""

    def __init__(self, tstar, values):
        tstar = np.asarray(tstar, dtype=np.float64)
        keep = tstar > 0.0
        self.x = np.log(tstar[keep])
        self.y = np.asarray(values, dtype=np.float64)[keep]
        self.m = self._second_derivatives(self.x, self.y)

    @staticmethod
    def _second_derivatives(x, y):
        """The natural spline's moments: y'' = 0 at both ends, continuity of y' inside."""
        n = len(x)
        h = np.diff(x)
        A = np.zeros((n, n))
        b = np.zeros((n,) + y.shape[1:])
        A[0, 0] = A[-1, -1] = 1.0
        for i in range(1, n - 1):
            A[i, i - 1], A[i, i], A[i, i + 1] = h[i - 1], 2.0 * (h[i - 1] + h[i]), h[i]
            b[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1])
        return np.linalg.solve(A, b)

    def column(self, xq, col):
        x, y, m = self.x, self.y, self.m
        i = np.clip(np.searchsorted(x, xq) - 1, 0, len(x) - 2)
        h = x[i + 1] - x[i]
        a, b = (x[i + 1] - xq) / h, (xq - x[i]) / h
        return a * y[i, col] + b * y[i + 1, col] + ((a**3 - a) * m[i, col] + (b**3 - b) * m[i + 1, col]) * h**2 / 6.0


class CollisionIntegrals:
    def __init__(self):
        t = database("integrals")
        self.delta = np.asarray(t["delta"], dtype=np.float64)
        self._omega22 = _Table(t["tstar22"], t["omega22"])
        self._astar = _Table(t["tstar"], t["astar"])

    def _at(self, table, tstar, delta):
        xq = np.log(np.asarray(tstar, dtype=np.float64))
        delta = np.asarray(delta, dtype=np.float64)
        d = self.delta
        j = np.clip(np.searchsorted(d, delta) - 1, 0, len(d) - 3)
        d0, d1, d2 = d[j], d[j + 1], d[j + 2]
        f0, f1, f2 = (table.column(xq, j + k) for k in range(3))
        # Lagrange's quadratic through the three columns
        return (f0 * (delta - d1) * (delta - d2) / ((d0 - d1) * (d0 - d2))
                + f1 * (delta - d0) * (delta - d2) / ((d1 - d0) * (d1 - d2))
                + f2 * (delta - d0) * (delta - d1) / ((d2 - d0) * (d2 - d1)))

    def omega22(self, tstar, delta):
        return self._at(self._omega22, tstar, delta)

    def astar(self, tstar, delta):
        return self._at(self._astar, tstar, delta)
