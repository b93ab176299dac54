"""The exact solution of the Riemann problem for a calorically perfect gas (Toro, Riemann Solvers
and Numerical Methods for Fluid Dynamics, 3rd ed., chapter 4): the star pressure by Newton's
method on the two waves' pressure functions, the star velocity from either side, and the
self-similar solution sampled at x/t. Toro's six tests (his Table 4.1, with test 0 Sod's) as the
reference's shock tubes run them (tests/riemann.py:7-14: rhoL, uL, pL, rhoR, uR, pR, x0, t and a
step)."""

import numpy as np

TESTS = {
    0: (1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 0.5, 0.2, 1e-4),
    1: (1.0, 0.75, 1.0, 0.125, 0.0, 0.1, 0.3, 0.2, 1e-4),
    2: (1.0, -2.0, 0.4, 1.0, 2.0, 0.4, 0.5, 0.15, 1e-4),
    3: (1.0, 0.0, 1000.0, 1.0, 0.0, 0.01, 0.5, 0.012, 5e-6),
    4: (5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950, 0.4, 0.035, 1e-5),
    5: (1.0, -19.5975, 1000.0, 1.0, -19.5975, 0.01, 0.8, 0.012, 1e-5),
}


class Riemann:
    def __init__(self, rhoL, uL, pL, rhoR, uR, pR, x0=0.5, t=0.2, dt=1e-4, gamma=1.4, R=287.0):
        self.rhoL, self.uL, self.pL, self.rhoR, self.uR, self.pR = rhoL, uL, pL, rhoR, uR, pR
        self.x0, self.t, self.dt, self.gamma, self.R = x0, t, dt, gamma, R
        self.cL, self.cR = np.sqrt(gamma * pL / rhoL), np.sqrt(gamma * pR / rhoR)
        self.TL, self.TR = pL / (rhoL * R), pR / (rhoR * R)

    @classmethod
    def toro(cls, test, gamma=1.4, R=287.0):
        return cls(*TESTS[test], gamma=gamma, R=R)

    def _wave(self, p, rho, pK, c):
        """The pressure function of one side and its derivative: a shock above pK, a rarefaction below."""
        g = self.gamma
        if p > pK:
            A, B = 2.0 / ((g + 1.0) * rho), (g - 1.0) / (g + 1.0) * pK
            s = np.sqrt(A / (p + B))
            return (p - pK) * s, s * (1.0 - 0.5 * (p - pK) / (p + B))
        r = (p / pK) ** ((g - 1.0) / (2.0 * g))
        return 2.0 * c / (g - 1.0) * (r - 1.0), r / (rho * c) * (pK / p)
    def star(self):
        """p* and u* (Toro 4.5): Newton on f_L(p) + f_R(p) + (u_R - u_L) = 0 from the linearised guess."""
        du = self.uR - self.uL
        p = max(1e-12, 0.5 * (self.pL + self.pR) - 0.125 * du * (self.rhoL + self.rhoR) * (self.cL + self.cR))
        for _ in range(100):
            fL, dL = self._wave(p, self.rhoL, self.pL, self.cL)
            fR, dR = self._wave(p, self.rhoR, self.pR, self.cR)
            pn = max(1e-12, p - (fL + fR + du) / (dL + dR))
            done = 2.0 * abs(pn - p) / (pn + p) < 1e-15
            p = pn
            if done:
                break
        fL, _ = self._wave(p, self.rhoL, self.pL, self.cL)
        fR, _ = self._wave(p, self.rhoR, self.pR, self.cR)
        return p, 0.5 * (self.uL + self.uR) + 0.5 * (fR - fL)

    def sample(self, pM, uM, S):
        """(p, u, rho) at x/t = S (Toro 4.7)."""
        g = self.gamma
        g1, g2 = (g - 1.0) / (2.0 * g), (g + 1.0) / (2.0 * g)
        if S <= uM:
            rho, u, p, c, sgn = self.rhoL, self.uL, self.pL, self.cL, 1.0
        else:
            rho, u, p, c, sgn = self.rhoR, self.uR, self.pR, self.cR, -1.0
        # mirror the right side onto the left: x/t and the velocities change sign
        S, u, uM = sgn * S, sgn * u, sgn * uM
        if pM > p:
            shock = u - c * np.sqrt(g2 * pM / p + g1)
            if S <= shock:
                out = (p, u, rho)
            else:
                ratio = pM / p
                out = (pM, uM, rho * (ratio + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * ratio + 1.0))
        else:
            head, cM = u - c, c * (pM / p) ** g1
            tail = uM - cM
            if S <= head:
                out = (p, u, rho)
            elif S >= tail:
                out = (pM, uM, rho * (pM / p) ** (1.0 / g))
            else:
                k = 2.0 / (g + 1.0) + (g - 1.0) / ((g + 1.0) * c) * (u - S)
                out = (p * k ** (2.0 * g / (g - 1.0)), 2.0 / (g + 1.0) * (c + 0.5 * (g - 1.0) * u + S), rho * k ** (2.0 / (g - 1.0)))
        return out[0], sgn * out[1], out[2]

    def solve(self, x):
        pM, uM = self.star()
        S = (np.asarray(x) - self.x0) / self.t
        cols = np.array([self.sample(pM, uM, s) for s in S])
        return dict(p=cols[:, 0], u=cols[:, 1], rho=cols[:, 2])

    def heads(self):
        """x/t of the two outermost waves: a shock's speed, or a rarefaction's head."""
        g = self.gamma
        pM, _ = self.star()
        g1, g2 = (g - 1.0) / (2.0 * g), (g + 1.0) / (2.0 * g)
        sL = self.uL - self.cL * (np.sqrt(g2 * pM / self.pL + g1) if pM > self.pL else 1.0)
        sR = self.uR + self.cR * (np.sqrt(g2 * pM / self.pR + g1) if pM > self.pR else 1.0)
        return sL, sR
