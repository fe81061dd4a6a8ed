"""Exact solution of the 1-D relativistic Riemann problem (ideal gas, no
tangential velocity), after Marti & Mueller, "Numerical hydrodynamics in
special relativity", Living Rev. Relativity 6 (2003), sect. 3.1 / appendix.

Test helper: gives the closed-form profile the first-order RHLLC scheme must
converge to.  Handles a left-going rarefaction or shock and a right-going
rarefaction or shock.
"""
from __future__ import annotations

import math

import numpy as np


def _enthalpy(g, rho, p):
    return 1.0 + g / (g - 1.0) * p / rho


def _cs(g, rho, p):
    h = _enthalpy(g, rho, p)
    return math.sqrt(g * p / (rho * h))


def _rarefaction_v(g, rho_a, p_a, v_a, p, sign):
    """Flow velocity behind a rarefaction connecting state a to pressure p.
    sign = +1 for a wave travelling to the left, -1 to the right."""
    s = math.sqrt(g - 1.0)
    rho = rho_a * (p / p_a) ** (1.0 / g)
    c = _cs(g, rho, p)
    ca = _cs(g, rho_a, p_a)
    A = ((s - c) / (s + c) * (s + ca) / (s - ca)) ** (sign * 2.0 / s)
    return ((1 + v_a) * A - (1 - v_a)) / ((1 + v_a) * A + (1 - v_a)), rho


def _shock(g, rho_a, p_a, v_a, p, sign):
    """Post-shock (v, rho) and shock speed for pressure p > p_a; sign +1 right-going."""
    h_a = _enthalpy(g, rho_a, p_a)
    # Taub adiabat with h = 1 + g/(g-1) p/rho, quadratic in h_b
    k = (g - 1.0) * (p_a - p) / (g * p)
    A = 1.0 + k
    B = -k
    C = h_a * (p_a - p) / rho_a - h_a * h_a
    h_b = (-B + math.sqrt(B * B - 4 * A * C)) / (2 * A)
    rho_b = g * p / ((g - 1.0) * (h_b - 1.0))
    j2 = (p - p_a) / (h_a / rho_a - h_b / rho_b)
    j = sign * math.sqrt(j2)
    W_a = 1.0 / math.sqrt(1.0 - v_a * v_a)
    r2w2 = rho_a * rho_a * W_a * W_a
    Vs = (r2w2 * v_a + j * math.sqrt(j2 + r2w2 * (1.0 - v_a * v_a))) / (r2w2 + j2)
    Ws = 1.0 / math.sqrt(1.0 - Vs * Vs)
    v_b = (h_a * W_a * v_a + Ws * (p - p_a) / j) / (h_a * W_a + (p - p_a) * (Ws * v_a / j + 1.0 / (rho_a * W_a)))
    return v_b, rho_b, Vs


class RiemannExact:
    def __init__(self, g, left, right):
        """left/right = (rho, v, p)."""
        self.g = g
        self.L, self.R = left, right
        self.p_star = self._solve()
        self.v_star, self.rhoL_star, self.wL = self._side(self.L, self.p_star, +1)
        vR, self.rhoR_star, self.wR = self._side(self.R, self.p_star, -1)

    def _side(self, st, p, wave_left):
        rho, v, pa = st
        if p < pa:  # rarefaction
            vb, rb = _rarefaction_v(self.g, rho, pa, v, p, wave_left)
            return vb, rb, ("rare", None)
        vb, rb, Vs = _shock(self.g, rho, pa, v, p, -wave_left)
        return vb, rb, ("shock", Vs)

    def _solve(self):
        f = lambda p: self._side(self.L, p, +1)[0] - self._side(self.R, p, -1)[0]
        lo, hi = min(self.L[2], self.R[2]) * 1e-6, max(self.L[2], self.R[2]) * 10
        flo, fhi = f(lo), f(hi)
        assert flo * fhi < 0
        for _ in range(200):
            mid = math.sqrt(lo * hi)
            fm = f(mid)
            if fm * flo > 0:
                lo, flo = mid, fm
            else:
                hi = mid
        return math.sqrt(lo * hi)

    def sample(self, xi):
        """(rho, v, p) at similarity coordinate xi = (x - x0)/t."""
        g = self.g
        out = np.zeros((len(xi), 3))
        for i, x in enumerate(np.atleast_1d(xi)):
            if x < self.v_star:  # left of the contact
                rho, v, p = self.L
                kind, Vs = self.wL
                if kind == "shock":
                    out[i] = (rho, v, p) if x < Vs else (self.rhoL_star, self.v_star, self.p_star)
                    continue
                ca = _cs(g, rho, p)
                head = (v - ca) / (1 - v * ca)
                cs_star = _cs(g, self.rhoL_star, self.p_star)
                tail = (self.v_star - cs_star) / (1 - self.v_star * cs_star)
                if x < head:
                    out[i] = (rho, v, p)
                elif x > tail:
                    out[i] = (self.rhoL_star, self.v_star, self.p_star)
                else:  # inside the fan: (v - c)/(1 - v c) = x
                    lo, hi = self.p_star, p
                    for _ in range(100):
                        pm = 0.5 * (lo + hi)
                        vm, rm = _rarefaction_v(g, rho, p, v, pm, +1)
                        cm = _cs(g, rm, pm)
                        if (vm - cm) / (1 - vm * cm) > x:
                            lo = pm
                        else:
                            hi = pm
                    vm, rm = _rarefaction_v(g, rho, p, v, 0.5 * (lo + hi), +1)
                    out[i] = (rm, vm, 0.5 * (lo + hi))
            else:
                rho, v, p = self.R
                kind, Vs = self.wR
                if kind == "shock":
                    out[i] = (self.rhoR_star, self.v_star, self.p_star) if x < Vs else (rho, v, p)
                    continue
                ca = _cs(g, rho, p)
                head = (v + ca) / (1 + v * ca)
                cs_star = _cs(g, self.rhoR_star, self.p_star)
                tail = (self.v_star + cs_star) / (1 + self.v_star * cs_star)
                if x > head:
                    out[i] = (rho, v, p)
                elif x < tail:
                    out[i] = (self.rhoR_star, self.v_star, self.p_star)
                else:
                    lo, hi = self.p_star, p
                    for _ in range(100):
                        pm = 0.5 * (lo + hi)
                        vm, rm = _rarefaction_v(g, rho, p, v, pm, -1)
                        cm = _cs(g, rm, pm)
                        if (vm + cm) / (1 + vm * cm) < x:
                            lo = pm
                        else:
                            hi = pm
                    vm, rm = _rarefaction_v(g, rho, p, v, 0.5 * (lo + hi), -1)
                    out[i] = (rm, vm, 0.5 * (lo + hi))
        return out
