"""Oracle: the method of moving asymptotes (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates loadpath/mma.py (Svanberg 1987; the 2007 "MMA and GCMMA"
formulation), the published method:
- asymptotes: init x -/+ asyinit (xmax - xmin) for the first two
  iterations, then widened by asyincr where the last two steps kept their
  sign, narrowed by asydecr where it flipped, clamped to
  [x - 10 span, x - 0.01 span] / [x + 0.01 span, x + 10 span]   mma.py:76-87
- move limits alfa / beta (albefa 0.1, move)                     mma.py:88-90
- convex approximation p, q with the 1e-3 / 1e-5 regularization  mma.py:92-107
- primal-dual interior point on the subproblem, epsi 1 -> 1e-7 by 10x,
  Newton on the (m + 1) reduced system, step to the boundary with 1.01
  margins, backtracking while the residual norm exceeds 2x       mma.py:115-214
- non-finite inputs read as large violations, steps clipped to [alfa, beta],
  tighten(factor) pulls the asymptotes toward the last design    mma.py:52-75, 106-112

Pinned by (tests/test_oracle_mma.py):
- the reference's own known answer: Svanberg's two-sphere toy problem
  (tests/test_loadpath.py:376-388 of the reference) converges to
  (2.0175, 1.7800, 1.2375);
- the subproblem solution satisfies its KKT conditions (stationarity of the
  convex approximation's Lagrangian, complementarity) to the final epsi;
- on a separable problem with a closed-form optimum (linear objective, one
  linear constraint) the iterates converge to it.
"""
from __future__ import annotations

import numpy as np

ALBEFA = 0.1
RAA0 = 1e-5
EPSIMIN = 1e-7


class MMA:
    def __init__(self, n, m, xmin=0.0, xmax=1.0, move=0.2, c=1000.0, asyinit=0.5, asyincr=1.2, asydecr=0.7):
        self.n, self.m = int(n), int(m)
        self.xmin = np.broadcast_to(np.asarray(xmin, float), (self.n,)).copy()
        self.xmax = np.broadcast_to(np.asarray(xmax, float), (self.n,)).copy()
        self.move = float(move)
        self.a0, self.a = 1.0, np.zeros(self.m)
        self.c, self.d = np.full(self.m, float(c)), np.ones(self.m)
        self.asyinit, self.asyincr, self.asydecr = asyinit, asyincr, asydecr
        self.xold1 = self.xold2 = self.low = self.upp = None
        self.iter = 0
        self.lam = None

    def tighten(self, factor):
        if self.xold1 is not None and self.low is not None:
            self.low = self.xold1 - factor * (self.xold1 - self.low)
            self.upp = self.xold1 + factor * (self.upp - self.xold1)

    def _asymptotes(self, x):
        span = self.xmax - self.xmin
        if self.iter <= 2 or self.low is None:
            self.low, self.upp = x - self.asyinit * span, x + self.asyinit * span
            return
        trend = (x - self.xold1) * (self.xold1 - self.xold2)
        fac = np.where(trend > 0, self.asyincr, np.where(trend < 0, self.asydecr, 1.0))
        low = x - fac * (self.xold1 - self.low)
        upp = x + fac * (self.upp - self.xold1)
        self.low = np.minimum(np.maximum(low, x - 10 * span), x - 0.01 * span)
        self.upp = np.minimum(np.maximum(upp, x + 0.01 * span), x + 10 * span)

    def approximation(self, x, df0, f, df):
        """(alfa, beta, p0, q0, P, Q, b) of the convex separable model."""
        span = self.xmax - self.xmin
        low, upp = self.low, self.upp
        alfa = np.maximum(np.maximum(low + ALBEFA * (x - low), x - self.move * span), self.xmin)
        beta = np.minimum(np.minimum(upp - ALBEFA * (upp - x), x + self.move * span), self.xmax)
        reg = RAA0 / np.maximum(span, 1e-5)
        u2, l2 = (upp - x) ** 2, (x - low) ** 2

        def split(g):
            gp, gm = np.maximum(g, 0.0), np.maximum(-g, 0.0)
            r = 0.001 * (gp + gm) + reg
            return (gp + r) * u2, (gm + r) * l2

        p0, q0 = split(df0)
        P, Q = split(df)
        b = P @ (1.0 / (upp - x)) + Q @ (1.0 / (x - low)) - f
        return alfa, beta, p0, q0, P, Q, b

    def update(self, x, f0, df0, f, df):
        x = np.clip(np.nan_to_num(np.asarray(x, float), nan=0.5), self.xmin, self.xmax)
        f = np.nan_to_num(np.asarray(f, float).reshape(self.m), nan=1e6, posinf=1e6, neginf=-1e6)
        df = np.nan_to_num(np.asarray(df, float).reshape(self.m, self.n), nan=0.0, posinf=1e6, neginf=-1e6)
        df0 = np.nan_to_num(np.asarray(df0, float), nan=0.0, posinf=1e6, neginf=-1e6)
        self.iter += 1
        self._asymptotes(x)
        alfa, beta, p0, q0, P, Q, b = self.approximation(x, df0, f, df)
        with np.errstate(divide="ignore", invalid="ignore"):
            xn = subsolve(self.low, self.upp, alfa, beta, p0, q0, P, Q, b, self.a0, self.a, self.c, self.d)
        self.lam = xn[1]
        xn = xn[0]
        xn = np.where(np.isfinite(xn), xn, x)
        xn = np.clip(xn, alfa, beta)
        self.xold2 = self.xold1 if self.xold1 is not None else x.copy()
        self.xold1 = x.copy()
        return xn


def _kkt(st, epsi, low, upp, alfa, beta, p0, q0, P, Q, b, a0, a, c, d):
    """The perturbed KKT residual of the subproblem at state st."""
    x, y, z, lam, xsi, eta, mu, zet, s = st
    ux, xl = upp - x, x - low
    plam, qlam = p0 + P.T @ lam, q0 + Q.T @ lam
    g = P @ (1.0 / ux) + Q @ (1.0 / xl)
    return np.concatenate([
        plam / ux ** 2 - qlam / xl ** 2 - xsi + eta,
        c + d * y - mu - lam,
        [a0 - zet - a @ lam],
        g - a * z - y + s - b,
        xsi * (x - alfa) - epsi,
        eta * (beta - x) - epsi,
        mu * y - epsi,
        [zet * z - epsi],
        lam * s - epsi,
    ])


def subsolve(low, upp, alfa, beta, p0, q0, P, Q, b, a0, a, c, d, epsimin=EPSIMIN, return_state=False):
    """Primal-dual interior point on the MMA subproblem (mma.py:115-214).
    Returns (x, lam) (and the whole final state when asked)."""
    m = len(b)
    x = 0.5 * (alfa + beta)
    y, z, lam = np.ones(m), 1.0, np.ones(m)
    xsi = np.maximum(1.0 / (x - alfa), 1.0)
    eta = np.maximum(1.0 / (beta - x), 1.0)
    mu, zet, s = np.maximum(1.0, 0.5 * c), 1.0, np.ones(m)
    args = (low, upp, alfa, beta, p0, q0, P, Q, b, a0, a, c, d)
    epsi = 1.0
    while epsi > epsimin:
        st = (x, y, z, lam, xsi, eta, mu, zet, s)
        r = _kkt(st, epsi, *args)
        norm, rmax = np.linalg.norm(r), np.abs(r).max()
        it = 0
        while rmax > 0.9 * epsi and it < 200:
            it += 1
            ux, xl = upp - x, x - low
            plam, qlam = p0 + P.T @ lam, q0 + Q.T @ lam
            g = P @ (1.0 / ux) + Q @ (1.0 / xl)
            GG = P / ux ** 2 - Q / xl ** 2
            delx = plam / ux ** 2 - qlam / xl ** 2 - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a @ lam - epsi / z
            dellam = g - a * z - y - b + epsi / lam
            diagx = 2.0 * (plam / ux ** 3 + qlam / xl ** 3) + xsi / (x - alfa) + eta / (beta - x)
            diagy = d + mu / y
            rhs_lam = dellam + dely / diagy - GG @ (delx / diagx)
            K = np.zeros((m + 1, m + 1))
            K[:m, :m] = np.diag(s / lam + 1.0 / diagy) + (GG / diagx) @ GG.T
            K[:m, m] = K[m, :m] = a
            K[m, m] = -zet / z
            sol = np.linalg.solve(K, np.append(rhs_lam, delz))
            dlam, dz = sol[:m], sol[m]
            dx = -(delx + GG.T @ dlam) / diagx
            dy = (dlam - dely) / diagy
            dxsi = -xsi + epsi / (x - alfa) - xsi * dx / (x - alfa)
            deta = -eta + epsi / (beta - x) + eta * dx / (beta - x)
            dmu = -mu + epsi / y - mu * dy / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - s * dlam / lam
            pos = np.concatenate([y, [z], lam, xsi, eta, mu, [zet], s])
            dpos = np.concatenate([dy, [dz], dlam, dxsi, deta, dmu, [dzet], ds])
            step = 1.0 / max(np.max(-1.01 * dpos / pos), np.max(-1.01 * dx / (x - alfa)),
                             np.max(1.01 * dx / (beta - x)), 1.0)
            base = (x, y, z, lam, xsi, eta, mu, zet, s)
            dirs = (dx, dy, dz, dlam, dxsi, deta, dmu, dzet, ds)
            for _ in range(50):
                st = tuple(v + step * dv for v, dv in zip(base, dirs))
                r = _kkt(st, epsi, *args)
                if np.linalg.norm(r) < 2.0 * norm:
                    break
                step /= 2.0
            x, y, z, lam, xsi, eta, mu, zet, s = st
            norm, rmax = np.linalg.norm(r), np.abs(r).max()
        epsi *= 0.1
    if return_state:
        return x, lam, (x, y, z, lam, xsi, eta, mu, zet, s)
    return x, lam
