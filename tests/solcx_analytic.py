"""Analytic solution of the SolCx problem, derived here from the Stokes
equations (the reference's own test, TESTS/SolCxBenchmarkTests.cpp, uses the
same problem; this module does not use its code).

Problem: unit square, viscosity eta_a for x < xc and eta_b for x > xc,
incompressible Stokes  div(sigma) + f = 0,  sigma = -p I + eta (grad u +
grad u^T),  with the body force f = (0, F0 sin(pi z) cos(pi x))  (MDOODZ's
roger_z = gz * rho with gz = 1 and rho = sin(pi z) cos(pi x)), and free slip
on every wall.

With a stream function psi = X(x) sin(pi z), u = psi_z, w = -psi_x, the
z-walls are free slip for any X, incompressibility holds identically, and
the curl of the momentum balance gives in each constant-viscosity strip

    eta (X'''' - 2 pi^2 X'' + pi^4 X) = -pi F0 sin(pi x),

whose solutions are X = A sin(pi x) + (a + b t) cosh(pi t) + (c + d t) sinh(pi t)
with A = -F0 / (4 pi^3 eta) and t = x (left strip) or x - 1 (right strip).
The x-momentum balance gives the pressure
    p = cos(pi z) [eta (X''' - pi^2 X') - F0 cos(pi x)] / pi  (+ a constant).
Eight conditions fix a, b, c, d in both strips: free slip at x = 0 and 1
(u = 0: X = 0; shear stress eta (u_z + w_x) = -eta (pi^2 X + X'') sin(pi z)
= 0: X'' = 0), and at x = xc continuity of X, X' (velocities), of
eta (pi^2 X + X'') (shear traction) and of eta (3 pi^2 X' - X''') (normal
traction, where cos(pi xc) = 0 for xc = 1/2 removes the forcing term)."""
import numpy as np

PI = np.pi


def _basis(t, k):
    """Values and first three derivatives (rows) of cosh, sinh, t cosh, t sinh at t."""
    ch, sh = np.cosh(k * t), np.sinh(k * t)
    return np.array([
        [ch, sh, t * ch, t * sh],
        [k * sh, k * ch, ch + k * t * sh, sh + k * t * ch],
        [k * k * ch, k * k * sh, 2 * k * sh + k * k * t * ch, 2 * k * ch + k * k * t * sh],
        [k ** 3 * sh, k ** 3 * ch, 3 * k * k * ch + k ** 3 * t * sh, 3 * k * k * sh + k ** 3 * t * ch],
    ])


def _particular(x, eta, F0):
    A = -F0 / (4.0 * PI ** 3 * eta)
    s, c = np.sin(PI * x), np.cos(PI * x)
    return np.array([A * s, A * PI * c, -A * PI ** 2 * s, -A * PI ** 3 * c])


class SolCx:
    def __init__(self, eta_a=1.0, eta_b=1e6, xc=0.5, F0=1.0):
        if abs(np.cos(PI * xc)) > 1e-15:
            raise ValueError("the traction condition used needs cos(pi xc) = 0 (xc = 1/2)")
        self.eta_a, self.eta_b, self.xc, self.F0 = eta_a, eta_b, xc, F0
        M = np.zeros((8, 8))
        r = np.zeros(8)
        # free slip at x = 0 (left strip, t = x) and x = 1 (right strip, t = x - 1)
        for row, (t, deriv, side) in enumerate(((0.0, 0, 0), (0.0, 2, 0), (0.0, 0, 1), (0.0, 2, 1))):
            x = t if side == 0 else 1.0 + t
            M[row, 4 * side:4 * side + 4] = _basis(t, PI)[deriv]
            r[row] = -_particular(x, eta_a if side == 0 else eta_b, F0)[deriv]
        # interface conditions: sum over strips of sign * weight * (combination of derivatives)
        BL, BR = _basis(xc, PI), _basis(xc - 1.0, PI)
        PL, PR = _particular(xc, eta_a, F0), _particular(xc, eta_b, F0)
        combos = [
            (lambda D: D[0], 1.0, 1.0),                            # X
            (lambda D: D[1], 1.0, 1.0),                            # X'
            (lambda D: PI ** 2 * D[0] + D[2], eta_a, eta_b),       # shear traction
            (lambda D: 3 * PI ** 2 * D[1] - D[3], eta_a, eta_b),   # normal traction
        ]
        for q, (comb, wl, wr) in enumerate(combos):
            row = 4 + q
            M[row, 0:4] = wl * comb(BL)
            M[row, 4:8] = -wr * comb(BR)
            r[row] = -(wl * comb(PL) - wr * comb(PR))
        self.coef = np.linalg.solve(M, r)

    def X(self, x):
        """X and its first three derivatives at x (array of shape (4, n))."""
        x = np.asarray(x, np.float64)
        out = np.zeros((4,) + x.shape)
        left = x < self.xc
        for mask, side, eta in ((left, 0, self.eta_a), (~left, 1, self.eta_b)):
            if not mask.any():
                continue
            t = x[mask] if side == 0 else x[mask] - 1.0
            B = _basis(t, PI)
            out[:, mask] = np.tensordot(self.coef[4 * side:4 * side + 4], B, axes=([0], [1])) + \
                _particular(x[mask], eta, self.F0)
        return out

    def eta(self, x):
        return np.where(np.asarray(x) < self.xc, self.eta_a, self.eta_b)

    def velocity(self, x, z):
        X = self.X(x)
        return PI * np.cos(PI * z) * X[0], -np.sin(PI * z) * X[1]

    def pressure(self, x, z):
        X = self.X(x)
        eta = self.eta(x)
        return np.cos(PI * z) * (eta * (X[3] - PI ** 2 * X[1]) - self.F0 * np.cos(PI * x)) / PI
