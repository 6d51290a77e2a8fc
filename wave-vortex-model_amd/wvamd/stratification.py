"""Vertical modes and projection operators of an arbitrary stratification N^2(z)
with a rigid lid (the host-side FP64 eigenproblem of the hydrostatic transform).

The reference obtains them from InternalModes (an external MATLAB package,
version 1.3.0, not vendored) through
  @WVStratification/quadraturePointsForStratifiedFlow.m:20-28   (quadrature points)
  @WVStratification/WVStratification.m:200-207                  (modes at omega = 0,
                                                                  geostrophic normalization)
  @WVStratification/verticalProjectionOperatorsWithRigidLid.m:17-72 (operators).
The eigenproblem solved here is the one those modes satisfy,
    G'' = -N^2 / (g h) G,  G(-Lz) = G(0) = 0,  F = h G',
    (1/g) int N^2 G^2 dz = 1,  F(0) > 0,
discretized by Chebyshev collocation in the WKB-stretched coordinate
s = int N dz, where it reads G_ss + (d ln N/ds) G_s = -G/(g h) (the WKB
spectral formulation the reference's InternalModesWKBSpectral is named for),
and the quadrature points are the Nz zeros of mode Nz-1 on [-Lz, 0] (both ends
included), which for constant N are the uniform grid of the DCT-I.  The
operators are then assembled exactly as verticalProjectionOperatorsWithRigidLid
does.
"""
import numpy as np
from numpy.polynomial import chebyshev as C
from scipy import linalg
from scipy.optimize import brentq


class Stratification:
    """N^2(z) on [-Lz, 0] and its vertical modes.

    N2: callable z -> N^2 (vectorized); dN2: optional callable z -> dN^2/dz.
    npoints: collocation points in s (the reference uses
    max(256, floor(2.1 Nz)) for its eigenproblem; more points resolve the
    highest retained modes to round-off).
    """

    def __init__(self, N2, Lz, g=9.81, dN2=None, npoints=513, ncheb=1025):
        self.N2f, self.dN2f, self.Lz, self.g = N2, dN2, float(Lz), float(g)
        # Chebyshev interpolants of N on [-Lz, 0] and its integral s(z)
        x = np.cos(np.pi * np.arange(ncheb) / (ncheb - 1))
        zc = 0.5 * self.Lz * (x - 1.0)
        Nc = np.sqrt(np.asarray(N2(zc), dtype=np.float64))
        if np.any(~np.isfinite(Nc)) or np.any(Nc <= 0):
            raise ValueError("N^2 must be positive and finite on [-Lz, 0]")
        self._Ncheb = C.Chebyshev.fit(x, Nc, ncheb - 1, domain=[-1, 1])
        self._scheb = self._Ncheb.integ(lbnd=-1) * (0.5 * self.Lz)  # ds/dx = N dz/dx
        self.S = float(self._scheb(1.0))
        self.n = int(npoints)
        self._solve()

    # -- coordinates ---------------------------------------------------------
    def _x(self, z):
        return 2.0 * np.asarray(z, dtype=np.float64) / self.Lz + 1.0

    def N(self, z):
        return np.sqrt(np.asarray(self.N2f(np.asarray(z, dtype=np.float64)), dtype=np.float64))

    def dNdz(self, z):
        z = np.asarray(z, dtype=np.float64)
        if self.dN2f is not None:
            return 0.5 * np.asarray(self.dN2f(z), dtype=np.float64) / self.N(z)
        return self._Ncheb.deriv()(self._x(z)) * (2.0 / self.Lz)

    def s(self, z):
        return self._scheb(self._x(z))

    def z_of_s(self, s):
        """Inverse of the stretched coordinate (Newton from a linear guess)."""
        s = np.atleast_1d(np.asarray(s, dtype=np.float64))
        z = -self.Lz + (s / self.S) * self.Lz
        for _ in range(100):
            dz = (self.s(z) - s) / self.N(z)
            z = np.clip(z - dz, -self.Lz, 0.0)
            if np.max(np.abs(dz)) < 1e-14 * self.Lz:
                break
        return z

    # -- eigenproblem ----------------------------------------------------------
    @staticmethod
    def _cheb(n):
        """Gauss-Lobatto points x_i = cos(pi i/(n-1)) and the differentiation matrix."""
        i = np.arange(n)
        x = np.cos(np.pi * i / (n - 1))
        c = np.where((i == 0) | (i == n - 1), 2.0, 1.0) * (-1.0) ** i
        X = x[:, None] - x[None, :]
        D = np.outer(c, 1.0 / c) / (X + np.eye(n))
        D -= np.diag(D.sum(axis=1))
        return x, D

    def _solve(self):
        n = self.n
        x, D = self._cheb(n)
        self._xs = x                                   # s = S (x + 1)/2
        sn = 0.5 * self.S * (x + 1.0)
        zn = self.z_of_s(sn)
        zn[0], zn[-1] = 0.0, -self.Lz
        Ds = D * (2.0 / self.S)
        a = self.dNdz(zn) / self.N(zn) ** 2          # d ln N / ds
        A = -(Ds @ Ds + a[:, None] * Ds)
        lam, V = linalg.eig(A[1:-1, 1:-1])
        lam = lam.real
        order = np.argsort(lam)
        keep = order[lam[order] > 0]
        lam, V = lam[keep], V[:, keep].real
        G = np.zeros((n, V.shape[1]))
        G[1:-1] = V
        # Clenshaw-Curtis weights in s: (1/g) int N^2 G^2 dz = (1/g) int N G^2 ds = 1
        w = self._ccweights(n) * (0.5 * self.S)
        norm = np.sqrt((w * self.N(zn)) @ (G * G) / self.g)
        G = G / norm[None, :]
        Gs = Ds @ G
        self.h = 1.0 / (self.g * lam)                  # equivalent depths, descending
        sign = np.where(Gs[0] < 0, -1.0, 1.0)           # F(0) = h N G_s at s = S (x = 1, node 0) > 0
        self._G, self._Gs = G * sign, Gs * sign
        # barycentric weights of the Gauss-Lobatto points
        self._bw = np.where((np.arange(n) == 0) | (np.arange(n) == n - 1), 0.5, 1.0) * (-1.0) ** np.arange(n)

    @staticmethod
    def _ccweights(n):
        N = n - 1
        theta = np.pi * np.arange(n) / N
        w = np.zeros(n)
        v = np.ones(n - 2)
        if N % 2 == 0:
            w[0] = w[N] = 1.0 / (N * N - 1)
            for k in range(1, N // 2):
                v -= 2.0 * np.cos(2 * k * theta[1:-1]) / (4 * k * k - 1)
            v -= np.cos(N * theta[1:-1]) / (N * N - 1)
        else:
            w[0] = w[N] = 1.0 / (N * N)
            for k in range(1, (N - 1) // 2 + 1):
                v -= 2.0 * np.cos(2 * k * theta[1:-1]) / (4 * k * k - 1)
        w[1:-1] = 2.0 * v / N
        return w

    def _interp(self, vals, z):
        """Barycentric interpolation of nodal values (columns) at points z."""
        return self._interp_x(vals, 2.0 * self.s(z) / self.S - 1.0)

    def _interp_x(self, vals, xs):
        xs = np.atleast_1d(np.asarray(xs, dtype=np.float64))
        X = xs[:, None] - self._xs[None, :]
        exact = np.isclose(X, 0.0, atol=1e-15, rtol=0.0)
        X = np.where(exact, 1.0, X)
        Wm = self._bw[None, :] / X
        out = (Wm @ vals) / Wm.sum(axis=1)[:, None]
        rows, cols = np.nonzero(exact)
        out[rows] = vals[cols]
        return out

    def G(self, z, modes=None):
        """G_j(z) for baroclinic modes j = 1.. (columns), [len(z) x nmodes]."""
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        vals = self._G if modes is None else self._G[:, :modes]
        return self._interp(vals, z)

    def F(self, z, modes=None):
        """F_j = h_j G_j'(z) = h_j N dG_j/ds, [len(z) x nmodes]."""
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        vals = self._Gs if modes is None else self._Gs[:, :modes]
        h = self.h if modes is None else self.h[:modes]
        return self._interp(vals, z) * self.N(z)[:, None] * h[None, :]

    def quadrature_points(self, Nz):
        """The Nz zeros of G_{Nz-1} on [-Lz, 0], ascending (ends included)."""
        n = Nz - 1
        sg = np.linspace(0.0, self.S, 64 * Nz + 1)
        col = self._G[:, n - 1:n]
        vals = self._interp_x(col, 2.0 * sg / self.S - 1.0)[:, 0]
        g = lambda s: float(self._interp_x(col, 2.0 * s / self.S - 1.0)[0, 0])
        roots = [0.0]
        for i in range(1, len(sg) - 2):
            if vals[i] == 0.0:
                roots.append(sg[i])
            elif vals[i] * vals[i + 1] < 0:
                roots.append(brentq(g, sg[i], sg[i + 1], xtol=1e-15 * self.S, rtol=1e-15))
        roots.append(self.S)
        if len(roots) != Nz:
            raise RuntimeError(f"found {len(roots)} quadrature points, expected {Nz}")
        z = self.z_of_s(np.array(roots))
        z[0], z[-1] = -self.Lz, 0.0
        return z

    def dLnN2(self, z):
        z = np.asarray(z, dtype=np.float64)
        if self.dN2f is not None:
            return np.asarray(self.dN2f(z), dtype=np.float64) / np.asarray(self.N2f(z), dtype=np.float64)
        # d ln N^2/dz = 2 d ln N/dz from the Chebyshev interpolant of N
        dN = self._Ncheb.deriv()(self._x(z)) * (2.0 / self.Lz)
        return 2.0 * dN / self._Ncheb(self._x(z))


def vertical_operators(strat: Stratification, Nz: int, Nj: int) -> dict:
    """Operators of verticalProjectionOperatorsWithRigidLid.m:17-72 at the
    quadrature points (MATLAB names; matrices [rows x cols] as numpy arrays)."""
    if not (1 <= Nj <= Nz - 1):
        raise ValueError("Nj must lie in [1, Nz-1]")
    z = strat.quadrature_points(Nz)
    nModes = Nz - 1
    Finv = strat.F(z, nModes)                     # [Nz x nModes]
    Ginv = strat.G(z, nModes)
    h = strat.h[:nModes].copy()
    # barotropic mode added to F, boundaries removed from G (:21-22)
    Finv = np.concatenate([np.ones((Nz, 1)), Finv], axis=1)
    Ginv = Ginv[1:-1, :-1]
    P = np.max(np.abs(Finv), axis=0)
    Q = np.max(np.abs(Ginv), axis=0)
    PFinv = Finv / P[None, :]
    QGinv = Ginv / Q[None, :]
    PF = np.linalg.inv(PFinv)
    QG = np.linalg.inv(QGinv)
    b = np.zeros(Nz)
    b[0] = strat.Lz
    w = np.linalg.solve(PFinv.T, b)               # quadrature weights (:35-37)
    PFinv = PFinv[:, :-1]
    PF = PF[:-1, :]
    QGinv = np.concatenate([np.zeros((Nz, 1)),
                            np.concatenate([np.zeros((1, nModes - 1)), QGinv, np.zeros((1, nModes - 1))], axis=0)],
                           axis=1)
    QG = np.concatenate([np.zeros((nModes, 1)), np.concatenate([np.zeros((1, Nz - 2)), QG], axis=0),
                         np.zeros((nModes, 1))], axis=1)
    h = np.concatenate([[1.0], h[:-1]])
    P = P[:-1]
    Q = np.concatenate([[1.0], Q])
    return {
        "z": z, "N2": np.asarray(strat.N2f(z), dtype=np.float64), "dLnN2": strat.dLnN2(z),
        "h": h[:Nj].copy(), "P": P[:Nj].copy(), "Q": Q[:Nj].copy(),
        "PFinv": np.ascontiguousarray(PFinv[:, :Nj]), "PF": np.ascontiguousarray(PF[:Nj, :]),
        "QGinv": np.ascontiguousarray(QGinv[:, :Nj]), "QG": np.ascontiguousarray(QG[:Nj, :]),
        "w": w,
    }


def exponential(N0=5.2e-3, b=1300.0):
    """N^2(z) = N0^2 exp(2 z / b) and its derivative."""
    return (lambda z: N0 * N0 * np.exp(2.0 * np.asarray(z) / b),
            lambda z: (2.0 / b) * N0 * N0 * np.exp(2.0 * np.asarray(z) / b))


def constant(N0=5.2e-3):
    return (lambda z: np.full(np.shape(z), N0 * N0, dtype=np.float64),
            lambda z: np.zeros(np.shape(z), dtype=np.float64))
