"""FP64 CPU restatement (numpy) of the hydrostatic wave-vortex transform of an
arbitrary stratification and its nonlinear flux.

TEST INFRASTRUCTURE ONLY (tests/ imports it; nothing in the library does).

It follows the reference's MATLAB implementation, which this container cannot
run, statement by statement:
  spatial fields     @WVGeometryDoublyPeriodicStratified/WVGeometryDoublyPeriodicStratified.m:207-223
                     (u = PF0inv*(P0.*(Apm+A0)), w = QG0inv*(Q0.*...)) with the
                     wave/geostrophic/inertial/MDA field coefficients of
                     FlowComponents/WVInternalGravityWaveComponent.m:284-301,
                     WVGeostrophicComponent, WVInertialOscillationComponent,
                     WVMeanDensityAnomalyComponent (h_pm = h_0 for hydrostatic modes)
  z derivatives      @WVGeometryDoublyPeriodicStratified/diffZF.m:30-35 and diffZG.m:26-28,
                     the dense operators applied to the spatial fields
  products           Forcing/WVNonlinearAdvection.m:103-107 (addHydrostaticSpatialForcing,
                     with eta * dLnN2 in the eta equation)
  projection         @WVTransformHydrostatic/transformUVEtaToWaveVortex.m:22-43 and
                     transformFromSpatialDomainWithFg/Gg (...Stratified.m:242-252)
  flux               @WVTransformHydrostatic/WVTransformHydrostatic.m:381-393
  RK4                the classic stage sequence of WVArrayIntegrator.m:84-111 /
                     PortableRuntime/src/WVRungeKutta.cpp:373-441
The retained horizontal modes (and their order) are the kernel descriptor's,
taken from the C++ oracle.  Parity with MATLAB itself is unpinned (MATLAB and
InternalModes are absent); the pins are the constant-N reduction to the
reference's constant-stratification kernel and the invariants in the tests.
"""
import numpy as np


class HydrostaticOracle:
    def __init__(self, cfg: dict, ops: dict, modes: dict):
        self.cfg = cfg
        self.ops = ops
        self.Nx, self.Ny, self.Nz, self.Nj = (int(cfg[k]) for k in ("Nx", "Ny", "Nz", "Nj"))
        self.g = float(cfg["g"])
        self.f = 2.0 * cfg["rotationRate"] * np.sin(np.deg2rad(cfg["latitude"]))
        self.md = modes
        self.Nkl = len(modes["k"])
        z = ops["z"]
        P, Q, h = ops["P"], ops["Q"], ops["h"]
        # synthesis (...Stratified.m:207-223)
        self.F = ops["PFinv"] * P[None, :]          # [Nz x Nj]
        self.G = ops["QGinv"] * Q[None, :]
        # dense z-derivative operators (diffZF.m:30-35, diffZG.m:26-28)
        self.DzF = -(ops["N2"] / self.g)[:, None] * (ops["QGinv"] @ ((Q / P)[:, None] * ops["PF"]))
        self.DzG = ops["PFinv"] @ ((P / (Q * h))[:, None] * ops["QG"])
        self.z = z
        self._tables()
        # full-grid indices of each mode and of its conjugate partner
        km, lm = modes["kMode"].astype(np.int64), modes["lMode"].astype(np.int64)
        self.ix, self.iy = km % self.Nx, lm % self.Ny
        self.jx, self.jy = (-km) % self.Nx, (-lm) % self.Ny
        self.selfc = (km == 0) & (lm == 0)
        kk = np.fft.fftfreq(self.Nx, 1.0 / self.Nx)
        ll = np.fft.fftfreq(self.Ny, 1.0 / self.Ny)
        self.kx = 2.0 * np.pi * kk / cfg["Lx"]
        self.ly = 2.0 * np.pi * ll / cfg["Ly"]
        if self.Nx % 2 == 0:
            self.kx[self.Nx // 2] = 0.0
        if self.Ny % 2 == 0:
            self.ly[self.Ny // 2] = 0.0

    # -- coefficient tables [Nkl x Nj] ------------------------------------------
    def _tables(self):
        md, g, f, Nj = self.md, self.g, self.f, self.Nj
        Kh = md["Kh"][:, None]
        k, l = md["k"][:, None], md["l"][:, None]
        ca, sa = md["cosAlpha"][:, None], md["sinAlpha"][:, None]
        j = np.arange(Nj)[None, :]
        h = np.where(j > 0, self.ops["h"][None, :], 1.0) * np.ones_like(Kh)
        om = np.sqrt(g * h * Kh ** 2 + f * f)
        wave = (Kh > 0) & (j > 0)
        inertial = (Kh == 0) & np.ones_like(j, dtype=bool)
        geo = (Kh > 0) & np.ones_like(j, dtype=bool)
        mda = (Kh == 0) & (j > 0)
        z0 = np.zeros((self.Nkl, Nj))
        with np.errstate(divide="ignore", invalid="ignore"):
            self.omega = om
            self.UAp = np.where(wave, ca - 1j * (f / om) * sa, np.where(inertial, 1.0 + 0j, 0j))
            self.VAp = np.where(wave, sa + 1j * (f / om) * ca, np.where(inertial, 1j, 0j))
            self.WAp = np.where(wave, -1j * Kh * h, 0j)
            self.NAp = np.where(wave, -Kh * h / om, 0.0)
            self.ApmD = np.where(wave, -1j / (2.0 * Kh * h), 0j)
            self.ApmN = np.where(wave, -om / (2.0 * Kh * h), 0.0)
            lr2 = np.where(j > 0, f * f / (g * h), 0.0)
            den = Kh ** 2 + lr2
            self.UA0 = np.where(geo, 1j * l / den, 0j)
            self.VA0 = np.where(geo, -1j * k / den, 0j)
            self.NA0 = np.where(geo, np.where(j > 0, -(f / g) / den, 0.0), np.where(mda, 1.0, 0.0))
            self.A0Z = np.where(geo, 1.0, np.where(mda, f * f / (2.0 * h), 0.0))
            self.A0N = np.where(geo, np.where(j > 0, -f / h, 0.0), np.where(mda, 1.0, 0.0))
        self.h0 = np.where(j > 0, self.ops["h"][None, :], 1.0)[0]
        # total-energy factors (as the constant-N kernel's, with h_pm = h_0)
        hpm = np.where(j > 0, h, 0.0)
        self.ApmTE = np.where(wave, 2.0 * hpm, np.where(inertial, np.where(j == 0, self.cfg["Lz"], hpm), 0.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            h0e = np.where(j == 0, self.cfg["Lz"], h)
            self.A0TE = np.where(geo, h0e / den, np.where(mda, g / 2.0, 0.0))
        del z0

    # -- transforms ---------------------------------------------------------------
    def _spectra(self, Ap, Am, A0, dt):
        p = np.exp(1j * self.omega * dt)
        ap, am = Ap * p, Am * np.conj(p)
        u = self.UAp * ap + np.conj(self.UAp) * am + self.UA0 * A0
        v = self.VAp * ap + np.conj(self.VAp) * am + self.VA0 * A0
        w = self.WAp * (ap + am)
        n = self.NAp * (ap - am) + self.NA0 * A0
        return u, v, w, n

    def _to_grid(self, spec, M):
        """sum_j M[z, j] spec[mode, j] on the full Hermitian spectrum, [Nz x Ny x Nx]."""
        vals = spec @ M.T                                   # [Nkl x Nz]
        out = np.zeros((self.Nz, self.Ny, self.Nx), np.complex128)
        vals_self = np.where(self.selfc[:, None], vals.real + 0j, vals)
        out[:, self.iy, self.ix] = vals_self.T
        nz = ~self.selfc
        out[:, self.jy[nz], self.jx[nz]] = np.conj(vals[nz]).T
        return out

    def _real(self, S):
        return np.real(np.fft.ifft2(S, axes=(1, 2))) * (self.Nx * self.Ny)

    def _dx(self, S):
        return self._real(1j * self.kx[None, None, :] * S)

    def _dy(self, S):
        return self._real(1j * self.ly[None, :, None] * S)

    def fields(self, Ap, Am, A0, t=0.0, t0=0.0):
        """(u, v, w, eta) [4 x Nz x Ny x Nx] (...Stratified.m:207-223)."""
        u, v, w, n = self._spectra(Ap, Am, A0, t - t0)
        return np.stack([self._real(self._to_grid(u, self.F)), self._real(self._to_grid(v, self.F)),
                         self._real(self._to_grid(w, self.G)), self._real(self._to_grid(n, self.G))])

    def forward(self, U, V, N, t=0.0, t0=0.0):
        """transformUVEtaToWaveVortex.m:22-43 for spatial (u, v, eta)."""
        sel = (slice(None), self.iy, self.ix)
        uh = (np.fft.fft2(U, axes=(1, 2)) / (self.Nx * self.Ny))[sel]   # [Nz x Nkl]
        vh = (np.fft.fft2(V, axes=(1, 2)) / (self.Nx * self.Ny))[sel]
        nh = (np.fft.fft2(N, axes=(1, 2)) / (self.Nx * self.Ny))[sel]
        P, Q = self.ops["P"][:, None], self.ops["Q"][:, None]
        ub = ((self.ops["PF"] @ uh) / P).T                              # [Nkl x Nj]
        vb = ((self.ops["PF"] @ vh) / P).T
        nb = ((self.ops["QG"] @ nh) / Q).T
        k, l = self.md["k"][:, None], self.md["l"][:, None]
        zeta = 1j * k * vb - 1j * l * ub
        A0 = self.A0Z * zeta + self.A0N * nb
        delta = self.h0[None, :] * (1j * k * ub + 1j * l * vb)
        nw = nb - self.NA0 * A0
        Ap = self.ApmD * delta + self.ApmN * nw
        Am = self.ApmD * delta - self.ApmN * nw
        io = self.md["Kh"] == 0
        Ap[io] = (ub[io] - 1j * vb[io]) / 2.0
        Am[io] = np.conj(Ap[io])
        p = np.exp(1j * self.omega * (t - t0))
        return Ap * np.conj(p), Am * p, A0

    def spatial_products(self, Ap, Am, A0, t=0.0, t0=0.0):
        """The advection products (Fu, Fv, Feta) [Nz x Ny x Nx] before the
        projection (WVNonlinearAdvection.m:103-107)."""
        u, v, w, n = self._spectra(Ap, Am, A0, t - t0)
        Su, Sv = self._to_grid(u, self.F), self._to_grid(v, self.F)
        Sw, Sn = self._to_grid(w, self.G), self._to_grid(n, self.G)
        U, V, W, Nn = self._real(Su), self._real(Sv), self._real(Sw), self._real(Sn)
        dzF = lambda X: np.tensordot(self.DzF, X, axes=(1, 0))
        dzG = lambda X: np.tensordot(self.DzG, X, axes=(1, 0))
        dl = self.ops["dLnN2"][:, None, None]
        Fu = -(U * self._dx(Su) + V * self._dy(Su) + W * dzF(U))
        Fv = -(U * self._dx(Sv) + V * self._dy(Sv) + W * dzF(V))
        Fn = -(U * self._dx(Sn) + V * self._dy(Sn) + W * (dzG(Nn) + Nn * dl))
        return Fu, Fv, Fn

    def nonlinear_flux(self, Ap, Am, A0, t=0.0, t0=0.0):
        return self.forward(*self.spatial_products(Ap, Am, A0, t, t0), t, t0)

    def rk4(self, Ap, Am, A0, t, dt, nsteps, t0=0.0):
        y = [np.array(a, np.complex128) for a in (Ap, Am, A0)]
        for _ in range(nsteps):
            k1 = self.nonlinear_flux(*y, t, t0)
            w = [a.copy() for a in k1]
            st = [yy + (0.5 * dt) * kk for yy, kk in zip(y, k1)]
            k2 = self.nonlinear_flux(*st, t + 0.5 * dt, t0)
            w = [ww + 2.0 * kk for ww, kk in zip(w, k2)]
            st = [yy + (0.5 * dt) * kk for yy, kk in zip(y, k2)]
            k3 = self.nonlinear_flux(*st, t + 0.5 * dt, t0)
            w = [ww + 2.0 * kk for ww, kk in zip(w, k3)]
            st = [yy + dt * kk for yy, kk in zip(y, k3)]
            k4 = self.nonlinear_flux(*st, t + dt, t0)
            w = [ww + kk for ww, kk in zip(w, k4)]
            y = [yy + (dt / 6.0) * ww for yy, ww in zip(y, w)]
            t = t + dt
        return y[0], y[1], y[2], t

    def energy(self, Ap, Am, A0):
        return float(np.sum(self.ApmTE * (np.abs(Ap) ** 2 + np.abs(Am) ** 2) + self.A0TE * np.abs(A0) ** 2))
