"""Independent numpy FP64 restatement of the OT potential and energy density.

Used only to cross-check the C oracle on small grids.  It is written
differently on purpose.  Kernels are sampled about index 0 (wrapped
coordinates) instead of at n/2 with a (-1)^(i+j+k) factor.  Transforms
are numpy's.  Derivatives of F are taken by np.roll on the sampled kernel.
Formulas follow src/ot.c:524-991 and src/ot-energy.c:38-288.
"""
import numpy as np

AUTOANG = 0.529177210903
AUTOK = 315775.02480407
AUTOAMU = 1.0 / 1822.888486209
AUTOMPS = 2187691.26364


def params():
    a3 = AUTOANG ** 3
    return dict(
        b=-718.99 / (AUTOK * a3), c2=-2.411857e4 / (AUTOK * a3 ** 2), c3=1.858496e6 / (AUTOK * a3 ** 3),
        rho0=0.0218360 * a3, h=2.1903 / AUTOANG, sigma=2.556 / AUTOANG, eps=10.22 / AUTOK,
        rho_0s=0.04 * a3, alpha_s=54.31 / a3, l_g=1.0 / AUTOANG, mass=4.0026 / AUTOAMU,
        g11=-19.7544, g12=12.5616 * AUTOANG ** 2, a1=1.023 * AUTOANG ** 2,
        g21=-0.2395, g22=0.0312 * AUTOANG ** 2, a2=0.14912 * AUTOANG ** 2,
        div_eps=1e-4, c_bf=1.0)


def _kfun(p, which, x, y, z):
    r2 = x * x + y * y + z * z
    if which == 0:
        with np.errstate(divide="ignore"):
            s = (p["sigma"] ** 2 / np.where(r2 > 0, r2, 1.0)) ** 3
        return np.where(r2 <= p["h"] ** 2, 0.0, 4.0 * p["eps"] * s * (s - 1.0))
    if which == 1:
        return np.where(r2 <= p["h"] ** 2, 3.0 / (4.0 * np.pi * p["h"] ** 3), 0.0)
    if which == 2:
        iw = 1.0 / p["l_g"]
        return (iw / np.sqrt(np.pi)) ** 3 * np.exp(-r2 * iw * iw)
    return ((p["g11"] + p["g12"] * r2) * np.exp(-p["a1"] * r2)
            + (p["g21"] + p["g22"] * r2) * np.exp(-p["a2"] * r2))


def kernel_map(p, which, shape, step, min_sub, max_sub):
    """Adaptive cell average (midpoint sub-grid, ns doubling until |change| < 0.01 K)."""
    ax = [np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n) * step for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    tol = 0.01 / AUTOK
    value = np.zeros(shape)
    done = np.zeros(shape, bool)
    prev = None
    ns = min_sub
    while ns <= max_sub:
        off = ((np.arange(ns) + 0.5) / ns - 0.5) * step
        cur = np.zeros(shape)
        for a in off:
            for b in off:
                for c in off:
                    cur += _kfun(p, which, X + a, Y + b, Z + c)
        cur /= ns ** 3
        if prev is not None:
            conv = (~done) & (np.abs(cur - prev) < tol)
            value[conv] = cur[conv]
            done |= conv
        prev = cur
        ns *= 2
    value[~done] = prev[~done]
    return value


class NumpyOT:
    def __init__(self, shape, step, kc=True, bf=True, min_sub=2, max_sub=8, kc_exact=False):
        self.p = p = params()
        self.shape, self.h = shape, step
        self.kc, self.bf = kc, bf
        h3 = step ** 3
        lj = kernel_map(p, 0, shape, step, min_sub, max_sub)
        self.L = np.fft.fftn(lj)
        self.L *= p["b"] / (h3 * self.L[0, 0, 0].real)
        sa = kernel_map(p, 1, shape, step, min_sub, max_sub)
        self.Pi = np.fft.fftn(sa)
        self.Pi /= h3 * self.Pi[0, 0, 0].real
        if kc:
            f = kernel_map(p, 2, shape, step, min_sub, max_sub)
            self.F = np.fft.fftn(f)
            self.dF = [np.fft.fftn(self.grad(f, d)) for d in range(3)]
        if bf:
            self.U = np.fft.fftn(kernel_map(p, 3, shape, step, min_sub, max_sub))
        self.k23 = (1.0 / p["rho_0s"]) if kc_exact else 1.0

    def grad(self, f, d):
        return (np.roll(f, -1, axis=d) - np.roll(f, 1, axis=d)) / (2.0 * self.h)

    def conv(self, f, K):
        return np.real(np.fft.ifftn(np.fft.fftn(f) * K)) * self.h ** 3

    def velocity(self, psi, mass, eps):
        rho = np.abs(psi) ** 2
        return [np.imag(np.conj(psi) * self.grad(psi, d)) / (mass * (rho + eps)) for d in range(3)]

    def potential(self, psi, wf_mass):
        p = self.p
        rho = np.abs(psi) ** 2
        rb = self.conv(rho, self.Pi)
        V = self.conv(rho, self.L) + p["c2"] / 2 * rb ** 2 + p["c3"] / 3 * rb ** 3
        V = V + self.conv((p["c2"] * rb + p["c3"] * rb ** 2) * rho, self.Pi)
        V = V.astype(np.complex128)
        if self.kc:
            c = p["alpha_s"] / (2.0 * p["mass"])
            rst = 1.0 - self.conv(rho, self.F) / p["rho_0s"]
            for d in range(3):
                g = self.grad(rho, d)
                G = g * rst
                J = self.conv(G, self.F)
                V += c * rst * self.conv(G, self.dF[d])
                V += self.k23 * c * self.conv(g * J, self.F)
                V -= self.k23 * c * J * self.conv(rho, self.dF[d])
        if self.bf:
            v = self.velocity(psi, wf_mass, p["div_eps"])
            v2 = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
            A = self.conv(rho, self.U)
            C = self.conv(rho * v2, self.U)
            B = [self.conv(rho * v[d], self.U) for d in range(3)]
            V += -0.5 * p["mass"] * (v2 * A - 2 * sum(v[d] * B[d] for d in range(3)) + C) * p["c_bf"]
            im = 0.0
            for d in range(3):
                W = v[d] * A - B[d]
                im = im + 0.5 * (self.grad(rho, d) / (rho + p["div_eps"])) * W + 0.5 * self.grad(W, d)
            V += 1j * im * p["c_bf"]
        return V

    def energy_density(self, psi, wf_mass):
        p = self.p
        rho = np.abs(psi) ** 2
        rb = self.conv(rho, self.Pi)
        e = 0.5 * rho * self.conv(rho, self.L) + p["c2"] / 2 * rho * rb ** 2 + p["c3"] / 3 * rho * rb ** 3
        if self.kc:
            rst = 1.0 - self.conv(rho, self.F) / p["rho_0s"]
            for d in range(3):
                G = self.grad(rho, d) * rst
                e = e - p["alpha_s"] / (4.0 * p["mass"]) * G * self.conv(G, self.F)
        if self.bf:
            v = self.velocity(psi, wf_mass, 1e-4)
            v2 = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
            m = p["mass"] * p["c_bf"]
            e = e - m / 4 * rho * v2 * self.conv(rho, self.U)
            e = e + m / 2 * sum(rho * v[d] * self.conv(rho * v[d], self.U) for d in range(3))
            e = e - m / 4 * rho * self.conv(rho * v2, self.U)
        return e
