"""GPU checks of the libgrid-compatible grid calls the reference drivers use:
value_at_index outside the grid (periodic boundary), the momentum frame of the
propagator and the kinetic energy (cgrid_set_momentum), grid_wf_energy with a
potential, and the driver helpers (Laplacian, conjugate products, |.|^p, curl,
power, product, division with eps, max and min, flux components, region and
weighted integrals), each against numpy on the same data.
"""
import ctypes as C

import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu

PRECS = [L.GRID_FP64, L.GRID_FP32]
PNAME = {L.GRID_FP64: "fp64", L.GRID_FP32: "fp32"}
TOL = {L.GRID_FP64: 1e-12, L.GRID_FP32: 2e-6}
SHAPE, STEP = (8, 16, 32), 0.7


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _data(seed=3, shape=SHAPE, prec=L.GRID_FP64):
    u = inputs.splitmix64_uniform(2 * int(np.prod(shape)), seed).reshape((2,) + shape)
    if prec == L.GRID_FP32:
        u = u.astype(np.float32).astype(np.float64)
    return u[0], u[1]


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_value_at_index_outside_is_periodic(prec):
    a, b = _data(prec=prec)
    r, c = L.RField(SHAPE, STEP, prec), L.CField(SHAPE, STEP, prec)
    r.set(a), c.set(a + 1j * b)
    lib = L.lib()
    nx, ny, nz = SHAPE
    for (i, j, k) in [(-1, 0, 0), (nx, 3, 5), (2, -1, nz), (-nx - 1, 2 * ny + 1, -3), (1, 2, 3)]:
        want = a[i % nx, j % ny, k % nz]
        assert lib.rgrid_value_at_index(r.h, i, j, k) == want
        z = lib.cgrid_value_at_index(c.h, i, j, k)
        assert (z.re, z.im) == (want, b[i % nx, j % ny, k % nz])
    r.free(), c.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("shape", [SHAPE, (32, 32, 16)], ids=["8x16x32", "32x32x16"])
def test_moving_frame(prec, shape):
    """With k0 = the plane wave's own k, the kinetic symbol (k - k0)^2 / 2M vanishes:
    free propagation leaves the wave unchanged and its kinetic energy is zero.
    With k0 != k the phase is exp(-i (k - k0)^2 dt / 2M).  (32 x 32 x 16 takes the
    small-grid block shapes.)"""
    SHAPE = shape
    x, y, z = np.meshgrid(*[np.arange(n) * STEP for n in SHAPE], indexing="ij")
    kv = np.array([2 * np.pi * 2 / (SHAPE[0] * STEP), 2 * np.pi * -3 / (SHAPE[1] * STEP), 2 * np.pi * 5 / (SHAPE[2] * STEP)])
    psi = np.exp(1j * (kv[0] * x + kv[1] * y + kv[2] * z))
    lib = L.lib()
    w = L.Wave(SHAPE, STEP, prec)
    pot = L.CField(SHAPE, STEP, prec)
    dt = 40.0 - 3.0j
    for k0 in (kv, np.array([0.0, 0.0, 2 * np.pi / (SHAPE[2] * STEP) * 1.5])):
        w.set(psi)
        lib.cgrid_set_momentum(w.grid, *k0)
        lib.grid_wf_propagate(w.h, pot.h, L.Complex.of(dt))
        kk = ((kv - k0) ** 2).sum()
        want = psi * np.exp(-1j * kk / (2 * L.HELIUM_MASS) * dt)
        assert relerr(w.get(), want) < (1e-12 if prec == L.GRID_FP64 else 3e-6)
        w.set(psi)
        ek = lib.grid_wf_energy(w.h, None)
        norm = np.prod(SHAPE) * STEP ** 3
        assert abs(ek - kk / (2 * L.HELIUM_MASS) * norm) < 1e-10 * (kv ** 2).sum() / (2 * L.HELIUM_MASS) * norm
    lib.cgrid_set_momentum(w.grid, 0.0, 0.0, 0.0)
    w.free(), pot.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_energy_with_potential(oracle, prec):
    """grid_wf_energy(gwf, V) = kinetic energy + h^3 sum |psi|^2 Re V."""
    psi = inputs.smooth_modulated(SHAPE, STEP, 1.0, 0.4)
    a, b = _data(5)
    V = 1e-3 * a + 1e-4j * b
    w, pot = L.Wave(SHAPE, STEP, prec), L.CField(SHAPE, STEP, prec)
    w.set(psi), pot.set(V)
    got = L.lib().grid_wf_energy(w.h, pot.h)
    want = oracle.kinetic_energy(psi, STEP) + (np.abs(psi) ** 2 * V.real).sum() * STEP ** 3
    assert abs(got / want - 1) < (1e-11 if prec == L.GRID_FP64 else 1e-5)
    w.free(), pot.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_driver_grid_calls(prec):
    a, b = _data(7, prec=prec)
    p, q = _data(8, prec=prec)
    f, g = a + 1j * b, p + 1j * q
    lib = L.lib()
    cf, cg, cd = L.CField(SHAPE, STEP, prec), L.CField(SHAPE, STEP, prec), L.CField(SHAPE, STEP, prec)
    cf.set(f), cg.set(g)
    # Laplacian (central differences, periodic)
    lib.cgrid_fd_laplace(cf.h, cd.h)
    lap = sum(np.roll(f, 1, d) + np.roll(f, -1, d) for d in range(3)) - 6 * f
    assert relerr(cd.get(), lap / STEP ** 2) < TOL[prec] * 10
    # conj(f) g
    lib.cgrid_conjugate_product(cd.h, cf.h, cg.h)
    assert relerr(cd.get(), np.conj(f) * g) < TOL[prec]
    # |f|^1.5
    lib.cgrid_abs_power(cd.h, cf.h, 1.5)
    assert relerr(cd.get(), np.abs(f) ** 1.5) < TOL[prec] * 10
    # integrals
    h3 = STEP ** 3
    z = lib.cgrid_integral_of_conjugate_product(cf.h, cg.h)
    want = (np.conj(f) * g).sum() * h3
    assert abs(complex(z.re, z.im) - want) < TOL[prec] * np.abs(f * g).sum() * h3
    z = lib.cgrid_integral(cf.h)
    assert abs(complex(z.re, z.im) - f.sum() * h3) < TOL[prec] * np.abs(f).sum() * h3
    # region: x_i = (i - n/2) h, bounds inclusive
    xs = [(np.arange(n) - n // 2) * STEP for n in SHAPE]
    lo, hi = (-1.5, -3.0, 0.0), (1.4, 2.1, 10.0)
    m = [(xs[d] >= lo[d] - 1e-12) & (xs[d] <= hi[d] + 1e-12) for d in range(3)]
    sel = m[0][:, None, None] & m[1][None, :, None] & m[2][None, None, :]
    z = lib.cgrid_integral_region(cf.h, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    assert abs(complex(z.re, z.im) - f[sel].sum() * h3) < TOL[prec] * np.abs(f).sum() * h3
    r = L.RField(SHAPE, STEP, prec)
    r.set(a)
    got = lib.rgrid_integral_region(r.h, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    assert abs(got - a[sel].sum() * h3) < TOL[prec] * np.abs(a).sum() * h3
    # |curl|, |r|^p
    fx, fy, fz = [L.RField(SHAPE, STEP, prec) for _ in range(3)]
    fx.set(a), fy.set(b), fz.set(p)
    D = lambda u, d: (np.roll(u, -1, d) - np.roll(u, 1, d)) / (2 * STEP)
    curl = np.sqrt((D(p, 1) - D(b, 2)) ** 2 + (D(a, 2) - D(p, 0)) ** 2 + (D(b, 0) - D(a, 1)) ** 2)
    lib.rgrid_abs_rot(r.h, fx.h, fy.h, fz.h)
    assert relerr(r.get(), curl) < TOL[prec] * 10
    lib.rgrid_abs_power(r.h, fx.h, 2.5)
    assert relerr(r.get(), np.abs(a) ** 2.5) < TOL[prec] * 10
    # power, product, division with eps, max / min, one flux component
    pos = np.abs(a) + 0.1
    fy.set(pos)
    lib.rgrid_power(r.h, fy.h, 0.5)
    assert relerr(r.get(), np.sqrt(pos)) < TOL[prec] * 10
    lib.rgrid_product(r.h, fx.h, fz.h)
    assert relerr(r.get(), a * p) < TOL[prec]
    lib.rgrid_division_eps(r.h, fx.h, fy.h, 0.05)
    assert relerr(r.get(), a / (pos + 0.05)) < TOL[prec] * 10
    assert lib.rgrid_max(fx.h) == a.max() and lib.rgrid_min(fx.h) == a.min()
    w = L.Wave(SHAPE, STEP, prec)
    w.set(f)
    jx = L.RField(SHAPE, STEP, prec)
    lib.grid_wf_probability_flux(w.h, r.h, None, None)
    lib.grid_wf_probability_flux_x(w.h, jx.h)
    assert np.array_equal(jx.get(), r.get())
    lib.grid_wf_probability_flux(w.h, None, None, r.h)
    lib.grid_wf_probability_flux_z(w.h, jx.h)
    assert np.array_equal(jx.get(), r.get())
    w.free(), jx.free()
    fy.set(b)
    # weighted integral with a host callback
    FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double, C.c_double)
    cb = FN(lambda _, x, y, z: 1.0 + 0.1 * x - 0.2 * y * z)
    lib.rgrid_weighted_integral.argtypes = [C.POINTER(L.RGrid), FN, C.c_void_p]
    lib.rgrid_weighted_integral.restype = C.c_double
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    got = lib.rgrid_weighted_integral(fx.h, cb, None)
    assert abs(got - (a * (1.0 + 0.1 * X - 0.2 * Y * Z)).sum() * h3) < 1e-12 * np.abs(a).sum() * h3 * 10
    for x in (cf, cg, cd, r, fx, fy, fz):
        x.free()


@pytest.mark.parametrize("angle", [1e-4, 3e-3, 0.7])
def test_fp32_phase_factors_keep_the_norm(angle):
    """A real potential makes e^{-iV dt/2} unitary, so grid_wf_propagate keeps the norm
    up to unbiased rounding.  The FP32 path applies the factor as its rounded value
    plus its residual; at angles below ~2.4e-4 the cosine rounds to exactly 1, and a
    residual added to an exact product is rounded away, always the same way: 1000
    steps at 1e-4 rad then gained (1 - cos) per half step, ~1e-5 of norm.  A mass of
    1e12 makes the kinetic factor's angles tiny too.  The state has a random phase per
    point, so that its norm is spread over all Fourier modes and the unbiased rounding
    of each transform averages out (a smooth state, carried by a few modes, shows
    ~1e-6 of rounding noise over these steps)."""
    shape, step = (32, 32, 32), 0.5 / L.AUTOANG
    lib = L.lib()
    rp = np.exp(2j * np.pi * np.random.default_rng(3).random(shape))
    psi0 = inputs.smooth_modulated(shape, step, 0.0218 * L.AUTOANG ** 3, 0.3, seed=11) * rp
    dt = 200.0
    V = 2.0 * angle / dt  # V dt / 2 = angle
    w = L.Wave(shape, step, L.GRID_FP32, mass=1e12)
    w.set(psi0.astype(np.complex64).astype(complex))
    pot = L.CField(shape, step, L.GRID_FP32)
    pot.set(np.full(shape, V + 0j))
    n0 = lib.grid_wf_norm(w.h)
    for _ in range(1000):
        lib.grid_wf_propagate(w.h, pot.h, L.Complex.of(dt))
    drift = lib.grid_wf_norm(w.h) / n0 - 1
    pot.free(), w.free()
    assert abs(drift) < 1e-6, drift
