"""Adjoint (reverse-mode) checks on the GPU: the transposed Jacobian of
V_OT against central finite differences of the potential (dot-product tests,
with respect to rho for PLAIN/KC and to psi for every model, backflow's
complex V included), the adjoint split-operator sweep against central finite
differences of the final energy with respect to the gaussian well's depth and
width (128^3 droplet with the -50 K / 2 A impurity well, full OT), and of the
bubble radius of a released cavity with respect to its initial radius
(512^3, 17 A cavity); complex64 sweeps against the FP64 sweep's gradient.
"""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _potential(model, shape, step, psi, kc_exact):
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    if kc_exact:
        f.set_kc_exact(True)
    pot = L.CField(shape, step, L.GRID_FP64)
    f.potential(pot)
    V = pot.get().real
    pot.free(), f.free(), w.free()
    return V


@pytest.mark.parametrize("model,kc_exact", [(0, False), (1, False), (1, True)], ids=["plain", "kc", "kc-exact"])
def test_potential_vjp_dot_product(model, kc_exact):
    """<vbar, dV/drho . drho> by central differences of dft_ot_potential equals
    <(dV/drho)^T vbar, drho> from dft_ot_potential_vjp."""
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    rho = np.abs(inputs.smooth_modulated(shape, step, rho0, 0.3)) ** 2
    drho = inputs.smooth_modulated(shape, step, rho0, 0.5, seed=77).real - np.sqrt(rho0)
    drho *= 1e-3 * rho0 / np.abs(drho).max()
    vbar = inputs.splitmix64_uniform(rho.size, 11).reshape(shape)
    eps = 1e-2
    Vp = _potential(model, shape, step, np.sqrt(rho + eps * drho).astype(complex), kc_exact)
    Vm = _potential(model, shape, step, np.sqrt(rho - eps * drho).astype(complex), kc_exact)
    fd = np.sum(vbar * (Vp - Vm)) / (2 * eps)
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(np.sqrt(rho).astype(complex))
    f = L.Functional(model, w, 2, 8)
    if kc_exact:
        f.set_kc_exact(True)
    vb = L.RField(shape, step, L.GRID_FP64)
    vb.set(vbar)
    rb = L.RField(shape, step, L.GRID_FP64)
    L.lib().dft_ot_potential_vjp(f.h, rb.h, vb.h, w.h)
    adj = np.sum(rb.get() * drho)
    _log("model %d exact %d: FD %.12e adjoint %.12e rel %.2e" % (model, kc_exact, fd, adj, abs(adj / fd - 1)))
    assert abs(adj / fd - 1) < 1e-7
    for x in (rb, vb, f, w):
        x.free()


def _complex_potential(model, shape, step, psi, cbf=1.0):
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    f.p.c_bfpot = cbf
    pot = L.CField(shape, step, L.GRID_FP64)
    f.potential(pot)
    V = pot.get()
    pot.free(), f.free(), w.free()
    return V


@pytest.mark.parametrize("model,cbf", [(0, 1.0), (1, 1.0), (8, 1.0), (9, 1.0), (9, 0.5)],
                         ids=["plain", "kc", "bf", "full", "full-cbf0.5"])
def test_potential_vjp_psi_dot_product(model, cbf):
    """<vbar, dV/dpsi . dpsi> (Re and Im of V, dpsi complex) by central differences
    of dft_ot_potential equals Re <psi_bar, dpsi> from dft_ot_potential_vjp_psi."""
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi = inputs.smooth_modulated(shape, step, rho0, 0.3)
    dpsi = inputs.smooth_modulated(shape, step, rho0, 0.5, seed=78) * np.exp(1j * 0.7) - np.sqrt(rho0)
    dpsi *= 1e-3 * np.sqrt(rho0) / np.abs(dpsi).max()
    u = inputs.splitmix64_uniform(2 * psi.size, 12).reshape((2,) + shape)
    vbar = u[0] + 1j * u[1]
    eps = 1e-2
    Vp = _complex_potential(model, shape, step, psi + eps * dpsi, cbf)
    Vm = _complex_potential(model, shape, step, psi - eps * dpsi, cbf)
    fd = np.sum(vbar.real * (Vp - Vm).real + vbar.imag * (Vp - Vm).imag) / (2 * eps)
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    f.p.c_bfpot = cbf
    vb, pb = L.CField(shape, step, L.GRID_FP64), L.CField(shape, step, L.GRID_FP64)
    vb.set(vbar)
    assert L.lib().dft_ot_potential_vjp_psi(f.h, pb.h, vb.h, w.h) == 0
    g = pb.get()
    adj = np.sum(g.real * dpsi.real + g.imag * dpsi.imag)
    _log("psi-vjp model %d cbf %.1f: FD %.12e adjoint %.12e rel %.2e" % (model, cbf, fd, adj, abs(adj / fd - 1)))
    assert abs(adj / fd - 1) < 1e-7
    for x in (pb, vb, f, w):
        x.free()


def test_potential_vjp_rho_refuses_backflow():
    """The density VJP refuses a backflow model (its V also depends on the phase)."""
    shape, step = (16, 16, 16), 0.6 / L.AUTOANG
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(inputs.smooth_modulated(shape, step, 0.0218 * L.AUTOANG ** 3, 0.2))
    f = L.Functional(9, w)
    rb, vb = L.RField(shape, step, L.GRID_FP64), L.RField(shape, step, L.GRID_FP64)
    assert L.lib().dft_ot_potential_vjp(f.h, rb.h, vb.h, w.h) == -1
    for x in (rb, vb, f, w):
        x.free()


def _well_gradient(model, shape, step, psi0, amp, width, mu, dt, nsteps, every, prec=L.GRID_FP64):
    w = L.Wave(shape, step, prec)
    w.set(psi0)
    f = L.Functional(model, w)
    E, g = f.adjoint_well(amp, width, (0.0, 0.0, 0.0), mu, dt, nsteps, every)
    f.free(), w.free()
    return E, g


def _fd_check(model, shape, step, psi0, amp, width, mu, dt, nsteps, every, tol):
    E, g = _well_gradient(model, shape, step, psi0, amp, width, mu, dt, nsteps, every)
    da, dw = 1e-4 * abs(amp), 1e-4 * width
    fa = (_well_gradient(model, shape, step, psi0, amp + da, width, mu, dt, nsteps, every)[0] -
          _well_gradient(model, shape, step, psi0, amp - da, width, mu, dt, nsteps, every)[0]) / (2 * da)
    fw = (_well_gradient(model, shape, step, psi0, amp, width + dw, mu, dt, nsteps, every)[0] -
          _well_gradient(model, shape, step, psi0, amp, width - dw, mu, dt, nsteps, every)[0]) / (2 * dw)
    _log("model %d %s K=%d every=%d: E %.10e; dE/damp adjoint %.10e FD %.10e (rel %.2e); "
         "dE/dwidth adjoint %.10e FD %.10e (rel %.2e)" % (model, shape, nsteps, every, E, g[0], fa,
                                                           abs(g[0] / fa - 1), g[1], fw, abs(g[1] / fw - 1)))
    assert abs(g[0] / fa - 1) < tol
    assert abs(g[1] / fw - 1) < tol


@pytest.mark.parametrize("model", [0, 1, 8, 9], ids=["plain", "kc", "bf", "full"])
def test_adjoint_well_small(model):
    """8 real-time steps at 32^3 with checkpoints every 3 steps (segments 3, 3, 2)."""
    shape, step = (32, 32, 32), 0.5 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.smooth_modulated(shape, step, rho0, 0.2)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(model, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 8, 3, 1e-6)


@pytest.mark.parametrize("model", [1, 9], ids=["kc", "full"])
def test_adjoint_well_imaginary_time(model):
    """Imaginary-time steps (complex dt) through the same sweep."""
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.smooth_modulated(shape, step, rho0, 0.2)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(model, shape, step, psi0, amp, width, 0.0, -1j * 2.0 / L.AUTOFS, 6, 4, 1e-6)


def test_adjoint_well_moving_frame():
    """The sweep in a momentum frame (cgrid_set_momentum): T and its adjoint use (k - k0)^2."""
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.smooth_modulated(shape, step, rho0, 0.2)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG

    def run(a, wd):
        w = L.Wave(shape, step, L.GRID_FP64)
        w.set(psi0)
        L.lib().cgrid_set_momentum(w.grid, 0.0, 0.0, 0.3)
        f = L.Functional(9, w)
        r = f.adjoint_well(a, wd, (0.0, 0.0, 0.0), 0.0, 5.0 / L.AUTOFS, 6, 4)
        f.free(), w.free()
        return r
    E, g = run(amp, width)
    da = 1e-4 * abs(amp)
    fa = (run(amp + da, width)[0] - run(amp - da, width)[0]) / (2 * da)
    _log("moving frame: dE/damp adjoint %.10e FD %.10e (rel %.2e)" % (g[0], fa, abs(g[0] / fa - 1)))
    assert abs(g[0] / fa - 1) < 1e-6


def test_adjoint_droplet128_well():
    """At full size (the differentiation workload): 128^3 droplet (N = 1000) with the
    -50 K, 2 A gaussian well, 500 real-time steps of 5 fs of the full OT functional
    (KC + backflow), checkpoints every 50.  Adjoint gradient vs central FD of
    E_final: <= 1e-4 (FP64)."""
    shape, step = (128, 128, 128), 0.4 / L.AUTOANG
    psi0 = inputs.droplet(shape, step)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(9, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 500, 50, 1e-4)


def test_adjoint_droplet128_well_complex64():
    """The same sweep in complex64 (FP64 reductions), started from the same
    FP32-representable state, against the FP64 sweep's gradient (which matches
    central differences): <= 1e-4 relative (measured 4e-6 and 2e-6).  FP32 central differences of E_final
    would be limited by the energy's rounding, so the FP64 adjoint is the reference."""
    shape, step = (128, 128, 128), 0.4 / L.AUTOANG
    psi0 = inputs.droplet(shape, step).astype(np.complex64).astype(np.complex128)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    E64, g64 = _well_gradient(9, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 500, 50)
    E32, g32 = _well_gradient(9, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 500, 50, prec=L.GRID_FP32)
    _log("complex64 droplet: E %.10e / %.10e; grad %s / %s" % (E32, E64, g32, g64))
    # E_final itself: 2.0e-6 after these 500 steps with the well
    assert abs(E32 / E64 - 1) < 1e-5
    for a, b in zip(g32, g64):
        assert abs(a / b - 1) < 1e-4


def _cavity_fd(model, shape, step, prec, rho0, R0, w, dt, nsteps, every, tol):
    def run(radius):
        wv = L.Wave(shape, step, prec)
        f = L.Functional(model, wv)
        mu = f.bulk("dEdRho", rho0)
        J, g = f.adjoint_cavity(rho0, radius, w, mu, dt, nsteps, every)
        f.free(), wv.free()
        return J, g
    J, g = run(R0)
    dR = 1e-4 * R0
    fd = (run(R0 + dR)[0] - run(R0 - dR)[0]) / (2 * dR)
    _log("cavity model %d %s K=%d every=%d: R_b %.10f (R0 %.6f); dR_b/dR0 adjoint %.10e FD %.10e (rel %.2e)"
         % (model, shape, nsteps, every, J, R0, g, fd, abs(g / fd - 1)))
    assert abs(g / fd - 1) < tol
    return J, g


@pytest.mark.parametrize("model", [0, 1, 9], ids=["plain", "kc", "full"])
def test_adjoint_cavity_small(model):
    """Released cavity at 64^3 (h = 0.8 A, R0 = 10 A, 1 A wall), 12 real-time steps of 5 fs,
    checkpoints every 5: bubble-radius gradient against central FD of R_b(R0)."""
    rho0 = 0.0218360 * L.AUTOANG ** 3
    _cavity_fd(model, (64, 64, 64), 0.8 / L.AUTOANG, L.GRID_FP64, rho0, 10.0 / L.AUTOANG, 1.0 / L.AUTOANG,
               5.0 / L.AUTOFS, 12, 5, 1e-6)


def test_adjoint_cavity_512():
    """At full size (512^3, h = 0.4 A, 17 A cavity with a 0.5 A wall, released;
    full OT), 40 steps of 2 fs with checkpoints every 10, FP64: dR_b/dR0 against
    central FD, <= 1e-4; then the complex64 sweep against the FP64 one, <= 1e-4."""
    rho0 = 0.0218360 * L.AUTOANG ** 3
    shape, step, R0, wd, dt = (512, 512, 512), 0.4 / L.AUTOANG, 17.0 / L.AUTOANG, 0.5 / L.AUTOANG, 2.0 / L.AUTOFS
    J, g = _cavity_fd(9, shape, step, L.GRID_FP64, rho0, R0, wd, dt, 40, 10, 1e-4)
    assert abs(J * L.AUTOANG - 17.0) < 1.0  # the bubble has barely moved in 80 fs
    wv = L.Wave(shape, step, L.GRID_FP32)
    f = L.Functional(9, wv)
    J32, g32 = f.adjoint_cavity(rho0, R0, wd, f.bulk("dEdRho", rho0), dt, 40, 10)
    f.free(), wv.free()
    _log("cavity complex64: R_b %.10f / %.10f; dR_b/dR0 %.10e / %.10e (rel %.2e)" % (J32, J, g32, g, abs(g32 / g - 1)))
    assert abs(g32 / g - 1) < 1e-4
