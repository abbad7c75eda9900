"""Adjoint (reverse-mode) checks on the GPU: the transposed Jacobian of
V_OT(rho) against central finite differences of the potential (dot-product
test), the adjoint split-operator sweep against central finite differences
of the final energy with respect to the gaussian well's depth and width
(128^3 droplet with the -50 K / 2 A impurity well), and of the
bubble radius of a released cavity with respect to its initial radius
(512^3, 17 A cavity).
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


@pytest.mark.parametrize("model", [0, 1], ids=["plain", "kc"])
def test_adjoint_well_small(model):
    """8 real-time steps at 32^3 with checkpoints every 3 steps (segments 3, 3, 2)."""
    shape, step = (32, 32, 32), 0.5 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.smooth_modulated(shape, step, rho0, 0.2)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(model, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 8, 3, 1e-6)


def test_adjoint_well_imaginary_time():
    """Imaginary-time steps (complex dt) through the same sweep."""
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.smooth_modulated(shape, step, rho0, 0.2)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(1, shape, step, psi0, amp, width, 0.0, -1j * 2.0 / L.AUTOFS, 6, 4, 1e-6)


def test_adjoint_droplet128_well():
    """At full size: 128^3 droplet (N = 1000) with the -50 K, 2 A gaussian well,
    500 real-time steps of 5 fs, checkpoints every 50; KC functional (the adjoint covers the
    PLAIN and KC models).  Adjoint gradient vs central FD of E_final: <= 1e-4."""
    shape, step = (128, 128, 128), 0.4 / L.AUTOANG
    psi0 = inputs.droplet(shape, step)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(1, shape, step, psi0, amp, width, 0.0, 5.0 / L.AUTOFS, 500, 50, 1e-4)


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


@pytest.mark.parametrize("model", [0, 1], ids=["plain", "kc"])
def test_adjoint_cavity_small(model):
    """Released cavity at 64^3 (h = 0.8 A, R0 = 10 A, 1 A wall), 12 real-time steps of 5 fs,
    checkpoints every 5: bubble-radius gradient against central FD of R_b(R0)."""
    rho0 = 0.0218360 * L.AUTOANG ** 3
    _cavity_fd(model, (64, 64, 64), 0.8 / L.AUTOANG, L.GRID_FP64, rho0, 10.0 / L.AUTOANG, 1.0 / L.AUTOANG,
               5.0 / L.AUTOFS, 12, 5, 1e-6)


def test_adjoint_cavity_512():
    """At full size (512^3, h = 0.4 A, 17 A cavity with a 0.5 A wall, released;
    KC functional: the adjoint covers PLAIN and KC), 40 steps of 2 fs with checkpoints
    every 10, FP64: dR_b/dR0 against central FD, <= 1e-4."""
    rho0 = 0.0218360 * L.AUTOANG ** 3
    J, g = _cavity_fd(1, (512, 512, 512), 0.4 / L.AUTOANG, L.GRID_FP64, rho0, 17.0 / L.AUTOANG, 0.5 / L.AUTOANG,
                      2.0 / L.AUTOFS, 40, 10, 1e-4)
    assert abs(J * L.AUTOANG - 17.0) < 1.0  # the bubble has barely moved in 80 fs


def test_adjoint_rejects_backflow():
    shape, step = (16, 16, 16), 0.6 / L.AUTOANG
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(inputs.smooth_modulated(shape, step, 0.0218 * L.AUTOANG ** 3, 0.2))
    f = L.Functional(9, w)
    with pytest.raises(RuntimeError):
        f.adjoint_well(-1e-4, 3.0, (0.0, 0.0, 0.0), 0.0, 1.0, 2, 1)
    f.free(), w.free()
