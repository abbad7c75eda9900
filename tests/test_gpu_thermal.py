"""GPU parity of the thermal OT models T400MK ... T3000MK against the FP64
oracle in both device precisions: their parameter rows (src/ot.c:1050-1216)
and the ideal Bose gas term (src/ot.c:402-407, src/ot-energy.c:101-104, in the
form of src/common.c:482-530), on a state whose density runs from 0.1 to 1.9
times the bulk one, so that the gas term is both saturated (x = rho lambda^3
above g_{3/2}(1)) and not.  The potential, the energy density and one fused
step in real and in imaginary time are compared.  Below saturation the gas
term's derivatives (c4 k_B T ln z for the energy, c4 k_B T lambda^3 / g_{1/2}(z)
for the potential, with the search's z constant between its steps) enter the
energy gradient, the transposed Jacobian and the adjoint sweep; those are
checked against central differences on densities that stay below saturation
(across it the reference's potential jumps).  The search's steps move each
point's energy and potential by ~1e-11 of their values, ~30 times within one
central-difference interval; summed over the grid they leave ~1e-7 of the
derivative, so these checks take 1e-6 (measured: 3.8e-7 for the gradient,
7e-8 for the transposed Jacobian, 2e-9 for the sweep).
"""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu

PRECS = [L.GRID_FP64, L.GRID_FP32]
PNAME = {L.GRID_FP64: "fp64", L.GRID_FP32: "fp32"}
TOL = {L.GRID_FP64: 1e-10, L.GRID_FP32: 2e-5}
SHAPE, STEP = (16, 16, 32), 0.6 / L.AUTOANG
# (t1200-hd-full: with the HD penalty and its backflow switch too; the state reaches 0.041 A^-3,
# across the switch, whose FP32 bar is 5x the usual one: see test_gpu_hd_backflow.py)
MODELS = {"t1200-full": 256 | 9, "t3000": 262144, "t400-kc": 32 | 1, "t1200-hd-full": 256 | 2 | 9}


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("name", list(MODELS))
def test_thermal_matches_oracle(oracle, prec, name):
    model = MODELS[name]
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.6, seed=41)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    assert f.p.c4 == o.p.c4 and f.p.temp == o.p.temp and f.p.b == o.p.b
    pot, eps = L.CField(SHAPE, STEP, prec), L.RField(SHAPE, STEP, prec)
    f.potential(pot)
    f.energy_density(eps)
    V, e = pot.get(), eps.get()
    Vo, eo = o.potential(psi), o.energy_density(psi)
    _log("%s prec %d: Re V %.2e, Im V %.2e, eps %.2e" % (name, prec, relerr(V.real, Vo.real),
                                                         relerr(V.imag, Vo.imag) if model & 8 else 0.0,
                                                         relerr(e, eo)))
    tol = TOL[prec] * (5 if prec == L.GRID_FP32 and model & 2 and model & 8 else 1)
    assert relerr(V.real, Vo.real) < tol
    if model & 8:
        assert relerr(V.imag, Vo.imag) < tol * 10
    assert relerr(e, eo) < tol
    E, Eo = f.energy(), eo.sum() * STEP ** 3
    assert abs(E - Eo) < tol * np.abs(eo).sum() * STEP ** 3
    mu0 = o.bulk("dEdRho", o.p.rho0)
    for dt in (0.5 / L.AUTOFS, -1j * 0.5 / L.AUTOFS):
        want = psi.copy()
        o.single_step(want, None, mu0, dt)
        w.set(psi)
        L.lib().dft_ot_propagate(f.h, w.h, None, mu0, L.Complex.of(dt))
        err = relerr(w.get(), want)
        _log("  step dt %s: rel %.2e" % (dt, err))
        assert err < tol
    pot.free(), eps.free(), f.free(), w.free()


def _below_saturation(o):
    """Densities from 0.11 to 0.79 rho0 at 3 K: every point below the saturation of
    the gas term (x = rho lambda^3 < g_{3/2}(1) up to 0.91 rho0), where it is smooth."""
    return inputs.smooth_modulated(SHAPE, STEP, 0.45 * o.p.rho0, 0.5, seed=43)


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_thermal_energy_gradient(oracle, prec):
    """dft_ot_energy_gradient (the gas term adds c4 k_B T ln z below saturation)
    against the oracle, and in FP64 against central differences of E[rho]."""
    model = 262144 | 1
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    rho = np.abs(_below_saturation(o)) ** 2
    w = L.Wave(SHAPE, STEP, prec)
    f = L.Functional(model, w, 2, 8)
    r, g = L.RField(SHAPE, STEP, prec), L.RField(SHAPE, STEP, prec)
    r.set(rho)
    L.lib().dft_ot_energy_gradient(f.h, g.h, r.h)
    assert relerr(g.get(), o.energy_gradient(rho)) < TOL[prec]
    if prec == L.GRID_FP64:
        eta = 1e-3 * o.p.rho0 * inputs.splitmix64_uniform(rho.size, 5).reshape(SHAPE)
        t = 1e-2
        r.set(rho + t * eta)
        ep = L.lib().dft_ot_energy_rho(f.h, r.h)
        r.set(rho - t * eta)
        em = L.lib().dft_ot_energy_rho(f.h, r.h)
        fd = (ep - em) / (2 * t)
        an = np.sum(g.get() * eta) * STEP ** 3
        _log("thermal gradient: FD %.12e analytic %.12e rel %.2e" % (fd, an, abs(an / fd - 1)))
        assert abs(an / fd - 1) < 1e-6
    for x in (r, g, f, w):
        x.free()


def _potential(model, psi):
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    pot = L.CField(SHAPE, STEP, L.GRID_FP64)
    f.potential(pot)
    V = pot.get()
    pot.free(), f.free(), w.free()
    return V


@pytest.mark.parametrize("model", [262144 | 1, 262144 | 9], ids=["t3000-kc", "t3000-full"])
def test_thermal_potential_vjp_psi_dot_product(oracle, model):
    """<vbar, dV/dpsi . dpsi> by central differences of dft_ot_potential equals
    Re <psi_bar, dpsi> from dft_ot_potential_vjp_psi (the gas term adds
    vbar c4 k_B T lambda^3 / g_{1/2}(z) to rho_bar)."""
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    psi = _below_saturation(o)
    dpsi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.5, seed=78) * np.exp(1j * 0.7) - np.sqrt(o.p.rho0)
    dpsi *= 1e-3 * np.sqrt(o.p.rho0) / np.abs(dpsi).max()
    u = inputs.splitmix64_uniform(2 * psi.size, 12).reshape((2,) + SHAPE)
    vbar = u[0] + 1j * u[1]
    eps = 1e-2
    dV = _potential(model, psi + eps * dpsi) - _potential(model, psi - eps * dpsi)
    fd = np.sum(vbar.real * dV.real + vbar.imag * dV.imag) / (2 * eps)
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    vb, pb = L.CField(SHAPE, STEP, L.GRID_FP64), L.CField(SHAPE, STEP, L.GRID_FP64)
    vb.set(vbar)
    assert L.lib().dft_ot_potential_vjp_psi(f.h, pb.h, vb.h, w.h) == 0
    g = pb.get()
    adj = np.sum(g.real * dpsi.real + g.imag * dpsi.imag)
    _log("thermal psi-vjp model %d: FD %.12e adjoint %.12e rel %.2e" % (model, fd, adj, abs(adj / fd - 1)))
    assert abs(adj / fd - 1) < 1e-6
    for x in (pb, vb, f, w):
        x.free()


def _well(model, psi0, amp, width, dt, nsteps, every):
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi0)
    f = L.Functional(model, w, 2, 8)
    E, g = f.adjoint_well(amp, width, (0.0, 0.0, 0.0), 0.0, dt, nsteps, every)
    f.free(), w.free()
    return E, g


def test_thermal_adjoint_well_matches_fd(oracle):
    """The checkpointed adjoint sweep through the gas term (T3000MK|KC, 6 real-time
    steps of 0.5 fs, checkpoints every 4) against central differences."""
    model, dt, K, every = 262144 | 1, 0.5 / L.AUTOFS, 6, 4
    psi0 = _below_saturation(oracle.OracleOT(model, SHAPE, STEP, 2, 8))
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    E, g = _well(model, psi0, amp, width, dt, K, every)
    da, dw = 1e-4 * abs(amp), 1e-4 * width
    fa = (_well(model, psi0, amp + da, width, dt, K, every)[0] - _well(model, psi0, amp - da, width, dt, K, every)[0]) / (2 * da)
    fw = (_well(model, psi0, amp, width + dw, dt, K, every)[0] - _well(model, psi0, amp, width - dw, dt, K, every)[0]) / (2 * dw)
    _log("thermal well: dE/damp %.10e FD %.10e; dE/dwidth %.10e FD %.10e" % (g[0], fa, g[1], fw))
    assert abs(g[0] / fa - 1) < 1e-6
    assert abs(g[1] / fw - 1) < 1e-6
