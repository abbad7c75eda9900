"""GPU parity of the high-density penalty (models HD = 2 and HD2 = 4, alone or
with KC) against the FP64 oracle, on a density bump that reaches the
penalty's onset (0.4 A^-3 at the centre): potential, density and energy
density, the fused split-operator step, the energy gradient, the transposed
Jacobian of the potential (dot-product test against central differences) and
the adjoint sweep (against central differences of the final energy).  HD
together with BACKFLOW (the backflow density switch) is covered by
test_gpu_hd_backflow.py.

The FP32 fused real-time step lands near 1.4e-5 of the oracle here (bar
2e-5): at 0.4 A^-3 the penalty is ~0.7 hartree and beta*rho ~ 16 amplifies
the FP32 rounding of rho into ~1e-6 of V, which a 0.5 fs step turns into
~1e-5 rad of phase.
"""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu

TOL = {L.GRID_FP64: 1e-10, L.GRID_FP32: 2e-5}
PRECS = [L.GRID_FP64, L.GRID_FP32]
PNAME = {L.GRID_FP64: "fp64", L.GRID_FP32: "fp32"}
MODELS = [2, 3, 4, 5]
MNAME = ["hd", "kc-hd", "hd2", "kc-hd2"]
SHAPE, STEP = (16, 16, 32), 0.6 / L.AUTOANG


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", MODELS, ids=MNAME)
def test_hd_potential_and_energy_match_oracle(oracle, prec, model):
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    psi = inputs.dense_bump(SHAPE, STEP, o.p.rho0)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    pot = L.CField(SHAPE, STEP, prec)
    f.potential(pot)
    V = pot.get()
    eps = L.RField(SHAPE, STEP, prec)
    f.energy_density(eps)
    e = eps.get()
    E = f.energy()
    Vo, eo = o.potential(psi), o.energy_density(psi)
    _log("model %d prec %d: V rel %.2e, eps rel %.2e" % (model, prec, relerr(V.real, Vo.real), relerr(e, eo)))
    assert relerr(V.real, Vo.real) < TOL[prec]
    assert np.abs(V.imag).max() == 0.0
    assert relerr(e, eo) < TOL[prec]
    Eo = eo.sum() * STEP ** 3
    assert abs(E - Eo) < TOL[prec] * np.abs(eo).max() * STEP ** 3 * eo.size
    pot.free(), eps.free(), f.free(), w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [2, 5], ids=["hd", "kc-hd2"])
def test_hd_fused_single_step_matches_oracle(oracle, prec, model):
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    psi0 = inputs.dense_bump(SHAPE, STEP, o.p.rho0)
    mu0 = o.bulk("dEdRho", o.p.rho0)
    w = L.Wave(SHAPE, STEP, prec)
    f = L.Functional(model, w, 2, 8)
    for dt in (-1j * 0.5 / L.AUTOFS, 0.5 / L.AUTOFS):
        want = psi0.copy()
        o.single_step(want, None, mu0, dt)
        w.set(psi0)
        L.lib().dft_ot_propagate(f.h, w.h, None, mu0, L.Complex.of(dt))
        err = relerr(w.get(), want)
        _log("model %d prec %d dt %s: step rel %.2e" % (model, prec, dt, err))
        assert err < TOL[prec]
    f.free(), w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [2, 3], ids=["hd", "kc-hd"])
def test_hd_energy_gradient_matches_oracle(oracle, prec, model):
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    rho = np.abs(inputs.dense_bump(SHAPE, STEP, o.p.rho0)) ** 2
    w = L.Wave(SHAPE, STEP, prec)
    f = L.Functional(model, w, 2, 8)
    r = L.RField(SHAPE, STEP, prec)
    g = L.RField(SHAPE, STEP, prec)
    r.set(rho)
    L.lib().dft_ot_energy_gradient(f.h, g.h, r.h)
    E = L.lib().dft_ot_energy_rho(f.h, r.h)
    assert relerr(g.get(), o.energy_gradient(rho)) < TOL[prec]
    Eo = o.energy_rho(rho)
    assert abs(E - Eo) < 10 * TOL[prec] * abs(Eo)
    for x in (g, r, f, w):
        x.free()


def _potential(model, psi):
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    pot = L.CField(SHAPE, STEP, L.GRID_FP64)
    f.potential(pot)
    V = pot.get().real
    pot.free(), f.free(), w.free()
    return V


@pytest.mark.parametrize("model", [2, 3], ids=["hd", "kc-hd"])
def test_hd_potential_vjp_dot_product(model):
    rho0 = 0.0218 * L.AUTOANG ** 3
    rho = np.abs(inputs.dense_bump(SHAPE, STEP, rho0)) ** 2
    drho = inputs.smooth_modulated(SHAPE, STEP, rho0, 0.5, seed=77).real - np.sqrt(rho0)
    drho *= 1e-3 * rho0 / np.abs(drho).max()
    vbar = inputs.splitmix64_uniform(rho.size, 11).reshape(SHAPE)
    eps = 1e-2
    Vp = _potential(model, np.sqrt(rho + eps * drho).astype(complex))
    Vm = _potential(model, np.sqrt(rho - eps * drho).astype(complex))
    fd = np.sum(vbar * (Vp - Vm)) / (2 * eps)
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(np.sqrt(rho).astype(complex))
    f = L.Functional(model, w, 2, 8)
    vb = L.RField(SHAPE, STEP, L.GRID_FP64)
    vb.set(vbar)
    rb = L.RField(SHAPE, STEP, L.GRID_FP64)
    L.lib().dft_ot_potential_vjp(f.h, rb.h, vb.h, w.h)
    adj = np.sum(rb.get() * drho)
    _log("model %d: FD %.12e adjoint %.12e rel %.2e" % (model, fd, adj, abs(adj / fd - 1)))
    assert abs(adj / fd - 1) < 1e-7
    for x in (rb, vb, f, w):
        x.free()


@pytest.mark.parametrize("model", [2, 3], ids=["hd", "kc-hd"])
def test_hd_adjoint_well_matches_fd(model):
    """The checkpointed adjoint sweep through the penalty (6 real-time steps of
    0.5 fs on the density bump, checkpoints every 4): dE/d(well depth, width)
    against central differences."""
    from tests.test_gpu_adjoint import _fd_check
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi0 = inputs.dense_bump(SHAPE, STEP, rho0)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    _fd_check(model, SHAPE, STEP, psi0, amp, width, 0.0, 0.5 / L.AUTOFS, 6, 4, 1e-6)
