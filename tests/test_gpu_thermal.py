"""GPU parity of the thermal OT models T400MK ... T3000MK against the FP64
oracle in both device precisions: their parameter rows (src/ot.c:1050-1216)
and the ideal Bose gas term (src/ot.c:402-407, src/ot-energy.c:101-104, in the
form of src/common.c:482-530), on a state whose density runs from 0.1 to 1.9
times the bulk one, so that the gas term is both saturated (x = rho lambda^3
above g_{3/2}(1)) and not.  The potential, the energy density and one fused
step in real and in imaginary time are compared, and the entry points whose
derivative does not cover the gas term (its fugacity is a bracketed search)
refuse these models.
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
MODELS = {"t1200-full": 256 | 9, "t3000": 262144, "t400-kc": 32 | 1}


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
    assert relerr(V.real, Vo.real) < TOL[prec]
    if model & 8:
        assert relerr(V.imag, Vo.imag) < TOL[prec] * 10
    assert relerr(e, eo) < TOL[prec]
    E, Eo = f.energy(), eo.sum() * STEP ** 3
    assert abs(E - Eo) < TOL[prec] * np.abs(eo).sum() * STEP ** 3
    mu0 = o.bulk("dEdRho", o.p.rho0)
    for dt in (0.5 / L.AUTOFS, -1j * 0.5 / L.AUTOFS):
        want = psi.copy()
        o.single_step(want, None, mu0, dt)
        w.set(psi)
        L.lib().dft_ot_propagate(f.h, w.h, None, mu0, L.Complex.of(dt))
        err = relerr(w.get(), want)
        _log("  step dt %s: rel %.2e" % (dt, err))
        assert err < TOL[prec]
    pot.free(), eps.free(), f.free(), w.free()


def test_thermal_refusals():
    """The energy gradient, the transposed Jacobians and the adjoint sweeps refuse
    the thermal models above 0 K (message, -1, outputs untouched)."""
    rho0 = 0.0218 * L.AUTOANG ** 3
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(inputs.smooth_modulated(SHAPE, STEP, rho0, 0.3))
    f = L.Functional(256 | 1, w, 2, 8)
    lib = L.lib()
    rb, vb = L.RField(SHAPE, STEP, L.GRID_FP64), L.RField(SHAPE, STEP, L.GRID_FP64)
    rb.set(np.full(SHAPE, 3.0))
    assert lib.dft_ot_potential_vjp(f.h, rb.h, vb.h, w.h) == -1
    lib.dft_ot_energy_gradient(f.h, rb.h, vb.h)
    assert np.all(rb.get() == 3.0)
    pb, cb = L.CField(SHAPE, STEP, L.GRID_FP64), L.CField(SHAPE, STEP, L.GRID_FP64)
    assert lib.dft_ot_potential_vjp_psi(f.h, pb.h, cb.h, w.h) == -1
    with pytest.raises(RuntimeError):
        f.adjoint_well(-50.0 / L.AUTOK, 2.0 / L.AUTOANG, (0.0, 0.0, 0.0), 0.0, 0.5 / L.AUTOFS, 2, 1)
    for x in (rb, vb, pb, cb, f, w):
        x.free()
