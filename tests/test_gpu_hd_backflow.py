"""GPU parity of the backflow density switch that HD / HD2 turn on together with
BACKFLOW (src/ot.c:843-982, src/ot-energy.c:221-227) and of the T0MK parameter
row (src/ot.c:1038-1048), against the FP64 oracle in both device precisions,
through the C ABI.

With the switch the backflow takes rho~ = G(rho) rho wherever it uses the
density, Re V_BF carries d(rho~)/d rho and the divergence half of Im V_BF
carries G.  Two settings of the switch are run: the published one (xi = 1e4
A^3, rho_bf = 0.033 A^-3: G drops from 1 to 0 within 1e-4 A^-3, so the test
states see G = 1 below rho_bf and G ~ 0 above it) and a widened one (xi = 300
A^3, rho_bf = 0.026 A^-3: G runs from ~1 to ~0.05 over the test densities, so
d(G rho)/d rho and the gradient of rho~ are exercised at every point).  The
states are smooth densities from 0.007 to 0.037 A^-3 with a smooth phase.

In FP32 the published switch is the one ill-conditioned step: x = xi (rho -
rho_bf) carries xi rho_bf ~ 330 times the rounding of rho (~2e-5 in x), and at
the points within ~1e-4 A^-3 of rho_bf, d(G rho)/d rho (up to xi rho / 2 ~ 165)
moves by about twice that, so Re V there agrees with the oracle to ~2.4e-5 of
its maximum; the bar is 5x the FP32 one (1e-4) for that case only.

The transposed Jacobian of the switched potential (dft_ot_potential_vjp_psi)
is checked against central differences of dft_ot_potential (dot-product test)
and the adjoint sweep through it against central differences of the final
energy, both with the widened switch in FP64.
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
SWITCH = {"published": None, "widened": (300.0, 0.026)}  # (xi A^3, rho_bf A^-3)


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _log(msg):
    print(msg, flush=True)


def _set_switch(sw, *params):
    if sw is None:
        return
    for p in params:
        if hasattr(p, "bf_xi"):
            p.bf_xi, p.bf_rhobf = sw[0] / L.AUTOANG ** 3, sw[1] * L.AUTOANG ** 3
        else:
            p.xi, p.rhobf = sw[0] / L.AUTOANG ** 3, sw[1] * L.AUTOANG ** 3


def _state(rho0, seed=21):
    return inputs.smooth_modulated(SHAPE, STEP, rho0, 0.45, seed=seed)


def _rfield(a, prec):
    r = L.RField(SHAPE, STEP, prec)
    r.set(a)
    return r


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("sw", list(SWITCH), ids=list(SWITCH))
@pytest.mark.parametrize("model", [10, 11, 13], ids=["hd-bf", "kc-hd-bf", "kc-hd2-bf"])
def test_switch_potential_energy_step_match_oracle(oracle, prec, sw, model):
    """The fused potential (Re and Im), the energy density and E, and one fused
    split-operator step in real and in imaginary time."""
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    _set_switch(SWITCH[sw], o.p)
    psi = _state(o.p.rho0)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    _set_switch(SWITCH[sw], f.p)
    pot, eps = L.CField(SHAPE, STEP, prec), L.RField(SHAPE, STEP, prec)
    f.potential(pot)
    f.energy_density(eps)
    V, e, E = pot.get(), eps.get(), f.energy()
    Vo, eo = o.potential(psi), o.energy_density(psi)
    tol = TOL[prec] * (5 if prec == L.GRID_FP32 and sw == "published" else 1)
    _log("model %d %s prec %d: Re V %.2e, Im V %.2e, eps %.2e" % (
        model, sw, prec, relerr(V.real, Vo.real), relerr(V.imag, Vo.imag), relerr(e, eo)))
    assert relerr(V.real, Vo.real) < tol
    assert relerr(V.imag, Vo.imag) < tol * 10
    assert relerr(e, eo) < tol
    Eo = eo.sum() * STEP ** 3
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


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_switch_unfused_entry_points(oracle, prec):
    """dft_ot_potential2 (W_d left in the velocity inputs), dft_ot_backflow_potential
    and dft_ot_energy_density_bf with the widened switch."""
    po = oracle
    sw = SWITCH["widened"]
    o11, o3, o10, o2 = [po.OracleOT(m, SHAPE, STEP, 2, 8) for m in (11, 3, 10, 2)]
    _set_switch(sw, o11.p, o3.p, o10.p, o2.p)
    psi = _state(o11.p.rho0, seed=23)
    rho = np.abs(psi) ** 2
    v = po.velocity(psi, STEP, L.HELIUM_MASS, L.DFT_EPS)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(11, w, 2, 8)
    _set_switch(sw, f.p)
    lib = L.lib()
    base = (1e-5 * inputs.splitmix64_uniform(psi.size, 9).reshape(SHAPE) + 2e-6j).astype(complex)
    pot = L.CField(SHAPE, STEP, prec)
    # potential2
    pot.set(base)
    gr, gv = _rfield(rho, prec), [_rfield(a, prec) for a in v]
    lib.dft_ot_potential2(f.h, pot.h, gr.h, *[g.h for g in gv])
    want = o11.potential2(rho, *v, pot=base)
    got = pot.get()
    assert relerr(got.real, want.real) < TOL[prec] * 10
    assert relerr(got.imag - base.imag, want.imag - base.imag) < TOL[prec] * 50
    W = [np.ascontiguousarray(a, np.float64).copy() for a in v]
    po.lib().orc_ot_potential2(o11._h, po._p(np.zeros(SHAPE, np.complex128)), po._p(np.ascontiguousarray(rho)),
                               *[po._p(a) for a in W])
    for d in range(3):
        assert relerr(gv[d].get(), W[d]) < TOL[prec] * 50
    # backflow part alone
    for g, a in zip(gv, v):
        g.set(a)
    pot.set(base)
    lib.dft_ot_backflow_potential(f.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    want = base + o11.potential2(rho, *v) - o3.potential2(rho, *v)
    got = pot.get()
    scale = np.abs(want - base).max()
    assert np.abs(got.real - want.real).max() < TOL[prec] * 50 * scale
    assert np.abs(got.imag - want.imag).max() < TOL[prec] * 50 * scale
    # backflow energy with a given density (velocity from psi)
    rho2 = np.abs(_state(o11.p.rho0, seed=24)) ** 2
    gr.set(rho2)
    eb = 1e-7 * inputs.splitmix64_uniform(psi.size, 4).reshape(SHAPE)
    eps = _rfield(eb, prec)
    lib.dft_ot_energy_density_bf(f.h, eps.h, w.h, gr.h)
    want_bf = eb + o10.energy_density_rv(rho2, *v) - o2.energy_density_rv(rho2, *v)
    s = np.abs(want_bf - eb).max()
    assert np.abs(eps.get() - want_bf).max() < TOL[prec] * 50 * s
    for x in gv + [gr, eps, pot, f, w]:
        x.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [16, 25], ids=["t0mk", "t0mk-kc-bf"])
def test_t0mk_matches_oracle(oracle, prec, model):
    """The T0MK row through the device path: potential, energy density and one step."""
    o = oracle.OracleOT(model, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.3, seed=5)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    assert f.p.b == o.p.b and f.p.lj_h == o.p.lj_h and f.p.rho0 == o.p.rho0
    pot, eps = L.CField(SHAPE, STEP, prec), L.RField(SHAPE, STEP, prec)
    f.potential(pot)
    f.energy_density(eps)
    Vo, eo = o.potential(psi), o.energy_density(psi)
    assert relerr(pot.get().real, Vo.real) < TOL[prec]
    assert relerr(eps.get(), eo) < TOL[prec]
    mu0 = o.bulk("dEdRho", o.p.rho0)
    want = psi.copy()
    o.single_step(want, None, mu0, 0.5 / L.AUTOFS)
    L.lib().dft_ot_propagate(f.h, w.h, None, mu0, L.Complex.of(0.5 / L.AUTOFS))
    assert relerr(w.get(), want) < TOL[prec]
    pot.free(), eps.free(), f.free(), w.free()


def _switch_potential(model, psi, sw):
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    _set_switch(sw, f.p)
    pot = L.CField(SHAPE, STEP, L.GRID_FP64)
    f.potential(pot)
    V = pot.get()
    pot.free(), f.free(), w.free()
    return V


@pytest.mark.parametrize("model", [10, 11], ids=["hd-bf", "kc-hd-bf"])
def test_switch_potential_vjp_psi_dot_product(model):
    """<vbar, dV/dpsi . dpsi> (Re and Im of V) by central differences of
    dft_ot_potential equals Re <psi_bar, dpsi> from dft_ot_potential_vjp_psi."""
    sw = SWITCH["widened"]
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi = _state(rho0)
    dpsi = inputs.smooth_modulated(SHAPE, STEP, rho0, 0.5, seed=78) * np.exp(1j * 0.7) - np.sqrt(rho0)
    dpsi *= 1e-3 * np.sqrt(rho0) / np.abs(dpsi).max()
    u = inputs.splitmix64_uniform(2 * psi.size, 12).reshape((2,) + SHAPE)
    vbar = u[0] + 1j * u[1]
    eps = 1e-2
    dV = _switch_potential(model, psi + eps * dpsi, sw) - _switch_potential(model, psi - eps * dpsi, sw)
    fd = np.sum(vbar.real * dV.real + vbar.imag * dV.imag) / (2 * eps)
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    _set_switch(sw, f.p)
    vb, pb = L.CField(SHAPE, STEP, L.GRID_FP64), L.CField(SHAPE, STEP, L.GRID_FP64)
    vb.set(vbar)
    assert L.lib().dft_ot_potential_vjp_psi(f.h, pb.h, vb.h, w.h) == 0
    g = pb.get()
    adj = np.sum(g.real * dpsi.real + g.imag * dpsi.imag)
    _log("switch psi-vjp model %d: FD %.12e adjoint %.12e rel %.2e" % (model, fd, adj, abs(adj / fd - 1)))
    assert abs(adj / fd - 1) < 1e-7
    for x in (pb, vb, f, w):
        x.free()


def _switch_well(model, psi0, amp, width, dt, nsteps, every):
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(psi0)
    f = L.Functional(model, w, 2, 8)
    _set_switch(SWITCH["widened"], f.p)
    E, g = f.adjoint_well(amp, width, (0.0, 0.0, 0.0), 0.0, dt, nsteps, every)
    f.free(), w.free()
    return E, g


def test_switch_adjoint_well_matches_fd():
    """The checkpointed adjoint sweep through the switch (KC|HD|BF, 6 real-time
    steps of 0.5 fs, checkpoints every 4): dE/d(well depth, width) against
    central differences."""
    model, dt, K, every = 11, 0.5 / L.AUTOFS, 6, 4
    psi0 = _state(0.0218 * L.AUTOANG ** 3)
    amp, width = -50.0 / L.AUTOK, 2.0 / L.AUTOANG
    E, g = _switch_well(model, psi0, amp, width, dt, K, every)
    da, dw = 1e-4 * abs(amp), 1e-4 * width
    fa = (_switch_well(model, psi0, amp + da, width, dt, K, every)[0]
          - _switch_well(model, psi0, amp - da, width, dt, K, every)[0]) / (2 * da)
    fw = (_switch_well(model, psi0, amp, width + dw, dt, K, every)[0]
          - _switch_well(model, psi0, amp, width - dw, dt, K, every)[0]) / (2 * dw)
    _log("switch well: E %.10e; dE/damp %.10e FD %.10e; dE/dwidth %.10e FD %.10e" % (E, g[0], fa, g[1], fw))
    assert abs(g[0] / fa - 1) < 1e-6
    assert abs(g[1] / fw - 1) < 1e-6


def test_models_accepted_and_refused():
    """HD / HD2 with BACKFLOW and T0MK are accepted; GP, DR and ZERO are refused at
    allocation."""
    w = L.Wave(SHAPE, STEP, L.GRID_FP64)
    w.set(_state(0.0218 * L.AUTOANG ** 3))
    lib = L.lib()
    for m in (524288, 1048576 | 9, 2097152):
        assert not lib.dft_ot_alloc(m, w.h, 2, 8), m
    for m in (2 | 9, 16 | 4 | 8, 16 | 1):
        f = L.Functional(m, w, 2, 8)
        f.free()
    w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_switch_on_512_point_lines(oracle, prec):
    """The switch's z-pass kernels on 512-point lines (two wavefronts per pair,
    block-barrier exchanges): KC|HD|BF potential and energy with the widened switch."""
    shape = (8, 8, 512)
    o = oracle.OracleOT(11, shape, STEP, 2, 8)
    _set_switch(SWITCH["widened"], o.p)
    psi = inputs.smooth_modulated(shape, STEP, o.p.rho0, 0.45, seed=29)
    w = L.Wave(shape, STEP, prec)
    w.set(psi)
    f = L.Functional(11, w, 2, 8)
    _set_switch(SWITCH["widened"], f.p)
    pot, eps = L.CField(shape, STEP, prec), L.RField(shape, STEP, prec)
    f.potential(pot)
    f.energy_density(eps)
    Vo, eo = o.potential(psi), o.energy_density(psi)
    V = pot.get()
    _log("512 lines prec %d: Re V %.2e, Im V %.2e, eps %.2e" % (prec, relerr(V.real, Vo.real),
                                                               relerr(V.imag, Vo.imag), relerr(eps.get(), eo)))
    assert relerr(V.real, Vo.real) < TOL[prec]
    assert relerr(V.imag, Vo.imag) < TOL[prec] * 10
    assert relerr(eps.get(), eo) < TOL[prec]
    pot.free(), eps.free(), f.free(), w.free()
