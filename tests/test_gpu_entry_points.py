"""GPU parity of the entry points beyond dft_ot_potential / dft_ot_energy_density,
against the FP64 oracle in both device precisions, through the C ABI:

  dft_ot_potential2 (all models; v_d A - B_d left in the velocity inputs,
    src/ot.c:923-931), dft_ot_backflow_potential (src/ot.c:840-991),
  dft_ot_add_lennard_jones_potential / _local_correlation_potential (src/ot.c:524, 547),
  dft_ot_energy_density_kc / _bf accumulating with a given density
    (src/ot-energy.c:123-288), grid_wf_velocity, rgrid_gradient_*,
  the max_veloc, max_bfpot and c_bfpot branches (src/ot.c:391-395, 911-917)
    in the potential entry points and the fused step,
and the backflow identities the oracle is pinned by (Galilean invariance,
Re V_BF = dE_BF/drho at fixed v, no norm moved by Im V_BF, the first-order
response of Im V_BF) run through the device path.
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


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _rfield(a, prec, shape=SHAPE, step=STEP):
    r = L.RField(shape, step, prec)
    r.set(a)
    return r


def _velocity_inputs(po, psi, step=STEP):
    return po.velocity(psi, step, L.HELIUM_MASS, L.DFT_EPS)


def _oracle_w(po, o, rho, v):
    """W_d = v_d A - B_d that potential2 leaves in its velocity inputs."""
    W = [np.ascontiguousarray(a, np.float64).copy() for a in v]
    po.lib().orc_ot_potential2(o._h, po._p(np.zeros(o.shape, np.complex128)), po._p(np.ascontiguousarray(rho)),
                               *[po._p(a) for a in W])
    return W


# ------------------------------------------------- potential2 / backflow --
@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [0, 1, 8, 9], ids=["plain", "kc", "bf", "full"])
def test_potential2_matches_oracle(oracle, prec, model):
    po = oracle
    o = po.OracleOT(model, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.3)
    rho = np.abs(psi) ** 2
    v = _velocity_inputs(po, psi)
    base = (1e-5 * inputs.splitmix64_uniform(psi.size, 9).reshape(SHAPE) + 2e-6j).astype(complex)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(model, w, 2, 8)
    pot = L.CField(SHAPE, STEP, prec)
    pot.set(base)
    gr, gv = _rfield(rho, prec), [_rfield(a, prec) for a in v]
    L.lib().dft_ot_potential2(f.h, pot.h, gr.h, *[g.h for g in gv])
    want = o.potential2(rho, *v, pot=base)
    got = pot.get()
    assert relerr(got.real, want.real) < TOL[prec] * 10
    if model & 8:
        assert relerr(got.imag - base.imag, want.imag - base.imag) < TOL[prec] * 50
    else:  # no imaginary part: the base is left as stored
        stored = base.imag if prec == L.GRID_FP64 else base.imag.astype(np.float32)
        assert np.array_equal(got.imag, stored)
    if model & 8:
        W = _oracle_w(po, o, rho, v)
        for d in range(3):
            assert relerr(gv[d].get(), W[d]) < TOL[prec] * 50
    else:
        for d in range(3):
            assert np.array_equal(gv[d].get(), v[d] if prec == L.GRID_FP64 else v[d].astype(np.float32))
    for x in gv + [gr, pot, f, w]:
        x.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_backflow_potential_matches_oracle(oracle, prec):
    """dft_ot_backflow_potential adds only the backflow part (Re and Im) and leaves W in v."""
    po = oracle
    o9, o1 = po.OracleOT(9, SHAPE, STEP, 2, 8), po.OracleOT(1, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o9.p.rho0, 0.3, seed=77)
    rho = np.abs(psi) ** 2
    v = _velocity_inputs(po, psi)
    base = (2e-6 * inputs.splitmix64_uniform(psi.size, 3).reshape(SHAPE)).astype(complex)
    want = base + o9.potential2(rho, *v) - o1.potential2(rho, *v)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(9, w, 2, 8)
    pot = L.CField(SHAPE, STEP, prec)
    pot.set(base)
    gr, gv = _rfield(rho, prec), [_rfield(a, prec) for a in v]
    L.lib().dft_ot_backflow_potential(f.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    got = pot.get()
    scale = np.abs(want - base).max()
    assert np.abs(got.real - want.real).max() < TOL[prec] * 50 * scale
    assert np.abs(got.imag - want.imag).max() < TOL[prec] * 50 * scale
    W = _oracle_w(po, o9, rho, v)
    for d in range(3):
        assert relerr(gv[d].get(), W[d]) < TOL[prec] * 50
    # a functional without backflow refuses (message, potential unchanged)
    f0 = L.Functional(1, w, 2, 8)
    pot.set(base)
    L.lib().dft_ot_backflow_potential(f0.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    assert np.array_equal(pot.get(), base if prec == L.GRID_FP64 else base.astype(np.complex64))
    for x in gv + [gr, pot, f0, f, w]:
        x.free()


# ---------------------------------------------------- _kc / _bf energies --
@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_lennard_jones_and_local_correlation_parts(oracle, prec):
    """dft_ot_add_lennard_jones_potential and dft_ot_add_local_correlation_potential
    (src/ot.c:524, 547) accumulate; their sum is the plain OT potential of the oracle.
    The split is pinned by the local part at uniform density, c2/2 rho^2 + c3/3 rho^3
    + (c2 rho + c3 rho^2) rho (the spherical average integrates to 1), and by the
    linearity of the Lennard-Jones part in rho."""
    o = oracle.OracleOT(0, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.3, seed=5)
    rho = np.abs(psi) ** 2
    base = (1e-5 * inputs.splitmix64_uniform(psi.size, 4).reshape(SHAPE) + 3e-6j).astype(complex)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(0, w, 2, 8)
    lib = L.lib()
    ws = [L.RField(SHAPE, STEP, prec) for _ in range(3)]

    def parts(r):
        gr = _rfield(r, prec)
        pl, pc = L.CField(SHAPE, STEP, prec), L.CField(SHAPE, STEP, prec)
        pl.set(base)
        pc.set(base)
        lib.dft_ot_add_lennard_jones_potential(f.h, pl.h, gr.h, ws[0].h, ws[1].h)
        lib.dft_ot_add_local_correlation_potential(f.h, pc.h, gr.h, ws[0].h, ws[1].h, ws[2].h)
        out = pl.get() - base, pc.get() - base
        for x in (gr, pl, pc):
            x.free()
        return out

    lj, loc = parts(rho)
    want = o.potential2(rho, *[np.zeros(SHAPE)] * 3)
    assert relerr((lj + loc).real, want.real) < TOL[prec] * 10
    assert np.abs(lj.imag).max() < 1e-10 and np.abs(loc.imag).max() < 1e-10
    # linearity of the LJ part; the local part is not linear
    lj2, loc2 = parts(2 * rho)
    assert relerr(lj2.real, 2 * lj.real) < TOL[prec] * 10
    assert relerr(loc2.real, 2 * loc.real) > 0.1
    # uniform density
    r0 = f.p.rho0
    _, locu = parts(np.full(SHAPE, r0))
    c2, c3 = f.p.c2, f.p.c3
    closed = c2 / 2 * r0 ** 2 + c3 / 3 * r0 ** 3 + (c2 * r0 + c3 * r0 ** 2) * r0
    assert np.abs(locu.real / closed - 1).max() < TOL[prec] * 10
    for x in ws + [f, w]:
        x.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_energy_density_kc_bf_accumulate(oracle, prec):
    """_kc and _bf add their parts to the energy density, with the density given
    (and, for _bf, the velocity of the wave function, src/ot-energy.c:231)."""
    po = oracle
    o9, o8, o1, o0 = [po.OracleOT(m, SHAPE, STEP, 2, 8) for m in (9, 8, 1, 0)]
    psi = inputs.smooth_modulated(SHAPE, STEP, o9.p.rho0, 0.3, seed=31)
    rho = np.abs(inputs.smooth_modulated(SHAPE, STEP, o9.p.rho0, 0.25, seed=32)) ** 2  # not |psi|^2
    v = _velocity_inputs(po, psi)
    zero = [np.zeros(SHAPE)] * 3
    base = 1e-7 * inputs.splitmix64_uniform(psi.size, 4).reshape(SHAPE)
    want_kc = base + o1.energy_density_rv(rho, *zero) - o0.energy_density_rv(rho, *zero)
    want_bf = base + o8.energy_density_rv(rho, *v) - o0.energy_density_rv(rho, *v)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(9, w, 2, 8)
    gr, eps = _rfield(rho, prec), _rfield(base, prec)
    L.lib().dft_ot_energy_density_kc(f.h, eps.h, w.h, gr.h)
    s = np.abs(want_kc - base).max()
    assert np.abs(eps.get() - want_kc).max() < TOL[prec] * 50 * s
    eps.set(base)
    L.lib().dft_ot_energy_density_bf(f.h, eps.h, w.h, gr.h)
    s = np.abs(want_bf - base).max()
    assert np.abs(eps.get() - want_bf).max() < TOL[prec] * 50 * s
    for x in (gr, eps, f, w):
        x.free()


# ------------------------------------------------ velocity and gradients --
@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_velocity_gradient_flux(oracle, prec):
    po = oracle
    rho0 = 0.0218 * L.AUTOANG ** 3
    psi = inputs.smooth_modulated(SHAPE, STEP, rho0, 0.3, seed=41)
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    g = [L.RField(SHAPE, STEP, prec) for _ in range(3)]
    for eps in (L.DFT_EPS, 1e-9):
        L.lib().grid_wf_velocity(w.h, *[x.h for x in g], eps)
        want = po.velocity(psi, STEP, L.HELIUM_MASS, eps)
        for d in range(3):
            assert relerr(g[d].get(), want[d]) < TOL[prec] * 10
    # flux = rho v at eps = 0
    L.lib().grid_wf_probability_flux(w.h, *[x.h for x in g])
    want = po.velocity(psi, STEP, L.HELIUM_MASS, 0.0)
    for d in range(3):
        assert relerr(g[d].get(), want[d] * np.abs(psi) ** 2) < TOL[prec] * 10
    rho = np.abs(psi) ** 2
    src = _rfield(rho, prec)
    want = po.gradient(rho, STEP)
    for d, fn in enumerate(("rgrid_gradient_x", "rgrid_gradient_y", "rgrid_gradient_z")):
        getattr(L.lib(), fn)(src.h, g[d].h)
        assert relerr(g[d].get(), want[d]) < TOL[prec] * 10
    for x in g + [src, w]:
        x.free()


# --------------------------------------------------- clamps and c_bfpot --
@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_clamps_and_backflow_weight(oracle, prec):
    """Lowered max_veloc and max_bfpot and c_bfpot = 0.5 (src/ot.c:391-395, 911-917)
    in dft_ot_potential, dft_ot_potential2 and the fused step."""
    po = oracle
    o = po.OracleOT(9, SHAPE, STEP, 2, 8)
    psi = inputs.smooth_modulated(SHAPE, STEP, o.p.rho0, 0.3, seed=51) * \
        np.exp(1j * 2.5 * inputs.coords(SHAPE, STEP)[2][None, None, :] * 0.3)
    v = _velocity_inputs(po, psi)
    vmax = 0.5 * max(np.abs(a).max() for a in v)
    o_free = po.OracleOT(9, SHAPE, STEP, 2, 8)
    vbf = (o_free.potential(psi) - po.OracleOT(1, SHAPE, STEP, 2, 8).potential(psi))
    bfmax = 0.3 * np.abs(vbf.real).max()
    w = L.Wave(SHAPE, STEP, prec)
    w.set(psi)
    f = L.Functional(9, w, 2, 8)
    for obj in (o.p, f.p):
        obj.max_veloc, obj.max_bfpot, obj.c_bfpot = vmax, bfmax, 0.5
    # the clamps are active: they change the potential
    want = o.potential(psi)
    assert np.abs(want - o_free.potential(psi)).max() > 1e-3 * np.abs(vbf).max()
    pot = L.CField(SHAPE, STEP, prec)
    f.potential(pot)
    assert relerr(pot.get().real, want.real) < TOL[prec] * 10
    assert np.abs(pot.get().imag - want.imag).max() < TOL[prec] * 50 * np.abs(want.imag).max()
    # potential2 with the same inputs
    rho = np.abs(psi) ** 2
    gr, gv = _rfield(rho, prec), [_rfield(a, prec) for a in v]
    pot.set(np.zeros(SHAPE))
    L.lib().dft_ot_potential2(f.h, pot.h, gr.h, *[g.h for g in gv])
    want2 = o.potential2(rho, *v)
    assert relerr(pot.get().real, want2.real) < TOL[prec] * 10
    # the fused step against the oracle's single step
    dt = 5.0 / L.AUTOFS
    mu = o.bulk("dEdRho", o.p.rho0)
    L.lib().dft_ot_propagate(f.h, w.h, None, mu, L.Complex.of(dt))
    ref = psi.copy()
    o.single_step(ref, None, mu, dt)
    assert relerr(w.get(), ref) < (1e-11 if prec == L.GRID_FP64 else 3e-6)
    for x in gv + [gr, pot, f, w]:
        x.free()


# ------------------------------------------ backflow identities on the device --
def _gpu_bf_energy(f, w, rho_field, eps_field):
    eps_field.set(np.zeros(SHAPE))
    L.lib().dft_ot_energy_density_bf(f.h, eps_field.h, w.h, rho_field.h)
    return L.lib().rgrid_integral(eps_field.h)


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_backflow_identities_on_device(oracle, prec):
    po = oracle
    o = po.OracleOT(8, SHAPE, STEP, 2, 8)
    rho0 = o.p.rho0
    w = L.Wave(SHAPE, STEP, prec)
    f = L.Functional(8, w, 2, 8)
    pot = L.CField(SHAPE, STEP, prec)
    # Galilean invariance: a plane wave on uniform density has no backflow
    Lz = SHAPE[2] * STEP
    x, y, z = np.meshgrid(*[np.arange(n) * STEP for n in SHAPE], indexing="ij")
    kv = [2 * np.pi / (SHAPE[0] * STEP), -2 * np.pi / (SHAPE[1] * STEP), 2 * np.pi * 3 / Lz]
    psi = np.sqrt(rho0) * np.exp(1j * (kv[0] * x + kv[1] * y + kv[2] * z))
    w.set(psi)
    gv = [L.RField(SHAPE, STEP, prec) for _ in range(3)]
    L.lib().grid_wf_velocity(w.h, *[g.h for g in gv], L.DFT_EPS)
    gr = _rfield(np.abs(psi) ** 2, prec)
    L.lib().dft_ot_backflow_potential(f.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    v2 = sum((np.sin(k * STEP) / STEP / L.HELIUM_MASS) ** 2 for k in kv)
    u0 = np.abs(o.spectrum("bf_tf")[0, 0, 0].real) * STEP ** 3
    scale = 0.5 * o.p.mass * v2 * rho0 * u0
    tol = 1e-8 if prec == L.GRID_FP64 else 1e-2
    assert np.abs(pot.get()).max() < tol * scale
    eps = L.RField(SHAPE, STEP, prec)
    assert abs(_gpu_bf_energy(f, w, gr, eps)) < tol * scale * rho0 * np.prod(SHAPE) * STEP ** 3
    # Re V_BF = (1/h^3) dE_BF/drho at fixed v (E_BF is quadratic in rho at fixed v)
    psi = inputs.smooth_modulated(SHAPE, STEP, rho0, 0.3, seed=61)
    w.set(psi)
    rho = np.abs(psi) ** 2
    eta = 0.1 * rho0 * inputs.splitmix64_uniform(rho.size, 62).reshape(SHAPE)
    t = 0.1
    gp, gm = _rfield(rho + t * eta, prec), _rfield(rho - t * eta, prec)
    fd = (_gpu_bf_energy(f, w, gp, eps) - _gpu_bf_energy(f, w, gm, eps)) / (2 * t)
    gr.set(rho)
    L.lib().grid_wf_velocity(w.h, *[g.h for g in gv], L.DFT_EPS)
    pot.set(np.zeros(SHAPE))
    L.lib().dft_ot_backflow_potential(f.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    an = (pot.get().real * eta).sum() * STEP ** 3
    # FP32: the difference of two FP32 energy sums (1.1e-4 measured)
    assert abs(fd - an) < (1e-9 if prec == L.GRID_FP64 else 5e-4) * abs(an)
    # sum rho Im V_BF = 0 with eps = 0 (summation by parts)
    f.p.div_epsilon = 0.0
    L.lib().grid_wf_velocity(w.h, *[g.h for g in gv], L.DFT_EPS)
    pot.set(np.zeros(SHAPE))
    L.lib().dft_ot_backflow_potential(f.h, pot.h, gr.h, *[g.h for g in gv], *([None] * 6))
    vim = pot.get().imag
    assert abs((rho * vim).sum()) < (1e-10 if prec == L.GRID_FP64 else 1e-4) * np.abs(rho * vim).sum()
    f.p.div_epsilon = L.DFT_EPS
    for g in gv + [gr, gp, gm, eps, pot, f, w]:
        g.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_backflow_imaginary_response_on_device(oracle, prec):
    """Im V_BF of a small phase wave on uniform rho0: -(rho0 c / 2M) (U^(0) - U^(k)) sin^2(kh)/h^2 phi_k."""
    po = oracle
    n, step = 32, 0.4 / L.AUTOANG
    shape = (n, 8, 8)
    o = po.OracleOT(8, shape, step, 2, 8)
    rho0 = o.p.rho0
    w = L.Wave(shape, step, prec)
    f = L.Functional(8, w, 2, 8)
    pot = L.CField(shape, step, prec)
    x = np.arange(n) * step
    spec = lambda m: o.spectrum("bf_tf")[m, 0, 0].real * (-1.0) ** m * step ** 3
    for m in (1, 3, 6):
        k = 2 * np.pi * m / (n * step)
        a = 1e-4 if prec == L.GRID_FP64 else 1e-2
        phi = a * np.cos(k * x)[:, None, None] * np.ones(shape)
        w.set(np.sqrt(rho0) * np.exp(1j * phi))
        pot.set(np.zeros(shape))
        f.potential(pot)
        amp = 2 * np.fft.fft(pot.get().imag[:, 0, 0])[m].real / n / a
        c = rho0 / (rho0 + L.DFT_EPS)
        pred = -(rho0 * c / (2 * L.HELIUM_MASS)) * (spec(0) - spec(m)) * np.sin(k * step) ** 2 / step ** 2
        assert abs(amp - pred) < (1e-6 if prec == L.GRID_FP64 else 2e-3) * abs(pred), (m, amp, pred)
    pot.free(), f.free(), w.free()
