"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

FP64 device paths must agree with the oracle to FFT rounding (1e-10 relative
to the field's max).  FP32 device paths are held to 2e-5 relative (single-
precision transforms of fields with a large uniform background).
"""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu

TOL = {L.GRID_FP64: 1e-10, L.GRID_FP32: 2e-5}
PRECS = [L.GRID_FP64, L.GRID_FP32]
PNAME = {L.GRID_FP64: "fp64", L.GRID_FP32: "fp32"}


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _lib():
    L.lib()
    yield


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_free_plane_wave_phase(prec):
    shape, step = (8, 16, 32), 0.7
    x, y, z = np.meshgrid(*[np.arange(s) * step for s in shape], indexing="ij")
    kv = [2 * np.pi * 2 / (shape[0] * step), 2 * np.pi * -3 / (shape[1] * step), 2 * np.pi * 5 / (shape[2] * step)]
    psi = np.exp(1j * (kv[0] * x + kv[1] * y + kv[2] * z))
    w = L.Wave(shape, step, prec)
    pot = L.CField(shape, step, prec)
    w.set(psi)
    dt = 40.0 - 3.0j
    L.lib().grid_wf_propagate(w.h, pot.h, L.Complex.of(dt))
    out = w.get()
    k2 = sum(k * k for k in kv)
    want = psi * np.exp(-1j * k2 / (2 * L.HELIUM_MASS) * dt)
    assert relerr(out, want) < (1e-12 if prec == L.GRID_FP64 else 3e-6)
    w.free(), pot.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 32, 16), (64, 32, 128), (32, 32, 16), (64, 64, 32)])
def test_propagate_matches_oracle(oracle, prec, shape):
    """(32, 32, 16) and (64, 64, 32) take the small-grid block shapes (64-thread line and column blocks)."""
    step = 0.9
    psi = inputs.smooth_modulated(shape, step, 1.0, 0.4)
    u = inputs.splitmix64_uniform(2 * psi.size, 5).reshape((2,) + shape)
    V = 1e-3 * u[0] + 2e-4j * u[1]
    dt = 30.0 - 2.0j
    want = oracle.propagate(psi, V, dt, step)
    w = L.Wave(shape, step, prec)
    pot = L.CField(shape, step, prec)
    w.set(psi), pot.set(V)
    L.lib().grid_wf_propagate(w.h, pot.h, L.Complex.of(dt))
    assert relerr(w.get(), want) < (1e-11 if prec == L.GRID_FP64 else 3e-6)
    assert abs(L.lib().grid_wf_norm(w.h) / oracle.norm(want, step) - 1) < (1e-12 if prec == L.GRID_FP64 else 1e-6)
    ek = L.lib().grid_wf_energy(w.h, None)
    assert abs(ek / oracle.kinetic_energy(want, step) - 1) < (1e-11 if prec == L.GRID_FP64 else 1e-5)
    w.free(), pot.free()


def _gpu_potential(model, shape, step, prec, psi, min_sub=2, max_sub=8, kc_exact=False):
    w = L.Wave(shape, step, prec)
    w.set(psi)
    f = L.Functional(model, w, min_sub, max_sub)
    if kc_exact:
        f.set_kc_exact(True)
    pot = L.CField(shape, step, prec)
    f.potential(pot)
    V = pot.get()
    rho = f.density()
    eps = L.RField(shape, step, prec)
    f.energy_density(eps)
    e = eps.get()
    E = f.energy()
    pot.free(), eps.free(), f.free(), w.free()
    return V, rho, e, E


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [0, 1, 8, 9], ids=["plain", "kc", "bf", "full"])
def test_potential_and_energy_match_oracle(oracle, prec, model):
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    o = oracle.OracleOT(model, shape, step, 2, 8)
    psi = inputs.smooth_modulated(shape, step, o.p.rho0, 0.3)
    V, rho, e, E = _gpu_potential(model, shape, step, prec, psi)
    Vo = o.potential(psi)
    eo = o.energy_density(psi)
    assert relerr(rho, np.abs(psi) ** 2) < TOL[prec]
    assert relerr(V.real, Vo.real) < TOL[prec]
    if model & 8:
        assert relerr(V.imag, Vo.imag) < TOL[prec] * 10
    else:
        assert np.abs(V.imag).max() == 0.0
    assert relerr(e, eo) < TOL[prec]
    Eo = eo.sum() * step ** 3
    assert abs(E - Eo) < TOL[prec] * np.abs(eo).max() * step ** 3 * eo.size


def test_potential_kc_exact_coefficient(oracle):
    shape, step = (16, 16, 16), 0.6 / L.AUTOANG
    o = oracle.OracleOT(1, shape, step, 2, 8)
    o.set_kc_exact(True)
    psi = inputs.smooth_modulated(shape, step, o.p.rho0, 0.3)
    V, _, _, _ = _gpu_potential(1, shape, step, L.GRID_FP64, psi, kc_exact=True)
    assert relerr(V.real, o.potential(psi).real) < 1e-10


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_potential_accumulates(oracle, prec):
    """dft_ot_potential adds to the potential grid (src/ot.c:296)."""
    shape, step = (16, 16, 16), 0.6 / L.AUTOANG
    o = oracle.OracleOT(9, shape, step, 2, 8)
    psi = inputs.smooth_modulated(shape, step, o.p.rho0, 0.3)
    base = (inputs.splitmix64_uniform(psi.size, 9).reshape(shape) * 1e-5 + 3e-6j).astype(complex)
    w = L.Wave(shape, step, prec)
    w.set(psi)
    f = L.Functional(9, w, 2, 8)
    pot = L.CField(shape, step, prec)
    pot.set(base)
    f.potential(pot)
    want = o.potential(psi, pot=base)
    assert relerr(pot.get(), want) < TOL[prec] * 10
    pot.free(), f.free(), w.free()


def test_kernel_spectra_match_oracle(oracle):
    """GPU kernel build (adaptive map + FFT in FP64, wrapped sampling) vs the oracle's
    centred sampling with the (-1)^(i+j+k) convolution sign."""
    shape, step = (16, 16, 16), 0.5 / L.AUTOANG
    o = oracle.OracleOT(9, shape, step, 4, 32)
    w = L.Wave(shape, step, L.GRID_FP64)
    f = L.Functional(9, w, 4, 32)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    sgn = (-1.0) ** (i + j + k)
    P = shape[2] // 2 + 1
    for name, field in (("lj_tf", "lennard_jones"), ("sa_tf", "spherical_avg"), ("g_tf", "gaussian_tf"),
                        ("bf_tf", "backflow_pot")):
        want = (o.spectrum(name) * sgn)[:, :, :P]
        got = np.empty(shape[0] * shape[1] * P)
        L.lib().rgrid_copy_to_host(getattr(f.p, field), got.ctypes.data_as(L.C.c_void_p))
        got = got.reshape(shape[0], shape[1], P)
        assert np.abs(want.imag).max() < 1e-12 * np.abs(want).max()
        assert relerr(got, want.real) < 1e-12, name
    f.free(), w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_uniform_invariant_full_128(prec):
    """psi = sqrt(rho0): V = dE/drho(rho0), eps = E_bulk(rho0) at 128^3 full OT."""
    shape, step = (128, 128, 128), 0.4 / L.AUTOANG
    w = L.Wave(shape, step, prec)
    f = L.Functional(9, w)
    rho0 = f.p.rho0
    L.lib().grid_wf_constant(w.h, L.Complex.of(np.sqrt(rho0)))
    pot = L.CField(shape, step, prec)
    f.potential(pot)
    V = pot.get()
    mu = f.bulk("dEdRho", rho0)
    tol = 1e-12 if prec == L.GRID_FP64 else 1e-5
    assert np.abs(V - mu).max() < tol * abs(mu)
    E = f.energy()
    Eb = f.bulk("energy", rho0) * np.prod(shape) * step ** 3
    assert abs(E / Eb - 1) < tol
    pot.free(), f.free(), w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_predict_correct_step_matches_oracle(oracle, prec):
    """One imaginary-time and one real-time predict-correct step with V_ext (CS2)."""
    shape, step = (16, 16, 32), 0.5 / L.AUTOANG
    model = 9
    o = oracle.OracleOT(model, shape, step, 2, 8)
    psi0 = inputs.smooth_modulated(shape, step, o.p.rho0, 0.2)
    vext = inputs.gaussian_well(shape, step)
    mu0 = o.bulk("dEdRho", o.p.rho0)
    lib = L.lib()
    w = L.Wave(shape, step, prec)
    wp = L.Wave(shape, step, prec)
    f = L.Functional(model, w, 2, 8)
    ve = L.RField(shape, step, prec)
    ve.set(vext)
    pot = L.CField(shape, step, prec)
    for dt in (-1j * 10.0 / L.AUTOFS, 5.0 / L.AUTOFS):
        want = psi0.copy()
        o.pc_step(want, vext, mu0, dt)
        w.set(psi0)
        z = L.Complex.of(dt)
        lib.grid_real_to_complex_re(pot.h, ve.h)
        lib.dft_ot_potential(f.h, pot.h, w.h)
        lib.cgrid_add(pot.h, L.Complex.of(-mu0))
        lib.grid_wf_propagate_predict(w.h, wp.h, pot.h, z)
        lib.grid_add_real_to_complex_re(pot.h, ve.h)
        lib.dft_ot_potential(f.h, pot.h, wp.h)
        lib.cgrid_add(pot.h, L.Complex.of(-mu0))
        lib.cgrid_multiply(pot.h, L.Complex.of(0.5))
        lib.grid_wf_propagate_correct(w.h, pot.h, z)
        assert relerr(w.get(), want) < (1e-10 if prec == L.GRID_FP64 else 2e-5)
    for x in (pot, ve, f, wp, w):
        x.free()


# ------------------------------------------------------- backward of E[rho] --
def _gpu_gradient(model, shape, step, prec, rho, min_sub=2, max_sub=8):
    w = L.Wave(shape, step, prec)
    f = L.Functional(model, w, min_sub, max_sub)
    r = L.RField(shape, step, prec)
    g = L.RField(shape, step, prec)
    r.set(rho)
    L.lib().dft_ot_energy_gradient(f.h, g.h, r.h)
    out = g.get()
    E = L.lib().dft_ot_energy_rho(f.h, r.h)
    for x in (g, r, f, w):
        x.free()
    return out, E


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [0, 1], ids=["plain", "kc"])
def test_energy_gradient_matches_oracle(oracle, prec, model):
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    o = oracle.OracleOT(model, shape, step, 2, 8)
    rho = np.abs(inputs.smooth_modulated(shape, step, o.p.rho0, 0.3)) ** 2
    g, E = _gpu_gradient(model, shape, step, prec, rho)
    assert relerr(g, o.energy_gradient(rho)) < TOL[prec]
    Eo = o.energy_rho(rho)
    assert abs(E - Eo) < 10 * TOL[prec] * abs(Eo)


def test_plain_gradient_equals_potential(oracle):
    """For PLAIN the backward and dft_ot_potential coincide (src/ot.c:524-586 is dE/drho)."""
    shape, step = (16, 16, 16), 0.6 / L.AUTOANG
    o = oracle.OracleOT(0, shape, step, 2, 8)
    psi = inputs.smooth_modulated(shape, step, o.p.rho0, 0.3)
    g, _ = _gpu_gradient(0, shape, step, L.GRID_FP64, np.abs(psi) ** 2)
    V, _, _, _ = _gpu_potential(0, shape, step, L.GRID_FP64, np.abs(psi).astype(complex))
    assert relerr(g, V.real) < 1e-12


def test_energy_gradient_fd_128(oracle):
    """Central FD of E[rho] (FP64 device) on 20 seeded perturbations of amplitude 1e-4 rho0 at 128^3,
    against the FP64 and FP32 backward; <= 1e-4 relative."""
    n, step = 128, 0.4 / L.AUTOANG
    shape = (n, n, n)
    lib = L.lib()
    w64 = L.Wave(shape, step, L.GRID_FP64)
    f64 = L.Functional(9, w64)
    rho0 = f64.p.rho0
    rho = np.abs(inputs.droplet(shape, step)) ** 2 + 0.05 * rho0
    r = L.RField(shape, step, L.GRID_FP64)
    g = L.RField(shape, step, L.GRID_FP64)
    r.set(rho)
    lib.dft_ot_energy_gradient(f64.h, g.h, r.h)
    grad64 = g.get()
    w32 = L.Wave(shape, step, L.GRID_FP32)
    f32 = L.Functional(9, w32)
    r32 = L.RField(shape, step, L.GRID_FP32)
    g32 = L.RField(shape, step, L.GRID_FP32)
    r32.set(rho)
    lib.dft_ot_energy_gradient(f32.h, g32.h, r32.h)
    grad32 = g32.get()
    worst64 = worst32 = 0.0
    for j in range(20):
        eta = 1e-4 * rho0 * inputs.splitmix64_uniform(rho.size, inputs.SEED + j).reshape(shape)
        r.set(rho + eta)
        ep = lib.dft_ot_energy_rho(f64.h, r.h)
        r.set(rho - eta)
        em = lib.dft_ot_energy_rho(f64.h, r.h)
        fd = (ep - em) / 2.0
        worst64 = max(worst64, abs(fd - (grad64 * eta).sum() * step ** 3) / abs(fd))
        worst32 = max(worst32, abs(fd - (grad32 * eta).sum() * step ** 3) / abs(fd))
    print("fd rel err fp64 %.2e fp32 %.2e" % (worst64, worst32))
    assert worst64 < 1e-6
    assert worst32 < 1e-4
    for x in (g, r, g32, r32, f64, f32, w64, w32):
        x.free()


# ---------------------------------------------------------- fused stepping --
@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("model", [0, 1, 8, 9], ids=["plain", "kc", "bf", "full"])
@pytest.mark.parametrize("shape", [(16, 32, 32), (32, 32, 16)], ids=["16x32x32", "32x32x16"])
def test_fused_single_step_matches_oracle(oracle, prec, model, shape):
    """dft_ot_propagate == zero/potential/-mu/propagate of the reference single-step driver
    (32 x 32 x 16: the small-grid block shapes)."""
    step = 0.5 / L.AUTOANG
    o = oracle.OracleOT(model, shape, step, 2, 8)
    psi0 = inputs.smooth_modulated(shape, step, o.p.rho0, 0.2)
    vext = inputs.gaussian_well(shape, step)
    mu0 = o.bulk("dEdRho", o.p.rho0)
    lib = L.lib()
    w = L.Wave(shape, step, prec)
    f = L.Functional(model, w, 2, 8)
    ve = L.RField(shape, step, prec)
    ve.set(vext)
    for dt in (-1j * 10.0 / L.AUTOFS, 5.0 / L.AUTOFS):
        want = psi0.copy()
        o.single_step(want, vext, mu0, dt)
        w.set(psi0)
        lib.dft_ot_propagate(f.h, w.h, ve.h, mu0, L.Complex.of(dt))
        assert relerr(w.get(), want) < (1e-10 if prec == L.GRID_FP64 else 2e-5)
        assert relerr(f.density(), np.abs(psi0) ** 2) < TOL[prec]
    for x in (ve, f, w):
        x.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("shape", [(8, 8, 512), (8, 512, 8), (512, 8, 8)], ids=["z512", "y512", "x512"])
def test_fused_single_step_512_lines(oracle, prec, shape):
    """Full OT single step with one axis of 512 points: the z-pass line of 512
    points spans two wavefronts (five radix stages), the column passes take the
    512-point plans."""
    step = 0.4 / L.AUTOANG
    o = oracle.OracleOT(9, shape, step, 2, 8)
    psi0 = inputs.smooth_modulated(shape, step, o.p.rho0, 0.2)
    mu0 = o.bulk("dEdRho", o.p.rho0)
    lib = L.lib()
    w = L.Wave(shape, step, prec)
    f = L.Functional(9, w, 2, 8)
    dt = 2.0 / L.AUTOFS
    want = psi0.copy()
    o.single_step(want, None, mu0, dt)
    w.set(psi0)
    lib.dft_ot_propagate(f.h, w.h, None, mu0, L.Complex.of(dt))
    assert relerr(w.get(), want) < (1e-10 if prec == L.GRID_FP64 else 2e-5)
    assert relerr(f.density(), np.abs(psi0) ** 2) < TOL[prec]
    f.free(), w.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_fused_pc_step_matches_oracle(oracle, prec):
    shape, step = (32, 16, 32), 0.5 / L.AUTOANG
    o = oracle.OracleOT(9, shape, step, 2, 8)
    psi0 = inputs.smooth_modulated(shape, step, o.p.rho0, 0.2)
    vext = inputs.gaussian_well(shape, step)
    mu0 = o.bulk("dEdRho", o.p.rho0)
    lib = L.lib()
    w = L.Wave(shape, step, prec)
    wp = L.Wave(shape, step, prec)
    f = L.Functional(9, w, 2, 8)
    ve = L.RField(shape, step, prec)
    ve.set(vext)
    for dt in (-1j * 10.0 / L.AUTOFS, 5.0 / L.AUTOFS):
        want = psi0.copy()
        o.pc_step(want, vext, mu0, dt)
        w.set(psi0)
        lib.dft_ot_propagate_pc(f.h, w.h, wp.h, ve.h, mu0, L.Complex.of(dt))
        assert relerr(w.get(), want) < (1e-10 if prec == L.GRID_FP64 else 2e-5)
    for x in (ve, f, wp, w):
        x.free()


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
def test_fused_steps_large_grid_consistent(prec):
    """256 x 128 x 64 (non-cubic, exercises every tile shape): fused step == grid API step."""
    shape, step = (64, 128, 256), 0.4 / L.AUTOANG
    lib = L.lib()
    w1 = L.Wave(shape, step, prec)
    w2 = L.Wave(shape, step, prec)
    f = L.Functional(9, w1)
    psi0 = inputs.vortex_line(shape, step, f.p.rho0)
    w1.set(psi0), w2.set(psi0)
    pot = L.CField(shape, step, prec)
    mu = f.bulk("dEdRho", f.p.rho0)
    dt = L.Complex.of(2.0 / L.AUTOFS)
    lib.dft_ot_propagate(f.h, w1.h, None, mu, dt)
    lib.cgrid_zero(pot.h)
    lib.dft_ot_potential(f.h, pot.h, w2.h)
    lib.cgrid_add(pot.h, L.Complex.of(-mu))
    lib.grid_wf_propagate(w2.h, pot.h, dt)
    assert relerr(w1.get(), w2.get()) < (1e-13 if prec == L.GRID_FP64 else 1e-6)
    for x in (pot, f, w1, w2):
        x.free()
