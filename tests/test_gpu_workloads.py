"""Workload-scale GPU checks (64^3 bulk, 128^3 droplet, 256^3 vortex line, 512^3 cavity), against the FP64 oracle
where the oracle finishes in seconds and through size-independent properties
(norm and energy conservation, FP32 vs FP64 agreement, quantised circulation)
where it does not.
"""
import json
import os
import subprocess
import time

import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _log(msg):
    print(msg, flush=True)


def test_bulk64_imaginary_time_matches_oracle(oracle):
    """Plain OT 64^3, 200 predict-correct imaginary-time steps (dt = 10 fs) with
    renormalisation; final E/N and mu against the CPU oracle run of the same steps."""
    n, step = 64, 0.5 / L.AUTOANG
    shape = (n, n, n)
    psi0 = inputs.noisy_bulk(shape, 0.0218 * L.AUTOANG ** 3, 0.01)
    dt = -1j * 10.0 / L.AUTOFS
    o = oracle.OracleOT(0, shape, step, 4, 32)
    mu0 = o.bulk("chempot_pressurized", 0.0)
    t0 = time.time()
    psi = psi0.copy()
    n0 = oracle.norm(psi, step)
    for _ in range(200):
        o.pc_step(psi, None, mu0, dt)
        psi *= np.sqrt(n0 / oracle.norm(psi, step))
    ref = o.observables(psi)
    _log("oracle 200 steps: %.1f s, E/N = %.12f K" % (time.time() - t0, ref["energy"] / ref["norm"] * L.AUTOK))
    lib = L.lib()
    for prec, tol in ((L.GRID_FP64, 1e-10), (L.GRID_FP32, 1e-6)):
        w = L.Wave(shape, step, prec)
        wp = L.Wave(shape, step, prec)
        w.set(psi0)
        w.h.contents.norm = lib.grid_wf_norm(w.h)
        f = L.Functional(0, w)
        for _ in range(200):
            lib.dft_ot_propagate_pc(f.h, w.h, wp.h, None, mu0, L.Complex.of(dt))
            lib.grid_wf_normalize(w.h)
        E = lib.grid_wf_energy(w.h, None) + f.energy()
        N = lib.grid_wf_norm(w.h)
        pot = L.CField(shape, step, prec)
        f.potential(pot)
        mu = (lib.grid_wf_energy(w.h, None) + lib.grid_wf_potential_energy(w.h, pot.h)) / N
        e_ref, mu_ref = ref["energy"] / ref["norm"], ref["mu"]
        _log("prec %d: E/N %.12f K (rel %.2e), mu %.12f K (rel %.2e)" % (
            prec, E / N * L.AUTOK, abs(E / N / e_ref - 1), mu * L.AUTOK, abs(mu / mu_ref - 1)))
        assert abs(E / N / e_ref - 1) < tol
        assert abs(mu / mu_ref - 1) < tol * 10
        assert abs(N / ref["norm"] - 1) < tol
        for x in (pot, f, wp, w):
            x.free()


def _energy(lib, f, w, vext=None, rho=None):
    E = lib.grid_wf_energy(w.h, None) + f.energy()
    if vext is not None:
        lib.grid_wf_density(w.h, rho.h)
        E += lib.rgrid_integral_of_product(rho.h, vext.h)
    return E


def test_droplet128_fp32_tracks_fp64():
    """Full length (1000 predict-correct steps): full OT 128^3 droplet with the gaussian well,
    real time dt = 5 fs.  FP32 fields + FP64 reductions stay within 1e-6 of the FP64 path in
    energy and norm; both keep energy and norm to 5e-3 (backflow is not norm-conserving)."""
    n, step = 128, 0.4 / L.AUTOANG
    shape = (n, n, n)
    # both precisions start from the same (FP32-representable) state and potential
    psi0 = inputs.droplet(shape, step).astype(np.complex64).astype(complex)
    vext = inputs.gaussian_well(shape, step).astype(np.float32).astype(float)
    dt = L.Complex.of(5.0 / L.AUTOFS)
    lib = L.lib()
    res = {}
    for prec in (L.GRID_FP64, L.GRID_FP32):
        w = L.Wave(shape, step, prec)
        wp = L.Wave(shape, step, prec)
        w.set(psi0)
        f = L.Functional(9, w)
        ve = L.RField(shape, step, prec)
        ve.set(vext)
        rho = L.RField(shape, step, prec)
        mu0 = f.bulk("chempot_pressurized", 0.0)
        E0, N0 = _energy(lib, f, w, ve, rho), lib.grid_wf_norm(w.h)
        for _ in range(1000):
            lib.dft_ot_propagate_pc(f.h, w.h, wp.h, ve.h, mu0, dt)
        res[prec] = (_energy(lib, f, w, ve, rho), lib.grid_wf_norm(w.h), E0, N0)
        _log("prec %d: E0 %.10f K E %.10f K N0 %.10f N %.10f" % (prec, E0 * L.AUTOK, res[prec][0] * L.AUTOK, N0,
                                                               res[prec][1]))
        for x in (rho, ve, f, wp, w):
            x.free()
    E64, N64, E064, N064 = res[L.GRID_FP64]
    E32, N32, _, _ = res[L.GRID_FP32]
    assert abs(E32 / E64 - 1) < 1e-6
    assert abs(N32 / N64 - 1) < 1e-6
    # the backflow potential has an imaginary part (src/ot.c:840-991), so even
    # real-time dynamics change the norm and the energy (1.2e-3 and 1.5e-3 over
    # these 1000 steps in FP64)
    assert abs(N64 / N064 - 1) < 5e-3
    assert abs(E64 / E064 - 1) < 5e-3


def test_vortex256_single_step_matches_oracle(oracle):
    """Vortex line at 256^3 (h = 0.3 A, full OT): one fused single step on the
    FP64 device path against the oracle (kernels sampled once per cell on both sides)."""
    n, step = 256, 0.3 / L.AUTOANG
    shape = (n, n, n)
    o = oracle.OracleOT(9, shape, step, 1, 1)
    psi0 = inputs.vortex_line(shape, step, o.p.rho0)
    mu = o.bulk("dEdRho", o.p.rho0)
    dt = 2.0 / L.AUTOFS
    t0 = time.time()
    want = psi0.copy()
    o.single_step(want, None, mu, dt)
    _log("oracle 256^3 step: %.1f s" % (time.time() - t0))
    o.close()
    lib = L.lib()
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(psi0)
    f = L.Functional(9, w, 1, 1)
    lib.dft_ot_propagate(f.h, w.h, None, mu, L.Complex.of(dt))
    got = w.get()
    err = np.abs(got - want).max() / np.abs(want).max()
    _log("256^3 single step rel err %.3e" % err)
    assert err < 1e-10
    f.free(), w.free()


def test_vortex256_circulation_and_conservation():
    """Vortex line at 256^3 over 500 steps: circulation around the core on a 10 A loop stays
    one quantum 2 pi hbar / M; FP32 energy within 1e-6 of FP64 from the same (FP32-representable)
    initial state.  Also logged: how far the FP64 energy moves when only the initial state is
    rounded to FP32 (the sensitivity of these dynamics to one rounding of psi)."""
    n, step = 256, 0.3 / L.AUTOANG
    shape = (n, n, n)
    lib = L.lib()
    dt = L.Complex.of(2.0 / L.AUTOFS)
    res = {}
    for prec, rounded in ((L.GRID_FP64, False), (L.GRID_FP64, True), (L.GRID_FP32, True)):
        w = L.Wave(shape, step, prec)
        f = L.Functional(9, w)
        psi0 = inputs.vortex_line(shape, step, f.p.rho0)
        if rounded:
            psi0 = psi0.astype(np.complex64).astype(complex)
        w.set(psi0)
        mu = f.bulk("dEdRho", f.p.rho0)
        E0, N0 = _energy(lib, f, w), lib.grid_wf_norm(w.h)
        for _ in range(500):
            lib.dft_ot_propagate(f.h, w.h, None, mu, dt)
        psi = w.get()
        # winding of the phase on a circle of radius 10 A in the mid plane
        R = 10.0 / L.AUTOANG
        th = np.linspace(0, 2 * np.pi, 721)
        ix = np.rint(R * np.cos(th) / step).astype(int) + n // 2
        iy = np.rint(R * np.sin(th) / step).astype(int) + n // 2
        ph = np.angle(psi[ix, iy, n // 2])
        dph = np.angle(np.exp(1j * np.diff(ph)))
        circ = dph.sum() / L.HELIUM_MASS
        # the core: centroid of the density deficit within 3 A of the axis in the mid plane
        rho = np.abs(psi[:, :, n // 2]) ** 2
        xs = (np.arange(n) - n // 2) * step
        X, Y = np.meshgrid(xs, xs, indexing="ij")
        near = X ** 2 + Y ** 2 < (3.0 / L.AUTOANG) ** 2
        dfc = np.where(near, np.clip(f.p.rho0 - rho, 0.0, None), 0.0)
        core = (np.sum(dfc * X) / dfc.sum(), np.sum(dfc * Y) / dfc.sum())
        res[(prec, rounded)] = (_energy(lib, f, w), lib.grid_wf_norm(w.h), E0, N0, circ, core, rho)
        _log("prec %d rounded-start %d: E0 %.10f E %.10f N %.10f circulation %.6f (quantum %.6f)" % (
            prec, rounded, E0 * L.AUTOK, res[(prec, rounded)][0] * L.AUTOK, res[(prec, rounded)][1], circ,
            2 * np.pi / L.HELIUM_MASS))
        f.free(), w.free()
    quantum = 2 * np.pi / L.HELIUM_MASS
    for k in res:
        assert abs(res[k][4] / quantum - 1) < 1e-9
    E64, N64, E064, N064, _, c64, rho64 = res[(L.GRID_FP64, True)]
    E64x = res[(L.GRID_FP64, False)][0]
    E32, N32, _, _, _, c32, rho32 = res[(L.GRID_FP32, True)]
    dcore = np.hypot(c32[0] - c64[0], c32[1] - c64[1]) * L.AUTOANG
    drho = np.abs(rho32 - rho64).max() / np.abs(rho64).max()
    _log("FP32 vs FP64 after 500 steps: energy %.3e, norm %.3e, core position %.3e A (core at %.4f, %.4f A), "
         "mid-plane density %.3e; FP64 sensitivity to rounding psi0 once: %.3e" % (
             abs(E32 / E64 - 1), abs(N32 / N64 - 1), dcore, c64[0] * L.AUTOANG, c64[1] * L.AUTOANG, drho,
             abs(E64x / E64 - 1)))
    assert abs(E32 / E64 - 1) < 1e-6
    assert abs(N32 / N64 - 1) < 1e-6
    # the dynamics themselves: where the core is and the density around it
    assert dcore < 1e-3
    assert drho < 1e-4
    # the backflow potential's imaginary part changes the norm (1.2e-3 over 500 steps)
    assert abs(N64 / N064 - 1) < 5e-3


def test_dftrun_driver_runs():
    """The C driver (host C on the C ABI) runs the plain-OT 64^3 imaginary-time workload and reports JSON lines."""
    exe = os.path.join(ROOT, "libdft_amd", "dftrun")
    out = subprocess.run([exe, "bulk64", "--steps", "20", "--report", "10"], capture_output=True, text=True,
                         timeout=300, check=True).stdout
    rows = [json.loads(x) for x in out.splitlines() if x.startswith("{")]
    assert [r["step"] for r in rows] == [0, 10, 20]
    assert abs(rows[-1]["energy_per_atom_K"] + 7.15) < 0.02
