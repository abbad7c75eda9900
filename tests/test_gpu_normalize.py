"""The renormalising steps dft_ot_propagate_normalize / _pc_normalize against the
reference driver's sequence, the step followed by grid_wf_normalize
(examples/imag_time_rot/imag_time.c:198-208), in both device precisions, single
and predict-correct steps, at 32^3 and 128^3."""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu

PRECS = [L.GRID_FP64, L.GRID_FP32]
PNAME = {L.GRID_FP64: "fp64", L.GRID_FP32: "fp32"}


@pytest.mark.parametrize("prec", PRECS, ids=PNAME.get)
@pytest.mark.parametrize("n,model,pc", [(32, 0, False), (32, 9, True), (128, 9, False), (128, 1, True)],
                         ids=["32-plain", "32-full-pc", "128-full", "128-kc-pc"])
def test_normalizing_step_matches_step_then_normalize(prec, n, model, pc):
    shape, step = (n, n, n), 0.5 / L.AUTOANG
    lib = L.lib()
    rho0 = 0.0218360 * L.AUTOANG ** 3
    psi0 = inputs.noisy_bulk(shape, rho0, 0.01) * np.exp(0.2j * inputs.splitmix64_uniform(n ** 3, 5).reshape(shape))
    psi0 = psi0.astype(np.complex64).astype(complex)
    target = 0.93 * rho0 * (n * step) ** 3
    dt = L.Complex.of(-1j * 10.0 / L.AUTOFS)
    a, b = L.Wave(shape, step, prec), L.Wave(shape, step, prec)
    ap, bp = L.Wave(shape, step, prec), L.Wave(shape, step, prec)
    for w in (a, b):
        w.set(psi0)
        w.h.contents.norm = target
    f = L.Functional(model, a)
    mu = f.bulk("dEdRho", rho0)
    for _ in range(3):
        if pc:
            assert lib.dft_ot_propagate_pc_normalize(f.h, a.h, ap.h, None, mu, dt) == 0
            lib.dft_ot_propagate_pc(f.h, b.h, bp.h, None, mu, dt)
        else:
            assert lib.dft_ot_propagate_normalize(f.h, a.h, None, mu, dt) == 0
            lib.dft_ot_propagate(f.h, b.h, None, mu, dt)
        lib.grid_wf_normalize(b.h)
    ga, gb = a.get(), b.get()
    tol = 1e-12 if prec == L.GRID_FP64 else 2e-6
    assert np.abs(ga - gb).max() <= tol * np.abs(gb).max()
    assert abs(lib.grid_wf_norm(a.h) / target - 1) < (1e-12 if prec == L.GRID_FP64 else 1e-6)
    for x in (f, a, b, ap, bp):
        x.free()


def test_normalizing_step_rejects_bad_arguments():
    shape, step = (16, 16, 16), 0.5 / L.AUTOANG
    lib = L.lib()
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(np.full(shape, 0.1 + 0j))
    f = L.Functional(0, w)
    w.h.contents.norm = 0.0
    assert lib.dft_ot_propagate_normalize(f.h, w.h, None, 0.0, L.Complex.of(-1j)) == -1
    assert lib.dft_ot_propagate_pc_normalize(f.h, w.h, None, None, 0.0, L.Complex.of(-1j)) == -1
    f.free(), w.free()
