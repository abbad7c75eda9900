"""grid_wf_normalize_noalloc (the entry point that allocates nothing) against
grid_wf_normalize on the same wave functions, through the C ABI: two calls in
a row, the second on a smaller grid, in both device precisions.  FP64 results
agree to the last bits of the norm's summation order; FP32 to one rounding of
the scale factor."""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [L.GRID_FP64, L.GRID_FP32], ids=["fp64", "fp32"])
def test_normalize_noalloc_matches_normalize(prec):
    lib = L.lib()
    tol = 1e-14 if prec == L.GRID_FP64 else 2e-7
    for shape, target, seed in (((32, 32, 64), 7.5, 3), ((16, 8, 16), 0.25, 4)):
        step = 0.6
        psi = inputs.smooth_modulated(shape, step, 1.3, 0.4, seed) * (1 + 0.1j)
        a, b = L.Wave(shape, step, prec), L.Wave(shape, step, prec)
        a.set(psi), b.set(psi)
        a.h.contents.norm = target
        b.h.contents.norm = target
        lib.grid_wf_normalize(a.h)
        assert lib.grid_wf_normalize_noalloc(b.h) == 0
        ga, gb = a.get(), b.get()
        assert np.abs(gb - ga).max() <= tol * np.abs(ga).max(), shape
        assert abs(lib.grid_wf_norm(b.h) / target - 1) < (1e-13 if prec == L.GRID_FP64 else 1e-6)
        a.free(), b.free()
