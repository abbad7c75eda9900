"""Grid files (.grd): write / read round trips through the C ABI, shape checks, and a
checkpoint-restart of the split-operator dynamics that continues bit for bit."""
import numpy as np
import pytest

import libdft_amd as L
from tests import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [L.GRID_FP64, L.GRID_FP32], ids=["fp64", "fp32"])
def test_grid_files_round_trip(tmp_path, prec):
    shape, step = (8, 16, 32), 0.7
    lib = L.lib()
    u = inputs.splitmix64_uniform(3 * int(np.prod(shape)), 3).reshape((3,) + shape)
    if prec == L.GRID_FP32:  # values the FP32 grid holds exactly
        u = u.astype(np.float32).astype(np.float64)
    c, r = L.CField(shape, step, prec), L.RField(shape, step, prec)
    c.set(u[0] + 1j * u[1])
    r.set(u[2])
    lib.cgrid_write_grid(str(tmp_path / "c").encode(), c.h)
    lib.rgrid_write_grid(str(tmp_path / "r").encode(), r.h)
    c2, r2 = L.CField(shape, step, prec), L.RField(shape, step, prec)
    lib.cgrid_read_grid(c2.h, str(tmp_path / "c.grd").encode())
    lib.rgrid_read_grid(r2.h, str(tmp_path / "r.grd").encode())
    assert np.array_equal(c2.get(), u[0] + 1j * u[1])
    assert np.array_equal(r2.get(), u[2])
    # a grid of another shape, or a file of the other kind, is left unchanged
    w = L.RField((8, 16, 16), step, prec)
    w.set(np.full((8, 16, 16), 7.0))
    lib.rgrid_read_grid(w.h, str(tmp_path / "r.grd").encode())
    lib.rgrid_read_grid(r2.h, str(tmp_path / "c.grd").encode())
    assert np.all(w.get() == 7.0)
    assert np.array_equal(r2.get(), u[2])
    for x in (c, r, c2, r2, w):
        x.free()


def test_checkpoint_restart_continues_exactly(tmp_path):
    shape, step = (16, 16, 32), 0.6 / L.AUTOANG
    lib = L.lib()
    w = L.Wave(shape, step, L.GRID_FP64)
    w.set(inputs.smooth_modulated(shape, step, 0.0218 * L.AUTOANG ** 3, 0.2))
    f = L.Functional(9, w)
    mu, dt = f.bulk("dEdRho", f.p.rho0), L.Complex.of(2.0 / L.AUTOFS)
    for _ in range(3):
        lib.dft_ot_propagate(f.h, w.h, None, mu, dt)
    lib.cgrid_write_grid(str(tmp_path / "psi").encode(), w.grid)
    for _ in range(3):
        lib.dft_ot_propagate(f.h, w.h, None, mu, dt)
    w2 = L.Wave(shape, step, L.GRID_FP64)
    lib.cgrid_read_grid(w2.grid, str(tmp_path / "psi.grd").encode())
    for _ in range(3):
        lib.dft_ot_propagate(f.h, w2.h, None, mu, dt)
    assert np.array_equal(w2.get(), w.get())
    f.free(), w.free(), w2.free()


@pytest.mark.parametrize("prec", [L.GRID_FP64, L.GRID_FP32], ids=["fp64", "fp32"])
def test_kernel_spectrum_files(tmp_path, prec):
    """A functional's kernel grids (real half spectra, rows at a padded pitch)
    round-trip through .grd files in place, at nz = 64 where the pitch (48)
    exceeds nz/2 + 1 (33); the potential is unchanged afterwards, so no
    neighbouring device memory was touched.  Files hold nx*ny*(nz/2+1) values,
    and a file of a half spectrum is refused by a real-field grid."""
    shape, step = (16, 16, 64), 0.6 / L.AUTOANG
    lib = L.lib()
    w = L.Wave(shape, step, prec)
    w.set(inputs.smooth_modulated(shape, step, 0.0218 * L.AUTOANG ** 3, 0.2))
    f = L.Functional(9, w, 2, 8)
    pot = L.CField(shape, step, prec)
    f.potential(pot)
    v0 = pot.get()
    half = (shape[0], shape[1], shape[2] // 2 + 1)
    lj0, sa0 = L.rgrid_get(f.p.lennard_jones, half), L.rgrid_get(f.p.spherical_avg, half)
    lib.rgrid_write_grid(str(tmp_path / "lj").encode(), f.p.lennard_jones)
    lib.rgrid_write_grid(str(tmp_path / "sa").encode(), f.p.spherical_avg)
    assert (tmp_path / "lj.grd").stat().st_size == 72 + 8 * int(np.prod(half))
    lib.rgrid_read_grid(f.p.spherical_avg, str(tmp_path / "lj.grd").encode())
    assert np.array_equal(L.rgrid_get(f.p.spherical_avg, half), lj0)
    lib.rgrid_read_grid(f.p.spherical_avg, str(tmp_path / "sa.grd").encode())
    assert np.array_equal(L.rgrid_get(f.p.spherical_avg, half), sa0)
    assert np.array_equal(L.rgrid_get(f.p.lennard_jones, half), lj0)
    r = L.RField(shape, step, prec)
    r.set(np.full(shape, 3.0))
    lib.rgrid_read_grid(r.h, str(tmp_path / "lj.grd").encode())
    assert np.all(r.get() == 3.0)
    pot.set(np.zeros(shape))
    f.potential(pot)
    assert np.array_equal(pot.get(), v0)
    for x in (r, pot, f, w):
        x.free()
