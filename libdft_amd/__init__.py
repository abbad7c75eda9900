"""Python mirror of the libdft_mi355x C ABI (ctypes).

The product is the C-ABI shared library ``libdft_amd/libdft_mi355x.so``
(hand-written HIP kernels for gfx950).  This module only binds it for
tests and benchmarks.  There is no fallback: if the library is missing,
``lib()`` raises.

Names and argument meanings follow the reference C API
(src/ot.c, src/ot-energy.c, src/helium-ot-bulk.c) and the libgrid subset
that its drivers use.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LIBDFT_MI355X_SO names another build of the library (e.g. one compiled with other tuning macros)
LIB_PATH = os.environ.get("LIBDFT_MI355X_SO") or os.path.join(_HERE, "libdft_mi355x.so")
_LIB = None

GRID_FP32, GRID_FP64 = 1, 2
DFT_OT_PLAIN, DFT_OT_KC, DFT_OT_HD, DFT_OT_HD2, DFT_OT_BACKFLOW, DFT_OT_T0MK = 0, 1, 2, 4, 8, 16
WF_PERIODIC_BOUNDARY, WF_2ND_ORDER_FFT = 1, 2
DFT_EPS = 1e-4

AUTOANG = 0.529177210903
AUTOK = 315775.02480407
AUTOFS = 0.024188843265857
AUTOAMU = 1.0 / 1822.888486209
AUTOMPS = 2187691.26364
AUTOBAR = 2.9421015697e8
HELIUM_MASS = 4.002602 / AUTOAMU


class Complex(C.Structure):
    """By-value REAL complex (two doubles; same SysV classification as double _Complex)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


class RGrid(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64), ("step", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("dev", C.c_void_p),
                ("dev_elems", C.c_int64), ("prec", C.c_int), ("flag", C.c_int), ("spectral", C.c_int),
                ("id", C.c_char * 32), ("value_outside", C.c_void_p), ("outside_params_ptr", C.c_void_p)]


class CGrid(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64), ("step", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("dev", C.c_void_p),
                ("dev_elems", C.c_int64), ("prec", C.c_int), ("flag", C.c_int), ("kx0", C.c_double),
                ("ky0", C.c_double), ("kz0", C.c_double), ("id", C.c_char * 32), ("value_outside", C.c_void_p),
                ("outside_params_ptr", C.c_void_p)]


class WF(C.Structure):
    _fields_ = [("grid", C.POINTER(CGrid)), ("mass", C.c_double), ("norm", C.c_double),
                ("boundary", C.c_char), ("propagator", C.c_char), ("cworkspace", C.POINTER(CGrid))]


_RG = C.POINTER(RGrid)


class OTF(C.Structure):
    _fields_ = ([("model", C.c_int64)] +
                [(n, C.c_double) for n in "b c2 c2_exp c3 c3_exp c4 rho_0s alpha_s l_g mass rho0 temp".split()] +
                [("lj_" + n, C.c_double) for n in "h sigma epsilon cval".split()] +
                [("bf_" + n, C.c_double) for n in "g11 g12 g21 g22 a1 a2".split()] +
                [(n, _RG) for n in ("lennard_jones spherical_avg gaussian_tf gaussian_x_tf gaussian_y_tf "
                                    "gaussian_z_tf backflow_pot").split()] +
                [(n, C.c_double) for n in "beta rhom C mu0 xi rhobf div_epsilon max_veloc max_bfpot c_bfpot".split()] +
                [("workspace%d" % i, _RG) for i in range(1, 10)] +
                [("density", _RG), ("kc_coef23", C.c_double), ("device", C.c_void_p)])


_SIGS = {
    "grid_set_precision": (None, ["i"]), "grid_get_precision": ("i", []),
    "grid_synchronize": (None, []), "grid_get_stream": ("p", []),
    "rgrid_alloc": ("rg", ["i64", "i64", "i64", "d", "p", "p", "s"]),
    "cgrid_alloc": ("cg", ["i64", "i64", "i64", "d", "p", "p", "s"]),
    "rgrid_free": (None, ["rg"]), "cgrid_free": (None, ["cg"]),
    "grid_wf_alloc": ("wf", ["i64", "i64", "i64", "d", "d", "c", "c", "s"]),
    "grid_wf_clone": ("wf", ["wf", "s"]), "grid_wf_free": (None, ["wf"]),
    "rgrid_copy_from_host": (None, ["rg", "p"]), "rgrid_copy_to_host": (None, ["rg", "p"]),
    "cgrid_copy_from_host": (None, ["cg", "p"]), "cgrid_copy_to_host": (None, ["cg", "p"]),
    "rgrid_zero": (None, ["rg"]), "cgrid_zero": (None, ["cg"]),
    "rgrid_constant": (None, ["rg", "d"]), "cgrid_constant": (None, ["cg", "z"]),
    "rgrid_copy": (None, ["rg", "rg"]), "cgrid_copy": (None, ["cg", "cg"]),
    "cgrid_add": (None, ["cg", "z"]), "cgrid_multiply": (None, ["cg", "z"]),
    "rgrid_add": (None, ["rg", "d"]), "rgrid_multiply": (None, ["rg", "d"]),
    "rgrid_add_scaled_product": (None, ["rg", "d", "rg", "rg"]),
    "grid_real_to_complex_re": (None, ["cg", "rg"]), "grid_add_real_to_complex_re": (None, ["cg", "rg"]),
    "grid_add_real_to_complex_im": (None, ["cg", "rg"]),
    "grid_wf_density": (None, ["wf", "rg"]), "grid_wf_velocity": (None, ["wf", "rg", "rg", "rg", "d"]),
    "rgrid_gradient_x": (None, ["rg", "rg"]), "rgrid_gradient_y": (None, ["rg", "rg"]),
    "rgrid_gradient_z": (None, ["rg", "rg"]),
    "rgrid_integral": ("d", ["rg"]), "rgrid_integral_of_product": ("d", ["rg", "rg"]),
    "grid_wf_potential_energy": ("d", ["wf", "cg"]), "grid_wf_norm": ("d", ["wf"]), "grid_wf_normalize": (None, ["wf"]),
    "grid_wf_energy": ("d", ["wf", "cg"]),
    "grid_wf_normalize_noalloc": ("i", ["wf"]),
    "grid_wf_propagate": (None, ["wf", "cg", "z"]),
    "grid_wf_propagate_predict": (None, ["wf", "wf", "cg", "z"]),
    "grid_wf_propagate_correct": (None, ["wf", "cg", "z"]),
    "dft_ot_alloc": ("otf", ["i64", "wf", "i64", "i64"]), "dft_ot_free": (None, ["otf"]),
    "dft_ot_init_params": (None, ["otf", "i64"]),
    "dft_ot_potential": (None, ["otf", "cg", "wf"]),
    "dft_ot_potential2": (None, ["otf", "cg", "rg", "rg", "rg", "rg"]),
    "dft_ot_backflow_potential": (None, ["otf", "cg", "rg", "rg", "rg", "rg"] + ["rg"] * 6),
    "dft_ot_add_lennard_jones_potential": (None, ["otf", "cg", "rg", "rg", "rg"]),
    "dft_ot_add_local_correlation_potential": (None, ["otf", "cg", "rg", "rg", "rg", "rg"]),
    "dft_ot_energy_density": (None, ["otf", "rg", "wf"]),
    "dft_ot_energy_density_kc": (None, ["otf", "rg", "wf", "rg"]),
    "dft_ot_energy_density_bf": (None, ["otf", "rg", "wf", "rg"]),
    "dft_ot_set_kc_mode": (None, ["otf", "i"]), "dft_ot_energy": ("d", ["otf", "wf"]),
    "dft_ot_energy_gradient": (None, ["otf", "rg", "rg"]), "dft_ot_energy_rho": ("d", ["otf", "rg"]),
    "dft_profile_enable": (None, ["i"]), "dft_profile_reset": (None, []), "dft_profile_collect": ("i", []),
    "grid_bench_copy": (None, ["rg", "rg"]),
    "dft_ot_propagate": (None, ["otf", "wf", "rg", "d", "z"]),
    "dft_ot_propagate_pc": (None, ["otf", "wf", "wf", "rg", "d", "z"]),
    "dft_ot_propagate_normalize": ("i", ["otf", "wf", "rg", "d", "z"]),
    "dft_ot_propagate_pc_normalize": ("i", ["otf", "wf", "wf", "rg", "d", "z"]),
    "dft_ot_potential_vjp": ("i", ["otf", "rg", "rg", "wf"]),
    "dft_ot_potential_vjp_psi": ("i", ["otf", "cg", "cg", "wf"]),
    "rgrid_value_at_index": ("d", ["rg", "i64", "i64", "i64"]),
    "cgrid_value_at_index": ("z", ["cg", "i64", "i64", "i64"]),
    "cgrid_set_momentum": (None, ["cg", "d", "d", "d"]),
    "cgrid_fd_laplace": (None, ["cg", "cg"]),
    "cgrid_conjugate_product": (None, ["cg", "cg", "cg"]),
    "cgrid_abs_power": (None, ["cg", "cg", "d"]),
    "rgrid_integral_region": ("d", ["rg", "d", "d", "d", "d", "d", "d"]),
    "cgrid_integral_region": ("z", ["cg", "d", "d", "d", "d", "d", "d"]),
    "cgrid_integral": ("z", ["cg"]),
    "cgrid_integral_of_conjugate_product": ("z", ["cg", "cg"]),
    "rgrid_abs_power": (None, ["rg", "rg", "d"]),
    "rgrid_abs_rot": (None, ["rg", "rg", "rg", "rg"]),
    "grid_wf_probability_flux": (None, ["wf", "rg", "rg", "rg"]),
    "grid_wf_probability_flux_x": (None, ["wf", "rg"]), "grid_wf_probability_flux_y": (None, ["wf", "rg"]),
    "grid_wf_probability_flux_z": (None, ["wf", "rg"]),
    "rgrid_power": (None, ["rg", "rg", "d"]), "rgrid_product": (None, ["rg", "rg", "rg"]),
    "rgrid_division_eps": (None, ["rg", "rg", "rg", "d"]), "rgrid_max": ("d", ["rg"]), "rgrid_min": ("d", ["rg"]),
    "dft_ot_adjoint_well": ("i", ["otf", "wf", "d", "d", "d", "d", "d", "d", "z", "i64", "i64", "p", "p"]),
    "dft_ot_adjoint_cavity": ("i", ["otf", "wf", "d", "d", "d", "d", "z", "i64", "i64", "p", "p"]),
    "dft_ot_dispersion": ("d", ["wf", "otf", "d", "p", "d", "b", "b"]),
    "dft_set_verbose": (None, ["i"]),
    "rgrid_write_grid": (None, ["s", "rg"]),
    "rgrid_read_grid": (None, ["rg", "s"]),
    "cgrid_write_grid": (None, ["s", "cg"]),
    "cgrid_read_grid": (None, ["cg", "s"]),
}
for _n in ("energy", "dEdRho", "pressure", "dPdRho", "density_pressurized", "chempot_pressurized",
           "compressibility", "sound_speed", "chempot3"):
    _SIGS["dft_ot_bulk_" + _n] = ("d", ["otf", "d"])
for _n in ("density", "chempot", "chempot2"):
    _SIGS["dft_ot_bulk_" + _n] = ("d", ["otf"])
for _n in ("dispersion", "istatic"):
    _SIGS["dft_ot_bulk_" + _n] = ("d", ["otf", "p", "d"])
_SIGS["dft_common_lwl3"] = ("d", ["d", "d"])
for _n in ("dft_common_cgaussian", "dft_common_planewave", "dft_initial_bubble") + tuple(
        "dft_initial_vortex_%s_n%d" % (a, q) for a in "xyz" for q in (1, 2)):
    _SIGS[_n] = ("z", ["p", "d", "d", "d"])
for _a in "xyz":
    _SIGS["dft_common_vortex_" + _a] = ("d", ["p", "d", "d", "d"])
_SIGS["dft_common_gaussian"] = ("d", ["p", "d", "d", "d"])
for _n in ("g12", "g32", "g52", "z"):
    _SIGS["dft_common_fit_" + _n] = ("d", ["d"])
for _n in ("bose_idealgas_energy", "bose_idealgas_dEdRho", "classical_idealgas_energy", "classical_idealgas_dEdRho"):
    _SIGS["dft_common_" + _n] = ("d", ["d", "p"])

_T = {"i": C.c_int, "i64": C.c_int64, "d": C.c_double, "p": C.c_void_p, "s": C.c_char_p, "c": C.c_char, "b": C.c_byte,
      "z": Complex, "rg": _RG, "cg": C.POINTER(CGrid), "wf": C.POINTER(WF), "otf": C.POINTER(OTF)}


def lib():
    """Load the HIP library.  Raises if it has not been built (no CPU fallback exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdft_mi355x.so is not built; run `make -C libdft_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = None if res is None else _T[res]
            f.argtypes = [_T[a] for a in args]
        _LIB = L
    return _LIB


def exported_symbols():
    return sorted(_SIGS)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Wave:
    """A wave function (libgrid wf) of the given shape, step (bohr) and precision."""

    def __init__(self, shape, step, prec=GRID_FP64, mass=HELIUM_MASS, name=b"gwf"):
        L = lib()
        L.grid_set_precision(prec)
        self.h = L.grid_wf_alloc(*shape, step, mass, WF_PERIODIC_BOUNDARY, WF_2ND_ORDER_FFT, name)
        if not self.h:
            raise RuntimeError("grid_wf_alloc failed")
        self.shape, self.step, self.prec = tuple(shape), step, prec

    @property
    def grid(self):
        return self.h.contents.grid

    def set(self, psi):
        psi = np.ascontiguousarray(psi, np.complex128)
        assert psi.shape == self.shape
        lib().cgrid_copy_from_host(self.grid, _ptr(psi))

    def get(self):
        out = np.empty(self.shape, np.complex128)
        lib().cgrid_copy_to_host(self.grid, _ptr(out))
        return out

    def free(self):
        if self.h:
            lib().grid_wf_free(self.h)
            self.h = None


class CField:
    def __init__(self, shape, step, prec=GRID_FP64, name=b"cgrid"):
        L = lib()
        L.grid_set_precision(prec)
        self.h = L.cgrid_alloc(*shape, step, None, None, name)
        if not self.h:
            raise RuntimeError("cgrid_alloc failed")
        self.shape = tuple(shape)

    def set(self, a):
        a = np.ascontiguousarray(a, np.complex128)
        lib().cgrid_copy_from_host(self.h, _ptr(a))

    def get(self):
        out = np.empty(self.shape, np.complex128)
        lib().cgrid_copy_to_host(self.h, _ptr(out))
        return out

    def free(self):
        if self.h:
            lib().cgrid_free(self.h)
            self.h = None


class RField:
    def __init__(self, shape, step, prec=GRID_FP64, name=b"rgrid"):
        L = lib()
        L.grid_set_precision(prec)
        self.h = L.rgrid_alloc(*shape, step, None, None, name)
        if not self.h:
            raise RuntimeError("rgrid_alloc failed")
        self.shape = tuple(shape)

    def set(self, a):
        a = np.ascontiguousarray(a, np.float64)
        lib().rgrid_copy_from_host(self.h, _ptr(a))

    def get(self):
        out = np.empty(self.shape, np.float64)
        lib().rgrid_copy_to_host(self.h, _ptr(out))
        return out

    def free(self):
        if self.h:
            lib().rgrid_free(self.h)
            self.h = None


def host_params(model):
    """A dft_ot_functional on the host holding only dft_ot_init_params(model)
    (enough for the bulk closed forms; no device state)."""
    otf = OTF()
    otf.model = model
    lib().dft_ot_init_params(C.byref(otf), model)
    return otf


def rgrid_get(h, shape):
    out = np.empty(shape, np.float64)
    lib().rgrid_copy_to_host(h, _ptr(out))
    return out


class Functional:
    """dft_ot_alloc / dft_ot_potential / dft_ot_energy_density on the GPU."""

    def __init__(self, model, wave, min_sub=4, max_sub=32):
        self.h = lib().dft_ot_alloc(model, wave.h, min_sub, max_sub)
        if not self.h:
            raise RuntimeError("dft_ot_alloc failed")
        self.wave = wave
        self.p = self.h.contents
        self.shape = wave.shape

    def potential(self, pot):
        lib().dft_ot_potential(self.h, pot.h, self.wave.h)

    def energy_density(self, eps):
        lib().dft_ot_energy_density(self.h, eps.h, self.wave.h)

    def energy(self):
        return lib().dft_ot_energy(self.h, self.wave.h)

    def density(self):
        return rgrid_get(self.p.density, self.shape)

    def set_kc_exact(self, exact):
        lib().dft_ot_set_kc_mode(self.h, 1 if exact else 0)

    def bulk(self, name, *args):
        return getattr(lib(), "dft_ot_bulk_" + name)(self.h, *[float(a) for a in args])

    def adjoint_well(self, amp, width, r0, mu, dt, nsteps, every):
        """dft_ot_adjoint_well: (E[psi_K], [dE/damp, dE/dwidth]) for nsteps steps from self.wave."""
        E = C.c_double(0.0)
        g = (C.c_double * 2)()
        rc = lib().dft_ot_adjoint_well(self.h, self.wave.h, amp, width, r0[0], r0[1], r0[2], mu, Complex.of(dt),
                                       nsteps, every, C.cast(C.byref(E), C.c_void_p), C.cast(g, C.c_void_p))
        if rc != 0:
            raise RuntimeError("dft_ot_adjoint_well failed")
        return E.value, [g[0], g[1]]

    def adjoint_cavity(self, rho0, radius, width, mu, dt, nsteps, every):
        """dft_ot_adjoint_cavity: (R_b[psi_K], dR_b/dradius); writes psi_0(radius) into self.wave."""
        J = C.c_double(0.0)
        g = C.c_double(0.0)
        rc = lib().dft_ot_adjoint_cavity(self.h, self.wave.h, rho0, radius, width, mu, Complex.of(dt), nsteps, every,
                                         C.cast(C.byref(J), C.c_void_p), C.cast(C.byref(g), C.c_void_p))
        if rc != 0:
            raise RuntimeError("dft_ot_adjoint_cavity failed")
        return J.value, g.value

    def dispersion(self, dt, k, amp, pred=False, direction=2):
        """dft_ot_dispersion on self.wave (overwritten): (energy in hartree, k rounded to the grid)."""
        kk = C.c_double(k)
        e = lib().dft_ot_dispersion(self.wave.h, self.h, dt, C.cast(C.byref(kk), C.c_void_p), amp, 1 if pred else 0,
                                    direction)
        if e < 0:
            raise RuntimeError("dft_ot_dispersion failed")
        return e, kk.value

    def free(self):
        if self.h:
            lib().dft_ot_free(self.h)
            self.h = None
