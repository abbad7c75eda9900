"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The product library never uses it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

AUTOANG = 0.529177210903
AUTOK = 315775.02480407
AUTOFS = 0.024188843265857
AUTOAMU = 1.0 / 1822.888486209
AUTOMPS = 2187691.26364
AUTOBAR = 2.9421015697e8
HELIUM_MASS = 4.002602 / AUTOAMU
PLAIN, KC, HD, HD2, BACKFLOW, T0MK = 0, 1, 2, 4, 8, 16


class _OrcOT(C.Structure):
    _fields_ = [("model", C.c_int64), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("step", C.c_double)] + [(n, C.c_double) for n in (
                    "b c2 c2_exp c3 c3_exp c4 rho_0s alpha_s l_g mass rho0 temp "
                    "lj_h lj_sigma lj_epsilon lj_cval g11 g12 g21 g22 a1 a2 "
                    "div_epsilon max_veloc max_bfpot c_bfpot kc_coef23").split()] + \
              [(n, C.c_void_p) for n in "lj_tf sa_tf g_tf gx_tf gy_tf gz_tf bf_tf".split()] + \
              [(n, C.c_double) for n in "hd_beta hd_rhom hd_C sa_h bf_xi bf_rhobf".split()]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        P, D, I64, I = C.c_void_p, C.c_double, C.c_int64, C.c_int
        sig = {
            "orc_ot_alloc": (C.POINTER(_OrcOT), [I64, I, I, I, D, I, I]),
            "orc_ig_energy": (D, [P, D]), "orc_ig_dEdRho": (D, [P, D]),
            "orc_ot_free": (None, [P]),
            "orc_ot_set_kc_mode": (None, [P, I]),
            "orc_ot_potential": (None, [P, P, P, D, P]),
            "orc_ot_potential2": (None, [P, P, P, P, P, P]),
            "orc_ot_energy_density": (None, [P, P, P, D]),
            "orc_ot_energy_density_rv": (None, [P, P, P, P, P, P]),
            "orc_set_plant": (None, [I]),
            "orc_ot_energy_gradient": (None, [P, P, P]),
            "orc_ot_energy_rho": (D, [P, P]),
            "orc_fft3d": (None, [P, I, I, I, I]),
            "orc_wf_velocity": (None, [P, P, P, P, I, I, I, D, D, D]),
            "orc_gradient": (None, [P, P, P, P, I, I, I, D]),
            "orc_wf_propagate": (None, [P, P, _Cplx, I, I, I, D, D]),
            "orc_wf_norm": (D, [P, I64, D]),
            "orc_wf_kinetic_energy": (D, [P, I, I, I, D, D]),
            "orc_integral": (D, [P, I64, D]),
            "orc_kernel_map": (None, [P, I, P, I, I]),
            "orc_pc_step": (None, [P, P, P, D, _Cplx, D, P]),
            "orc_single_step": (None, [P, P, P, D, _Cplx, D, P]),
            "orc_observables": (None, [P, P, P, D, P]),
            "orc_wall_time": (D, []),
            "orc_num_threads": (I, []),
        }
        for name in ("energy", "dEdRho", "pressure", "dPdRho", "density_pressurized",
                     "chempot_pressurized", "compressibility", "sound_speed"):
            sig["orc_bulk_" + name] = (D, [P, D])
        sig["orc_bulk_istatic"] = (D, [P, D, D])
        sig["orc_bulk_dispersion"] = (D, [P, D, D])
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _Cplx(C.Structure):
    """double complex passed by value (two SSE eightbytes, like struct {double, double})."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _cdt(dt):
    return _Cplx(complex(dt).real, complex(dt).imag)


class OracleOT:
    """FP64 CPU restatement of dft_ot_alloc / dft_ot_potential / dft_ot_energy_density."""

    def __init__(self, model, shape, step, min_sub=4, max_sub=32):
        self.shape = tuple(int(s) for s in shape)
        self.step = float(step)
        self.model = int(model)
        self._h = lib().orc_ot_alloc(model, *self.shape, self.step, min_sub, max_sub)
        if not self._h:
            raise MemoryError("orc_ot_alloc failed")
        self.p = self._h.contents

    def close(self):
        if self._h:
            lib().orc_ot_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def set_kc_exact(self, exact):
        lib().orc_ot_set_kc_mode(self._h, 1 if exact else 0)

    def spectrum(self, name):
        ptr = getattr(self.p, name)
        if not ptr:
            return None
        buf = (C.c_double * (2 * self.n)).from_address(ptr)
        return np.frombuffer(buf, dtype=np.complex128).reshape(self.shape).copy()

    def potential(self, psi, pot=None, wf_mass=HELIUM_MASS, want_rho=False):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        pot = np.zeros(self.shape, np.complex128) if pot is None else np.ascontiguousarray(pot, np.complex128).copy()
        rho = np.empty(self.shape) if want_rho else None
        lib().orc_ot_potential(self._h, _p(pot), _p(psi), wf_mass, _p(rho) if want_rho else None)
        return (pot, rho) if want_rho else pot

    def potential2(self, rho, vx, vy, vz, pot=None):
        rho = np.ascontiguousarray(rho, np.float64)
        v = [np.ascontiguousarray(a, np.float64).copy() for a in (vx, vy, vz)]
        pot = np.zeros(self.shape, np.complex128) if pot is None else np.ascontiguousarray(pot, np.complex128).copy()
        lib().orc_ot_potential2(self._h, _p(pot), _p(rho), *[_p(a) for a in v])
        return pot

    def energy_density(self, psi, wf_mass=HELIUM_MASS):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        eps = np.empty(self.shape)
        lib().orc_ot_energy_density(self._h, _p(eps), _p(psi), wf_mass)
        return eps

    def energy_density_rv(self, rho, vx, vy, vz):
        """Energy density with the density and the (backflow) velocity given."""
        a = [np.ascontiguousarray(x, np.float64) for x in (rho, vx, vy, vz)]
        eps = np.empty(self.shape)
        lib().orc_ot_energy_density_rv(self._h, _p(eps), *[_p(x) for x in a])
        return eps

    def energy_gradient(self, rho):
        """Backward of E[rho]: exact discrete dE/drho (LJ, c2/c3, KC)."""
        rho = np.ascontiguousarray(rho, np.float64)
        g = np.empty(self.shape)
        lib().orc_ot_energy_gradient(self._h, _p(g), _p(rho))
        return g

    def energy_rho(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        return lib().orc_ot_energy_rho(self._h, _p(rho))

    def kernel_map(self, which, min_sub=4, max_sub=32):
        out = np.empty(self.shape)
        lib().orc_kernel_map(self._h, which, _p(out), min_sub, max_sub)
        return out

    def pc_step(self, psi, vext, mu0, dt, wf_mass=HELIUM_MASS):
        work = np.empty(2 * self.n, np.complex128)
        lib().orc_pc_step(self._h, _p(psi), _p(vext) if vext is not None else None, mu0, _cdt(dt), wf_mass, _p(work))

    def single_step(self, psi, vext, mu0, dt, wf_mass=HELIUM_MASS):
        work = np.empty(self.n, np.complex128)
        lib().orc_single_step(self._h, _p(psi), _p(vext) if vext is not None else None, mu0, _cdt(dt), wf_mass, _p(work))

    def observables(self, psi, vext=None, wf_mass=HELIUM_MASS):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = np.empty(6)
        lib().orc_observables(self._h, _p(psi), _p(vext) if vext is not None else None, wf_mass, _p(out))
        return dict(zip(("energy", "norm", "ekin", "eot", "eext", "mu"), out))

    def bulk(self, name, *args):
        return getattr(lib(), "orc_bulk_" + name)(self._h, *[float(a) for a in args])


def fft3d(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft3d(_p(a), *a.shape, sign)
    return a


def velocity(psi, step, mass=HELIUM_MASS, eps=1e-4):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    v = [np.empty(psi.shape) for _ in range(3)]
    lib().orc_wf_velocity(_p(psi), *[_p(a) for a in v], *psi.shape, step, mass, eps)
    return v


def gradient(f, step):
    f = np.ascontiguousarray(f, np.float64)
    g = [np.empty(f.shape) for _ in range(3)]
    lib().orc_gradient(_p(f), *[_p(a) for a in g], *f.shape, step)
    return g


def propagate(psi, pot, dt, step, mass=HELIUM_MASS):
    psi = np.ascontiguousarray(psi, np.complex128).copy()
    pot = np.ascontiguousarray(pot, np.complex128)
    lib().orc_wf_propagate(_p(psi), _p(pot), _cdt(dt), *psi.shape, step, mass)
    return psi


def norm(psi, step):
    psi = np.ascontiguousarray(psi, np.complex128)
    return lib().orc_wf_norm(_p(psi), psi.size, step)


def kinetic_energy(psi, step, mass=HELIUM_MASS):
    psi = np.ascontiguousarray(psi, np.complex128)
    return lib().orc_wf_kinetic_energy(_p(psi), *psi.shape, step, mass)


def set_plant(which):
    """Tests only: plant mistake number `which` in the oracle (0 = none)."""
    lib().orc_set_plant(int(which))


def wall_time():
    return lib().orc_wall_time()


def num_threads():
    return lib().orc_num_threads()
