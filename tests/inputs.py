"""Seeded synthetic inputs shared by the tests, the bench and the drivers.

Uniform deviates come from splitmix64 (seed 20261202 unless stated) so the
CPU oracle and the GPU path see bit-identical inputs.
"""
import numpy as np

AUTOANG = 0.529177210903
SEED = 20261202
_M64 = (1 << 64) - 1


def splitmix64_uniform(n, seed=SEED, lo=-1.0, hi=1.0):
    """n deviates in [lo, hi) from splitmix64 (53-bit mantissas)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def coords(shape, step):
    """x_i = (i - n/2) h for each axis (libgrid convention, origin 0)."""
    return [(np.arange(n) - n // 2) * step for n in shape]


def noisy_bulk(shape, rho, amp, seed=SEED):
    """psi = sqrt(rho) (1 + amp U(-1,1)), zero phase (the imaginary-time bulk start)."""
    u = splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    return (np.sqrt(rho) * (1.0 + amp * u)).astype(np.complex128)


def droplet(shape, step, natoms=1000.0, rho0_ang=0.0218, width_ang=1.0):
    """rho0/2 (1 - tanh((r - R)/w)), R = (3N / 4 pi rho0)^(1/3), zero phase."""
    rho0 = rho0_ang * AUTOANG ** 3
    R = (3.0 * natoms / (4.0 * np.pi * rho0)) ** (1.0 / 3.0)
    x, y, z = np.meshgrid(*coords(shape, step), indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    rho = 0.5 * rho0 * (1.0 - np.tanh((r - R) / (width_ang / AUTOANG)))
    return np.sqrt(rho).astype(np.complex128)


def dense_bump(shape, step, rho0, peak_ang=0.4, width_ang=1.0):
    """Bulk rho0 with a gaussian density bump rising to peak_ang (A^-3) at the
    origin, zero phase: reaches the densities where the high-density penalty acts."""
    x, y, z = np.meshgrid(*coords(shape, step), indexing="ij")
    w = width_ang / AUTOANG
    peak = peak_ang * AUTOANG ** 3
    rho = rho0 + (peak - rho0) * np.exp(-(x * x + y * y + z * z) / (2.0 * w * w))
    return np.sqrt(rho).astype(np.complex128)


def vortex_line(shape, step, rho0, noise=1e-3, seed=SEED):
    """Bulk rho0 with phase atan2(y, x) along z and seeded density noise."""
    x, y, _ = np.meshgrid(*coords(shape, step), indexing="ij")
    u = splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    amp = np.sqrt(rho0 * (1.0 + noise * u))
    return amp * np.exp(1j * np.arctan2(y, x))


def gaussian_well(shape, step, depth_K=-50.0, width_ang=2.0):
    """V_ext = depth exp(-r^2 / (2 w^2)) in hartree."""
    x, y, z = np.meshgrid(*coords(shape, step), indexing="ij")
    w = width_ang / AUTOANG
    return (depth_K / 315775.02480407) * np.exp(-(x * x + y * y + z * z) / (2.0 * w * w))


def smooth_modulated(shape, step, rho0, amp=0.3, seed=SEED):
    """Smooth, strongly modulated density with a smooth phase (exercises KC and BF)."""
    L = [n * step for n in shape]
    x, y, z = np.meshgrid(*[np.arange(n) * step for n in shape], indexing="ij")
    u = splitmix64_uniform(12, seed)
    rho = rho0 * (1.0 + amp * (np.sin(2 * np.pi * x / L[0] + 3 * u[0]) * np.cos(2 * np.pi * y / L[1] + 3 * u[1])
                               + 0.5 * np.sin(4 * np.pi * z / L[2] + 3 * u[2])))
    phase = (np.sin(2 * np.pi * x / L[0] + 3 * u[3]) + 0.7 * np.cos(2 * np.pi * z / L[2] + 3 * u[4])
             + 0.4 * np.sin(2 * np.pi * (y / L[1] + z / L[2]) + 3 * u[5]))
    return np.sqrt(rho) * np.exp(1j * phase)


def cavity(shape, step, rho0, radius_ang=17.0, width_ang=0.5):
    """Bulk rho0 with a spherical cavity: rho = rho0 (1 + tanh((r - R)/w)) / 2,
    zero phase (the state a 1000 K (1 - tanh((r - 17 A)/0.5 A))/2 barrier holds)."""
    x, y, z = np.meshgrid(*coords(shape, step), indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    rho = 0.5 * rho0 * (1.0 + np.tanh((r - radius_ang / AUTOANG) / (width_ang / AUTOANG)))
    return np.sqrt(rho).astype(np.complex128)
