#!/usr/bin/env python3
"""Headline benchmark: Orsay-Trento split-operator steps/s at 256^3 on one MI355X.

Workload: full OT (PLAIN|KC|BACKFLOW), periodic 256^3 grid, h = 0.3 A, bulk
rho0 with a straight vortex line (phase atan2(y, x) along z) and 1e-3 seeded
density noise, real-time step dt = 2 fs.  One step is the reference driver's
single-step pattern (doc/libdft-manual.tex:496-504): zero the potential,
dft_ot_potential, add -mu, grid_wf_propagate; the value is measured on the
library's fused form of that sequence (dft_ot_propagate), and the four-call
form is timed beside it.  Fields are FP32 on the device
with FP64 reductions; the FP64 device path runs the same steps for the
precision comparison.  The CPU baseline is the FP64 oracle (OpenMP on the host
cores) timed on one step of the same workload.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import libdft_amd as L  # noqa: E402
from tests import inputs  # noqa: E402

METRIC = "OT split-op steps/s at 256³ (FP32 vs FP64 GPU, ≤1e-6 rel energy) + HBM GB/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of the 256^3 FP32 step's kernels, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) of this script, written by --traffic (the command is in the file)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_ot_step_256_fp32.md")
# device symbol (as the profiler names it) of each profiled launch label, FP32 256^3
SYMBOL = {"ot_xconv_single": "dft::fused::k_xconv<float, 256, false, true>",
          "ot_xconv_multi": "dft::fused::k_xconv<float, 256, false, false>",
          "ot_xmid_fused": "dft::fused::k_xmid<float, 256, 1, true>",
          "fft_y_c2c_fwd": "dft::fft::k_strided<float, 256, -1>",
          "fft_y_c2c_inv": "dft::fft::k_strided<float, 256, 1>",
          "ot_zAB_c2r_local_kc_bf_r2c": "dft::fused::k_zAB<float, 256, true, true, 0>",
          "ot_zBC_c2r_kc_terms_r2c": "dft::fused::k_zBC<float, 256, true>",
          "ot_zA_density_r2c": "dft::fused::k_zA<float, 256, true>",
          "ot_zC_c2r_phase_z_fwd": "dft::fused::k_zC<float, 256, true>",
          "fft_z_c2c_phase_inv": "dft::fft::k_zc2c_pot<float, 256, 1>",
          "fft_x_kinetic_c2c": "dft::fft::k_xkin<float, 256>"}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def profile_table():
    lib = L.lib()
    lib.dft_profile_collect.restype = C.c_int
    lib.dft_profile_entry.argtypes = [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    n = lib.dft_profile_collect()
    rows = []
    for i in range(n):
        name = C.create_string_buffer(128)
        cnt, ms, by = C.c_int64(), C.c_double(), C.c_double()
        lib.dft_profile_entry(i, name, 128, C.byref(cnt), C.byref(ms), C.byref(by))
        rows.append({"kernel": name.value.decode(), "launches": cnt.value, "total_ms": ms.value,
                     "avg_us": 1e3 * ms.value / max(cnt.value, 1), "bytes_per_launch": by.value / max(cnt.value, 1),
                     "GBps": by.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0})
    rows.sort(key=lambda r: -r["total_ms"])
    return rows


class Run:
    """One device precision of the workload."""

    def __init__(self, n, step, prec, psi0, model=9):
        self.lib = L.lib()
        self.w = L.Wave((n, n, n), step, prec)
        self.w.set(psi0)
        self.f = L.Functional(model, self.w)
        self.pot = L.CField((n, n, n), step, prec)
        self.mu = self.f.bulk("dEdRho", self.f.p.rho0)
        self.n, self.step = n, step

    def steps(self, k, dt, pattern=False):
        """k single steps: the fused dft_ot_propagate, or (pattern) the reference driver's
        cgrid_zero / dft_ot_potential / cgrid_add(-mu) / grid_wf_propagate sequence."""
        lib, z = self.lib, L.Complex.of(dt)
        if not pattern:
            for _ in range(k):
                lib.dft_ot_propagate(self.f.h, self.w.h, None, self.mu, z)
            return
        mmu = L.Complex.of(-self.mu)
        for _ in range(k):
            lib.cgrid_zero(self.pot.h)
            lib.dft_ot_potential(self.f.h, self.pot.h, self.w.h)
            lib.cgrid_add(self.pot.h, mmu)
            lib.grid_wf_propagate(self.w.h, self.pot.h, z)

    def timed(self, k, dt, pattern=False):
        self.lib.grid_synchronize()
        t0 = time.perf_counter()
        self.steps(k, dt, pattern)
        self.lib.grid_synchronize()
        return (time.perf_counter() - t0) / k

    def energy(self):
        """Total energy (kinetic + OT) and norm, FP64 reductions."""
        ekin = self.lib.grid_wf_energy(self.w.h, None)
        eot = self.f.energy()
        return ekin + eot, self.lib.grid_wf_norm(self.w.h)

    def free(self):
        self.pot.free()
        self.f.free()
        self.w.free()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, step, psi0, dt, timed=3):
    """The same workload on the FP64 oracle (OpenMP over the host cores): one
    warm-up step, then `timed` steps timed with the wall clock.  Returns the
    baseline record and the oracle's total energy after the first step."""
    from oracle import pyoracle as po
    o = po.OracleOT(9, (n, n, n), step, 1, 1)   # kernel sampling cost is setup, not part of a step
    psi = psi0.copy()
    mu = o.bulk("dEdRho", o.p.rho0)
    o.single_step(psi, None, mu, dt)
    e_oracle = o.observables(psi)["energy"]
    t0 = time.perf_counter()
    for _ in range(timed):
        o.single_step(psi, None, mu, dt)
    sec = (time.perf_counter() - t0) / timed
    o.close()
    return ({"value": 1.0 / sec, "unit": "steps/s", "cores": po.num_threads(), "kind": "port",
             "cpu": cpu_model(), "nproc": os.cpu_count(),
             "sample": "%d full-OT single steps of this workload (%d^3, FP64) after 1 warm-up step, on the "
                       "C/OpenMP oracle; kernels sampled once per cell (setup cost excluded, step cost "
                       "unchanged); %.2f s per step" % (timed, n, sec)}, e_oracle)


def copy_calibration(n, step, reps=3):
    """A plain device copy of one FP32 field (n^3 reads + n^3 writes), for scaling HBM
    traffic counters (FETCH_SIZE / WRITE_SIZE) read under a profiler: kernel 'k_copy_d'."""
    lib = L.lib()
    a = L.RField((n, n, n), step, L.GRID_FP32)
    b = L.RField((n, n, n), step, L.GRID_FP32)
    lib.rgrid_constant(a.h, 1.0)
    for _ in range(reps):
        lib.grid_bench_copy(b.h, a.h)
    lib.grid_synchronize()
    a.free(), b.free()
    return {"kernel": "k_copy_d", "read_bytes": 4.0 * n ** 3, "write_bytes": 4.0 * n ** 3, "launches": reps}


def _kernel_symbol(name):
    """Kernel name from a profiler CSV without its argument list or 'void ' prefix."""
    name = name.split("(")[0].strip()
    return name[5:] if name.startswith("void ") else name


def summarize_traffic(fetch_csv, write_csv):
    """Markdown table of HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    runs of this script: each counter (KiB per dispatch) is scaled by what it reads for the
    copy kernel k_copy_d, which reads and writes one 256^3 FP32 field (64 MiB each way)."""
    import csv
    import collections

    def load(path, counter):
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                agg[_kernel_symbol(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return agg

    f, w = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_kib = 4.0 * 256 ** 3 / 1024
    sf = copy_kib / np.mean(f["k_copy_d"])
    sw = copy_kib / np.mean(w["k_copy_d"])
    lines = ["copy kernel k_copy_d (64 MiB read + 64 MiB written): FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB "
             "per dispatch -> scale fetch x%.4f, write x%.4f" % (np.mean(f["k_copy_d"]), np.mean(w["k_copy_d"]), sf, sw),
             "", "| kernel | dispatches | fetch MB | write MB | HBM MB per dispatch |", "|---|---|---|---|---|"]
    for k in sorted(f, key=lambda k: -np.sum(f[k])):
        if k not in w or k == "k_copy_d":
            continue
        fb, wb = np.mean(f[k]) * sf * 1024, np.mean(w[k]) * sw * 1024
        lines.append("| `%s` | %d | %.1f | %.1f | %.1f |" % (k, len(f[k]), fb / 1e6, wb / 1e6, (fb + wb) / 1e6))
    return "\n".join(lines)


def summarize_kernel_stats(stats_csv, top=40):
    """Markdown table of a rocprofv3 --kernel-trace --stats kernel_stats.csv."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    lines = ["| kernel | calls | avg us | total ms | % |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        lines.append("| `%s` | %s | %.1f | %.2f | %s |" % (_kernel_symbol(r["Name"]), r["Calls"],
                                                         float(r["AverageNs"]) / 1e3,
                                                         float(r["TotalDurationNs"]) / 1e6, r["Percentage"]))
    return "\n".join(lines)


def summarize_counters(csvs, kernels=None):
    """Markdown table of per-dispatch counter averages from rocprofv3 --pmc
    counter_collection.csv files (one pass each), one row per kernel of the
    256^3 FP32 step, with derived ratios: wave lifetime (SQ_WAVE_CYCLES /
    SQ_WAVES, in SQ clock cycles per wave), the shares of wave cycles spent
    waiting for a dependency (SQ_WAIT_ANY: memory or LDS results) and waiting
    to issue (SQ_WAIT_INST_ANY: barriers, issue stalls), and issuing VALU and
    LDS instructions, instruction counts per wave, LDS bank-conflict cycles
    per LDS instruction and the L2 hit rate."""
    import csv
    import collections
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in csvs:
        for r in csv.DictReader(open(path)):
            vals[_kernel_symbol(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    wanted = kernels or [SYMBOL[k] for k in SYMBOL]
    cols = [("wave cycles", lambda a: a["SQ_WAVE_CYCLES"] / a["SQ_WAVES"]),
            ("dependency wait %", lambda a: 100 * a["SQ_WAIT_ANY"] / a["SQ_WAVE_CYCLES"]),
            ("issue wait %", lambda a: 100 * a["SQ_WAIT_INST_ANY"] / a["SQ_WAVE_CYCLES"]),
            ("VALU %", lambda a: 100 * a["SQ_ACTIVE_INST_VALU"] / a["SQ_WAVE_CYCLES"]),
            ("LDS %", lambda a: 100 * a["SQ_ACTIVE_INST_LDS"] / a["SQ_WAVE_CYCLES"]),
            ("VALU insts/wave", lambda a: a["SQ_INSTS_VALU"] / a["SQ_WAVES"]),
            ("VMEM rd/wave", lambda a: a["SQ_INSTS_VMEM_RD"] / a["SQ_WAVES"]),
            ("VMEM wr/wave", lambda a: a["SQ_INSTS_VMEM_WR"] / a["SQ_WAVES"]),
            ("LDS insts/wave", lambda a: a["SQ_INSTS_LDS"] / a["SQ_WAVES"]),
            ("bank conflict/LDS inst", lambda a: a["SQ_LDS_BANK_CONFLICT"] / a["SQ_INSTS_LDS"]),
            ("L2 hit %", lambda a: 100 * a["TCC_HIT_sum"] / (a["TCC_HIT_sum"] + a["TCC_MISS_sum"]))]
    lines = ["| kernel | " + " | ".join(c for c, _ in cols) + " |", "|---" * (len(cols) + 1) + "|"]
    for k in wanted:
        if k not in vals:
            continue
        a = {c: float(np.mean(v)) for c, v in vals[k].items()}
        cells = []
        for _, f in cols:
            try:
                cells.append("%.1f" % f(a))
            except (KeyError, ZeroDivisionError):
                cells.append("")
        lines.append("| `%s` | " % k + " | ".join(cells) + " |")
    return "\n".join(lines)


def measured_traffic(label, n):
    """HBM bytes per dispatch of a profiled launch label from TRAFFIC_FILE (256^3 FP32 only)."""
    sym = SYMBOL.get(label)
    if n != 256 or sym is None or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == "`%s`" % sym:
            return float(cells[4]) * 1e6
    return None


# step traffic in units of one real field R (4N B in FP32, 8N B in FP64):
# 134 R for the full functional, 34 R for plain OT (algorithmic schedule counts)
# algorithmic bytes per step in real-field units R: plain OT, full OT, and full OT with
# the HD / HD2 backflow switch (the penalty is pointwise; the switch transforms one more
# field, rho~ = G rho: its z-pass spectrum write, the y pass over it and the x-middle read)
STEP_R = {0: 34, 9: 134, 11: 138}


def size_sweep(steps=10, warmup=3):
    """Steps/s and the whole-step HBM fraction at each workload size: plain OT at
    64^3 (FP64, h = 0.5 A, noisy bulk), full OT at 128^3 (droplet, h = 0.4 A) in
    FP32 and FP64, full OT at 512^3 (bulk with a 17 A cavity, h = 0.4 A) in
    FP32, and KC|HD|BACKFLOW (the high-density penalty and its backflow switch) on
    the 256^3 vortex workload in FP32; real-time single steps (dt = 2 fs) except
    imaginary time at 64^3."""
    rho0 = 0.0218360 * L.AUTOANG ** 3
    cases = [
        ("plain OT 64^3 imaginary time", 64, 0.5, L.GRID_FP64, 0,
         lambda n, h: inputs.noisy_bulk((n, n, n), rho0, 0.01), -1j * 10.0 / L.AUTOFS),
        ("full OT 128^3 droplet", 128, 0.4, L.GRID_FP32, 9, lambda n, h: inputs.droplet((n, n, n), h), 2.0 / L.AUTOFS),
        ("full OT 128^3 droplet", 128, 0.4, L.GRID_FP64, 9, lambda n, h: inputs.droplet((n, n, n), h), 2.0 / L.AUTOFS),
        ("full OT 512^3 cavity", 512, 0.4, L.GRID_FP32, 9, lambda n, h: inputs.cavity((n, n, n), h, rho0),
         2.0 / L.AUTOFS),
        ("KC|HD|BACKFLOW 256^3 vortex", 256, 0.3, L.GRID_FP32, 11,
         lambda n, h: inputs.vortex_line((n, n, n), h, rho0), 2.0 / L.AUTOFS),
    ]
    out = []
    # the plain-OT 64^3 workload's pattern: imaginary-time steps renormalised every step, fused and as the
    # reference driver does it (the step, then grid_wf_normalize)
    n, step, prec = 64, 0.5 / L.AUTOANG, L.GRID_FP64
    log("size sweep: plain OT 64^3 imaginary time, renormalised every step")
    r = Run(n, step, prec, inputs.noisy_bulk((n, n, n), rho0, 0.01), 0)
    r.w.h.contents.norm = L.lib().grid_wf_norm(r.w.h)
    z = L.Complex.of(-1j * 10.0 / L.AUTOFS)
    t = {}
    k64 = max(steps, 200)  # 64^3 steps take ~0.1 ms
    for kind in ("fused", "step_then_normalize"):
        for it in range(warmup + k64):
            if it == warmup:
                L.lib().grid_synchronize()
                t0 = time.perf_counter()
            if kind == "fused":
                L.lib().dft_ot_propagate_normalize(r.f.h, r.w.h, None, r.mu, z)
            else:
                L.lib().dft_ot_propagate(r.f.h, r.w.h, None, r.mu, z)
                L.lib().grid_wf_normalize(r.w.h)
        L.lib().grid_synchronize()
        t[kind] = (time.perf_counter() - t0) / k64
    r.free()
    out.append({"workload": "plain OT 64^3 imaginary time, renormalised every step", "grid": n, "dtype": "f64",
                "steps_per_s": 1.0 / t["fused"], "steps_per_s_step_then_normalize": 1.0 / t["step_then_normalize"]})
    for name, n, h_ang, prec, model, make, dt in cases:
        step = h_ang / L.AUTOANG
        log("size sweep: %s (%s)" % (name, "FP32" if prec == L.GRID_FP32 else "FP64"))
        r = Run(n, step, prec, make(n, step), model)
        r.steps(warmup, dt)
        t = r.timed(steps, dt)
        r.free()
        R = (4.0 if prec == L.GRID_FP32 else 8.0) * n ** 3
        gbs = STEP_R[model] * R / t / 1e9
        out.append({"workload": name, "grid": n, "dtype": "f32" if prec == L.GRID_FP32 else "f64",
                    "steps_per_s": 1.0 / t, "step_GBps": gbs, "step_frac": gbs / HBM_PEAK_GBS})
    return out


def cavity512_adjoint(steps=2000, every=45, prec=L.GRID_FP32, rel_fd=1e-2, model=9, fd_prec=None):
    """The 512^3 released-cavity workload (h = 0.4 A, 17 A cavity with a 0.5 A wall, 2 fs
    steps; full OT): bubble radius (missing volume within a
    25.5 A window, as dft_ot_adjoint_cavity defines it) after `steps` steps and its
    derivative with respect to the initial cavity radius from the checkpointed adjoint sweep
    (ceil(steps/every) + every stored states), against central differences of two forward runs
    (in the precision fd_prec, default the adjoint's)."""
    n, step = 512, 0.4 / L.AUTOANG
    rho0 = 0.0218360 * L.AUTOANG ** 3
    R0, wd, dt = 17.0 / L.AUTOANG, 0.5 / L.AUTOANG, 2.0 / L.AUTOFS
    lib = L.lib()
    w = L.Wave((n, n, n), step, prec)
    f = L.Functional(model, w)
    mu = f.bulk("dEdRho", rho0)
    log("512^3 cavity adjoint: %d steps, checkpoints every %d" % (steps, every))
    lib.dft_set_verbose(1)  # a progress line per checkpoint
    lib.grid_synchronize()
    t0 = time.perf_counter()
    J, g = f.adjoint_cavity(rho0, R0, wd, mu, dt, steps, every)
    t_adj = time.perf_counter() - t0
    lib.dft_set_verbose(0)

    def radius(R):
        w.set(inputs.cavity((n, n, n), step, rho0, R * L.AUTOANG, wd * L.AUTOANG))
        for it in range(steps):
            lib.dft_ot_propagate(f.h, w.h, None, mu, L.Complex.of(dt))
            if it % 500 == 0:
                lib.grid_synchronize()
                log("512^3 cavity forward R0 = %.4f A: step %d" % (R * L.AUTOANG, it))
        # the bubble measure of dft_ot_adjoint_cavity: the missing volume inside the window
        x = (np.arange(n) - n // 2) * step
        r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
        win = 0.5 * (1.0 - np.tanh((r - 1.5 * R) / (4.0 * wd)))  # the window moves with the cavity radius
        del r
        rho = np.abs(w.get()) ** 2
        vb = step ** 3 * np.sum(win * (1.0 - rho / rho0))
        return (3.0 * vb / (4.0 * np.pi)) ** (1.0 / 3.0)

    if fd_prec is not None and fd_prec != prec:
        f.free(), w.free()
        w = L.Wave((n, n, n), step, fd_prec)
        f = L.Functional(model, w)
    log("512^3 cavity: forward runs for the finite difference")
    t0 = time.perf_counter()
    d = rel_fd * R0
    fd = (radius(R0 + d) - radius(R0 - d)) / (2 * d)
    t_fwd = (time.perf_counter() - t0) / 2
    f.free(), w.free()
    states = -(-steps // every) + every
    return {"workload": "released 17 A cavity, 512^3, h = 0.4 A, model %d" % model, "dtype": "f32" if prec == L.GRID_FP32
            else "f64", "steps": steps, "every": every, "stored_states": states,
            "stored_GB": states * n ** 3 * (8 if prec == L.GRID_FP32 else 16) / 1e9,
            "R_b_A": J * L.AUTOANG, "dRb_dR0_adjoint": g, "dRb_dR0_fd": fd, "rel_err": abs(g / fd - 1),
            "fd_step_A": d * L.AUTOANG, "fd_dtype": "f64" if (fd_prec or prec) == L.GRID_FP64 else "f32",
            "adjoint_s": t_adj, "forward_s": t_fwd}


def accuracy(n, step, psi0, dt, k):
    """FP32 fields + FP64 reductions against the FP64 path: k steps from the same
    (FP32-representable) state; and one step against the FP64 oracle (kernels sampled
    once per cell on both sides)."""
    out = {}
    res = {}
    for prec in (L.GRID_FP64, L.GRID_FP32):
        r = Run(n, step, prec, psi0)
        r.steps(k, dt)
        res[prec] = r.energy()
        r.free()
    (e64, n64), (e32, n32) = res[L.GRID_FP64], res[L.GRID_FP32]
    out["energy_rel_err_fp32_vs_fp64"] = abs(e32 - e64) / abs(e64)
    out["norm_rel_err_fp32_vs_fp64"] = abs(n32 - n64) / abs(n64)
    out["energy_K"] = e64 * L.AUTOK
    out["steps_compared"] = k
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--acc-steps", type=int, default=50,
                    help="steps of the FP32 vs FP64 energy comparison (0: skip)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the per-size steps/s table")
    ap.add_argument("--cavity512-adjoint", action="store_true",
                    help="run the 512^3 released-cavity adjoint (2000 steps) against finite differences and exit")
    ap.add_argument("--cavity-fd-f64", action="store_true",
                    help="with --cavity512-adjoint: central differences from FP64 forward runs with a 0.1%% step")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the HBM traffic table from two rocprofv3 --pmc counter_collection.csv files and exit")
    ap.add_argument("--kernel-stats", metavar="STATS_CSV",
                    help="print a rocprofv3 kernel_stats.csv as a table and exit")
    ap.add_argument("--counters", nargs="+", metavar="COUNTER_CSV",
                    help="print per-kernel counter ratios from rocprofv3 --pmc counter_collection.csv files and exit")
    args = ap.parse_args()
    if args.steps < 1 or args.profile_steps < 1 or args.warmup < 0:
        ap.error("--steps and --profile-steps take at least 1 (the roofline needs the per-launch events), "
                 "--warmup at least 0")
    if args.counters:
        print(summarize_counters(args.counters))
        return
    if args.traffic:
        print(summarize_traffic(*args.traffic))
        return
    if args.kernel_stats:
        print(summarize_kernel_stats(args.kernel_stats))
        return
    if args.cavity512_adjoint:
        if args.cavity_fd_f64:
            print(json.dumps(cavity512_adjoint(rel_fd=1e-3, fd_prec=L.GRID_FP64)))
        else:
            print(json.dumps(cavity512_adjoint()))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    n = args.grid
    step = 0.3 / L.AUTOANG
    dt = 2.0 / L.AUTOFS
    rho0 = 0.0218360 * L.AUTOANG ** 3
    # the initial state is FP32-representable, so both precisions start from the same state
    psi0 = inputs.vortex_line((n, n, n), step, rho0, noise=1e-3).astype(np.complex64).astype(complex)
    lib = L.lib()
    calib = copy_calibration(n, step)

    # FP32 fields, FP64 reductions: the measured path
    log("fp32 setup %d^3" % n)
    r32 = Run(n, step, L.GRID_FP32, psi0)
    log("fp32 warmup + %d timed steps" % args.steps)
    r32.steps(args.warmup, dt)
    t32 = r32.timed(args.steps, dt)
    # same launches again with per-launch HIP events, for the per-kernel roofline
    lib.dft_profile_reset()
    lib.dft_profile_enable(1)
    tprof = r32.timed(args.profile_steps, dt)
    lib.dft_profile_enable(0)
    rows = profile_table()
    # the same steps through the reference driver's four calls (unfused potential accumulate + propagate)
    tpat = r32.timed(max(5, args.steps // 5), dt, pattern=True)
    # predict-correct steps (two potential evaluations per step)
    wp = L.Wave((n, n, n), step, L.GRID_FP32)
    kpc = max(5, args.steps // 5)
    lib.grid_synchronize()
    t0 = time.perf_counter()
    for _ in range(kpc):
        lib.dft_ot_propagate_pc(r32.f.h, r32.w.h, wp.h, None, r32.mu, L.Complex.of(dt))
    lib.grid_synchronize()
    tpc = (time.perf_counter() - t0) / kpc
    wp.free()
    r32.free()

    out = {"metric": METRIC, "value": 1.0 / t32, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * t32, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "full OT (KC+BF) real-time Strang step, vortex line in bulk He, %d^3, "
                                  "h=0.3 A, dt=2 fs, FP32 fields + FP64 reductions" % n,
                      "grid": [n, n, n], "model": "PLAIN|KC|BACKFLOW", "step_kind": "single Strang step"}}

    # roofline of the dominant kernel (largest total time in the profiled steps)
    dom = rows[0]
    out["roofline"] = {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS,
                       "traffic": measured_traffic(dom["kernel"], n),
                       "avg_us": dom["avg_us"], "bytes_per_launch": dom["bytes_per_launch"]}
    # whole-step roofline against the canonical 134 R per full-OT step (R = 4N bytes)
    step_bytes = 134 * 4.0 * n ** 3
    out["step_roofline"] = {"algorithmic_bytes": step_bytes, "achieved": step_bytes / t32 / 1e9,
                            "frac": step_bytes / t32 / 1e9 / HBM_PEAK_GBS, "unit": "GB/s",
                            "moved_bytes_per_step": sum(r["bytes_per_launch"] * r["launches"] for r in rows)
                            / args.profile_steps}
    out["profiled_ms_per_step"] = 1e3 * tprof
    out["driver_pattern_ms_per_step"] = 1e3 * tpat
    out["predict_correct_steps_per_s"] = 1.0 / tpc
    out["copy_calibration"] = calib
    out["kernels"] = [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()} for r in rows[:12]]

    if not args.no_fp64:
        log("fp64 setup and %d timed steps" % args.steps)
        r64 = Run(n, step, L.GRID_FP64, psi0)
        r64.steps(args.warmup, dt)
        t64 = r64.timed(args.steps, dt)
        r64.free()
        out["fp64_steps_per_s"] = 1.0 / t64
        out["fp32_over_fp64_speedup"] = t64 / t32
        if args.acc_steps > 0:
            log("accuracy: %d steps in FP32 and FP64" % args.acc_steps)
            out.update(accuracy(n, step, psi0, dt, args.acc_steps))

    if not args.no_sweep:
        out["sizes"] = size_sweep()

    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 warm-up + 3 timed steps)")
        out["cpu_baseline"], e_oracle = cpu_baseline(n, step, psi0, dt)
        # one FP32 step from the same state, kernels sampled once per cell as in the oracle run
        w = L.Wave((n, n, n), step, L.GRID_FP32)
        w.set(psi0)
        f = L.Functional(9, w, 1, 1)
        lib.dft_ot_propagate(f.h, w.h, None, f.bulk("dEdRho", f.p.rho0), L.Complex.of(dt))
        e1 = lib.grid_wf_energy(w.h, None) + f.energy()
        f.free(), w.free()
        out["energy_rel_err_fp32_vs_oracle_1step"] = abs(e1 - e_oracle) / abs(e_oracle)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
