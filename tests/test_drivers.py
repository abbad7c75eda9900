"""Drop-in boundary: reference drivers compile and link unchanged against
include/ and libdft_mi355x.so (compile and link only; nothing is run).

Skipped where the reference tree is absent.  The drivers take <grid/grid.h>,
<grid/au.h>, <dft/dft.h> and <dft/ot.h> and use <complex.h>'s I, libgrid's
REAL math names (SQRT, EXP, CEXP, CREAL, ...), HBAR and the unit
conversions, and the libgrid and libdft calls listed in the header.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/examples"
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "libdft_amd")

# each entry: the sources of one driver program
DRIVERS = [
    ["imag_time/imag_time.c"],
    ["dynamics/dynamics.c"],
    ["dispersion/dispersion.c"],
    ["dispersion/dispersion2.c"],
    ["dispersion/dispersion3.c"],
    ["dispersion/static-resp.c"],
    ["bulk/bulk_properties.c"],
    ["vortex-ring/vring.c"],
    ["gas-bubble/bubble.c", "gas-bubble/pot.c", "gas-bubble/analyze.c"],
    ["piston/piston.c"],
    ["shock/shock.c"],
    ["added_mass_viscosity/analyze4.c"],
    ["soliton/soliton.c"],
    ["1d-soliton/soliton-1d.c"],
    ["vortex-trace/trace.c"],
]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference examples not present")
@pytest.mark.parametrize("srcs", DRIVERS, ids=[d[0] for d in DRIVERS])
def test_reference_driver_compiles_and_links(tmp_path, srcs):
    assert os.path.exists(os.path.join(LIBDIR, "libdft_mi355x.so")), "run `make -C libdft_amd`"
    exe = tmp_path / "driver"
    cmd = (["gcc", "-std=gnu99", "-Werror=implicit-function-declaration", "-Werror=incompatible-pointer-types",
            "-I", INC] + [os.path.join(REF, s) for s in srcs] +
           ["-o", str(exe), "-L", LIBDIR, "-ldft_mi355x", "-lm"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()
