"""The C-ABI library loads and exports every function that include/*.h declares (no GPU calls)."""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "libdft_amd", "libdft_mi355x.so")


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"#[^\n]*", "", src)
        # function declarations: "<type> name(args);" at top level
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_built():
    assert os.path.exists(LIB), "run `make -C libdft_amd`"


def test_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = _declared()
    assert len(declared) > 60
    missing = [n for n in declared if n not in exported]
    assert not missing, missing


def test_ctypes_binding_loads():
    import libdft_amd as L
    lib = L.lib()
    for name in L.exported_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the sizes the C compiler gives the structs."""
    import ctypes

    import libdft_amd as L
    code = r'''
#include <stdio.h>
#include <stddef.h>
#include "libdft_mi355x.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rgrid), sizeof(cgrid), sizeof(wf), sizeof(dft_ot_functional),
         offsetof(dft_ot_functional, density), offsetof(dft_ot_functional, device),
         offsetof(rgrid, value_outside), offsetof(cgrid, outside_params_ptr));
  return 0;
}
'''
    src, exe = str(tmp_path / "layout.c"), str(tmp_path / "layout")
    open(src, "w").write(code)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(L.RGrid), ctypes.sizeof(L.CGrid), ctypes.sizeof(L.WF), ctypes.sizeof(L.OTF),
            L.OTF.density.offset, L.OTF.device.offset, L.RGrid.value_outside.offset,
            L.CGrid.outside_params_ptr.offset]
    assert got == want


def test_oracle_is_not_linked():
    """The product library does not depend on the oracle."""
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out
    out = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "orc_" not in out
