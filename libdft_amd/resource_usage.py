"""Per-kernel register, scratch, LDS and occupancy figures of the gfx950 build,
from the compiler's kernel-resource-usage remarks (no GPU needed).

    python3 libdft_amd/resource_usage.py [--filter SUBSTRING] [--src csrc/ot.hip ...]

Compiles each source for the device only with the library's flags and
-Rpass-analysis=kernel-resource-usage, demangles the kernel names and prints
one markdown row per kernel: VGPRs, AGPRs, spilled VGPRs, scratch bytes per
lane, LDS bytes per block, and the occupancy in waves per SIMD.
"""
import argparse
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# EXTRA_FLAGS: further compiler options, e.g. tuning macros (-DDFT_ZA_MINB=5)
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage", "--cuda-device-only"] + \
    os.environ.get("EXTRA_FLAGS", "").split()
KEYS = ["VGPRs", "AGPRs", "VGPRs Spill", "ScratchSize [bytes/lane]", "LDS Size [bytes/block]", "Occupancy [waves/SIMD]"]


def remarks(src):
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([HIPCC] + FLAGS + ["-c", src, "-o", os.path.join(tmp, "dev.o")], cwd=HERE,
                           capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stderr[-4000:])
        raise SystemExit(r.returncode)
    rows, cur = [], None
    for line in r.stderr.splitlines():
        m = re.search(r"remark:\s+(.*?)\s*\[-Rpass-analysis=kernel-resource-usage\]", line)
        if not m:
            continue
        key, _, val = m.group(1).partition(":")
        key, val = key.strip(), val.strip()
        if key == "Function Name":
            cur = {"name": val}
            rows.append(cur)
        elif cur is not None:
            cur[key] = val
    names = subprocess.run(["c++filt"], input="\n".join(r["name"] for r in rows), capture_output=True,
                           text=True).stdout.splitlines()
    for r, n in zip(rows, names):
        r["name"] = n
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", nargs="*", default=["csrc/ot.hip", "csrc/fft.hip"])
    ap.add_argument("--filter", default="")
    a = ap.parse_args()
    print("| kernel | VGPRs | AGPRs | VGPR spills | scratch B/lane | LDS B/block | waves/SIMD |")
    print("|---|---|---|---|---|---|---|")
    for src in a.src:
        for r in remarks(src):
            if a.filter and a.filter not in r["name"]:
                continue
            name = re.sub(r"\(.*\)$", "", r["name"])
            print("| `%s` | %s |" % (name, " | ".join(r.get(k, "") for k in KEYS)))


if __name__ == "__main__":
    main()
