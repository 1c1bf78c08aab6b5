"""Build recipes for the native parts of ntlp_amd.

``build_library()`` compiles the HIP droplet-step library for gfx950 in-tree
(ntlp_amd/libntlp_gpu.so); ``build_fortran()`` compiles the ISO_C_BINDING
module and the Fortran drop-in driver against it when a Fortran compiler is
present.  Both always recompile.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libntlp_gpu.so")
CSRC = os.path.join(HERE, "csrc")
FORTRAN_DIR = os.path.join(ROOT, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-shared",
    "-fPIC",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused only within one source expression, before any loop
    # hoisting: every kernel that inlines the same droplet arithmetic (the
    # multi-step kernel, the single-call update kernels, the serving step)
    # rounds it identically
    "-ffp-contract=on",
    # no machine-level hoisting of loop invariants: the droplet kernels' Newton
    # loops otherwise keep hoisted copies of constants live in vector
    # registers (fewer registers, measured faster step)
    "-mllvm", "-disable-machine-licm",
]


def _run(cmd: list[str], cwd: str | None = None) -> str:
    """Runs cmd; raises with its output on failure, echoes anything it printed
    except the host pass's packed-FP32 notice and the resource remarks, and
    returns the full output."""
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized" not in l and "kernel-resource-usage" not in l
                    and not l.startswith("In file included from"))
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")
    return proc.stdout


# per-kernel register, scratch and occupancy figures of the last library build
RESOURCES = os.path.join(HERE, "kernel_resources.json")
_REMARK = re.compile(r"remark:\s+(Function Name|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|"
                     r"Occupancy \[waves/SIMD\]|VGPRs Spill|SGPRs Spill|LDS Size \[bytes/block\]):\s+(\S+)")


def parse_resources(text: str) -> dict:
    """{mangled kernel name: {field: int}} from -Rpass-analysis=kernel-resource-usage."""
    res, cur = {}, None
    for line in text.splitlines():
        m = _REMARK.search(line)
        if not m:
            continue
        key, val = m.group(1), m.group(2)
        if key == "Function Name":
            cur = res.setdefault(val, {})
        elif cur is not None:
            cur[key.split(" [")[0]] = int(val)
    return res


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build_library() -> str:
    """Compile csrc/ntlp_gpu.hip into ntlp_amd/libntlp_gpu.so (gfx950)."""
    tmp = LIB + ".tmp"
    out = _run([hipcc(), *HIPCC_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-fno-caret-diagnostics", "-o", tmp,
                os.path.join(CSRC, "ntlp_gpu.hip")])
    os.replace(tmp, LIB)
    with open(RESOURCES, "w") as f:
        json.dump(parse_resources(out), f, indent=1, sort_keys=True)
    return LIB


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang"),
                 shutil.which("gfortran")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran() -> str | None:
    """Compile fortran/ntlp_gpu_mod.f90 and the drop-in driver; returns the
    driver path, or None when no Fortran compiler exists."""
    fc = flang()
    if fc is None:
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    exe = os.path.join(FORTRAN_BUILD, "ntlp_dropin_driver")
    _run([fc, "-O2", "-J", FORTRAN_BUILD, "-o", exe,
          os.path.join(FORTRAN_DIR, "ntlp_gpu_mod.f90"),
          os.path.join(FORTRAN_DIR, "ntlp_dropin_driver.f90"),
          "-L" + HERE, "-lntlp_gpu", "-Wl,-rpath," + HERE])
    return exe


if __name__ == "__main__":
    print(build_library())
    print(build_fortran())
