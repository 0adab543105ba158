"""Build the native libraries of craton_bolt_amd in-tree.

* ``libbolt_hip.so``     — every ``csrc/*.hip`` file, compiled by hipcc for
  gfx950 only and linked with hipRTC.
* ``libbolt_datagen.so`` — the host-only seeded table generators (gcc, OpenMP).

Objects go to ``_build/`` at the repository root; the shared libraries are
written next to this file so they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{ROOT / '_build'}",
]

LIB_HIP = PKG / "libbolt_hip.so"
LIB_GEN = PKG / "libbolt_datagen.so"


def _run(cmd: list[str], quiet: bool) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in p.stdout.splitlines() if "packed-fp32-ops" not in ln  # host-pass notice
    )
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and not quiet:
        print(out)


def _write_jit_prelude() -> None:
    """Embed device_acc.h, gb_lds.h and jp_tile.h as a string literal: they are
    prepended to the kernels the library generates per query and compiles with
    hipRTC."""
    text = "".join((CSRC / h).read_text().replace("#pragma once\n", "")
                   for h in ("device_acc.h", "gb_lds.h", "jp_tile.h"))
    out = OBJ / "jit_prelude.inc"
    body = 'static const char kJitPrelude[] = R"BOLTJIT(\n' + text + '\n)BOLTJIT";\n'
    if not out.exists() or out.read_text() != body:
        out.write_text(body)


def _compile_hip(src: Path, quiet: bool) -> Path:
    obj = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)], quiet)
    return obj


def build_hip(quiet: bool = True, jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    _write_jit_prelude()
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile_hip(s, quiet), srcs))
    tmp = LIB_HIP.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"], quiet)
    os.replace(tmp, LIB_HIP)
    return LIB_HIP


TEST_KERNEL = ROOT / "tests" / "kernels" / "abi_fill.hip"


def build_test_kernels(quiet: bool = True) -> Path:
    """The stand-alone code object the runtime-ABI tests load through
    bolt_hip_module_load_data_ex."""
    out = TEST_KERNEL.with_suffix(".hsaco")
    tmp = out.with_suffix(".hsaco.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "--genco", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-o", str(tmp), str(TEST_KERNEL)], quiet)
    os.replace(tmp, out)
    return out


def build_datagen(quiet: bool = True) -> Path:
    cc = shutil.which("gcc") or "cc"
    tmp = LIB_GEN.with_suffix(".so.tmp")
    _run([cc, "-O3", "-fopenmp", "-shared", "-fPIC", "-std=c11", "-o", str(tmp), str(CSRC / "datagen.c"), "-lm"],
         quiet)
    os.replace(tmp, LIB_GEN)
    return LIB_GEN


def build_all(quiet: bool = True) -> None:
    build_datagen(quiet)
    build_hip(quiet)
    build_test_kernels(quiet)


if __name__ == "__main__":
    build_all(quiet="-v" not in sys.argv)
    print(f"built {LIB_HIP} and {LIB_GEN}")
