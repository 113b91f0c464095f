"""Compiles the gfx950 HIP library `libsttopt_hip.so` in place, next to this file.

The library is plain `hipcc` output with a C ABI (`include/sttopt_hip.h`); Python
binds it with ctypes (`_lib.py`), so no torch headers are involved in the build.
"""

from __future__ import annotations

import os
import pathlib
import subprocess

PKG_DIR = pathlib.Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_NAME = "libsttopt_hip.so"
LIB_PATH = PKG_DIR / LIB_NAME
ARCH = os.environ.get("STTOPT_OFFLOAD_ARCH", "gfx950")

SOURCES = ["stt_apply.hip", "stt_mg.hip", "stt_pcg.hip", "stt_grad.hip", "stt_dense.hip",
           "stt_smooth.hip", "stt_coarse_pair.hip", "stt_heat.hip",
           "stt_filter.hip", "stt_mma.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Packed fp32 instructions are kept out of every kernel.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wl,--no-undefined",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(pathlib.Path(rocm) / "bin" / "hipcc")


def build_library(verbose: bool = False) -> pathlib.Path:
    """Compile every source into `libsttopt_hip.so` (always recompiles).

    :return: the path of the library
    """
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *(str(CSRC / s) for s in SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            f"hipcc failed ({proc.returncode}):\n{proc.stdout}\n{proc.stderr}"
        )
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
