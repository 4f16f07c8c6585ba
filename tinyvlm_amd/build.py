"""Compiles the native parts in-tree: the HIP backend (gfx950) and the C++ host layer."""
import os
import subprocess

from . import PKG_DIR


def build_hip(jobs: int = 8) -> str:
    """hipcc --offload-arch=gfx950 via csrc/Makefile -> tinyvlm_amd/libtinyvlm_hip.so (always rebuilds stale objects)."""
    csrc = os.path.join(PKG_DIR, "csrc")
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", csrc], check=True)
    return os.path.join(PKG_DIR, "libtinyvlm_hip.so")


def build_all(jobs: int = 8):
    return [build_hip(jobs)]
