"""Compiles the native parts in-tree: the HIP backend (gfx950) and the C++ host layer."""
import os
import subprocess

from . import PKG_DIR


def _make(subdir: str, jobs: int, force: bool):
    cmd = ["make", "-s", f"-j{jobs}", "-C", os.path.join(PKG_DIR, subdir)]
    subprocess.run(cmd + (["-B"] if force else []), check=True)


def build_hip(jobs: int = 8, force: bool = False) -> str:
    """hipcc --offload-arch=gfx950 via csrc/Makefile -> tinyvlm_amd/libtinyvlm_hip.so (force: rebuild everything)."""
    _make("csrc", jobs, force)
    return os.path.join(PKG_DIR, "libtinyvlm_hip.so")


def build_vlm_cpp(jobs: int = 8, force: bool = False) -> str:
    """g++ via vlm/Makefile -> tinyvlm_amd/libtinyvlm_vlm.so (vlm::Backend over the C ABI) and vlm/bin/*."""
    _make("vlm", jobs, force)
    return os.path.join(PKG_DIR, "libtinyvlm_vlm.so")


def build_all(jobs: int = 8, force: bool = False):
    return [build_hip(jobs, force), build_vlm_cpp(jobs, force)]
