"""The C++ vlm::Backend layer (tinyvlm_amd/libtinyvlm_vlm.so) and its test programs.

CPU: the tensor data model (slice / copy / fill / reshape strides of the reference's
tests/ut_tensor.cpp) on host tensors, and the library's exported Backend entry points.
GPU: the same data-model cases through device tensors, and the reference's steady known-answer
programs (tests/vlm_elliptic_coeffs.cpp: CL within 3% of lifting-line theory on elliptic_45x45;
tests/vlm_square_coeffs.cpp: CL within 1% of 2 pi alpha on four infinite-wing meshes) run by the
C++ VLM driver over BackendHIP, with every CL/CM also compared to the oracle's VLM::run on the
same mesh and angle (FP64 driver: 1e-10 relative to the FP64 oracle; FP32 driver: no further from
the FP64 oracle than 3x the FP32 oracle's own distance from it).
"""
import os
import subprocess

import numpy as np
import pytest

from tinyvlm_amd import PKG_DIR

BIN = os.path.join(PKG_DIR, "vlm", "bin")
LIB = os.path.join(PKG_DIR, "libtinyvlm_vlm.so")


def _run(args, timeout=120):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print("running", " ".join(args), flush=True)
    return subprocess.run([exe, *args[1:]], capture_output=True, text=True, timeout=timeout)


def test_ut_tensor_host():
    r = _run(["ut_tensor"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all passed" in r.stdout


def test_library_exports_backend_factory():
    r = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB], capture_output=True, text=True, check=True)
    for sym in ("vlm::create_backend", "vlm::get_available_backends", "vlm::Backend::coeff_cl_multibody",
                "vlm::Backend::displace_wing", "vlm::Backend::wake_shed", "vlm::MeshIO::read",
                "vlm::VLMT<float>::run", "vlm::VLMT<double>::run"):
        assert sym in r.stdout, sym


def test_backend_without_device_fails_loudly():
    """No GPU: creating the HIP backend must raise, never fall back to host arithmetic."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = _run(["vlm_coeffs", "square", os.path.join(os.path.dirname(PKG_DIR), "mesh")])
    assert r.returncode == 3 and "error:" in r.stdout, r.stdout


def _results(stdout):
    out = {}
    for line in stdout.splitlines():
        if line.startswith("RESULT "):
            _, label, deg, cl, cmx, cmy, cmz = line.split()
            out[(label, float(deg))] = (float(cl), np.array([float(cmx), float(cmy), float(cmz)]))
    return out


MESH_SETS = {
    "elliptic_45x45": ["elliptic_45x45.x"],
    "infinite_rectangular_2x2_parts": ["infinite_rectangular_2x2_part0.x", "infinite_rectangular_2x2_part1.x"],
    "infinite_rectangular_2x4": ["infinite_rectangular_2x4.x"],
    "infinite_rectangular_10x5": ["infinite_rectangular_10x5.x"],
    "infinite_rectangular_20x5": ["infinite_rectangular_20x5.x"],
}


@pytest.mark.gpu
def test_ut_tensor_device():
    r = _run(["ut_tensor", "--device"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "device cases run" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["square", "elliptic"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_cpp_vlm_known_answers_and_oracle(root, oracle, which, prec):
    mesh_dir = os.path.join(root, "mesh")
    r = _run(["vlm_coeffs", which, mesh_dir, prec], timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr
    res = _results(r.stdout)
    labels = ["elliptic_45x45"] if which == "elliptic" else [k for k in MESH_SETS if k.startswith("infinite")]
    for label in labels:
        verts64 = [oracle.read_mesh(os.path.join(mesh_dir, f)) for f in MESH_SETS[label]]
        verts32 = [v.astype(np.float32) for v in verts64]
        for deg in (0, 1, 2, 3, 4, 5, 10, 15):
            # FlowData's alpha: degrees * PI_f / 180 in FP32 (vlm/include/vlm_utils.hpp:68-70)
            alpha = float(np.float32(deg) * np.float32(3.141593) / np.float32(180))
            _, cl64, cm64, info = oracle.vlm_run(verts64, alpha)
            assert info == 0
            cl, cm = res[(label, float(deg))]
            scale = max(abs(cl64), 1e-3)
            cscale = max(float(np.max(np.abs(cm64))), 1e-3)
            if prec == "f64":
                assert abs(cl - cl64) <= 1e-10 * scale, (label, deg, cl, cl64)
                assert np.max(np.abs(cm - cm64)) <= 1e-10 * cscale, (label, deg, cm, cm64)
            else:
                # FP32: the GPU's distance from the FP64 answer may be at most 3x the FP32
                # oracle's own distance from it (plus a floor of a few FP32 ulps)
                _, cl32, cm32, info32 = oracle.vlm_run(verts32, alpha)
                assert info32 == 0
                tol_cl = 3 * abs(cl32 - cl64) + 1e-6 * scale
                tol_cm = 3 * float(np.max(np.abs(cm32 - cm64))) + 1e-6 * cscale
                assert abs(cl - cl64) <= tol_cl, (label, deg, cl, cl32, cl64)
                assert np.max(np.abs(cm - cm64)) <= tol_cm, (label, deg, cm, cm32, cm64)
