"""CPU checks of the C-ABI library: it loads and exports every symbol include/tinyvlm_hip.h declares.

No compute entry point is called here (there is no GPU in the CPU suite).
"""
import ctypes


def test_header_declares_backend_surface():
    from tinyvlm_amd import capi

    names = capi.header_symbols()
    # one entry point per Backend / LU / BLAS / Memory method, in both precisions
    for base in ("lhs_assemble", "rhs_assemble_velocities", "rhs_assemble_wake_influence", "forces_steady",
                 "forces_unsteady", "coeff_cl", "coeff_cm", "mesh_metrics", "mesh_mac", "gamma_wake_from_coeffs",
                 "sum", "wake_shed", "displace_wing", "getrf", "getrs", "gemv", "gemm", "axpy", "scal", "nrm2"):
        for s in ("f64", "f32"):
            assert f"tvlm_{base}_{s}" in names
    for n in ("tvlm_init", "tvlm_malloc", "tvlm_free", "tvlm_copy", "tvlm_fill_f32", "tvlm_fill_f64",
              "tvlm_last_error", "tvlm_backend_name", "tvlm_stream_synchronize"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from tinyvlm_amd import capi

    lib = capi.lib()
    missing = [n for n in capi.header_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.tvlm_backend_name() == b"hip"


def test_library_is_gfx950_code_object():
    """The library built here carries a gfx950 device code object (our own build output)."""
    from tinyvlm_amd import capi

    blob = open(capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_error_path_without_gpu_work():
    """Argument validation happens on the host and reports through tvlm_last_error."""
    from tinyvlm_amd import capi

    lib = capi.lib()
    v = capi.View()
    st = lib.tvlm_lhs_assemble_f64(ctypes.byref(v), ctypes.c_int32(0), None, None, None, None, ctypes.c_int32(1),
                                   ctypes.c_double(0.0), None)
    assert st != 0
    assert b"null" in lib.tvlm_last_error() or b"n_surf" in lib.tvlm_last_error()
    views = (capi.View * 1)()
    st = lib.tvlm_lhs_assemble_f64(ctypes.byref(v), ctypes.c_int32(17), views, views, views, views,
                                   ctypes.c_int32(1), ctypes.c_double(0.0), None)
    assert st == 1 and b"n_surf" in lib.tvlm_last_error()


def test_native_plot3d_reader_equals_oracle(root, oracle):
    """The product-side reader (C ABI, host only) gives the oracle's mesh bit for bit, with and
    without the quarter-chord shift, for every mesh fixture."""
    import glob
    import os

    import numpy as np

    from tinyvlm_amd.mesh_io import read_plot3d

    files = sorted(glob.glob(os.path.join(root, "mesh", "*.x")))
    assert len(files) >= 10
    for f in files:
        for qc in (True, False):
            mine = read_plot3d(f, qc)
            ref = oracle.read_mesh(f, qc=qc)
            assert mine.shape == ref.shape
            assert np.array_equal(mine, ref), (f, qc)


def test_native_plot3d_reader_errors(tmp_path):
    import pytest

    from tinyvlm_amd import capi
    from tinyvlm_amd.mesh_io import read_plot3d

    bad = tmp_path / "two_blocks.x"
    bad.write_text("2\n2 2 1\n")
    with pytest.raises(capi.TvlmError):
        read_plot3d(str(bad))
    with pytest.raises(capi.TvlmError):
        read_plot3d(str(tmp_path / "missing.x"))
