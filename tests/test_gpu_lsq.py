"""GPU parity of LSQ::solve (tvlm_gels_*) and of Anderson acceleration over BackendHIP.

gels: the device factors, Q^T B and X are compared with the oracle's ?gels restatement (itself
pinned to LAPACK's dgels/sgels layout) on the same inputs: FP64 within 1e-12 relative, FP32 within
2e-4; rank deficiency and all-zero A as LAPACK reports them; sizes outside the kernel's range fail
with an error instead of running.  Anderson: the C++ program ut_anderson (vlm_solvers over
BackendHIP) against the oracle's restatement of the same iteration and the closed-form fixed point.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(m, n, nrhs, dt, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, n)).astype(dt), rng.standard_normal((m, nrhs)).astype(dt)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nrhs", [(1, 1, 1), (7, 7, 1), (50, 6, 2), (1025, 10, 1), (6600, 10, 1),
                                      (4000, 3, 5), (300, 31, 1), (2049, 16, 16)])
def test_gels_matches_oracle(oracle, dt, m, n, nrhs):
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    A, B = _case(m, n, nrhs, dt, 1000 * n + m)
    a_ref, b_ref, info_ref = oracle.gels(A, B)
    tdt = torch.float64 if dt == np.float64 else torch.float32
    a = H.to_device(A, tdt)
    b = H.to_device(B, tdt)
    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    print(f"gels m={m} n={n} nrhs={nrhs} {dt.__name__}", flush=True)
    H.gels(a, b, info)
    torch.cuda.synchronize()
    assert int(info.item()) == info_ref == 0
    tol = 1e-12 if dt == np.float64 else 2e-4
    ga, gb = H.to_host(a), H.to_host(b)
    assert np.max(np.abs(ga - a_ref)) <= tol * max(1.0, float(np.max(np.abs(a_ref))))
    assert np.max(np.abs(gb - b_ref)) <= tol * max(1.0, float(np.max(np.abs(b_ref))))
    x = np.linalg.lstsq(A.astype(np.float64), B.astype(np.float64), rcond=None)[0]
    assert np.max(np.abs(gb[:n] - x)) <= 10 * tol * max(1.0, float(np.max(np.abs(x))))


def test_gels_vector_rhs_and_strided_columns(oracle):
    """B as a 1-D vector (the reference reshapes gamma to {n,1}) and A as a column slice of a wider
    history buffer (leading dimension > M is not used; a slice keeps stride[1] = M)."""
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    A, B = _case(500, 10, 1, np.float64, 3)
    hist = H.to_device(np.asfortranarray(np.hstack([A, np.ones((500, 4))])), torch.float64)
    a = hist[:, :10]
    b = H.to_device(B[:, 0].copy(), torch.float64)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.gels(a, b, info)
    _, b_ref, _ = oracle.gels(A, B)
    assert int(info.item()) == 0
    assert np.max(np.abs(H.to_host(b) - b_ref[:, 0])) <= 1e-12 * float(np.max(np.abs(b_ref)))
    assert np.all(H.to_host(hist[:, 10:]) == 1.0)  # columns beyond the slice untouched


def test_gels_rank_deficient_zero_and_errors(oracle):
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    A = np.zeros((8, 3))
    A[:, 0] = np.arange(8.0)
    A[:, 2] = 1.0
    a = H.to_device(A, torch.float64)
    b = H.to_device(np.ones((8, 1)), torch.float64)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.gels(a, b, info)
    assert int(info.item()) == 2 == oracle.gels(A, np.ones((8, 1)))[2]
    z = H.to_device(np.zeros((5, 2)), torch.float64)
    bz = H.to_device(np.ones((5, 1)), torch.float64)
    H.gels(z, bz, info)
    assert int(info.item()) == 0 and np.all(H.to_host(bz) == 0)
    with pytest.raises(H.TvlmError):  # underdetermined
        H.gels(H.to_device(np.ones((3, 5)), torch.float64), H.to_device(np.ones((3, 1)), torch.float64), info)
    with pytest.raises(H.TvlmError):  # N + NRHS > 32
        H.gels(H.to_device(np.ones((100, 30)), torch.float64), H.to_device(np.ones((100, 3)), torch.float64), info)


def _anderson_problem(n=64):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    M = 0.9 * np.sin(1.0 + i + 3.0 * j) / n
    c = np.cos(np.arange(n, dtype=np.float64))
    return M, c


def _parse(out, tag):
    k = None
    x = {}
    for line in out.splitlines():
        if line.startswith(tag + "ITER "):
            k = int(line.split()[1])
        elif line.startswith(tag + "X "):
            _, i, v = line.split()
            x[int(i)] = float(v)
    return k, np.array([x[i] for i in range(len(x))])


@pytest.mark.parametrize("m", [1, 3, 10])
def test_cpp_anderson_matches_oracle(oracle, m):
    from tinyvlm_amd import PKG_DIR

    exe = os.path.join(PKG_DIR, "vlm", "bin", "ut_anderson")
    print("running ut_anderson", m, flush=True)
    r = subprocess.run([exe, str(m), "1e-12", "1e-5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    M, c = _anderson_problem()
    exact = np.linalg.solve(np.eye(64) - M, c)
    k, x = _parse(r.stdout, "")
    xo, ko = oracle.anderson_acceleration(np.zeros(64), lambda v: M @ v + c, 100, 1e-12, m)
    assert abs(k - ko) <= 1, (k, ko)
    assert np.max(np.abs(x - exact)) < 1e-10
    assert np.max(np.abs(x - xo)) < 1e-10
    k32, x32 = _parse(r.stdout, "F32")
    M32, c32 = M.astype(np.float32), c.astype(np.float32)
    xo32, ko32 = oracle.anderson_acceleration(np.zeros(64, dtype=np.float32), lambda v: M32 @ v + c32, 100,
                                              np.float32(1e-5), m)
    assert k32 < 100 and ko32 < 100
    assert np.max(np.abs(x32 - exact)) < 1e-4


@pytest.mark.parametrize("m", [1, 3, 10])
def test_python_anderson_matches_oracle(oracle, m):
    """hb.AndersonSolver (device vectors, the HB driver's loop) on the same contraction: the oracle's
    iteration count +-1 and the closed-form fixed point within 1e-10."""
    import torch

    from tinyvlm_amd import capi as H
    from tinyvlm_amd.hb import AndersonSolver

    H.init(0)
    M, c = _anderson_problem()
    exact = np.linalg.solve(np.eye(64) - M, c)
    Md, cd = H.to_device(np.asfortranarray(M), torch.float64), H.to_device(c, torch.float64)

    def f(x, y):
        y.copy_(Md @ x + cd)

    print(f"python anderson m={m}", flush=True)
    s = AndersonSolver(64, m)
    x = torch.zeros(64, dtype=torch.float64, device="cuda")
    k = s.run(x, f, 100, 1e-12)
    xo, ko = oracle.anderson_acceleration(np.zeros(64), lambda v: M @ v + c, 100, 1e-12, m)
    assert abs(k - ko) <= 1, (k, ko)
    assert np.max(np.abs(H.to_host(x) - exact)) < 1e-10


def test_python_anderson_rank_deficient_history_raises():
    """A history whose first column is zero and second is not (the map leaves the residual unchanged
    once, then changes it) makes R(0,0) = 0: gels reports info = 1 and the solver raises."""
    import torch

    from tinyvlm_amd import capi as H
    from tinyvlm_amd.hb import AndersonSolver

    H.init(0)
    c = torch.linspace(0.5, 1.5, 64, dtype=torch.float64, device="cuda")
    calls = [0]

    def f(x, y):
        calls[0] += 1
        # x0 = 0 -> c -> 2c with residual c twice (first history column 0); then f(2c) = 4c, residual
        # 2c, second column c: the k = 2 least squares has A = [0, c]
        y.copy_(x + c if calls[0] <= 2 else 0.5 * x + 3.0 * c)

    s = AndersonSolver(64, 3)
    with pytest.raises(H.TvlmError, match="rank-deficient"):
        s.run(torch.zeros(64, dtype=torch.float64, device="cuda"), f, 100, 1e-12)
