"""GPU checks of getrf / getrs at the largest served size (N = 16384, the 128 x 128 mesh of the
steady-solve mix), where the oracle (LAPACK through SciPy) would take minutes: size-independent
properties instead.

- every pivot is a row at or below its column (1-based, LAPACK's convention) and info = 0;
- backward error of the factorisation: with P the row interchanges applied in order,
  ||P A - L U||_F / ||A||_F <= 4 N eps (computed in FP64 on the device; a dropped update, a wrong
  interchange or a stale panel gives O(1));
- the solve: ||A x - b|| / (||A|| ||x||) <= 8 sqrt(N) eps for one right-hand side (measured
  1.1e-15 FP64, 5.8e-7 FP32, against bounds 2.3e-13 and 1.2e-4).
FP32 and FP64; both take the cooperative panels (several work-groups per tall panel) and the
multi-work-group triangular solves.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 16384


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_getrf_getrs_full_size_backward_error(prec):
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    print(f"[case] lu full size n={N} {prec}", flush=True)
    dt = torch.float32 if prec == "f32" else torch.float64
    eps = float(torch.finfo(dt).eps)
    g = torch.Generator(device="cuda")
    g.manual_seed(123)
    a0 = torch.randn(N, N, dtype=torch.float64, device="cuda", generator=g)
    a = a0.to(dt).t().contiguous().t()  # column-major
    a0 = a.to(torch.float64).clone()  # the matrix actually factored (a copy: getrf works in place)
    ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    b = torch.randn(N, 1, dtype=torch.float64, device="cuda", generator=g).to(dt).clone()
    x = b.clone()  # an N x 1 column is column-major as it is; getrs overwrites x only
    H.getrs(a, ipiv, x)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    piv = ipiv.cpu().numpy().astype(np.int64) - 1
    assert np.all(piv >= np.arange(N)) and np.all(piv < N)
    # P A: the interchanges applied in order to the rows of A
    perm = np.arange(N)
    for k in range(N):
        p = piv[k]
        if p != k:
            perm[k], perm[p] = perm[p], perm[k]
    pa = a0[torch.as_tensor(perm, device="cuda")]
    lu = a.to(torch.float64)
    L = torch.tril(lu, -1) + torch.eye(N, dtype=torch.float64, device="cuda")
    U = torch.triu(lu)
    del lu
    r = torch.linalg.norm(pa - L @ U) / torch.linalg.norm(a0)
    del L, U, pa
    bound = 4 * N * eps
    assert float(r) <= bound, (float(r), bound)
    xs = x.to(torch.float64)
    res = torch.linalg.norm(a0 @ xs - b.to(torch.float64)) / (torch.linalg.norm(a0) * torch.linalg.norm(xs))
    bound_solve = 8 * N ** 0.5 * eps
    assert float(res) <= bound_solve, (float(res), bound_solve)
