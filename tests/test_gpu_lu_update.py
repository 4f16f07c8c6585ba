"""GPU check of getrf's trailing update (tvlm_lu_update: C -= A B on the matrix cores, batched)
against a vendor GEMM (torch.matmul in FP64) at the shapes getrf uses and at ragged edges.

Bound, elementwise: |C_gpu - (C - A B)| <= (k + 2) eps (|A| |B| + |C|) (the standard error bound of
a dot product of length k plus the final subtraction); the padding rows below m inside each
column (ld > m) and everything outside the m x n window stay untouched, bit for bit.
"""
import ctypes

import numpy as np
import pytest

CASES = [
    # prec, m, n, k, batch, ld
    ("f32", 300, 170, 64, 3, 320),     # ragged edges, 64 x 64 tiles
    ("f32", 130, 120, 128, 2, 160),    # two 64-deep slabs
    ("f32", 1000, 64, 64, 5, 1024),    # narrow update of the look-ahead path
    ("f32", 100, 900, 37, 1, 128),     # k not a multiple of the 16-deep slice
    ("f32", 4032, 4032, 64, 4, 4096),  # served batch shape, 128 x 64 tiles
    ("f64", 300, 170, 64, 3, 320),
    ("f64", 100, 77, 100, 2, 101),
    ("f64", 2049, 1500, 128, 1, 2112),
    ("f64", 16320, 256, 64, 1, 16384),
]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,m,n,k,batch,ld", CASES)
def test_trailing_update_matches_vendor_gemm(prec, m, n, k, batch, ld):
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    print(f"[case] lu update {prec} m={m} n={n} k={k} batch={batch} ld={ld}", flush=True)
    dt = torch.float32 if prec == "f32" else torch.float64
    eb = 4 if prec == "f32" else 8
    eps = float(torch.finfo(dt).eps)
    cols = max(n, k) + 3
    stride = ld * cols + 5  # matrices of one batch do not start on aligned boundaries
    g = torch.Generator(device="cuda")
    g.manual_seed(m * 7 + n * 13 + k)
    buf = torch.randn(batch * stride + 64, dtype=torch.float64, device="cuda", generator=g).to(dt)
    # C, A and B: column-major, leading dimension ld, matrix z at offset z * stride of its buffer
    abuf = torch.randn(batch * stride + 64, dtype=torch.float64, device="cuda", generator=g).to(dt)
    bbuf = torch.randn(batch * stride + 64, dtype=torch.float64, device="cuda", generator=g).to(dt)
    before = buf.clone()

    def view(t, z, rows, ncols):
        return t.as_strided((rows, ncols), (1, ld), z * stride)

    ref = []
    for z in range(batch):
        A = view(abuf, z, m, k).to(torch.float64)
        B = view(bbuf, z, k, n).to(torch.float64)
        C = view(buf, z, m, n).to(torch.float64)
        ref.append((C - A.contiguous() @ B.contiguous(), A.abs().contiguous() @ B.abs().contiguous() + C.abs()))
    H.check(H.lib().tvlm_lu_update(ctypes.c_int32(eb), ctypes.c_void_p(buf.data_ptr()), ctypes.c_int64(ld),
                                   ctypes.c_void_p(abuf.data_ptr()), ctypes.c_int64(ld),
                                   ctypes.c_void_p(bbuf.data_ptr()), ctypes.c_int64(ld), ctypes.c_int64(m),
                                   ctypes.c_int64(n), ctypes.c_int64(k), ctypes.c_int32(batch), ctypes.c_int64(stride),
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "tvlm_lu_update")
    torch.cuda.synchronize()
    mask = torch.ones_like(buf, dtype=torch.bool)
    for z in range(batch):
        got = view(buf, z, m, n).to(torch.float64)
        want, scale = ref[z]
        err = (got - want).abs()
        bound = (k + 2) * eps * scale
        worst = float((err - bound).max())
        assert worst <= 0.0, (z, float(err.max()), float((err / scale.clamp_min(1e-300)).max()))
        view(mask, z, m, n).fill_(False)
    assert torch.equal(buf[mask], before[mask])


def test_trailing_update_rejects_bad_arguments():
    """Argument checks only: no GPU work, runs without a GPU."""
    from tinyvlm_amd import capi as H

    lib = H.lib()
    p = ctypes.c_void_p(1)
    assert lib.tvlm_lu_update(ctypes.c_int32(2), p, ctypes.c_int64(10), p, ctypes.c_int64(10), p, ctypes.c_int64(10),
                              ctypes.c_int64(10), ctypes.c_int64(10), ctypes.c_int64(10), ctypes.c_int32(1),
                              ctypes.c_int64(0), None) != 0
    # leading dimension of C below m
    assert lib.tvlm_lu_update(ctypes.c_int32(4), p, ctypes.c_int64(5), p, ctypes.c_int64(10), p, ctypes.c_int64(10),
                              ctypes.c_int64(10), ctypes.c_int64(10), ctypes.c_int64(10), ctypes.c_int32(1),
                              ctypes.c_int64(0), None) != 0
    # an empty update does nothing and succeeds, whatever the pointers
    assert lib.tvlm_lu_update(ctypes.c_int32(4), None, ctypes.c_int64(1), None, ctypes.c_int64(1), None,
                              ctypes.c_int64(1), ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0),
                              ctypes.c_int32(1), ctypes.c_int64(0), None) == 0
