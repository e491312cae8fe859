"""Lane layouts of the MFMA builtins and the 64x64 tile GEMM used by the Cholesky kernels."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(gpu, a):
    return gpu.DeviceBuffer(np.ascontiguousarray(a))


@pytest.mark.parametrize("prec,dt", [(0, np.float64), (1, np.float32)])
def test_mfma_16x16x4_layout(gpu, prec, dt):
    L = gpu.lib()
    L.lkd_probe_mfma.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (16, 4)).astype(dt)
    B = rng.integers(-4, 5, (4, 16)).astype(dt)
    lanes = np.arange(64)
    a_in = A[lanes % 16, lanes // 16]        # A[i][k] in lane i + 16 k
    b_in = B[lanes // 16, lanes % 16]        # B[k][j] in lane j + 16 k
    da, db = _dev(gpu, a_in), _dev(gpu, b_in)
    out = gpu.DeviceBuffer(256 * np.dtype(dt).itemsize, dt)
    assert L.lkd_probe_mfma(prec, da.addr, db.addr, out.addr, None) == 0
    got = out.download().reshape(64, 4)
    D = A @ B
    row = (lambda l, e: l // 16 + 4 * e) if prec == 0 else (lambda l, e: 4 * (l // 16) + e)
    exp = np.array([[D[row(l, e), l % 16] for e in range(4)] for l in range(64)])
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("prec,dt,K,ra,rb", [(0, np.float64, 64, 64, 64), (0, np.float64, 37, 50, 21),
                                             (1, np.float32, 128, 64, 64), (0, np.float64, 200, 64, 9)])
def test_tile_gemm(gpu, prec, dt, K, ra, rb):
    L = gpu.lib()
    L.lkd_probe_tile_gemm.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                      C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(K)
    lda, ldb = 70, 90
    A = rng.standard_normal((K, lda)).astype(dt)     # column k of the panel = A[k, :]
    B = rng.standard_normal((K, ldb)).astype(dt)
    da, db = _dev(gpu, A), _dev(gpu, B)
    out = gpu.DeviceBuffer(64 * 64 * np.dtype(dt).itemsize, dt)
    assert L.lkd_probe_tile_gemm(prec, da.addr, lda, ra, db.addr, ldb, rb, K, out.addr, None) == 0
    got = out.download().reshape(64, 64).T            # [r][c]
    Am = np.zeros((64, K)); Am[:ra] = A[:, :ra].T
    Bm = np.zeros((64, K)); Bm[:rb] = B[:, :rb].T
    exp = Am @ Bm.T
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert np.abs(got - exp).max() <= tol * max(1.0, np.abs(exp).max())
