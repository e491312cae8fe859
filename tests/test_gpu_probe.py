"""Lane layouts of the MFMA builtins, the 128x64 tile GEMM (NT and transposed TN accumulator
layouts), its FP64 direct-to-LDS forms (128x64 and 128x128) and the 64x64 diagonal
factorization used by the Cholesky kernels."""
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


@pytest.mark.parametrize("prec,dt,K,ra,rb", [(0, np.float64, 64, 128, 64), (0, np.float64, 37, 50, 21),
                                             (1, np.float32, 128, 128, 64), (0, np.float64, 200, 100, 9),
                                             (1, np.float32, 70, 77, 33), (0, np.float64, 3, 128, 64)])
def test_tile_gemm(gpu, prec, dt, K, ra, rb):
    L = gpu.lib()
    L.lkd_probe_tile_gemm.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                      C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(K)
    lda, ldb = 130, 90
    A = rng.standard_normal((K, lda)).astype(dt)     # column k of the panel = A[k, :]
    B = rng.standard_normal((K, ldb)).astype(dt)
    da, db = _dev(gpu, A), _dev(gpu, B)
    out = gpu.DeviceBuffer(2 * 128 * 64 * np.dtype(dt).itemsize, dt)
    assert L.lkd_probe_tile_gemm(prec, da.addr, lda, ra, db.addr, ldb, rb, K, out.addr, None) == 0
    both = out.download().reshape(2, 64, 128)
    got, got_tn = both[0].T, both[1].T                 # [r][c] from the NT and the TN core
    Am = np.zeros((128, K)); Am[:ra] = A[:, :ra].T
    Bm = np.zeros((64, K)); Bm[:rb] = B[:, :rb].T
    exp = Am @ Bm.T
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert np.abs(got - exp).max() <= tol * max(1.0, np.abs(exp).max())
    assert np.abs(got_tn - exp).max() <= tol * max(1.0, np.abs(exp).max())


@pytest.mark.parametrize("K,ra,rb", [(64, 128, 128), (48, 77, 99), (256, 1, 65), (16, 128, 64), (160, 63, 127)])
def test_tile_gemm_direct_to_lds(gpu, K, ra, rb):
    """The FP64 STEP / PART products staged with direct-to-LDS loads (one 128 x 64 and one
    128 x 128 product) equal the register-staged TN core bit for bit, and the product."""
    L = gpu.lib()
    L.lkd_probe_tile_gemm_dma.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64, C.c_int,
                                          C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(K + ra)
    lda, ldb = 130, 136
    A = rng.standard_normal((K, lda))
    B = rng.standard_normal((K, ldb))
    da, db = _dev(gpu, A), _dev(gpu, B)
    out = gpu.DeviceBuffer(4 * 128 * 64 * 8, np.float64)
    assert L.lkd_probe_tile_gemm_dma(da.addr, lda, ra, db.addr, ldb, rb, K, out.addr, None) == 0
    reg, dma, h0, h1 = (m.T for m in out.download().reshape(4, 64, 128))
    np.testing.assert_array_equal(dma, reg)
    np.testing.assert_array_equal(h0, reg)
    exp = A[:, :ra].T @ B[:, :rb]
    got = np.concatenate([h0, h1], axis=1)[:ra, :rb]
    assert np.abs(got - exp).max() <= 1e-12 * max(1.0, np.abs(exp).max())


@pytest.mark.parametrize("prec,dt,kb", [(0, np.float64, 64), (0, np.float64, 17), (1, np.float32, 64)])
def test_diag_potrf(gpu, prec, dt, kb):
    """The in-LDS factorization of DIAG (identity padding past kb), against numpy."""
    L = gpu.lib()
    L.lkd_probe_potrf.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(kb)
    G = rng.standard_normal((kb, kb + 8))
    A = np.eye(64)
    A[:kb, :kb] = G @ G.T / kb + np.eye(kb)
    da = _dev(gpu, A.T.astype(dt))                   # column-major
    ok = gpu.DeviceBuffer(np.zeros(1, np.int32))
    assert L.lkd_probe_potrf(prec, da.addr, ok.addr, None) == 0
    assert ok.download()[0] == 1
    got = da.download().reshape(64, 64).T
    exp = np.linalg.cholesky(A)
    tol = 1e-13 if dt == np.float64 else 2e-5
    assert np.abs(got - exp).max() <= tol * np.abs(exp).max()
    # a matrix that is not positive definite is reported, not factored
    bad = A.copy()
    bad[3, 3] = -1.0
    db = _dev(gpu, bad.T.astype(dt))
    assert L.lkd_probe_potrf(prec, db.addr, ok.addr, None) == 0
    assert ok.download()[0] == 0
