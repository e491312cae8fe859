"""Warm-start network kernels on the device against the FP64 oracle (oracle/pymlp.py): the
BF16 MFMA lane layout, forward, gradients (BF16 operands, FP32 accumulation), the Adam step and
training on a teacher network."""
import ctypes as C

import numpy as np
import pytest

import pymlp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_mfma_bf16_16x16x16_layout(gpu):
    L = gpu.lib()
    L.lkd_probe_mfma_bf16.argtypes = [C.c_void_p] * 4
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32)     # small integers: exact in bf16
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32)
    lanes = np.arange(64)
    # lane l holds A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16], i = 0..3
    a_in = np.stack([A[lanes % 16, 4 * (lanes // 16) + i] for i in range(4)], 1).ravel()
    b_in = np.stack([B[4 * (lanes // 16) + i, lanes % 16] for i in range(4)], 1).ravel()
    da, db = gpu.DeviceBuffer(a_in), gpu.DeviceBuffer(b_in)
    out = gpu.DeviceBuffer(256 * 4, np.float32)
    assert L.lkd_probe_mfma_bf16(da.addr, db.addr, out.addr, None) == 0
    got = out.download().reshape(64, 4)
    D = A @ B
    exp = np.array([[D[4 * (l // 16) + e, l % 16] for e in range(4)] for l in range(64)])
    np.testing.assert_array_equal(got, exp)


CASES = [((37, 70, 50, 45), 33), ((100, 128, 96, 300), 64), ((256, 512, 512, 130), 200)]


@pytest.mark.parametrize("dims,batch", CASES)
def test_forward_matches_oracle(gpu, dims, batch):
    net = gpu.MLP(*dims, max_batch=batch, seed=11)
    p = net.params().astype(np.float64)
    u = np.random.default_rng(1).normal(size=(batch, dims[0])).astype(np.float32)
    y = net.forward(u)
    y_bf = pymlp.forward(p, dims, u, bf=True)[2]
    y_64 = pymlp.forward(p, dims, u, bf=False)[2]
    assert _rel(y, y_bf) < 2e-3          # FP32 accumulation against exact sums of BF16 products
    assert _rel(y, y_64) < 3e-2          # BF16 operands against exact FP64
    net.close()


@pytest.mark.parametrize("dims,batch", CASES)
def test_gradients_match_oracle(gpu, dims, batch):
    net = gpu.MLP(*dims, max_batch=batch, seed=12)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(2)
    u = rng.normal(size=(batch, dims[0])).astype(np.float32)
    t = rng.normal(size=(batch, dims[3])).astype(np.float32)
    g, loss = net.gradients(u, t)
    g_bf, loss_bf = pymlp.gradients(p, dims, u, t, bf=True)
    assert abs(loss - loss_bf) < 1e-4 * loss_bf
    assert _rel(g, g_bf) < 5e-3
    # block by block, so that no parameter group hides behind a larger one
    o = 0
    for blk in pymlp.split(g_bf, dims):
        k = blk.size
        assert _rel(g[o:o + k], blk.ravel()) < 1e-2
        o += k
    net.close()


def test_adam_first_step(gpu):
    dims = (40, 64, 48, 30)
    net = gpu.MLP(*dims, max_batch=32, seed=5)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(32, 40)).astype(np.float32)
    t = rng.normal(size=(32, 30)).astype(np.float32)
    p0 = net.params().astype(np.float64)
    g, _ = net.gradients(u, t)
    net.set_dataset(u, t)
    lr = 1e-3
    net.train(1, 32, lr)           # one full batch: the rows in a shuffled order
    p1 = net.params().astype(np.float64)
    g = g.astype(np.float64)
    # first Adam step with bias correction: m_hat = g, v_hat = g^2
    want = p0 - lr * g / (np.abs(g) + 1e-8)
    big = np.abs(g) > 1e-4         # where the summation order cannot flip the update's size
    np.testing.assert_allclose(p1[big], want[big], rtol=0, atol=2e-6 * lr + 1e-6)
    assert np.all(np.abs(p1 - p0) <= lr * 1.0001 + 1e-7)
    net.close()


def test_training_fits_a_teacher_and_is_deterministic(gpu):
    dims = (32, 128, 128, 64)
    rng = np.random.default_rng(4)
    teacher = gpu.MLP(*dims, max_batch=256, seed=99)
    u = rng.normal(size=(2048, 32)).astype(np.float32)
    t = teacher.forward(u)
    teacher.close()
    runs = []
    for _ in range(2):
        net = gpu.MLP(*dims, max_batch=256, seed=7)
        net.set_dataset(u, t)
        hist = net.train(40, 256, 2e-3, seed=1)
        runs.append((hist, net.params()))
        net.close()
    hist = runs[0][0]
    assert hist[-1] < 0.2 * hist[0], hist
    assert np.array_equal(runs[0][1], runs[1][1])       # fixed summation order: bitwise repeatable
    assert np.array_equal(runs[0][0], runs[1][0])
