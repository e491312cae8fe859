"""Normal-equations assembly (M0_ij = sum_k A_ik Theta_k A_jk, src/cuda/kkt.cu:177-215).

CPU: the host planner's pair map holds exactly the products of A -- every segment (i, j) of the
permuted lower triangle lists each column k with A[P_i, k] and A[P_j, k] nonzero once, with
pp = A[P_i, k] A[P_j, k] -- and the oracle's fixed-order segment sums (oracle.c
or_segment_sums, exact fma) equal A Theta A^T to rounding.
GPU: the device passes (lkd_assemble_diag, lkd_assemble_pairs) reproduce the oracle bit for
bit: M0's diagonal, D, every equilibrated entry, and the panel chunks (zeros elsewhere)."""
import ctypes as C

import numpy as np
import pytest


def _cases():
    from lakshya import gen
    return {
        "random_sparse": gen.random_sparse(60, 150, 4, seed=3),
        # linking rows at 20% density: diagonal and cross segments of every length class
        "block_angular": gen.block_angular(4, 60, 150, 5, 6, 0.2, seed=3),
        "dense_columns": gen.dense_columns(80, 200, 3, 2, seed=4),
    }


def _plan(native, lp):
    rp, ci, va = lp.csr()
    return rp, ci, va, native.plan_export(rp, ci, va, lp.n)


@pytest.mark.parametrize("name", ["random_sparse", "block_angular", "dense_columns"])
def test_pair_map_holds_the_products_of_A(native, name):
    lp = _cases()[name]
    rp, ci, va, p = _plan(native, lp)
    A = lp.to_scipy().toarray()
    P = p["perm"]
    Ap = A[P]
    nz = Ap != 0
    seg = p["segptr"]
    L = np.diff(seg)
    assert (L >= 1).all()
    seen = set()
    for t in range(len(L)):
        i, j = int(p["segi"][t]), int(p["segj"][t])
        assert i >= j and (i, j) not in seen
        seen.add((i, j))
        ks = np.flatnonzero(nz[i] & nz[j])
        got = sorted(zip(p["pk"][seg[t]:seg[t + 1]].tolist(), p["pp"][seg[t]:seg[t + 1]].tolist()))
        assert [k for k, _ in got] == ks.tolist()
        np.testing.assert_array_equal([v for _, v in got], Ap[i, ks] * Ap[j, ks])
    # every structural nonzero of the lower triangle of A A^T has its segment
    S = (nz.astype(int) @ nz.T.astype(int)) > 0
    assert seen == {(i, j) for i, j in zip(*np.nonzero(np.tril(S)))}
    for i in range(lp.m):
        d = int(p["dseg"][i])
        assert (d < 0) == (not nz[i].any())
        if d >= 0:
            assert (p["segi"][d], p["segj"][d]) == (i, i)
    # all three summation classes occur where the instance has them
    if name == "block_angular":
        assert (L <= 4).any() and ((L > 4) & (L <= 64)).any() and (L > 64).any()


@pytest.mark.parametrize("name", ["random_sparse", "block_angular", "dense_columns"])
def test_fixed_order_sums_equal_A_theta_At(native, oracle, name):
    lp = _cases()[name]
    rp, ci, va, p = _plan(native, lp)
    rng = np.random.default_rng(5)
    theta = np.exp(rng.uniform(-3, 3, lp.n))
    sums = oracle.segment_sums(p["segptr"], p["pp"], p["pk"], theta)
    A = lp.to_scipy().toarray()[p["perm"]]
    M = (A * theta) @ A.T
    exp = M[p["segi"], p["segj"]]
    scale = np.abs(A).__pow__(2) @ theta
    bound = 1e-13 * np.sqrt(scale[p["segi"]] * scale[p["segj"]]) + 1e-300
    assert (np.abs(sums - exp) <= bound).all()


def _long(native, idx):
    class Long(C.Structure):
        _fields_ = [("n", C.c_int32), ("idx", C.c_void_p), ("split", C.c_int32), ("part", C.c_void_p)]
    buf = native.DeviceBuffer(np.ascontiguousarray(idx, np.int32)) if len(idx) else None
    return Long(len(idx), buf.ptr.value if buf else None, 0, None), buf


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["random_sparse", "block_angular", "dense_columns"])
@pytest.mark.parametrize("delta", [0.0, 1e-9])
def test_device_assembly_is_bitwise_the_fixed_order(gpu, oracle, name, delta):
    native = gpu
    lp = _cases()[name]
    rp, ci, va, p = _plan(native, lp)
    m = lp.m
    nseg = len(p["segptr"]) - 1
    rng = np.random.default_rng(11)
    theta = np.exp(rng.uniform(-3, 3, lp.n))
    L = native.lib()
    keep = []

    def dev(a, dt):
        b = native.DeviceBuffer(np.ascontiguousarray(a, dt))
        keep.append(b)
        return b.ptr

    d_seg, d_pp, d_pk = dev(p["segptr"], np.int64), dev(p["pp"], np.float64), dev(p["pk"], np.int32)
    d_theta, d_dseg = dev(theta, np.float64), dev(p["dseg"], np.int64)
    D = native.DeviceBuffer(8 * m, np.float64)
    m0d = native.DeviceBuffer(8 * m, np.float64)
    mid_d, b1 = _long(native, p["mid_diag"])
    long_d, b2 = _long(native, p["long_diag"])
    vp = C.c_void_p
    L.lkd_assemble_diag.argtypes = [C.c_int32, vp, vp, vp, vp, vp, C.c_int, vp, vp, type(mid_d), type(long_d), vp]
    L.lkd_assemble_diag.restype = C.c_int
    assert L.lkd_assemble_diag(m, d_dseg, d_seg, d_pp, d_pk, d_theta, 1, D.ptr, m0d.ptr, mid_d, long_d, None) == 0
    panel_elems = int(p["sn_panel"][-1])
    panel = native.DeviceBuffer(8 * panel_elems, np.float64)
    mval = native.DeviceBuffer(8 * max(nseg, 1), np.float64)
    mid_s, b3 = _long(native, p["mid_segs"])
    long_s, b4 = _long(native, p["long_segs"])
    L.lkd_assemble_pairs.argtypes = [C.c_int, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                     type(mid_s), type(long_s), vp, C.c_int64, vp, vp, vp]
    L.lkd_assemble_pairs.restype = C.c_int
    assert L.lkd_assemble_pairs(0, nseg, d_seg, dev(p["segdst"], np.int64), dev(p["segi"], np.int32),
                                dev(p["segj"], np.int32), d_pp, d_pk, d_theta, D.ptr,
                                dev(p["empty_row"], np.uint8), delta, mid_s, long_s, mval.ptr,
                                panel_elems, dev(p["chunk_seg"], np.int64), panel.ptr, None) == 0
    sums = oracle.segment_sums(p["segptr"], p["pp"], p["pk"], theta)
    dseg = p["dseg"]
    m0_exp = np.where(dseg >= 0, sums[np.maximum(dseg, 0)], 0.0)
    D_exp = 1.0 / np.sqrt(np.maximum(np.abs(m0_exp), 1e-12))
    np.testing.assert_array_equal(m0d.download(), m0_exp)
    np.testing.assert_array_equal(D.download(), D_exp)
    si, sj = p["segi"], p["segj"]
    v = sums * D_exp[si] * D_exp[sj]
    got = mval.download()[:nseg]
    diag = (si == sj) & (p["empty_row"][si] == 0)
    if delta == 0.0:
        np.testing.assert_array_equal(got, v)
    else:
        np.testing.assert_array_equal(got[~diag], v[~diag])
        # the delta add may be contracted with the last product into one fma
        prod = sums[diag] * D_exp[si[diag]]
        fused = oracle.fma(prod, D_exp[sj[diag]], delta)
        ok = (got[diag] == v[diag] + delta) | (got[diag] == fused)
        assert ok.all()
    pan = panel.download()
    exp_panel = np.zeros(panel_elems)
    exp_panel[p["segdst"]] = got
    np.testing.assert_array_equal(pan, exp_panel)
