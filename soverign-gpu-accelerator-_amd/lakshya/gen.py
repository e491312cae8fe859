"""Seeded synthetic LP families (standard form: min c^T x, A x = b, x >= 0).

Every instance is a pure function of (family, sizes, seed): a numpy Generator is seeded with
SeedSequence([FAMILY_TAG, seed, attempt]) and an attempt is rejected (and the next one drawn)
when it contains an empty row, so no row of A is ever empty.

"Standard construction": x* ~ U(0.5, 1.5)^n, b = A x*, y ~ N(0,1)^m, s ~ U(0.1, 1.0)^n,
c = A^T y + s -- strictly feasible primal and dual, so the LP has a finite optimum.

Families (sizes of the full workloads):
  random_sparse    m=2,000  n=5,000, k=6 distinct rows per column, N(0,1)
  block_angular    64 blocks 1,500 x 4,000 (k=8) + 50 linking rows at 5% density, U(-1,1)
  transport        700 x 700 arcs + 20,000 side rows (10 arcs, N(0,1)); one redundant
                   sink row is dropped (the transport block has rank m-1)
  batch            4,096 LPs, m ~ U{50..800}, n = round(2.5 m), Bernoulli(0.05) N(0,1)
  parametric       fixed A (m=20,000, n=50,000, k=8); b = b0 (1 + 0.1 N), c = c0 (1 + 0.1 N)
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

TAGS = {"random_sparse": 1, "block_angular": 2, "transport": 3, "batch": 4,
        "parametric": 5, "dense_small": 6, "dense_columns": 7}


@dataclass
class LP:
    """CSC standard-form LP.  indptr/indices/data follow scipy's csc layout (rows ascending)."""
    m: int
    n: int
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    b: np.ndarray
    c: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.indptr, self.indices, self.data, self.b, self.c):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.data, self.indices, self.indptr), shape=(self.m, self.n))

    def csr(self):
        """(indptr, indices, data) of A in CSR with ascending columns."""
        a = self.to_scipy().tocsr()
        a.sort_indices()
        return (a.indptr.astype(np.int32), a.indices.astype(np.int32),
                a.data.astype(np.float64))


def _rng(family: str, seed: int, attempt: int = 0) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([TAGS[family], int(seed), attempt]))


def _distinct_rows(rng, m: int, n: int, k: int) -> np.ndarray:
    """n x k matrix of row indices, each line k distinct values from [0, m), sorted."""
    rows = np.sort(rng.integers(0, m, size=(n, k)), axis=1)
    while True:
        dup = np.any(rows[:, 1:] == rows[:, :-1], axis=1) if k > 1 else np.zeros(n, bool)
        if not dup.any():
            return rows
        idx = np.nonzero(dup)[0]
        rows[idx] = np.sort(rng.integers(0, m, size=(idx.size, k)), axis=1)


def _csc_from_cols(m, n, rows, vals):
    """rows/vals: n x k arrays (sorted rows per column)."""
    k = rows.shape[1]
    indptr = np.arange(0, (n + 1) * k, k, dtype=np.int64)
    return indptr.astype(np.int32), rows.reshape(-1).astype(np.int32), vals.reshape(-1)


def _standard_rhs(rng, A, m, n):
    xs = rng.uniform(0.5, 1.5, n)
    b = A @ xs
    y = rng.standard_normal(m)
    s = rng.uniform(0.1, 1.0, n)
    c = A.T @ y + s
    return b, c


def _has_empty_row(m, indices) -> bool:
    return np.bincount(indices, minlength=m).min() == 0 if m > 0 else False


def random_sparse(m=2000, n=5000, k=6, seed=0) -> LP:
    import scipy.sparse as sp
    for attempt in range(1000):
        rng = _rng("random_sparse", seed, attempt)
        rows = _distinct_rows(rng, m, n, k)
        vals = rng.standard_normal((n, k))
        indptr, indices, data = _csc_from_cols(m, n, rows, vals)
        if _has_empty_row(m, indices):
            continue
        A = sp.csc_matrix((data, indices, indptr), shape=(m, n))
        b, c = _standard_rhs(rng, A, m, n)
        return LP(m, n, indptr, indices, data, b, c, f"random_sparse_m{m}_n{n}_k{k}_s{seed}",
                  {"attempt": attempt})
    raise RuntimeError("random_sparse: no instance without empty rows")


def dense_columns(m=600, n=1500, k=4, ndense=3, seed=0) -> LP:
    """random_sparse-style columns plus `ndense` fully dense columns (entries N(0, 1/m))."""
    import scipy.sparse as sp
    for attempt in range(1000):
        rng = _rng("dense_columns", seed, attempt)
        rows = _distinct_rows(rng, m, n - ndense, k)
        vals = rng.standard_normal((n - ndense, k))
        indptr, indices, data = _csc_from_cols(m, n - ndense, rows, vals)
        if _has_empty_row(m, indices):
            continue
        S = sp.csc_matrix((data, indices, indptr), shape=(m, n - ndense))
        Dn = sp.csc_matrix(rng.standard_normal((m, ndense)) / np.sqrt(m))
        A = sp.hstack([S, Dn]).tocsc()
        A.sort_indices()
        b, c = _standard_rhs(rng, A, m, n)
        return LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                  A.data.astype(np.float64), b, c, f"dense_columns_m{m}_n{n}_k{k}_d{ndense}_s{seed}",
                  {"attempt": attempt})
    raise RuntimeError("dense_columns: no instance without empty rows")


def block_angular(nblocks=64, mb=1500, nb=4000, k=8, nlink=50, dens=0.05, seed=0) -> LP:
    import scipy.sparse as sp
    m, n = nblocks * mb + nlink, nblocks * nb
    for attempt in range(1000):
        rng = _rng("block_angular", seed, attempt)
        rows = _distinct_rows(rng, mb, n, k) + (np.arange(n) // nb)[:, None] * mb
        vals = rng.standard_normal((n, k))
        blk = sp.csc_matrix((vals.reshape(-1), rows.reshape(-1),
                             np.arange(0, (n + 1) * k, k)), shape=(m, n))
        link = sp.random(nlink, n, density=dens, format="csc", random_state=rng,
                         data_rvs=lambda s: rng.uniform(-1.0, 1.0, s))
        link = sp.vstack([sp.csc_matrix((nblocks * mb, n)), link]).tocsc()
        A = (blk + link).tocsc()
        A.sort_indices()
        if _has_empty_row(m, A.indices):
            continue
        b, c = _standard_rhs(rng, A, m, n)
        return LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                  A.data.astype(np.float64), b, c,
                  f"block_angular_{nblocks}x{mb}x{nb}_k{k}_L{nlink}_s{seed}",
                  {"attempt": attempt, "nblocks": nblocks, "mb": mb, "nb": nb, "nlink": nlink})
    raise RuntimeError("block_angular: no instance without empty rows")


def transport(ns=700, nd=700, nside=20000, kside=10, seed=0) -> LP:
    """Balanced transportation LP plus side constraints; the last sink row is dropped."""
    import scipy.sparse as sp
    n = ns * nd
    for attempt in range(1000):
        rng = _rng("transport", seed, attempt)
        sup = rng.uniform(50, 150, ns)
        dem = rng.uniform(50, 150, nd)
        dem *= sup.sum() / dem.sum()
        cost = rng.uniform(1, 100, n)
        arcs = np.arange(n)
        src, snk = arcs // nd, arcs % nd
        # transport rows: sources 0..ns-1, sinks ns..ns+nd-2 (sink nd-1 dropped)
        keep = snk < nd - 1
        r_t = np.concatenate([src, ns + snk[keep]])
        c_t = np.concatenate([arcs, arcs[keep]])
        side_cols = _distinct_rows(rng, n, nside, kside)          # nside x kside arcs
        side_vals = rng.standard_normal((nside, kside))
        m = ns + nd - 1 + nside
        r_s = np.repeat(np.arange(nside) + ns + nd - 1, kside)
        A = sp.csc_matrix((np.concatenate([np.ones(r_t.size), side_vals.reshape(-1)]),
                           (np.concatenate([r_t, r_s]),
                            np.concatenate([c_t, side_cols.reshape(-1)]))), shape=(m, n))
        A.sum_duplicates()
        A.sort_indices()
        if _has_empty_row(m, A.indices):
            continue
        xf = np.outer(sup, dem).reshape(-1) / sup.sum()
        b = np.concatenate([sup, dem[:-1], (A[ns + nd - 1:, :] @ xf)])
        return LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                  A.data.astype(np.float64), b, cost,
                  f"transport_{ns}x{nd}_side{nside}_s{seed}",
                  {"attempt": attempt, "dropped_row": "sink %d" % (nd - 1)})
    raise RuntimeError("transport: no instance without empty rows")


def dense_small(m: int, n: int | None = None, density=0.05, seed=0, family="batch") -> LP:
    """One batch-family LP: Bernoulli(density) N(0,1) entries, standard construction."""
    import scipy.sparse as sp
    n = int(round(2.5 * m)) if n is None else n
    for attempt in range(1000):
        rng = _rng(family, seed, attempt)
        mask = rng.random((m, n)) < density
        vals = rng.standard_normal((m, n))
        A = sp.csc_matrix(np.where(mask, vals, 0.0))
        A.sort_indices()
        if _has_empty_row(m, A.indices):
            continue
        b, c = _standard_rhs(rng, A, m, n)
        return LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                  A.data.astype(np.float64), b, c, f"{family}_m{m}_s{seed}", {"attempt": attempt})
    raise RuntimeError("dense_small: no instance without empty rows")


def batch_sizes(count=4096, lo=50, hi=800, seed=0) -> np.ndarray:
    rng = _rng("batch", 10_000_000 + seed)
    return rng.integers(lo, hi + 1, size=count)


def batch_lp(m: int, seed: int, density=0.05) -> LP:
    """One LP of the batch family: n = round(2.5 m), every entry present with probability
    `density` (the mask drawn column by column), N(0,1) values, standard construction."""
    import scipy.sparse as sp
    n = int(round(2.5 * m))
    for attempt in range(1000):
        rng = _rng("batch", seed, attempt)
        mask = rng.random((n, m), dtype=np.float32) < density   # line j = column j of A
        cols, rows = np.nonzero(mask)                              # column-major, rows ascending
        if np.bincount(rows, minlength=m).min() == 0:
            continue
        indptr = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(cols, minlength=n), out=indptr[1:])
        data = rng.standard_normal(rows.size)
        A = sp.csc_matrix((data, rows.astype(np.int32), indptr.astype(np.int32)), shape=(m, n))
        b, c = _standard_rhs(rng, A, m, n)
        return LP(m, n, indptr.astype(np.int32), rows.astype(np.int32), data, b, c,
                  f"batch_m{m}_s{seed}", {"attempt": attempt})
    raise RuntimeError("batch_lp: no instance without empty rows")


def batch(count=4096, lo=50, hi=800, seed=0, indices=None):
    """The batch family: LP i has m_i ~ U{lo..hi} and instance seed seed * 1_000_003 + i
    (`indices`: only those members, in that order)."""
    sizes = batch_sizes(count, lo, hi, seed)
    idx = range(count) if indices is None else indices
    return [batch_lp(int(sizes[i]), seed * 1_000_003 + i) for i in idx]


def parametric_base(m=20000, n=50000, k=8, seed=0) -> LP:
    lp = random_sparse(m, n, k, seed=seed + 500_000)
    lp.name = f"parametric_base_m{m}_n{n}_k{k}_s{seed}"
    return lp


def parametric_instance(base: LP, inst_seed: int, scale=0.1):
    rng = _rng("parametric", inst_seed)
    eb = rng.standard_normal(base.m)
    ec = rng.standard_normal(base.n)
    b = base.b * (1.0 + scale * eb)
    c = base.c * (1.0 + scale * ec)
    return b, c, eb, ec
