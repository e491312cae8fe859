"""Learned warm start for a parametric LP family.

The family keeps A fixed (m x n, k distinct rows per column, N(0,1) values; b0 and c0 from the
standard construction) and perturbs the data per instance seed:
    b = b0 (1 + scale e_b),  c = c0 (1 + scale e_c),  e_b, e_c ~ N(0, 1)   (scale 0.1).
The pipeline:
  1. solve training instances with the plain FP64 batched IPM from the standard start
     (x = s = 1, y = 0) and keep, per instance, its iterate after `target_iter` iterations
     (target_iter None: the certified solution);
  2. compress each instance's perturbation (b - b0, c - c0) to `dims` principal components
     (PCA fitted on the training perturbations, whitened);
  3. standardize the targets (log x, y, log s of the kept points) per coordinate;
  4. train learn::WarmStartMLP (dims -> hidden -> hidden -> n + m + n, BF16 MFMA products, FP32
     accumulation, Adam on the squared error) on the device;
  5. for new instances, predict (x, y, s) = (exp, id, exp) of the network output and solve from
     it: the device rejects a point whose relative infeasibility
     max(||A x - b||_inf / (1 + ||b||_inf), ||A^T y + s - c||_inf / (1 + ||c||_inf)) exceeds
     `max_infeas` and takes the standard start instead, and lifts each accepted pair to
     x_j s_j >= warm_mu (scaled space).
Every start must reach the same certificate; the gain is counted in IPM iterations and wall
time against the standard start and against the family's mean point (the kept points'
coordinate-wise mean in the target space: a start that needs no network).

Why an early iterate and not the solution: a predicted optimum puts most x_j or s_j near zero,
and wherever the perturbation moves the optimal basis the IPM must cross from one boundary to
the other; such starts took 1.5-12x the iterations of the standard start on this family.  An
early iterate is well centred (x_j s_j ~ mu for all j), so prediction errors cost little.

The start point this replaces is the reference's x = s = 1, y = 0 (src/ipm/mehrotra_device.cu:
93-102); the reference has no learned start.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import gen
from . import native as N


# ------------------------------------------------------------------ family
class Family:
    """Fixed A, b0, c0 (gen.parametric_base) and seeded instances (gen.parametric_instance)."""

    def __init__(self, m, n, k=8, seed=0, scale=0.1):
        self.base = gen.parametric_base(m, n, k, seed)
        self.m, self.n, self.scale = m, n, scale

    def instances(self, seeds):
        """(B, C): the b and c of each seed, stacked (len x m, len x n)."""
        B = np.empty((len(seeds), self.m))
        Cc = np.empty((len(seeds), self.n))
        for i, sd in enumerate(seeds):
            b, c, _, _ = gen.parametric_instance(self.base, int(sd), self.scale)
            B[i], Cc[i] = b, c
        return B, Cc

    def perturbation(self, B, Cc):
        """The network's raw input: (b - b0, c - c0) per instance."""
        return np.hstack([B - self.base.b, Cc - self.base.c])


@dataclass
class BatchRun:
    status: np.ndarray
    iterations: np.ndarray
    objective: np.ndarray
    rel_gap: np.ndarray
    rel_primal: np.ndarray
    rel_dual: np.ndarray
    warm: np.ndarray                 # 0 standard start, 1 warm point used, 2 rejected
    warm_infeas: np.ndarray
    ms: float                        # device time of the solves (sum over chunks)
    wall_s: float                    # host wall time around create + solve + fetch
    x: np.ndarray | None = None
    y: np.ndarray | None = None
    s: np.ndarray | None = None
    info: dict = field(default_factory=dict)

    @property
    def certified(self):
        return (self.status == 0) & (self.rel_gap <= 1e-8) & (self.rel_primal <= 1e-8) & \
            (self.rel_dual <= 1e-8)


def solve(fam: Family, B, Cc, warm=None, chunk=4096, keep=True, **opts) -> BatchRun:
    """Solve the instances (rows of B, Cc) with the batched device IPM, `chunk` LPs per batch.
    warm: optional (X, Y, S) original-space start points, one row per instance."""
    base = fam.base
    m, n, cnt = fam.m, fam.n, B.shape[0]
    opts.setdefault("stop_rule", 1)
    fields = ("status", "iterations", "objective", "rel_gap", "rel_primal", "rel_dual", "warm",
              "warm_infeas")
    out = {f: np.zeros(cnt) for f in fields}
    X = np.zeros((cnt, n)) if keep else None
    Y = np.zeros((cnt, m)) if keep else None
    S = np.zeros((cnt, n)) if keep else None
    ms = 0.0
    t0 = time.perf_counter()
    for c0 in range(0, cnt, chunk):
        k = min(chunk, cnt - c0)
        bt = N.Batch.from_arrays(
            np.full(k, m), np.full(k, n), np.tile(base.indptr.astype(np.int32), k),
            np.tile(base.indices.astype(np.int32), k), np.tile(base.data, k),
            B[c0:c0 + k].ravel(), Cc[c0:c0 + k].ravel(), **opts)
        try:
            if warm is not None:
                bt.set_warm_start(warm[0][c0:c0 + k].ravel(), warm[1][c0:c0 + k].ravel(),
                                  warm[2][c0:c0 + k].ravel())
            ms += bt.solve()["ms"]
            for i, r in enumerate(bt.results()):
                for f in fields:
                    out[f][c0 + i] = r[f]
            if keep:
                x, y, s = bt.solutions()
                X[c0:c0 + k] = x.reshape(k, n)
                Y[c0:c0 + k] = y.reshape(k, m)
                S[c0:c0 + k] = s.reshape(k, n)
        finally:
            bt.close()
    wall = time.perf_counter() - t0
    return BatchRun(out["status"].astype(int), out["iterations"].astype(int), out["objective"],
                    out["rel_gap"], out["rel_primal"], out["rel_dual"], out["warm"].astype(int),
                    out["warm_infeas"], ms, wall, X, Y, S)


# ------------------------------------------------------------------ compression and targets
class PCA:
    """Whitened principal components of the training perturbations (host, float64)."""

    def __init__(self, P, dims):
        self.mean = P.mean(0)
        Z = P - self.mean
        cov = Z.T @ Z / max(1, Z.shape[0] - 1)
        w, V = np.linalg.eigh(cov)
        top = np.argsort(w)[::-1][:dims]
        self.V = V[:, top]
        self.scale = np.sqrt(np.maximum(w[top], 1e-30))
        self.explained = float(w[top].sum() / max(w.sum(), 1e-30))

    def __call__(self, P):
        return ((P - self.mean) @ self.V) / self.scale


X_FLOOR = 1e-12          # log of x and s taken at max(v, X_FLOOR)


def targets(X, Y, S):
    return np.hstack([np.log(np.maximum(X, X_FLOOR)), Y, np.log(np.maximum(S, X_FLOOR))])


class Standardizer:
    def __init__(self, T):
        self.mu = T.mean(0)
        self.sd = np.maximum(T.std(0), 1e-6)

    def fwd(self, T):
        return (T - self.mu) / self.sd

    def inv(self, Z):
        return Z * self.sd + self.mu


# ------------------------------------------------------------------ the predictor
class WarmStartModel:
    """PCA + standardization + the device network; predict() returns original-space (X, Y, S)."""

    def __init__(self, fam: Family, dims=256, hidden=512, seed=0, max_batch=256):
        self.fam, self.dims, self.hidden, self.seed, self.max_batch = fam, dims, hidden, seed, max_batch
        self.net = None
        self.history = []

    def fit(self, B, Cc, run: BatchRun, epochs=30, lr=1e-3, batch=256, val_frac=0.05,
            lr_decay=(1.0, 0.3, 0.1), certified_only=True):
        ok = run.certified if certified_only else np.ones(run.status.size, bool)
        P = self.fam.perturbation(B[ok], Cc[ok])
        T = targets(run.x[ok], run.y[ok], run.s[ok])
        self.pca = PCA(P, min(self.dims, P.shape[1]))
        U = self.pca(P).astype(np.float32)
        self.std = Standardizer(T)
        Z = self.std.fwd(T).astype(np.float32)
        nval = max(1, int(val_frac * U.shape[0]))
        self.Uv, self.Zv = U[:nval], Z[:nval]
        m, n = self.fam.m, self.fam.n
        self.net = N.MLP(U.shape[1], self.hidden, self.hidden, 2 * n + m, self.max_batch, self.seed)
        self.net.set_dataset(U[nval:], Z[nval:])
        t0 = time.perf_counter()
        per = max(1, epochs // len(lr_decay))
        for i, f in enumerate(lr_decay):
            e = per if i < len(lr_decay) - 1 else epochs - per * (len(lr_decay) - 1)
            self.history += list(self.net.train(e, batch, lr * f, seed=self.seed + i))
        self.train_s = time.perf_counter() - t0
        self.val_loss = float(np.mean((self.net.forward(self.Uv) - self.Zv) ** 2))
        self.mean_loss = float(np.mean(self.Zv ** 2))      # predicting the training mean
        return self

    def predict(self, B, Cc):
        U = self.pca(self.fam.perturbation(B, Cc)).astype(np.float32)
        T = self.std.inv(self.net.forward(U).astype(np.float64))
        n, m = self.fam.n, self.fam.m
        X = np.exp(np.clip(T[:, :n], np.log(X_FLOOR), 30.0))
        Y = T[:, n:n + m]
        S = np.exp(np.clip(T[:, n + m:], np.log(X_FLOOR), 30.0))
        return X, Y, S

    def close(self):
        if self.net is not None:
            self.net.close()
            self.net = None


def mean_point(model: "WarmStartModel", count: int):
    """The family's mean point (the network-free start): the kept points' mean in the target
    space, mapped back like a prediction."""
    n, m = model.fam.n, model.fam.m
    T = model.std.mu[None, :]
    X = np.exp(np.clip(T[:, :n], np.log(X_FLOOR), 30.0))
    S = np.exp(np.clip(T[:, n + m:], np.log(X_FLOOR), 30.0))
    return (np.repeat(X, count, 0), np.repeat(T[:, n:n + m], count, 0), np.repeat(S, count, 0))


def train(fam: Family, seeds, target_iter=2, epochs=20, lr=1e-3, dims=256, hidden=512, seed=0):
    """Training data from the plain batched IPM, then the network.  Returns (model, info)."""
    B, Cc = fam.instances(seeds)
    t0 = time.perf_counter()
    run = solve(fam, B, Cc, **({} if target_iter is None else {"max_iter": int(target_iter)}))
    t_data = time.perf_counter() - t0
    model = WarmStartModel(fam, dims=dims, hidden=hidden, seed=seed)
    model.fit(B, Cc, run, epochs=epochs, lr=lr, certified_only=target_iter is None)
    info = {"instances": len(seeds), "target_iter": target_iter, "data_s": t_data,
            "data_device_ms": run.ms, "train_s": model.train_s, "epochs": epochs,
            "loss_first_epoch": model.history[0], "loss_last_epoch": model.history[-1],
            "val_loss": model.val_loss, "val_loss_of_mean": model.mean_loss,
            "pca_explained": model.pca.explained, "params": model.net.nparams}
    return model, info


def compare(fam, model, B, Cc, warm_mu=0.0, max_infeas=1e-2, **opts):
    """Held-out comparison on the same instances and options: the standard start, the network's
    start and the family's mean point.  Returns {"cold", "net", "mean": BatchRun} and the
    seconds spent predicting (PCA on the host + the network on the device)."""
    out = {"cold": solve(fam, B, Cc, keep=False, **opts)}
    t0 = time.perf_counter()
    W = model.predict(B, Cc)
    t_pred = time.perf_counter() - t0
    kw = dict(keep=False, warm_mu=warm_mu, warm_max_infeas=max_infeas, **opts)
    out["net"] = solve(fam, B, Cc, warm=W, **kw)
    out["mean"] = solve(fam, B, Cc, warm=mean_point(model, B.shape[0]), **kw)
    return out, t_pred


def summary(cold: BatchRun, warm: BatchRun, t_pred: float = 0.0) -> dict:
    both = cold.certified & warm.certified
    rel_obj = np.abs(cold.objective - warm.objective) / np.maximum(1.0, np.abs(cold.objective))
    return {
        "instances": int(cold.status.size),
        "certified_cold": int(cold.certified.sum()),
        "certified_warm": int(warm.certified.sum()),
        "iters_cold_mean": float(cold.iterations.mean()),
        "iters_warm_mean": float(warm.iterations.mean()),
        "warm_used": int((warm.warm == 1).sum()),
        "warm_rejected": int((warm.warm == 2).sum()),
        "warm_infeas_median": float(np.median(warm.warm_infeas)),
        "max_rel_objective_diff": float(rel_obj[both].max()) if both.any() else None,
        "device_ms_cold": cold.ms,
        "device_ms_warm": warm.ms,
        "predict_s": t_pred,
        "solves_per_s_cold": cold.status.size / (cold.ms / 1e3),
        "solves_per_s_warm": warm.status.size / ((warm.ms / 1e3) + t_pred),
    }


def experiment(m=400, n=1000, scale=0.1, ntrain=50_000, ntest=1000, target_iter=2, epochs=20,
               lr=1e-3, max_infeas=(1e-2, 1e30)) -> dict:
    """The whole pipeline on one family: training on seeds 100000.., held-out seeds 1..ntest;
    one comparison per fallback threshold."""
    fam = Family(m, n, scale=scale)
    model, info = train(fam, range(100_000, 100_000 + ntrain), target_iter, epochs, lr)
    B, Cc = fam.instances(range(1, 1 + ntest))
    solve(fam, B[:64], Cc[:64], keep=False)          # first-launch costs out of the timings
    res = {"family": {"m": m, "n": n, "k": 8, "scale": scale}, "training": info, "held_out": {}}
    try:
        for thr in max_infeas:
            runs, t_pred = compare(fam, model, B, Cc, max_infeas=thr)
            res["held_out"][f"max_infeas={thr:g}"] = {
                "net": summary(runs["cold"], runs["net"], t_pred),
                "mean_point": summary(runs["cold"], runs["mean"]),
            }
    finally:
        model.close()
    return res


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="learned warm start: train on a parametric family, "
                                             "compare starts on held-out instances")
    ap.add_argument("--m", type=int, default=400)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--scale", type=float, default=0.1)
    ap.add_argument("--train", type=int, default=50_000)
    ap.add_argument("--test", type=int, default=1000)
    ap.add_argument("--target-iter", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    a = ap.parse_args()
    print(json.dumps(experiment(a.m, a.n, a.scale, a.train, a.test, a.target_iter, a.epochs, a.lr)))
