"""Learned warm start on a parametric family (fixed A 120 x 300, 8 nonzeros per column;
b = b0 (1 + 0.1 e_b), c = c0 (1 + 0.1 e_c)): the network trained on the device from plain-IPM
iterates, the held-out starts against the standard start and the family's mean point, and the
device's infeasibility fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(gpu):
    from lakshya import warmstart as W
    fam = W.Family(120, 300)
    model, info = W.train(fam, range(100_000, 150_000), target_iter=2, epochs=20, lr=1e-3)
    yield W, fam, model, info
    model.close()


def _infeasibility(fam, B, Cc, X, Y, S):
    A = fam.base.to_scipy()
    rp = np.abs((A @ X.T).T - B).max(1) / (1 + np.abs(B).max(1))
    rd = np.abs((A.T @ Y.T).T + S - Cc).max(1) / (1 + np.abs(Cc).max(1))
    return np.maximum(rp, rd)


def test_network_learns_beyond_the_family_mean(trained):
    W, fam, model, info = trained
    assert info["loss_last_epoch"] < 0.7 * info["loss_first_epoch"], info
    assert info["val_loss"] < 0.7 * info["val_loss_of_mean"], info


def test_learned_start_reaches_the_same_certificate_in_fewer_iterations(trained):
    W, fam, model, info = trained
    B, Cc = fam.instances(range(1, 257))
    runs, _ = W.compare(fam, model, B, Cc, max_infeas=1e30)
    cold, net, mean = runs["cold"], runs["net"], runs["mean"]
    assert cold.certified.all() and net.certified.all() and mean.certified.all()
    assert (net.warm == 1).all() and (cold.warm == 0).all()
    # two certified points of one LP: objectives within the gap tolerance of each other
    rel = np.abs(net.objective - cold.objective) / np.maximum(1.0, np.abs(cold.objective))
    assert rel.max() <= 2e-8
    assert net.iterations.mean() <= cold.iterations.mean() - 0.5, (net.iterations.mean(), cold.iterations.mean())
    assert net.iterations.mean() < mean.iterations.mean()


def test_fallback_rejects_infeasible_predictions(trained):
    W, fam, model, info = trained
    B, Cc = fam.instances(range(1, 129))
    X, Y, S = model.predict(B, Cc)
    host = _infeasibility(fam, B, Cc, X, Y, S)
    thr = float(np.median(host))           # half of the points on either side
    cold = W.solve(fam, B, Cc, keep=False)
    warm = W.solve(fam, B, Cc, warm=(X, Y, S), keep=False, warm_max_infeas=thr)
    np.testing.assert_allclose(warm.warm_infeas, host, rtol=1e-10)
    rejected = host > thr
    assert 0 < rejected.sum() < rejected.size
    assert np.array_equal(warm.warm, np.where(rejected, 2, 1))
    # a rejected point means the standard start: the same iterations and bits as the cold solve
    assert np.array_equal(warm.iterations[rejected], cold.iterations[rejected])
    assert np.array_equal(warm.objective[rejected], cold.objective[rejected])
    assert warm.certified.all()
    # at the default threshold 1e-2 these predictions are all rejected
    spec = W.solve(fam, B, Cc, warm=(X, Y, S), keep=False, warm_max_infeas=1e-2)
    assert np.array_equal(spec.warm, np.where(host > 1e-2, 2, 1))
