"""The learned part end to end: the 5-member BF16 MLP ensemble trained on the GPU (the
library's own forward/backward/AdamW kernels) on the committed training set of 2048
batched-LBM runs of the generic F1 body, and what it gains over the method without
it.

The training set's labels repeat only to a few percent (64 setups were run twice from
perturbed starts; the wake of the body makes 2400-step force averages scatter), so
accuracy is stated against that floor and against a linear least-squares fit, and the
fallback rule (ensemble spread > 2% -> batched LBM) is exercised and checked to
return exactly the LBM's answer."""
import time

import numpy as np
import pytest

from aerolbm import surrogate as S

pytestmark = pytest.mark.gpu


def r2(p, y):
    return 1.0 - ((p - y) ** 2).mean(0) / y.var(0)


@pytest.fixture(scope="module")
def trained():
    d = S.load_dataset()
    X, Y = d["X"], d["Y"]
    test, val, train = np.arange(0, 256), np.arange(256, 512), np.arange(512, 2048)
    sur = S.Surrogate(members=5, seed=0)
    t0 = time.perf_counter()
    curves = sur.fit(X[train], Y[train], val=(X[val], Y[val]), steps=3000)
    train_s = time.perf_counter() - t0
    print(f"trained 5 members x 3000 steps in {train_s:.2f} s; val loss {curves['val'][0]:.3f} -> "
          f"{curves['val'].min():.3f}", flush=True)
    yield d, sur, (test, val, train), train_s
    sur.close()


def test_surrogate_accuracy_against_lbm_labels(trained):
    d, sur, (test, val, train), _ = trained
    X, Y, Yr = d["X"], d["Y"], d["Y_repeat"]
    mean, spread, _ = sur.predict(X[test])
    r2_net = r2(mean, Y[test])
    A = np.hstack([X, np.ones((len(X), 1))])
    coef, *_ = np.linalg.lstsq(A[train], Y[train], rcond=None)
    r2_lin = r2(A[test] @ coef, Y[test])
    noise = np.median(np.abs(Yr - Y[:64]), 0)       # LBM vs LBM (perturbed start)
    err = np.median(np.abs(mean[:64] - Y[:64]), 0)  # surrogate vs LBM, same 64 setups
    print("outputs      ", S.OUT_NAMES)
    print("R2 network   ", r2_net.round(3), "\nR2 linear    ", r2_lin.round(3))
    print("median |err| surrogate vs LBM", err.round(4), "\nmedian |LBM - LBM repeat|   ", noise.round(4), flush=True)
    assert r2_net[1] > 0.8                      # rear-wing Cl*A, the least noisy label
    assert r2_net.mean() >= r2_lin.mean()       # no worse than the linear fit overall
    assert (r2_net > 0).all()                   # better than the training mean everywhere


def test_surrogate_gain_and_lbm_fallback(trained):
    """Gain: answering 4096 setups costs well under a millisecond of GPU time against
    ~0.24 s of batched LBM per setup.  Fallback: setups whose ensemble spread exceeds
    2% are answered by the batched LBM with the training protocol, which reproduces
    their training labels."""
    d, sur, (test, _, _), _ = trained
    X, Y = d["X"], d["Y"]
    q = S.lhs_setups(4096, seed=7)
    mean, spread, ms = sur.predict(q)
    lbm_s = float(d["wall_s"]) / len(X)
    gain = lbm_s * len(q) / (ms / 1e3)
    acc = {t: float((spread <= t).mean()) for t in (0.02, 0.05, 0.10, 0.20)}
    print(f"4096 answers in {ms:.3f} ms device; LBM {lbm_s:.3f} s per setup; gain {gain:.3g}x; "
          f"accepted fraction by spread threshold {acc}", flush=True)
    assert gain > 1e4
    _, test_spread, _ = sur.predict(X[test])
    sel = test[np.argsort(test_spread)[-4:]]  # the 4 least certain test setups
    t0 = time.perf_counter()
    ans = sur.answer(X[sel], settle=int(d["settle"]), measure=int(d["measure"]), max_spread=0.02)
    print(f"fallback: {int((~ans.from_surrogate).sum())} of {len(sel)} setups to LBM, "
          f"{time.perf_counter() - t0:.2f} s", flush=True)
    assert (~ans.from_surrogate).all()
    assert np.allclose(ans.y[:, :3], Y[sel, :3], rtol=1e-6, atol=1e-7)
