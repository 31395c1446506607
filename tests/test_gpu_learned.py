"""GPU checks of the learned dispatcher's cost model kernels (bcsr_cm_*)
against the float64 MLP oracle (oracle/mlp.py), and of the trained
dispatcher against exhaustive autotuning on held-out matrices."""
import os
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle.mlp as om
from matrixmultiplicationcuda_amd import learned

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "matrixmultiplicationcuda_amd", "data")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def problem(n=257, n_in=24, hid=64, out=8, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n_in)).astype(np.float32)
    Y = rng.standard_normal((n, out)).astype(np.float32)
    M = (rng.uniform(size=(n, out)) > 0.25).astype(np.float32)
    return X, Y, M, 1.0 / float(M.sum())


def dev(*arrays):
    return [mm.DeviceArray.from_numpy(np.ascontiguousarray(a, np.float32)) for a in arrays]


@pytest.mark.parametrize("n,seed", [(1, 1), (257, 2), (3000, 3)])
def test_forward_and_gradients_match_oracle(n, seed):
    mm.check(mm.lib().bcsr_set_device(0))
    cm = learned.CostModel(24, 64, 8, seed=seed)
    p0 = cm.get_params().astype(np.float64)
    X, Y, M, ic = problem(n=n, seed=seed)
    dX, dY, dM = dev(X, Y, M)
    dP = mm.DeviceArray((n, 8), np.float32)
    cm.forward_dev(dX, n, dP)
    mm.synchronize()
    P_ref, _ = om.forward(p0, X, 24, 64, 8)
    P = dP.numpy()
    assert np.max(np.abs(P - P_ref)) <= 1e-5 * max(1.0, np.abs(P_ref).max())
    loss = cm.backward_dev(dX, dY, dM, n, ic)
    g = cm.get_grads().astype(np.float64)
    L_ref, g_ref = om.loss_and_grad(p0, X, Y, M, ic, 24, 64, 8)
    log(f"n={n}: loss {loss:.6g} vs {L_ref:.6g}; max |dg| {np.max(np.abs(g - g_ref)):.3e} "
        f"of max |g| {np.max(np.abs(g_ref)):.3e}")
    assert abs(loss - L_ref) <= 1e-5 * abs(L_ref) + 1e-7
    # fp32 accumulation over up to 3000 rows: bound relative to the gradient scale
    assert np.max(np.abs(g - g_ref)) <= 2e-5 * np.max(np.abs(g_ref))
    for a in (dX, dY, dM, dP):
        a.free()
    cm.close()


def test_adam_steps_match_oracle():
    mm.check(mm.lib().bcsr_set_device(0))
    cm = learned.CostModel(24, 64, 8, seed=5)
    p = cm.get_params().astype(np.float64)
    X, Y, M, ic = problem(n=300, seed=5)
    dX, dY, dM = dev(X, Y, M)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        cm.train_dev(dX, dY, dM, 300, ic, 1, 1e-2)
        _, g = om.loss_and_grad(p, X, Y, M, ic, 24, 64, 8)
        p, m, v = om.adam_step(p, g, m, v, step, 1e-2)
        got = cm.get_params().astype(np.float64)
        err = np.max(np.abs(got - p))
        log(f"adam step {step}: max |dp| {err:.3e}")
        assert err <= 2e-4  # three steps of lr 1e-2 on fp32 gradients
    for a in (dX, dY, dM):
        a.free()
    cm.close()


def test_training_fits_a_smooth_target():
    mm.check(mm.lib().bcsr_set_device(0))
    rng = np.random.default_rng(9)
    F = rng.standard_normal((2000, 24)).astype(np.float32)
    W = rng.standard_normal((24, 8)).astype(np.float32) * 0.3
    Y = np.tanh(F @ W) + 0.1 * (F[:, :8] ** 2)
    M = np.ones_like(Y)
    cm = learned.CostModel(24, 64, 8, seed=1)
    losses = cm.fit(F, Y, M, steps=1500, lr=3e-3, log_every=500)
    log(f"losses {losses}")
    P = cm.predict(F)
    mse = float(np.mean((P - Y) ** 2))
    assert mse < 0.05 * float(np.var(Y)) and losses[-1] < losses[0]
    cm.close()


@pytest.mark.parametrize("precision", ["f16", "fp8"])
def test_learned_dispatch_regret_on_held_out(precision):
    """Trained on the database's training matrices by the GPU kernels (the
    last 500 of them held back to choose the margin by which a configuration
    must beat the predicted automatic choice), evaluated on its held-out
    matrices (disjoint seeds): mean regret of the dispatcher vs exhaustive
    autotuning <= 10%, and below the library's static choice."""
    from matrixmultiplicationcuda_amd import autotune
    mm.check(mm.lib().bcsr_set_device(0))
    path = os.path.join(DATA, f"timings_{precision}.jsonl")
    if not os.path.exists(path):
        pytest.skip("no timing database")
    recs = autotune.TimingDB(path).load()
    train = [r for r in recs if int(r["key"].split(":")[1]) < 200000]
    test = [r for r in recs if int(r["key"].split(":")[1]) >= 200000]
    assert len(train) >= 500 and len(test) >= 100
    train.sort(key=lambda r: int(r["key"].split(":")[1]))
    fit, val = train[:-500], train[-500:]
    F, Y, M, ms, auto = learned.dataset(fit)
    assert Y.shape[1] == ms.shape[1] + 1  # one output per configuration and the automatic choice
    cm = learned.CostModel(24, 64, Y.shape[1], seed=0)
    cm.fit(F, Y, M, steps=3000, lr=3e-3)
    Fv, _, Mv, msv, autov = learned.dataset(val)
    margin = learned.tune_margin(cm.predict(Fv), Mv, msv, autov)
    Ft, Yt, Mt, mst, autot = learned.dataset(test)
    rep = learned.regret_report(cm.predict(Ft), Mt, mst, autot, margin=margin)
    log(f"{precision} regret report (margin {margin}): {rep}")
    assert rep["dispatch_mean_regret"] <= 0.10
    assert rep["learned_mean_regret"] < rep["static_mean_regret"]
    cm.close()
