"""The autotuner's cost model: forward/backward/Adam kernels of csrc/mlp.hip against the
float64 numpy restatement (oracle/mlp_ref.py), finite differences on that restatement, and
the gain of the trained network over tuning without it on the measured MI355X dataset."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def test_gradients_match_numpy_oracle():
    import torch
    from cellato_amd.costmodel import CostModel
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, 15)).astype(np.float32)
    y = rng.normal(size=300).astype(np.float32)
    m = CostModel(15, hidden=(32, 16), seed=3)
    loss, grads = m.loss_and_grads(torch.from_numpy(X).cuda(), torch.from_numpy(y.reshape(-1, 1)).cuda())
    ref_loss, ref_grads = mlp_ref.loss_and_grads(m.params(), X.astype(np.float64), y.astype(np.float64))
    assert abs(float(loss.item()) - ref_loss) <= 1e-5 * max(1.0, abs(ref_loss))     # fp32 vs fp64
    for (dW, db), (rW, rb) in zip(grads, ref_grads):
        np.testing.assert_allclose(dW.cpu().numpy(), rW, rtol=2e-4, atol=2e-6)
        np.testing.assert_allclose(db.cpu().numpy(), rb, rtol=2e-4, atol=2e-6)
    acts = m.forward(torch.from_numpy(X).cuda())
    np.testing.assert_allclose(acts[-1].cpu().numpy()[:, 0], mlp_ref.forward(m.params(), X)[-1][:, 0], rtol=1e-4, atol=1e-5)


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = [(rng.normal(size=(5, 4)), rng.normal(size=4)), (rng.normal(size=(4, 1)), rng.normal(size=1))]
    X, y = rng.normal(size=(20, 5)), rng.normal(size=20)
    _, grads = mlp_ref.loss_and_grads(params, X, y)
    eps = 1e-6
    for k in range(2):
        for idx in [(0, 0), (2, 1)] if k == 0 else [(1, 0), (3, 0)]:
            p = [(W.copy(), b.copy()) for W, b in params]
            p[k][0][idx] += eps
            lp, _ = mlp_ref.loss_and_grads(p, X, y)
            p[k][0][idx] -= 2 * eps
            lm, _ = mlp_ref.loss_and_grads(p, X, y)
            assert abs((lp - lm) / (2 * eps) - grads[k][0][idx]) < 1e-6


def test_training_reduces_loss():
    from cellato_amd.costmodel import CostModel
    rng = np.random.default_rng(2)
    X = rng.normal(size=(512, 8)).astype(np.float32)
    y = (np.sin(X[:, 0]) + X[:, 1] * X[:, 2]).astype(np.float32)
    hist = CostModel(8, hidden=(32, 32), seed=1).fit(X, y, epochs=300, lr=3e-3)
    assert hist[-1] < 0.25 * hist[0]


def test_network_beats_tuning_without_it():
    from cellato_amd.tuning_eval import evaluate
    by_size = evaluate(budget=2, epochs=600, holdout="size")
    print(by_size)
    # Two measurements guided by the network, on grid sizes the network never saw, land
    # clearly closer to the best launch shape than two random measurements or the fixed shape
    # T=4, and within 7 % of measuring all 21 shapes.
    assert by_size["model_topk"] > by_size["random_k"] + 0.05
    assert by_size["model_topk"] > by_size["default_T4"] + 0.03
    assert by_size["model_topk"] >= 0.93


def test_network_transfers_to_an_unseen_automaton():
    # Trained on nine automata, the network ranks the launch shapes (depth x rows per task) of
    # the tenth, never seen, from what its rule is made of and what the compiler made of it
    # (planes, AST node kinds, registers per lane, waves per SIMD): its two picks land closer to
    # the best than two random picks, and about as close as the fixed shape T=4 (which the
    # network does not know to be a good default).
    from cellato_amd.tuning_eval import evaluate
    by_rule = evaluate(budget=2, epochs=600, holdout="rule")
    print(by_rule)
    assert by_rule["model_topk"] > by_rule["random_k"] + 0.03
    assert by_rule["model_topk"] > by_rule["default_T4"] - 0.02
