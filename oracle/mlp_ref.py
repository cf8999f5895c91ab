"""ORACLE — test infrastructure only (used by tests/, never by the product).

Float64 numpy restatement of the cost-model network the autotuner trains on the GPU
(csrc/mlp.hip): fully connected layers y = act(x W + b) with ReLU on hidden layers, mean
squared error, and the gradients of all parameters by the chain rule. Used to check the
GPU forward/backward kernels (and finite differences check this file)."""
import numpy as np


def forward(params, X):
    """params: list of (W, b); returns the activations of every layer (input first)."""
    acts = [np.asarray(X, np.float64)]
    for k, (W, b) in enumerate(params):
        z = acts[-1] @ W + b
        acts.append(np.maximum(z, 0.0) if k < len(params) - 1 else z)
    return acts


def loss_and_grads(params, X, y):
    acts = forward(params, X)
    pred = acts[-1][:, 0]
    n = len(y)
    loss = float(np.mean((pred - y) ** 2))
    g = (2.0 * (pred - y) / n)[:, None]
    grads = [None] * len(params)
    for k in range(len(params) - 1, -1, -1):
        W, b = params[k]
        if k < len(params) - 1:
            g = g * (acts[k + 1] > 0)
        grads[k] = (acts[k].T @ g, g.sum(axis=0))
        g = g @ W.T
    return loss, grads
