"""Learned cost model of the autotuner: a small fully connected network (features of a
configuration -> log of nanoseconds per cell update), trained on the GPU with the
project's own forward/backward/Adam kernels (C ABI `cellato_mlp_*`, csrc/mlp.hip).

The tuner ranks the candidate (depth, chunk) pairs of a size bucket with it and measures
only the best-ranked few instead of all of them."""
import ctypes

import numpy as np

from . import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class CostModel:
    def __init__(self, n_in, hidden=(64, 64), seed=0, device="cuda"):
        import torch
        g = torch.Generator().manual_seed(seed)
        dims = [n_in, *hidden, 1]
        self.layers = []
        for i in range(len(dims) - 1):
            W = (torch.randn(dims[i], dims[i + 1], generator=g) * (2.0 / dims[i]) ** 0.5).float().to(device)
            b = torch.zeros(dims[i + 1], dtype=torch.float32, device=device)
            self.layers.append([W, b])
        self.dims = dims
        self.device = device
        self.mu = np.zeros(n_in, np.float32)
        self.sd = np.ones(n_in, np.float32)
        self.ymu, self.ysd = 0.0, 1.0
        self.adam_state = None
        self.step = 0

    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def params(self):
        return [(W.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)) for W, b in self.layers]

    def forward(self, X):
        """X: torch float32 (N, n_in) on the device -> list of activations."""
        L = _lib.lib()
        acts = [X]
        for k, (W, b) in enumerate(self.layers):
            import torch
            Y = torch.empty((X.shape[0], W.shape[1]), dtype=torch.float32, device=self.device)
            relu = 1 if k < len(self.layers) - 1 else 0
            _lib.check(L.cellato_mlp_linear_forward(_p(acts[-1]), _p(W), _p(b), _p(Y), X.shape[0], W.shape[0],
                                                     W.shape[1], relu, self._stream()))
            acts.append(Y)
        return acts

    def loss_and_grads(self, X, y):
        """MSE loss and parameter gradients, all on the device."""
        import torch
        L = _lib.lib()
        acts = self.forward(X)
        N = X.shape[0]
        dP = torch.empty((N, 1), dtype=torch.float32, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _lib.check(L.cellato_mlp_mse(_p(acts[-1]), _p(y), _p(dP), _p(loss), N, self._stream()))
        grads = [None] * len(self.layers)
        g = dP
        for k in range(len(self.layers) - 1, -1, -1):
            W, b = self.layers[k]
            relu = 1 if k < len(self.layers) - 1 else 0
            G = torch.empty_like(acts[k + 1])
            dW = torch.empty_like(W)
            db = torch.empty_like(b)
            dX = torch.empty_like(acts[k]) if k > 0 else None
            _lib.check(L.cellato_mlp_linear_backward(_p(acts[k]), _p(W), _p(acts[k + 1]), _p(g), _p(G),
                                                      _p(dX) if dX is not None else None, _p(dW), _p(db),
                                                      N, W.shape[0], W.shape[1], relu, self._stream()))
            grads[k] = (dW, db)
            g = dX
        return loss, grads

    def fit(self, X, y, epochs=400, lr=3e-3):
        """Train on numpy features X (N, n_in) and targets y (N,); returns the loss history."""
        import torch
        L = _lib.lib()
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float32)
        self.mu, self.sd = X.mean(0), X.std(0) + 1e-6
        self.ymu, self.ysd = float(y.mean()), float(y.std() + 1e-6)
        Xd = torch.from_numpy((X - self.mu) / self.sd).to(self.device)
        yd = torch.from_numpy(((y - self.ymu) / self.ysd).reshape(-1, 1)).to(self.device)
        if self.adam_state is None:
            self.adam_state = [[torch.zeros_like(W), torch.zeros_like(W), torch.zeros_like(b), torch.zeros_like(b)]
                               for W, b in self.layers]
        hist = []
        for _ in range(epochs):
            loss, grads = self.loss_and_grads(Xd, yd)
            self.step += 1
            for (W, b), (dW, db), (mW, vW, mb, vb) in zip(self.layers, grads, self.adam_state):
                _lib.check(L.cellato_mlp_adam(_p(W), _p(dW), _p(mW), _p(vW), W.numel(), lr, 0.9, 0.999, 1e-8,
                                              self.step, self._stream()))
                _lib.check(L.cellato_mlp_adam(_p(b), _p(db), _p(mb), _p(vb), b.numel(), lr, 0.9, 0.999, 1e-8,
                                              self.step, self._stream()))
            hist.append(loss)
        return [float(l.item()) for l in hist[:: max(1, len(hist) // 20)]] + [float(hist[-1].item())]

    def predict(self, X):
        import torch
        Xd = torch.from_numpy((np.asarray(X, np.float32) - self.mu) / self.sd).to(self.device)
        out = self.forward(Xd)[-1]
        return out.cpu().numpy()[:, 0] * self.ysd + self.ymu
