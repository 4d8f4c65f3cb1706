"""The MLP of the learned warm start on librumgpu's bf16 matrix-core kernels (rumgpu_mlp_*).

MLP        in -> hidden -> hidden -> out with ReLU; forward, backward (mean 1/2 squared error) and
           Adam run on the GPU; parameters are a float32 vector (W1, b1, W2, b2, W3, b3).
WarmStart  the model behind BatchedSolver(model=...): trained on pairs (dp, dx2, dz) of certified
           draws (dp = P_b - p_hat on the observed cells, dz = the draw's final dual z there minus
           p_hat's), it predicts dz for a new draw, which the batch object adds to the slot's warm
           state (rumgpu_bm_slot_add_obs). Inputs and targets are scaled by their RMS. The draw's
           bracket is still its own float64 certificate; a poor prediction costs iterations only.
"""
import ctypes
import math

import numpy as np

from .. import _lib


class MLP:
    def __init__(self, in_dim, hidden=512, out_dim=None, max_batch=256, seed=0, device=0):
        self.ctx = _lib.context(device)
        self.lib = self.ctx.lib
        self.in_dim, self.hidden = int(in_dim), int(hidden)
        self.out_dim = int(in_dim if out_dim is None else out_dim)
        self.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        _lib.check(self.lib.rumgpu_mlp_create(self.ctx.h, self.in_dim, self.hidden, self.out_dim, self.max_batch,
                                              ctypes.byref(h)), "mlp_create")
        self.h = h
        c = ctypes.c_int64()
        _lib.check(self.lib.rumgpu_mlp_param_count(self.h, ctypes.byref(c)), "mlp_param_count")
        self.n_params = c.value
        self.set_params(self.init_params(seed))

    def init_params(self, seed):
        """He-uniform hidden layers, a small last layer (the first predictions stay near zero)"""
        rng = np.random.default_rng(seed)
        i, h, o = self.in_dim, self.hidden, self.out_dim
        parts = [rng.uniform(-1, 1, (i, h)) * math.sqrt(6.0 / i), np.zeros(h),
                 rng.uniform(-1, 1, (h, h)) * math.sqrt(6.0 / h), np.zeros(h),
                 rng.uniform(-1, 1, (h, o)) * 0.1 * math.sqrt(3.0 / h), np.zeros(o)]
        return np.concatenate([q.ravel() for q in parts]).astype(np.float32)

    def set_params(self, theta):
        t = _lib.c_array(theta, np.float32)
        if t.size != self.n_params:
            raise ValueError("expected %d parameters" % self.n_params)
        _lib.check(self.lib.rumgpu_mlp_set_params(self.h, _lib.ptr(t)), "mlp_set_params")

    def get_params(self):
        t = np.empty(self.n_params, dtype=np.float32)
        _lib.check(self.lib.rumgpu_mlp_get_params(self.h, _lib.ptr(t)), "mlp_get_params")
        return t

    def forward(self, X):
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float32)
        Y = np.empty((X.shape[0], self.out_dim), dtype=np.float32)
        for lo in range(0, X.shape[0], self.max_batch):
            xb = np.ascontiguousarray(X[lo:lo + self.max_batch])
            yb = np.empty((xb.shape[0], self.out_dim), dtype=np.float32)
            _lib.check(self.lib.rumgpu_mlp_forward(self.h, xb.shape[0], _lib.ptr(xb), _lib.ptr(yb)), "mlp_forward")
            Y[lo:lo + xb.shape[0]] = yb
        return Y

    def backward(self, X, T, want_grad=False):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        if X.shape[0] > self.max_batch or T.shape != (X.shape[0], self.out_dim):
            raise ValueError("batch of at most %d rows, targets (N, %d)" % (self.max_batch, self.out_dim))
        loss = ctypes.c_float()
        g = np.empty(self.n_params, dtype=np.float32) if want_grad else None
        _lib.check(self.lib.rumgpu_mlp_backward(self.h, X.shape[0], _lib.ptr(X), _lib.ptr(T), ctypes.byref(loss),
                                                _lib.ptr(g)), "mlp_backward")
        return (loss.value, g) if want_grad else loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _lib.check(self.lib.rumgpu_mlp_adam(self.h, lr, beta1, beta2, eps), "mlp_adam")

    def fit(self, X, T, epochs=200, batch=None, lr=1e-3, seed=0):
        """minibatch Adam on the GPU; returns the mean loss per epoch"""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        batch = min(self.max_batch, X.shape[0]) if batch is None else int(batch)
        rng = np.random.default_rng(seed)
        hist = []
        for _ in range(int(epochs)):
            order = rng.permutation(X.shape[0])
            tot, cnt = 0.0, 0
            for lo in range(0, X.shape[0], batch):
                sel = order[lo:lo + batch]
                tot += self.backward(X[sel], T[sel]) * sel.size
                cnt += sel.size
                self.adam(lr)
            hist.append(tot / max(1, cnt))
        return hist

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStart:
    """dz on the observed cells from dp there (both scaled by their RMS over the training set)."""

    def __init__(self, M, hidden=512, seed=0, max_batch=256):
        self.M = int(M)
        self.net = MLP(self.M, hidden, self.M, max_batch=max_batch, seed=seed)
        self.sx = self.sz = 1.0
        self.trained_on = 0
        self.history = []

    def train(self, pairs, epochs=300, lr=1e-3, batch=64, seed=0):
        """pairs: (dp, dx2, dz, iterations) as BatchedSolver(collect=True) keeps them"""
        if not pairs:
            raise ValueError("no certified draws to train on")
        X = np.stack([q[0] for q in pairs])
        Z = np.stack([q[2] for q in pairs])
        self.sx = float(np.sqrt((X ** 2).mean())) or 1.0
        self.sz = float(np.sqrt((Z ** 2).mean())) or 1.0
        self.history = self.net.fit(X / self.sx, Z / self.sz, epochs=epochs, batch=batch, lr=lr, seed=seed)
        self.trained_on = len(pairs)
        return self.history

    def correction(self, dp):
        z = self.net.forward((np.asarray(dp) / self.sx)[None, :])[0].astype(np.float64) * self.sz
        return None, z

    def corrections(self, dP):
        """correction() of many draws at once (rows of dP): one forward per max_batch rows"""
        Z = self.net.forward(np.asarray(dP) / self.sx).astype(np.float64) * self.sz
        return None, Z

    def close(self):
        self.net.close()
