"""Seeded synthetic L2-logistic FedNL problems.

Each problem is a dense design of ``n_clients * m`` samples by ``d`` features
(intercept, when present, is the last column of ones), labels in {-1, +1}
drawn from a planted logistic model, split contiguously: client ``c`` owns
samples ``[c*m, (c+1)*m)``. This is the split dopt's loader makes without
``--reshuffle_train`` (dopt/bin_fednl_local/include/TrainLoop.h:212-250).

Arrays are float64 with values that are exact in float32 (binary features, or
N(0,1) draws rounded to float32) so that the same problem can be fed to the
FP64, FP32 and split-BF16 Hessian forms.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Problem:
    name: str
    features: np.ndarray   # (n, m, d) float64, C-contiguous
    labels: np.ndarray     # (n, m) float64 in {-1, +1}
    lam: float
    w_true: np.ndarray     # planted weights (d,)

    @property
    def n(self) -> int:
        return self.features.shape[0]

    @property
    def m(self) -> int:
        return self.features.shape[1]

    @property
    def d(self) -> int:
        return self.features.shape[2]


def _sigmoid(t):
    return 1.0 / (1.0 + np.exp(-t))


def _planted_labels(rng, a, w, target_pos=None):
    """Labels +-1 with P(+1) = sigmoid(a.w + b); b is solved so that the mean
    positive probability equals ``target_pos`` (when given)."""
    t = a @ w
    b = 0.0
    if target_pos is not None:
        lo, hi = -60.0, 60.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if _sigmoid(t + mid).mean() > target_pos:
                hi = mid
            else:
                lo = mid
        b = 0.5 * (lo + hi)
    p = _sigmoid(t + b)
    y = np.where(rng.random(t.shape[0]) < p, 1.0, -1.0)
    return y, b


def binary_problem(name, n, m, d_feat, density, lam=1e-3, seed=0, target_pos=None,
                   intercept=True) -> Problem:
    """Bernoulli(density) binary features in {0,1} (+ intercept column)."""
    rng = np.random.default_rng(seed)
    N = n * m
    a = (rng.random((N, d_feat)) < density).astype(np.float64)
    w = rng.standard_normal(d_feat)
    y, b = _planted_labels(rng, a, w, target_pos)
    if intercept:
        a = np.concatenate([a, np.ones((N, 1))], axis=1)
        w = np.concatenate([w, [b]])
    return Problem(name, np.ascontiguousarray(a.reshape(n, m, -1)),
                   np.ascontiguousarray(y.reshape(n, m)), lam, w)


def gaussian_problem(name, n, m, d, lam=1e-3, seed=0, target_pos=0.5) -> Problem:
    """Dense N(0,1) features rounded to float32 (exact in fp32 and fp64)."""
    rng = np.random.default_rng(seed)
    N = n * m
    a = rng.standard_normal((N, d)).astype(np.float32).astype(np.float64)
    w = rng.standard_normal(d) / math.sqrt(d)
    y, _ = _planted_labels(rng, a, w, target_pos)
    return Problem(name, np.ascontiguousarray(a.reshape(n, m, d)),
                   np.ascontiguousarray(y.reshape(n, m)), lam, w)


def gaussian_pool(size, seed=0) -> np.ndarray:
    """``size`` N(0,1) draws rounded to float32, as float64: one buffer many
    serving requests take their features from (a request is a window of it, so
    each request's entries are i.i.d. N(0,1); requests may share entries)."""
    rng = np.random.default_rng(seed)
    out = np.empty(size, dtype=np.float64)
    step = 1 << 26
    for a in range(0, size, step):
        b = min(size, a + step)
        out[a:b] = rng.standard_normal(b - a, dtype=np.float32)
    return out


def pooled_problem(name, pool, offset, n, m, d, lam=1e-3, seed=0, target_pos=0.5) -> Problem:
    """A dense N(0,1) problem whose features are the window
    pool[offset : offset + n*m*d] (no copy); labels from a planted model."""
    rng = np.random.default_rng(seed)
    N = n * m
    a = pool[offset:offset + N * d].reshape(N, d)
    w = rng.standard_normal(d) / math.sqrt(d)
    y, _ = _planted_labels(rng, a, w, target_pos)
    return Problem(name, a.reshape(n, m, d), np.ascontiguousarray(y.reshape(n, m)), lam, w)


# The five workloads the engine is measured on. Seeds are fixed.
def config1(seed=1):
    """50 clients x 221 samples, 67 Bernoulli(0.5) features + intercept (d=68)."""
    return binary_problem("c1_bern05_n50_m221_d68", 50, 221, 67, 0.5, seed=seed)


def config2(seed=2):
    """142 clients x 350 samples, 299 binary features at ~4% + intercept (d=300), ~3% positives."""
    return binary_problem("c2_bin04_n142_m350_d300", 142, 350, 299, 0.04, seed=seed,
                          target_pos=0.03)


def config3(seed=3):
    """80 clients x 407 samples, 122 binary features at ~11% + intercept (d=123), balanced."""
    return binary_problem("c3_bin11_n80_m407_d123", 80, 407, 122, 0.11, seed=seed,
                          target_pos=0.5)


def batched_shapes(count=64, seed=4):
    """(d, n, m) of the batched workload: d~U[32,512], n~U[16,256], m~U[100,600]."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(32, 513)), int(rng.integers(16, 257)), int(rng.integers(100, 601)))
            for _ in range(count)]


def serving_shapes(count, seed=5):
    """(d, n, m) of serving requests: d log-uniform in [64,2048], n in [32,512], m in [200,1000]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
        out.append((d, int(rng.integers(32, 513)), int(rng.integers(200, 1001))))
    return out


def small_problem(n=4, m=37, d_feat=9, density=0.3, seed=7, target_pos=0.4, lam=1e-3):
    """Tiny binary problem for fast parity checks."""
    return binary_problem(f"small_n{n}_m{m}_d{d_feat + 1}", n, m, d_feat, density, lam=lam,
                          seed=seed, target_pos=target_pos)
