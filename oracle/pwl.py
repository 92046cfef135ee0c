"""CPU oracle of the learned model's training step -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of the device training kernel
(k_pwl_fit).  The model has no reference counterpart; its definition:

  s_j = softplus(theta_j)                 (j = 0 .. K-1, K = 1024 intervals)
  c_k = s_0 + ... + s_k,  C = c_{K-1}
  y_0 = 0,  y_k = c_{k-1} / C             (k = 1 .. K; knot heights, non-decreasing)
  F(u) = y_i + (y_{i+1} - y_i) (u - i),   i = min(K-1, floor(u))
  L = mean over points of (F(u) - t)^2

and the gradient dL/dtheta.  Computed here in float64 by explicit
differentiation; tests pin it with central finite differences.
"""
from __future__ import annotations

import numpy as np

K = 1024


def softplus(x):
    return np.where(x > 20.0, x, np.log1p(np.exp(np.minimum(x, 20.0))))


def knots(theta: np.ndarray) -> np.ndarray:
    s = softplus(theta.astype(np.float64))
    c = np.cumsum(s)
    return np.concatenate([[0.0], c / c[-1]])


def loss(theta: np.ndarray, u: np.ndarray, t: np.ndarray) -> float:
    y = knots(theta)
    u = u.astype(np.float64)
    i = np.minimum(K - 1, np.floor(u).astype(np.int64))
    f = u - i
    r = y[i] + (y[i + 1] - y[i]) * f - t.astype(np.float64)
    return float(np.mean(r * r))


def loss_grad(theta: np.ndarray, u: np.ndarray, t: np.ndarray):
    """(L, dL/dtheta) in float64."""
    th = theta.astype(np.float64)
    s = softplus(th)
    c = np.cumsum(s)
    C = c[-1]
    y = np.concatenate([[0.0], c / C])
    u = u.astype(np.float64)
    i = np.minimum(K - 1, np.floor(u).astype(np.int64))
    f = u - i
    r = y[i] + (y[i + 1] - y[i]) * f - t.astype(np.float64)
    T = u.size
    g = np.zeros(K + 1)
    np.add.at(g, i, 2.0 * r * (1.0 - f) / T)
    np.add.at(g, i + 1, 2.0 * r * f / T)
    # y_k = c_{k-1} / C for k >= 1: dy_k/ds_j = ([j < k] - y_k) / C
    dot = float(np.dot(g[1:], y[1:]))
    suffix = np.cumsum(g[1:][::-1])[::-1]  # suffix[j] = sum_{k >= j+1} g_k
    ds = (suffix - dot) / C
    sig = 1.0 / (1.0 + np.exp(-th))
    return float(np.mean(r * r)), ds * sig
