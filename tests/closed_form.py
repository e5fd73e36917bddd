"""Closed forms of one FedNL round that the mathematics fixes.

L2-regularised logistic loss of client c with samples (a_s, b_s), s < m:
    f_c(x) = (1/m) sum_s log(1 + exp(-b_s a_s.x)) + (lambda/2) |x|^2
    grad f_c(x) = lambda x - (1/m) sum_s b_s a_s (1 - sigma(b_s a_s.x))
    H_c(x) = (1/m) sum_s sigma(t_s)(1 - sigma(t_s)) a_s a_s^T + lambda I
(logistic_regression.h:26-547 computes these with the sign folded into the
features). With the Identical compressor and alpha = 1 every client sends all
of Delta_c = H_c(x_r) - learned_c, so the master's estimate after the round is
the client mean of H_c(x_r), and the round is the Newton step
    x_{r+1} = x_r - (mean_c H_c(x_r) + shift I)^{-1} grad f(x_r)
with shift = lambda (option a) or the client mean of |Delta_c|_F (option b)
(TrainLoop.h:1197-1274, 1380-1466). numpy computes it here independently of
the orchestration code under test: prescale, client order, shift option.
"""
import numpy as np


def _sigmoid(t):
    return 1.0 / (1.0 + np.exp(-t))


def client_grad_hess(feat_c, lab_c, lam, x):
    t = lab_c * (feat_c @ x)
    s = _sigmoid(t)
    m, d = feat_c.shape
    g = lam * x - (feat_c.T @ (lab_c * (1.0 - s))) / m
    H = (feat_c.T * (s * (1.0 - s))) @ feat_c / m + lam * np.eye(d)
    return g, H


def unpack_sym(packed, d):
    M = np.zeros((d, d))
    # packed order: column by column, i <= j
    cols = np.concatenate([np.full(j + 1, j) for j in range(d)])
    rows = np.concatenate([np.arange(j + 1) for j in range(d)])
    M[rows, cols] = packed
    M[cols, rows] = packed
    return M


def newton_round(prob, x, learned_packed, option_b):
    """(dx, x_next, shift) of an Identical, alpha = 1 FedNL round at x with
    the clients' learned Hessians `learned_packed` (n x S)."""
    n, m, d = prob.features.shape
    gs, Hs, lks = [], [], []
    for c in range(n):
        g, H = client_grad_hess(prob.features[c], prob.labels[c], prob.lam, x)
        gs.append(g)
        Hs.append(H)
        lks.append(np.linalg.norm(H - unpack_sym(learned_packed[c], d), "fro"))
    g = np.mean(gs, axis=0)
    Hm = np.mean(Hs, axis=0)
    shift = float(np.mean(lks)) if option_b else prob.lam
    dx = np.linalg.solve(Hm + shift * np.eye(d), g)
    return dx, x - dx, shift
