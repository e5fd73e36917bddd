"""Shared numerical criteria of the parity tests.

Gradient-norm trajectories: |g - g_ref| <= REL * g_ref + FLOOR on ||grad f||,
i.e. 1e-10 relative in ||grad f||^2 away from the floor. FLOOR is the FP64
evaluation noise of a gradient whose terms are O(1/m) products summed over
m*d entries (measured reference-vs-reference drift stays far below it).
"""
import numpy as np

GRAD_NORM_REL = 5e-11
GRAD_NORM_FLOOR = 1e-16


def assert_grad_norms_close(g, g_ref, feature_scale=1.0):
    """feature_scale: the largest |feature| when it exceeds 1 -- the gradient's
    terms, and with them the evaluation noise, grow with it."""
    g = np.asarray(g, dtype=np.float64)
    g_ref = np.asarray(g_ref, dtype=np.float64)
    bound = GRAD_NORM_REL * g_ref + GRAD_NORM_FLOOR * max(1.0, feature_scale)
    bad = np.abs(g - g_ref) > bound
    assert not bad.any(), (np.nonzero(bad)[0], g[bad], g_ref[bad])
