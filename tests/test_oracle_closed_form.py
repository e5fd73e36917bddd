"""The oracle's round orchestration against closed forms (CPU).

Identical compressor with alpha = 1: the round is an exact Newton step with
the option's shift (tests/closed_form.py). A dropped 1/n, a wrong alpha
prescale, a shift taken from the wrong option or a transposed Hessian fails
it. Theoretical alpha option 1 satisfies (1 - alpha)^2 = 1 - k/d^2
(TrainLoop.h:409-518, SpecialMathRoutinesForMatrix.h:208-218)."""
import numpy as np
import pytest

from tests.closed_form import newton_round
from ufednl_amd import datagen


@pytest.mark.parametrize("option_b", [False, True])
@pytest.mark.parametrize("which", ["small", "config1"])
def test_identical_round_is_newton_step(oracle, option_b, which):
    prob = (datagen.small_problem(n=5, m=41, d_feat=11, density=0.35, seed=11) if which == "small"
            else datagen.config1())
    orc = oracle.FedNL(prob, "identical", option_b=option_b)
    assert orc.alpha == 1.0
    for r in range(4):
        st = orc.get_state()
        dx_cf, x_cf, shift = newton_round(prob, st["x"], st["learned"], option_b)
        out = orc.round(tracking=False)
        x_next = orc.get_state()["x"]
        s = np.abs(dx_cf).max()
        np.testing.assert_allclose(out["dx"], dx_cf, rtol=1e-9, atol=1e-11 * s, err_msg=f"round {r}")
        np.testing.assert_allclose(x_next, x_cf, rtol=1e-12, atol=1e-11 * s, err_msg=f"round {r}")
        if option_b:
            np.testing.assert_allclose(np.mean(out["lk"]), shift, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("d,kmult", [(68, 8.0), (300, 8.0), (123, 8.0), (10, 0.5)])
def test_theoretical_alpha_option1(oracle, d, kmult):
    k = int(kmult * d)
    a = oracle.theoretical_alpha("topk", k, d, option=1)
    assert 0.0 < a < 1.0
    assert abs((1.0 - a) ** 2 - (1.0 - k / (d * d))) <= 4e-16
