"""Full FedNL rounds on the GPU vs the oracle (same seeded problem).

RandK / RandSeqK select indices independently of the values, so the whole
trajectory is compared (||grad f||^2 within 1e-10 relative while above the
floating-point floor). TopK / TopLEK are compared in lockstep: the oracle's
state (x, learned Hessians, H, generator states) is loaded into the engine
before every round and one round's outputs are compared.
"""
import numpy as np
import pytest

from tests.lockstep import lockstep_round
from tests.tolerances import assert_grad_norms_close
from ufednl_amd import datagen

pytestmark = pytest.mark.gpu

FLOOR = 1e-24


def _engine(prob, comp, **kw):
    from ufednl_amd.engine import Engine
    return Engine.from_problem(prob, compressor=comp, **kw)


@pytest.mark.parametrize("comp,kmult", [("randk", 2.0), ("seqk", 2.0), ("randk", 5.0), ("natural", 0.0),
                                       ("identical", 0.0)])
def test_trajectory_value_independent(oracle, comp, kmult):
    prob = datagen.small_problem(n=5, m=41, d_feat=11, density=0.35, seed=11)
    R = 30
    print(f"trajectory {comp} k_mult={kmult}", flush=True)
    orc = oracle.FedNL(prob, comp, k_mult=kmult)
    eng = _engine(prob, comp, k_mult=kmult)
    eng.run(R)
    tr = eng.trace()
    g_o = np.array([orc.round()["grad_norm"] for _ in range(R)])
    assert_grad_norms_close(tr["grad_norm"], g_o)
    assert np.all(tr["chol_ok"] == 1)


@pytest.mark.parametrize("comp", ["topk", "toplek", "randk", "seqk", "natural", "identical"])
def test_lockstep_rounds(oracle, comp):
    """Selection, message values, learned Hessians, master H and generator
    states bit-exact on the engine's Delta; g, Lk, Delta, dx and the next
    iterate against the oracle's round (tests/lockstep.py)."""
    prob = datagen.small_problem(n=6, m=45, d_feat=13, density=0.3, seed=5)
    kmult = 2.0
    print(f"lockstep {comp}", flush=True)
    orc = oracle.FedNL(prob, comp, k_mult=kmult, workers=2)
    eng = _engine(prob, comp, k_mult=kmult, workers=2)
    for r in range(8):
        lockstep_round(eng, orc, comp, workers=2)


@pytest.mark.parametrize("comp,alpha_mode,alpha", [("topk", 1, None), ("toplek", 1, None), ("topk", 0, 0.37),
                                                   ("randk", 0, 0.37), ("seqk", 0, 1.0)])
def test_alpha_option1_and_explicit(oracle, comp, alpha_mode, alpha):
    """--use_theoretical_alpha_option_1 (alpha = 1 - sqrt(1 - k/d^2)) and an
    explicit --alpha_step_size (TrainLoop.h:358-518)."""
    prob = datagen.small_problem(n=6, m=45, d_feat=13, density=0.3, seed=5)
    kmult = 2.0
    d = prob.d
    k = int(kmult * d)
    a_ref = oracle.theoretical_alpha(comp, k, d, option=1) if alpha_mode == 1 else alpha
    orc = oracle.FedNL(prob, comp, k_mult=kmult, alpha=a_ref, workers=2)
    eng = _engine(prob, comp, k_mult=kmult, workers=2, alpha_mode=alpha_mode, alpha=alpha or 0.0)
    assert eng.alpha == a_ref
    for r in range(6):
        lockstep_round(eng, orc, comp, workers=2)


@pytest.mark.parametrize("option_b", [False, True])
def test_identical_round_is_newton_step(option_b):
    """Identical with alpha = 1 makes each round the Newton step with the
    option's shift, computed in numpy from the closed-form logistic
    derivatives (tests/closed_form.py)."""
    from tests.closed_form import newton_round
    for prob in (datagen.small_problem(n=5, m=41, d_feat=11, density=0.35, seed=11), datagen.config1()):
        eng = _engine(prob, "identical", option_b=option_b)
        assert eng.alpha == 1.0
        for r in range(4):
            st = eng.get_state()
            dx_cf, x_cf, shift = newton_round(prob, st["x"], st["learned"], option_b)
            eng.run(1)
            out = eng.last_round()
            s = np.abs(dx_cf).max()
            np.testing.assert_allclose(out["dx"], dx_cf, rtol=1e-9, atol=1e-11 * s, err_msg=f"round {r}")
            np.testing.assert_allclose(eng.x(), x_cf, rtol=1e-12, atol=1e-11 * s, err_msg=f"round {r}")
            if option_b:
                np.testing.assert_allclose(np.mean(out["lk"]), shift, rtol=1e-9, atol=1e-12)
        eng.close()


def test_round0_selects_first_k_for_topk():
    """Delta_0 = H(x0) - learned(x0) = 0 exactly, so the heap keeps the first k list positions."""
    prob = datagen.small_problem(n=3, m=30, d_feat=9, seed=2)
    eng = _engine(prob, "topk", k_mult=2.0)
    eng.run(1)
    out = eng.last_round()
    for c in range(eng.n):
        np.testing.assert_array_equal(out["pos"][c], np.arange(eng.k, dtype=np.uint32))
        assert np.all(out["vals"][c] == 0.0)


def test_config1_topk_lockstep_first_rounds(oracle):
    prob = datagen.config1()
    print("config1 topk lockstep", flush=True)
    orc = oracle.FedNL(prob, "topk", k_mult=8.0)
    eng = _engine(prob, "topk", k_mult=8.0)
    for r in range(4):
        diag = lockstep_round(eng, orc, "topk", max_flip_frac=1e-3)
        print(f"config1 round {r}: {diag}", flush=True)


def test_config1_converges_like_oracle(oracle):
    prob = datagen.config1()
    R = 40
    eng = _engine(prob, "topk", k_mult=8.0)
    eng.run(R)
    tr = eng.trace()
    orc = oracle.FedNL(prob, "topk", k_mult=8.0)
    g_o = np.array([orc.round()["grad_norm"] for _ in range(R)])
    # same trajectory to ~1e-12 over the early rounds, same decay afterwards
    rel = np.abs(tr["grad_norm"][:10] ** 2 - g_o[:10] ** 2) / g_o[:10] ** 2
    assert rel.max() < 1e-8, rel
    assert tr["grad_norm"][-1] ** 2 < 1.5 * g_o[-1] ** 2 + 1e-20
