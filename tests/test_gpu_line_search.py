"""FedNL-LS on the GPU engine (TrainLoop.h:344-355, 1546-1640) against the
oracle: the Newton direction from the round, t = 1, acceptance when
f(x_k - t dx) <= f(x_k) + (c t)(-(dx, g)), else t *= gamma, the last trial
kept after max_line_search_steps.

The cases are the ones the oracle is checked on against the reference build
(backtracking and forced stops happen in them). Trial counts are compared
while ||grad f|| is above 1e-7; below it the acceptance test compares f
values that differ by less than their own rounding, on either side.
"""
import numpy as np
import pytest

from tests.test_oracle_vs_ref import FALLBACK_CASES, LS_CASES
from tests.tolerances import assert_grad_norms_close

pytestmark = pytest.mark.gpu


def _problem(seed, scale):
    from ufednl_amd import datagen
    prob = datagen.small_problem(n=5, m=40, d_feat=11, seed=seed)
    prob.features[...] *= scale
    return prob


def _live(g_ref):
    below = np.asarray(g_ref) <= 1e-7
    return int(np.argmax(below)) if below.any() else len(g_ref)


@pytest.mark.parametrize("comp,seed,scale,option_b,ls", LS_CASES)
def test_line_search_trajectory_matches_oracle(oracle, comp, seed, scale, option_b, ls):
    from ufednl_amd.engine import Engine
    prob = _problem(seed, scale)
    R = 25
    f = oracle.FedNL(prob, comp, k_mult=0.5, option_b=option_b, line_search=ls)
    outs = [f.round(tracking=False) for _ in range(R)]
    g_ref = np.array([o["grad_norm"] for o in outs])
    cnt_ref = np.array([o["ls_count"] for o in outs])
    eng = Engine.from_problem(prob, compressor=comp, k_mult=0.5, option_b=option_b, tracking=False,
                              line_search=ls)
    eng.run(R)
    tr = eng.trace()
    live = _live(g_ref)
    assert live >= 5
    assert_grad_norms_close(tr["grad_norm"][:live + 1], g_ref[:live + 1], feature_scale=scale)
    np.testing.assert_array_equal(tr["ls_trials"][:live], cnt_ref[:live])
    assert np.all((tr["ls_trials"] >= 1) & (tr["ls_trials"] <= ls[2]))
    # f(x_k) of each round is reported with line search on, as with tracking
    fo = np.array([o["f"] for o in outs])
    np.testing.assert_allclose(tr["f"][:live], fo[:live], rtol=1e-13)
    eng.close()


def test_line_search_first_round_state(oracle):
    """Round 0 in lockstep: x after the accepted trial equals the oracle's."""
    from ufednl_amd.engine import Engine
    comp, seed, scale, option_b, ls = LS_CASES[0]
    prob = _problem(seed, scale)
    f = oracle.FedNL(prob, comp, k_mult=0.5, option_b=option_b, line_search=ls)
    eng = Engine.from_problem(prob, compressor=comp, k_mult=0.5, option_b=option_b, tracking=False,
                              line_search=ls)
    for r in range(3):
        o = f.round(tracking=False)
        eng.run(1)
        assert eng.trace()["ls_trials"][r] == o["ls_count"]
        np.testing.assert_allclose(eng.x(), f.get_state()["x"], rtol=1e-12, atol=1e-14)
    eng.close()


def test_zero_trials_keeps_iterate():
    """max_line_search_steps = 0: the trial loop never runs and x stays x_k
    (TrainLoop.h:1566, 1627-1637)."""
    from ufednl_amd.engine import Engine
    prob = _problem(3, 2.0)
    eng = Engine.from_problem(prob, compressor="topk", k_mult=0.5, tracking=False, line_search=(0.5, 0.5, 0))
    eng.run(3)
    assert np.all(eng.x() == 0.0)
    tr = eng.trace()
    assert np.all(tr["ls_trials"] == 0)
    assert tr["grad_norm"][0] == tr["grad_norm"][2]
    eng.close()


@pytest.mark.parametrize("bad", [(0.0, 0.5, 5), (0.5, 1.0, 5), (0.5, 0.0, 5), (0.5, 0.5, -1)])
def test_line_search_argument_errors(bad):
    from ufednl_amd.engine import Engine, FednlError
    prob = _problem(3, 1.0)
    with pytest.raises(FednlError):
        Engine.from_problem(prob, compressor="topk", k_mult=0.5, line_search=bad)


@pytest.mark.parametrize("use_graph", [True, False])
def test_line_search_in_batch_matches_solo(use_graph):
    """Members with and without line search, different max steps, in one
    batch (direct launches or a captured graph): each follows its solo run
    exactly."""
    from ufednl_amd.engine import Batch, Engine
    cases = [(LS_CASES[0], True), (LS_CASES[1], True), (LS_CASES[3], False), (LS_CASES[2], True)]
    rounds = 12

    def make():
        out = []
        for (comp, seed, scale, option_b, ls), on in cases:
            out.append(Engine.from_problem(_problem(seed, scale), compressor=comp, k_mult=0.5, option_b=option_b,
                                           tracking=False, line_search=ls if on else None))
        return out

    solo = []
    for e in make():
        e.run(rounds, check=False)
        tr = e.trace()
        solo.append((tr["grad_norm"].copy(), tr["ls_trials"].copy(), e.x().copy()))
        e.close()
    engs = make()
    b = Batch(engs, use_graph=use_graph)
    b.run(5)
    b.run(rounds - 5)
    b.close()
    for i, e in enumerate(engs):
        tr = e.trace()
        g, cnt, x = solo[i]
        np.testing.assert_array_equal(tr["grad_norm"], g, err_msg=f"member {i}")
        np.testing.assert_array_equal(tr["ls_trials"], cnt, err_msg=f"member {i}")
        np.testing.assert_array_equal(e.x(), x, err_msg=f"member {i}")
        e.close()


@pytest.mark.parametrize("comp,seed,scale,ls", FALLBACK_CASES)
def test_factorization_fallback_matches_oracle(oracle, comp, seed, scale, ls):
    """Rounds whose Cholesky factorization fails are solved by Gauss-Jordan
    elimination on the master's full H (upper part current, lower part as at
    creation), with or without line search; the failing rounds and the
    trajectory follow the oracle."""
    from ufednl_amd.engine import Engine
    prob = _problem(seed, scale)
    R = 25
    f = oracle.FedNL(prob, comp, k_mult=0.5, option_b=False, line_search=ls)
    outs = [f.round(tracking=False) for _ in range(R)]
    g_ref = np.array([o["grad_norm"] for o in outs])
    ok_ref = np.array([o["ok"] for o in outs], dtype=np.int32)
    assert (ok_ref == 0).any()
    eng = Engine.from_problem(prob, compressor=comp, k_mult=0.5, option_b=False, tracking=False, line_search=ls)
    eng.run(R)
    tr = eng.trace()
    live = _live(g_ref)
    np.testing.assert_array_equal(tr["chol_ok"][:live + 1], ok_ref[:live + 1])
    assert_grad_norms_close(tr["grad_norm"][:live + 1], g_ref[:live + 1], feature_scale=scale)
    eng.close()


def test_fallback_step_equals_oracle_solve(oracle):
    """The fallback's dx for a failing round, from the engine's own H and g,
    equals the oracle's Gauss-Jordan solve of the same system bit for bit
    (both round every product and sum separately)."""
    from ufednl_amd.engine import Engine
    comp, seed, scale, ls = FALLBACK_CASES[0]
    prob = _problem(seed, scale)
    eng = Engine.from_problem(prob, compressor=comp, k_mult=0.5, option_b=False, tracking=False)
    n, m, d = prob.features.shape
    H0 = eng.get_state()["H"].copy()
    failed = False
    for r in range(10):
        eng.run(1)
        if eng.trace()["chol_ok"][r] == 0:
            failed = True
            break
    assert failed
    H = eng.get_state()["H"]
    last = eng.last_round()
    M = np.zeros((d, d))
    for j in range(d):
        for i in range(j + 1):
            p = j * (j + 1) // 2 + i
            M[i, j] = H[p] + (prob.lam if i == j else 0.0)
            if i != j:
                M[j, i] = H0[p]
    x = np.zeros(d)
    oracle.lib().orc_gauss_solve(oracle.dp(M), oracle.dp(last["grad"].copy()), d, oracle.dp(x))
    np.testing.assert_array_equal(last["dx"], x)
    eng.close()
