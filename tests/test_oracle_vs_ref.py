"""The oracle against the reference's own code compiled in oracle/_ref
(skipped where the reference sources were not available at build time)."""
import numpy as np
import pytest

from tests.tolerances import assert_grad_norms_close
from ufednl_amd import datagen


@pytest.fixture(scope="module")
def ref(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref/libdopt_ref.so not built (reference sources absent)")
    return oracle.ref()


def _ref_topk(oracle, ref, vals, d, k):
    full = np.zeros(d * oracle.lda(d))
    full[oracle.upper_list(d)] = vals
    out = np.zeros(max(k, len(vals)), dtype=np.uint32)
    c = ref.ref_topk(oracle.dp(full), d, k, oracle.up(out))
    return out[:c]


def test_topk_random(oracle, ref):
    rng = np.random.default_rng(99)
    for trial in range(300):
        d = int(rng.integers(1, 45))
        S = d * (d + 1) // 2
        k = int(rng.integers(1, S + 2))
        kind = trial % 5
        if kind == 0:
            v = rng.standard_normal(S)
        elif kind == 1:
            v = rng.integers(-2, 3, S).astype(float)
        elif kind == 2:
            v = np.where(rng.random(S) < 0.5, 0.0, -0.0)
        elif kind == 3:
            v = np.round(rng.standard_normal(S), 1)
        else:
            v = np.repeat(rng.standard_normal((S + 3) // 4), 4)[:S]
        got = oracle.pos_to_flat(oracle.topk_packed(v, k), d)
        np.testing.assert_array_equal(got, _ref_topk(oracle, ref, v, d, k), err_msg=f"trial {trial}")


def test_client_eval_and_newton(oracle, ref):
    prob = datagen.config3()
    d, m = prob.d, prob.m
    rng = np.random.default_rng(1)
    x = rng.standard_normal(d) * 0.2
    o = oracle.client_eval(prob.features[3], prob.labels[3], prob.lam, x)
    t = np.zeros(m); s = np.zeros(m); g = np.zeros(d); h = np.zeros(d * oracle.lda(d)); f = np.zeros(1)
    ref.ref_client_eval(oracle.dp(np.ascontiguousarray(prob.features[3])), oracle.dp(np.ascontiguousarray(prob.labels[3])),
                        d, m, prob.lam, oracle.dp(x), oracle.dp(t), oracle.dp(s), oracle.dp(g), oracle.dp(h),
                        oracle.dp(f))
    np.testing.assert_allclose(o["grad"], g, rtol=1e-13, atol=1e-17)
    # binary features: products are exact, so the Hessian sums agree bit for bit
    # whenever the sigmoids do (they differ by at most an ulp of exp)
    np.testing.assert_allclose(o["hess_up"], oracle.pack_upper(h, d), rtol=1e-13, atol=1e-18)
    H = o["hess_up"]
    b = rng.standard_normal(d)
    ok, dx, _ = oracle.cholesky_solve(H, d, 1e-3, b)
    dxr = np.zeros(d)
    okr = ref.ref_newton_solve(oracle.dp(oracle.unpack_full(H, d)), d, 1e-3, oracle.dp(b), oracle.dp(dxr), None)
    assert ok and okr
    np.testing.assert_allclose(dx, dxr, rtol=1e-11)


@pytest.mark.parametrize("comp,seed", [("randk", 21), ("seqk", 22), ("topk", 23), ("toplek", 24),
                                       ("natural", 25), ("identical", 26)])
def test_fednl_trajectories(oracle, comp, seed):
    prob = datagen.small_problem(n=7, m=50, d_feat=15, density=0.3, seed=seed)
    R = 20
    r = oracle.ref_fednl_run(prob, comp, 3.0, rounds=R, workers=3)
    f = oracle.FedNL(prob, comp, k_mult=3.0, workers=3)
    g = np.array([f.round()["grad_norm"] for _ in range(R)])
    assert_grad_norms_close(g, r["grad_norm"])
    np.testing.assert_allclose(f.get_state()["x"], r["x"], rtol=1e-7, atol=1e-9)


# FedNL-LS (TrainLoop.h:1546-1640). With option a (shift lambda) and a small
# k the first Newton steps overshoot, so backtracking, and with a low
# max_steps the forced stop at the last trial, happen; c stays within the
# reference's (0, 0.5]. Trial counts are compared while ||grad f|| is above
# 1e-7: below it the acceptance test compares f values that differ by less
# than their own rounding.
LS_CASES = [("topk", 3, 2.0, False, (0.49, 0.5, 5)), ("topk", 9, 2.0, False, (0.49, 0.5, 2)),
            ("toplek", 3, 4.0, False, (0.49, 0.5, 5)), ("seqk", 9, 4.0, False, (0.3, 0.7, 4)),
            ("randk", 10, 4.0, False, (0.49, 0.5, 5)), ("natural", 7, 4.0, True, (0.5, 0.5, 5)),
            ("topk", 5, 1.0, True, (0.5, 0.5, 5))]


@pytest.mark.parametrize("comp,seed,scale,option_b,ls", LS_CASES)
def test_fednl_line_search_trajectories(oracle, comp, seed, scale, option_b, ls):
    prob = datagen.small_problem(n=5, m=40, d_feat=11, seed=seed)
    prob.features[...] *= scale
    R = 25
    r = oracle.ref_fednl_run(prob, comp, 0.5, option_b=option_b, rounds=R, tracking=False, line_search=ls)
    f = oracle.FedNL(prob, comp, k_mult=0.5, option_b=option_b, line_search=ls)
    outs = [f.round(tracking=False) for _ in range(R)]
    g = np.array([o["grad_norm"] for o in outs])
    cnt = np.array([o["ls_count"] for o in outs])
    # rounds up to the first one at ||grad f|| <= 1e-7 (its own trial count
    # excluded) follow acceptance decisions that are not rounding-bound
    live = int(np.argmin(r["grad_norm"] > 1e-7)) if np.any(r["grad_norm"] <= 1e-7) else R
    assert live >= 5 and all(o["ok"] for o in outs)
    assert_grad_norms_close(g[:live + 1], r["grad_norm"][:live + 1], feature_scale=scale)
    np.testing.assert_array_equal(cnt[:live], np.asarray(r["ls_counts"])[:live])
    assert np.all((cnt >= 1) & (cnt <= ls[2]))
    return cnt[:live]


def test_config1_topk_first_rounds(oracle, ref):
    prob = datagen.config1()
    R = 12
    r = oracle.ref_fednl_run(prob, "topk", 8.0, rounds=R, workers=4, dump_rounds=2)
    f = oracle.FedNL(prob, "topk", k_mult=8.0)
    outs = [f.round() for _ in range(R)]
    assert_grad_norms_close(np.array([o["grad_norm"] for o in outs]), r["grad_norm"])
    for rr in range(2):
        for c in range(prob.n):
            np.testing.assert_array_equal(oracle.pos_to_flat(outs[rr]["pos"][c], prob.d), r["dump"]["idx"][rr, c])


# A failed Cholesky factorization (a negative pivot of H + shift I, here with
# option a and a small k) falls back to Gauss-Jordan elimination on the
# master's full H, whose lower triangle keeps its creation-time values
# (TrainLoop.h:1410-1420, 1476-1486; GaussEliminationSolvers.h:27-120).
FALLBACK_CASES = [("topk", 4, 1.0, None), ("topk", 10, 2.0, None), ("toplek", 6, 4.0, (0.49, 0.5, 5)),
                  ("topk", 8, 2.0, (0.49, 0.5, 2))]


@pytest.mark.parametrize("comp,seed,scale,ls", FALLBACK_CASES)
def test_fednl_factorization_fallback(oracle, comp, seed, scale, ls):
    prob = datagen.small_problem(n=5, m=40, d_feat=11, seed=seed)
    prob.features[...] *= scale
    R = 25
    r = oracle.ref_fednl_run(prob, comp, 0.5, option_b=False, rounds=R, tracking=False, line_search=ls)
    f = oracle.FedNL(prob, comp, k_mult=0.5, option_b=False, line_search=ls)
    outs = [f.round(tracking=False) for _ in range(R)]
    ok = np.array([o["ok"] for o in outs], dtype=np.int32)
    assert (ok == 0).any()
    np.testing.assert_array_equal(ok, r["chol_ok"])
    g = np.array([o["grad_norm"] for o in outs])
    live = int(np.argmin(r["grad_norm"] > 1e-7)) if np.any(r["grad_norm"] <= 1e-7) else R
    assert_grad_norms_close(g[:live + 1], r["grad_norm"][:live + 1], feature_scale=scale)


def test_gauss_solve_known_answers(oracle):
    """Gauss-Jordan restatement: exact on an integer system that needs row
    swaps (zero leading entry) and equal to numpy's solve on random
    non-symmetric matrices."""
    M = np.array([[0.0, 2.0, 1.0], [1.0, -2.0, -3.0], [-1.0, 1.0, 2.0]])
    xs = np.array([1.0, -2.0, 3.0])
    b = M @ xs
    x = np.zeros(3)
    oracle.lib().orc_gauss_solve(oracle.dp(M.copy()), oracle.dp(b.copy()), 3, oracle.dp(x))
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-14)
    rng = np.random.default_rng(3)
    for d in (1, 7, 40):
        A = rng.standard_normal((d, d))
        bb = rng.standard_normal(d)
        x = np.zeros(d)
        oracle.lib().orc_gauss_solve(oracle.dp(A.copy()), oracle.dp(bb.copy()), d, oracle.dp(x))
        np.testing.assert_allclose(x, np.linalg.solve(A, bb), rtol=1e-10, atol=1e-12)


# FedNL-PP (bin_fednl_local_pp): tau clients per round from the client
# sampler (seed 1234, Args.h:242-244), running averages of the clients'
# direction and Lk changes, x_{k+1} = (H + shift I)^-1 g.
PP_CASES = [("topk", 3, 2, True), ("randk", 4, 3, True), ("seqk", 5, 1, False), ("toplek", 6, 4, True),
            ("natural", 7, 2, True), ("identical", 8, 3, False), ("topk", 9, 5, True)]


@pytest.mark.parametrize("comp,seed,tau,option_b", PP_CASES)
def test_fednl_pp_trajectories(oracle, comp, seed, tau, option_b):
    prob = datagen.small_problem(n=5, m=40, d_feat=11, seed=seed)
    R = 30
    r = oracle.ref_fednl_pp_run(prob, comp, 1.0, tau=tau, seed=1234, option_b=option_b, rounds=R, workers=2)
    f = oracle.FedNL(prob, comp, k_mult=1.0, option_b=option_b, workers=2, pp=(tau, 1234))
    outs = [f.pp_round(tracking=True) for _ in range(R)]
    np.testing.assert_array_equal(np.array([o["sel"] for o in outs]), r["sel"])
    assert np.all(r["sel"].sum(axis=1) == tau)
    np.testing.assert_array_equal([int(o["ok"]) for o in outs], r["chol_ok"])
    assert_grad_norms_close(np.array([o["grad_norm"] for o in outs]), r["grad_norm"])
    np.testing.assert_allclose([o["f"] for o in outs], r["f"], rtol=1e-13)
    np.testing.assert_allclose([o["x_norm"] for o in outs], r["x_norm"], rtol=1e-10, atol=1e-14)


def test_pp_client_sampler_is_the_reference_shuffle(oracle):
    """The clients of a round are the first tau entries of dopt::shuffle of
    0..n-1 (Shuffle.h:31-93), one generator across rounds."""
    n, tau, R = 9, 4, 6
    prob = datagen.small_problem(n=n, m=20, d_feat=5, seed=2)
    r = oracle.ref_fednl_pp_run(prob, "topk", 1.0, tau=tau, seed=1234, rounds=R)
    st = np.array([1234], dtype=np.uint32)
    for rr in range(R):
        v = np.arange(n, dtype=np.uint32)
        oracle.lib().orc_shuffle_partial(oracle.up(v), n, tau, oracle.up(st))
        want = np.zeros(n, dtype=np.int32)
        want[v[:tau]] = 1
        np.testing.assert_array_equal(r["sel"][rr], want)


@pytest.mark.parametrize("tracking", [True, False])
def test_gd_matches_reference(oracle, tracking):
    """GD (--algorithm gd): the oracle's rounds against the reference's own
    client oracle and vector arithmetic (LocalWorkerForGD.h, TrainLoop.h:993-1052),
    including the recorded f aggregate (each client adds f_c and then
    sizeof(double) before the division by n, so f(x) reads as the mean + 8)."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built")
    from ufednl_amd import datagen
    prob = datagen.small_problem(n=6, m=41, d_feat=15, density=0.3, seed=12)
    a = oracle.gd_run(prob, 1.5, 40, tracking)
    b = oracle.ref_gd_run(prob, 1.5, 40, tracking)
    for key in ("grad_norm", "x_norm"):
        np.testing.assert_allclose(a[key], b[key], rtol=1e-13, atol=1e-300, err_msg=key)
    np.testing.assert_allclose(a["f"], b["f"], rtol=1e-14, atol=0.0)
    np.testing.assert_allclose(a["x"], b["x"], rtol=1e-12, atol=1e-15)
    if tracking:
        # round 0 at x = 0: every sigma is 1/2, f_c = log 2
        assert abs(a["f"][0] - (np.log(2.0) + 8.0)) < 1e-14
    else:
        assert not a["f"].any()
    # the step: x_1 = -step * g(0), so ||x_1|| = step * ||g_0||
    assert abs(a["x_norm"][1] - 1.5 * a["grad_norm"][0]) <= 1e-14 * a["x_norm"][1]
