"""The sparse FP64 Hessian form (binary designs) against the oracle and
against the dense MFMA form.

With 0/1 features and +-1 labels every nonzero term A[s,i] (D_s A[s,j]) of
the reference's Hessian sum (logistic_regression.h:365-547) is D_s exactly,
and every nonzero gradient term A[s,i] g_s is +-g_s, so the sparse form's
ascending-order sums over the samples that hold the features are the
oracle's sequential sums bit for bit whenever the sample weights agree. At
x = 0 they do (sigma(0) = 1/2 exactly on both sides): the learned Hessians
made at creation and the first round's gradient are compared bit for bit.
Elsewhere the weights carry the device exp's rounding and the forms agree
within FP64 tolerance; whole runs are compared with the dense form and the
oracle.
"""
import numpy as np
import pytest

from tests.lockstep import lockstep_round
from tests.tolerances import assert_grad_norms_close
from ufednl_amd import datagen

pytestmark = pytest.mark.gpu

# (n, m, d_feat, density): one segment of 4096 entries; two segments; eleven
# segments (d = 300, C2's density) with empty trailing tile slots; a client
# count above one XCD round
SHAPES = [(5, 120, 40, 0.08), (6, 200, 119, 0.04), (3, 90, 299, 0.04), (17, 64, 60, 0.1)]


def _binary(shape, seed):
    n, m, df, rho = shape
    return datagen.binary_problem(f"sp_n{n}_m{m}_d{df + 1}", n, m, df, rho, seed=seed, target_pos=0.2)


def _engine(prob, form, **kw):
    from ufednl_amd.engine import Engine
    kw.setdefault("compressor", "topk")
    kw.setdefault("k_mult", 8.0)
    return Engine.from_problem(prob, hessian_form=form, **kw)


@pytest.mark.parametrize("shape", SHAPES)
def test_learned_hessians_at_x0_bit_exact(oracle, shape):
    prob = _binary(shape, seed=shape[0] * shape[1])
    print(f"sparse H(0) {shape}", flush=True)
    eng = _engine(prob, "sparse")
    assert eng.hessian_form == "sparse"
    x0 = np.zeros(eng.d)
    H = eng.hessians_at(x0)
    for c in range(eng.n):
        ref = oracle.client_eval(prob.features[c], prob.labels[c], prob.lam, x0)["hess_up"]
        bad = np.nonzero(H[c] != ref)[0]
        assert len(bad) == 0, (c, bad[:5], H[c][bad[:5]], ref[bad[:5]])
    eng.close()


@pytest.mark.parametrize("shape", SHAPES[1:3])
def test_first_round_gradient_bit_exact(oracle, shape):
    """Round 0 runs at x0 = 0: client gradients lambda x + sum_s b_s g_s in
    sample order, their client-order sum times 1/n, as the reference."""
    prob = _binary(shape, seed=7 + shape[0])
    eng = _engine(prob, "sparse")
    orc = oracle.FedNL(prob, "topk", k_mult=8.0)
    eng.run(1)
    g = eng.last_round()["grad"]
    g_ref = orc.round()["grad"]
    assert np.array_equal(g, g_ref), np.abs(g - g_ref).max()
    eng.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_hessians_at_x_sparse_vs_dense_and_oracle(oracle, shape):
    prob = _binary(shape, seed=3 * shape[0] + shape[2])
    rng = np.random.default_rng(11)
    es = _engine(prob, "sparse")
    ed = _engine(prob, "dense")
    assert ed.hessian_form == "dense"
    x = 0.3 * rng.standard_normal(es.d)
    Hs, Hd = es.hessians_at(x), ed.hessians_at(x)
    scale = np.abs(Hd).max()
    err_d = np.abs(Hs - Hd).max() / scale
    err_o = 0.0
    for c in range(es.n):
        ref = oracle.client_eval(prob.features[c], prob.labels[c], prob.lam, x)["hess_up"]
        err_o = max(err_o, np.abs(Hs[c] - ref).max() / np.abs(ref).max())
    print(f"{shape}: sparse vs dense {err_d:.2e}, vs oracle {err_o:.2e}", flush=True)
    # the weights D differ from the oracle's by the exp's rounding only
    assert err_d <= 1e-13 and err_o <= 1e-13
    es.close()
    ed.close()


@pytest.mark.parametrize("comp", ["randk", "seqk"])
def test_value_independent_trajectory_vs_oracle(oracle, comp):
    prob = _binary((8, 150, 119, 0.05), seed=21)
    R = 40
    eng = _engine(prob, "sparse", compressor=comp, k_mult=2.0)
    orc = oracle.FedNL(prob, comp, k_mult=2.0)
    eng.run(R)
    g = eng.trace()["grad_norm"]
    g_o = np.array([orc.round()["grad_norm"] for _ in range(R)])
    assert_grad_norms_close(g, g_o)
    eng.close()


@pytest.mark.parametrize("comp", ["topk", "toplek"])
def test_lockstep_sparse(oracle, comp):
    """Selection, messages, learned and master H bit-exact on the engine's
    own Delta; Delta, gradient, Lk and the step against the oracle's round."""
    prob = _binary((6, 200, 119, 0.04), seed=5)
    orc = oracle.FedNL(prob, comp, k_mult=8.0, workers=2)
    eng = _engine(prob, "sparse", compressor=comp, workers=2)
    for _ in range(2):
        orc.round()
    for r in range(5):
        lockstep_round(eng, orc, comp, workers=2, max_flip_frac=1e-3)
    eng.close()


@pytest.mark.parametrize("opts", [dict(), dict(line_search=(0.5, 0.5, 5)), dict(pp=(3, 1234))])
def test_sparse_matches_dense_runs(opts):
    """FedNL, FedNL-LS and FedNL-PP (option b, which records H_c(x)) reach
    the same optimum with either form. (FedNL-PP with option a diverges on
    this problem in the reference as well.)"""
    prob = _binary((6, 200, 119, 0.04), seed=9)
    R = 250
    xs = []
    for form in ("sparse", "dense"):
        eng = _engine(prob, form, tracking=True, **opts)
        assert eng.hessian_form == form
        eng.run(R)
        g = eng.trace()["grad_norm"]
        xs.append(eng.x())
        print(f"{form} {opts}: ||g||^2 at the end {g[-1] ** 2:.3e}", flush=True)
        assert g[-1] ** 2 <= 1e-17
        eng.close()
    assert np.abs(xs[0] - xs[1]).max() <= 1e-8


def test_batch_member_bit_identical_to_solo():
    """A sparse member and a dense member stepped together in one batch give
    the bits of their solo runs."""
    from ufednl_amd.engine import Batch
    pa = _binary((6, 200, 119, 0.04), seed=31)
    pb = datagen.gaussian_problem("g_n5_m80_d40", 5, 80, 40, seed=32)
    R = 12
    solo = []
    for prob, form in ((pa, "sparse"), (pb, "auto")):
        eng = _engine(prob, form)
        eng.run(R)
        solo.append((eng.trace()["grad_norm"], eng.x()))
        eng.close()
    ea, eb = _engine(pa, "sparse"), _engine(pb, "auto")
    assert ea.hessian_form == "sparse" and eb.hessian_form == "dense"
    b = Batch([ea, eb])
    b.run(R)
    for eng, (g0, x0) in zip((ea, eb), solo):
        assert np.array_equal(eng.trace()["grad_norm"], g0)
        assert np.array_equal(eng.x(), x0)
    b.close()
    ea.close()
    eb.close()


def test_auto_choice():
    """Auto: sparse for the binary low-density designs, dense for C1's
    Bernoulli(1/2) design, for Gaussian features and in the reduced-precision
    modes."""
    sp = _binary((4, 100, 99, 0.04), seed=2)
    cases = [(sp, {}, "sparse"), (sp, dict(hessian="fp32"), "dense"),
             (datagen.binary_problem("b05", 4, 100, 30, 0.5, seed=3), {}, "dense"),
             (datagen.gaussian_problem("g", 4, 60, 20, seed=4), {}, "dense")]
    for prob, kw, want in cases:
        eng = _engine(prob, "auto", **kw)
        assert eng.hessian_form == want, (prob.name, kw)
        eng.close()


def test_forced_sparse_rejected_where_it_does_not_apply():
    from ufednl_amd.engine import FednlError
    g = datagen.gaussian_problem("g", 3, 40, 12, seed=5)
    with pytest.raises(FednlError, match="sparse Hessian form"):
        _engine(g, "sparse", k_mult=2.0)
    sp = _binary((3, 50, 20, 0.1), seed=6)
    with pytest.raises(FednlError, match="sparse Hessian form"):
        _engine(sp, "sparse", hessian="bf16x2")


def _with_dense_column(prob, col):
    feats = prob.features.copy()
    feats[:, :, col] = 1.0
    return datagen.Problem(prob.name + f"_dense{col}", feats, prob.labels, prob.lam, prob.w_true)


@pytest.mark.parametrize("case", ["second_dense_column", "intercept_only", "empty_features"])
def test_runs_and_edge_designs_bit_exact(oracle, case):
    """Runs (consecutive samples: a dense column's diagonal and gradient, and
    the off-diagonal entry of two dense columns), a design of the intercept
    alone (d = 1) and one whose features are all zero: H(0) against the
    oracle bit for bit."""
    if case == "second_dense_column":
        prob = _with_dense_column(_binary((4, 120, 60, 0.06), seed=41), 10)
        kw = {}
    elif case == "intercept_only":
        prob = _binary((3, 50, 0, 0.0), seed=42)
        kw = dict(k_mult=1.0)
    else:
        prob = _binary((3, 70, 20, 0.0), seed=43)
        kw = dict(k_mult=2.0)
    eng = _engine(prob, "sparse", **kw)
    assert eng.hessian_form == "sparse"
    x0 = np.zeros(eng.d)
    H = eng.hessians_at(x0)
    for c in range(eng.n):
        ref = oracle.client_eval(prob.features[c], prob.labels[c], prob.lam, x0)["hess_up"]
        assert np.array_equal(H[c], ref), (case, c)
    eng.close()


def test_runs_in_rounds_match_dense():
    """A design with two dense columns: runs in every round (Delta, the
    gradient through them) against the dense form."""
    prob = _with_dense_column(_binary((5, 160, 80, 0.05), seed=44), 7)
    R = 40
    out = []
    for form in ("sparse", "dense"):
        eng = _engine(prob, form, compressor="randk", k_mult=2.0, tracking=True)
        eng.run(R)
        out.append(eng.trace()["grad_norm"])
        eng.close()
    assert_grad_norms_close(out[0], out[1])


def test_line_search_batch_member_bit_identical_to_solo():
    """FedNL-LS members (a sparse and a dense one) in one batch: the trial
    margins of the sparse member run in the batch's sparse margin launch;
    both give their solo runs' bits, trial counts included."""
    from ufednl_amd.engine import Batch
    pa = _binary((6, 150, 90, 0.05), seed=51)
    pb = datagen.gaussian_problem("g_n4_m70_d30", 4, 70, 30, seed=52)
    ls = (0.5, 0.5, 5)
    R = 10
    solo = []
    for prob, form in ((pa, "sparse"), (pb, "dense")):
        eng = _engine(prob, form, line_search=ls)
        eng.run(R)
        solo.append((eng.trace()["grad_norm"], eng.x(), eng.trace()["ls_trials"]))
        eng.close()
    ea, eb = _engine(pa, "sparse", line_search=ls), _engine(pb, "dense", line_search=ls)
    b = Batch([ea, eb])
    b.run(R)
    for eng, (g0, x0, t0) in zip((ea, eb), solo):
        assert np.array_equal(eng.trace()["grad_norm"], g0)
        assert np.array_equal(eng.x(), x0)
        assert np.array_equal(eng.trace()["ls_trials"], t0)
    b.close()
    ea.close()
    eb.close()
