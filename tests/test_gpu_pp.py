"""FedNL-PP on the GPU engine (dopt/bin_fednl_local_pp) against the oracle.

Each round the engine draws tau clients with the reference's client sampler
(partial dopt::shuffle of 0..n-1, seed 1234); the drawn sets must be
identical to the oracle's, round by round. The selected clients do the FedNL
client step and send the changes of their direction and (option b) of their
Lk; the master keeps the running averages and solves for the next iterate.
Trajectories (the full gradient norm, f over the selected clients, ||x||)
follow the oracle within the FP64 tolerance. The cases are the ones the
oracle is checked on against the reference build.
"""
import numpy as np
import pytest

from tests.test_oracle_vs_ref import PP_CASES
from tests.tolerances import assert_grad_norms_close

pytestmark = pytest.mark.gpu


def _problem(seed):
    from ufednl_amd import datagen
    return datagen.small_problem(n=5, m=40, d_feat=11, seed=seed)


@pytest.mark.parametrize("comp,seed,tau,option_b", PP_CASES)
def test_pp_trajectory_matches_oracle(oracle, comp, seed, tau, option_b):
    from ufednl_amd.engine import Engine
    prob = _problem(seed)
    R = 30
    f = oracle.FedNL(prob, comp, k_mult=1.0, option_b=option_b, workers=2, pp=(tau, 1234))
    eng = Engine.from_problem(prob, compressor=comp, k_mult=1.0, option_b=option_b, workers=2, tracking=True,
                              pp=(tau, 1234))
    outs = []
    for r in range(R):
        outs.append(f.pp_round(tracking=True))
        eng.run(1)
        np.testing.assert_array_equal(eng.pp_selection(), outs[-1]["sel"], err_msg=f"round {r}")
    tr = eng.trace()
    assert np.all(tr["participants"] == tau)
    items_ref = np.array([o["cnt"].sum() for o in outs], dtype=np.float64)
    # TopLEK's random rounding compares energies that the engine sums in
    # another order than the oracle: a draw on the edge may flip. The rounds
    # up to the first differing count are compared (x_{r+1} is the first
    # iterate the flip touches); the other compressors match to the end.
    diff = np.nonzero(tr["items"] != items_ref)[0]
    upto = R if diff.size == 0 else int(diff[0]) + 1
    if comp != "toplek":
        assert upto == R
    assert upto >= 10
    assert_grad_norms_close(tr["grad_norm"][:upto], np.array([o["grad_norm"] for o in outs])[:upto])
    np.testing.assert_allclose(tr["f"][:upto], [o["f"] for o in outs][:upto], rtol=1e-12)
    np.testing.assert_allclose(tr["x_norm"][:upto], [o["x_norm"] for o in outs][:upto], rtol=1e-9, atol=1e-14)
    np.testing.assert_array_equal(tr["chol_ok"][:upto], [int(o["ok"]) for o in outs][:upto])
    if upto == R:
        np.testing.assert_allclose(eng.last_round()["grad"], outs[-1]["grad"], rtol=1e-9, atol=1e-15)
    eng.close()


def test_pp_c1_shape_converges(oracle):
    """A C1-sized problem (50 clients x 221 samples, d=68), 10 clients per
    round, TopK k=8d: the engine follows the oracle for 40 rounds while
    ||grad f|| falls by more than an order of magnitude."""
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    prob = datagen.config1()
    R = 40
    f = oracle.FedNL(prob, "topk", k_mult=8.0, pp=(10, 1234))
    g_ref = np.array([f.pp_round(tracking=False)["grad_norm"] for _ in range(R)])
    eng = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False, pp=(10, 1234))
    eng.run(R)
    g = eng.trace()["grad_norm"]
    assert_grad_norms_close(g, g_ref)
    assert g[-1] < 0.1 * g[0]
    eng.close()


@pytest.mark.parametrize("mode", ["fp32", "bf16x2"])
def test_pp_reduced_precision_hessian(mode):
    """FedNL-PP with the FP32 / BF16-split Hessian reaches ||grad f||^2 <=
    1e-17 at the FP64 run's optimum (every client each round: tau = n)."""
    from ufednl_amd.engine import Engine
    prob = _problem(9)
    xs = {}
    for h in ("fp64", mode):
        eng = Engine.from_problem(prob, compressor="topk", k_mult=1.0, tracking=False, hessian=h, pp=(5, 1234))
        eng.run(40, check=False)
        g2 = eng.trace()["grad_norm"] ** 2
        assert g2.min() <= 1e-17, (h, g2.min())
        xs[h] = eng.x()
        eng.close()
    assert np.max(np.abs(xs[mode] - xs["fp64"])) <= 1e-8 * (1.0 + np.max(np.abs(xs["fp64"])))


def test_pp_argument_errors():
    from ufednl_amd.engine import Batch, Engine, FednlError
    prob = _problem(3)
    with pytest.raises(FednlError):
        Engine.from_problem(prob, compressor="topk", k_mult=1.0, pp=(6, 1234))   # tau > n
    with pytest.raises(FednlError):
        Engine.from_problem(prob, compressor="topk", k_mult=1.0, pp=(2, 1234), line_search=(0.5, 0.5, 5))
    eng = Engine.from_problem(prob, compressor="topk", k_mult=1.0, pp=(2, 1234))
    with pytest.raises(FednlError):
        Batch([eng])
    eng.close()
