"""Serving loop: requests that arrive over time are admitted by device-memory
footprint and stepped together (continuous batching); each result must be
the one the problem gives alone, and admission must respect the budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REQS = [
    # (arrival_s, n, m, d_feat, density, compressor, k_mult, hessian, seed)
    (0.00, 6, 50, 24, 0.3, "topk", 8.0, "fp64", 41),
    (0.00, 9, 41, 40, 0.2, "seqk", 3.0, "fp64", 42),
    (0.05, 5, 70, 61, 0.15, "randk", 2.0, "fp32", 43),
    (0.10, 12, 33, 17, 0.3, "toplek", 4.0, "bf16x2", 44),
    (0.20, 7, 80, 90, 0.1, "topk", 8.0, "fp64", 45),
]


def _requests(max_rounds=60):
    from ufednl_amd import datagen
    out = []
    for (t, n, m, df, dens, comp, km, hs, seed) in REQS:
        prob = datagen.binary_problem(f"s{seed}", n, m, df, dens, seed=seed, target_pos=0.3)
        out.append(dict(arrival_s=t, problem=prob, max_rounds=max_rounds, tol=1e-17,
                        opts=dict(compressor=comp, k_mult=km, hessian=hs, workers=2)))
    return out


def test_served_results_equal_solo_runs():
    from ufednl_amd import engine as E
    reqs = _requests()
    res, xs = E.serve(reqs, hbm_budget_bytes=8e9, rounds_per_step=3)
    for i, (rq, r) in enumerate(zip(reqs, res)):
        assert r["status"] in (0, 2), (i, r)
        assert r["rounds"] > 0 and r["finished_s"] >= r["admitted_s"] >= rq["arrival_s"] - 1e-9
        eng = E.Engine.from_problem(rq["problem"], tracking=False, **rq["opts"])
        eng.run(r["rounds"], check=False)
        np.testing.assert_array_equal(eng.x(), xs[i], err_msg=f"request {i}")
        g2 = eng.trace()["grad_norm"] ** 2
        hit = np.nonzero(g2 <= 1e-17)[0]
        assert r["rounds_to_tol"] == (int(hit[0]) if hit.size else -1), i
        eng.close()


def test_admission_respects_budget():
    from ufednl_amd import engine as E
    reqs = _requests(max_rounds=9)
    foot = [E.footprint(rq["problem"].features.shape[0], rq["problem"].features.shape[1],
                        rq["problem"].features.shape[2], **rq["opts"]) for rq in reqs]
    budget = max(foot) * 1.5  # at most one large member beside a small one
    res, _ = E.serve(reqs, hbm_budget_bytes=budget, rounds_per_step=3)
    # intervals [admitted, finished) of concurrently resident requests stay within the budget
    for i, r in enumerate(res):
        t = r["admitted_s"]
        live = [j for j, q in enumerate(res) if q["admitted_s"] <= t < q["finished_s"]]
        assert sum(foot[j] for j in live) <= budget or len(live) == 1, (i, live)
    assert all(r["status"] in (0, 2) for r in res)
