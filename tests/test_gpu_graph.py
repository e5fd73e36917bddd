"""A solo engine's round replayed from a captured HIP graph gives the same
bits as the same round launched kernel by kernel: trajectories, f, ||x||,
factorization flags, line-search trials and the iterate, over several
replays with the program's own eager work (state reads, Hessians at a point)
between them. The per-round timestamps the trace records on the device clock
increase round by round and span the host's wall time of the run.
"""
import os
import time

import numpy as np
import pytest

from ufednl_amd import datagen

pytestmark = pytest.mark.gpu


def _engine(prob, graph, **kw):
    from ufednl_amd.engine import Engine
    old = os.environ.get("FEDNL_ROUND_GRAPH")
    os.environ["FEDNL_ROUND_GRAPH"] = "1" if graph else "0"
    try:
        return Engine.from_problem(prob, **kw)
    finally:
        if old is None:
            del os.environ["FEDNL_ROUND_GRAPH"]
        else:
            os.environ["FEDNL_ROUND_GRAPH"] = old


CASES = {
    "topk": dict(compressor="topk", k_mult=8.0),
    "randk": dict(compressor="randk", k_mult=2.0),
    "toplek": dict(compressor="toplek", k_mult=4.0),
    "seqk_ls": dict(compressor="seqk", k_mult=2.0, line_search=(0.5, 0.5, 6)),
    "natural_pp": dict(compressor="natural", k_mult=1.0, pp=(3, 1234)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replay_equals_plain_launches(case):
    prob = datagen.small_problem(n=8, m=41, d_feat=17, density=0.3, seed=11)
    kw = dict(CASES[case], tracking=True)
    a = _engine(prob, False, **kw)
    b = _engine(prob, True, **kw)
    x = np.zeros(prob.features.shape[2])
    for chunk in (1, 4, 7):
        print(f"{case}: chunk of {chunk} rounds", flush=True)
        a.run(chunk)
        b.run(chunk)
        # eager work between replays
        xa, xb = a.x(), b.x()
        assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
        ha, hb = a.hessians_at(x), b.hessians_at(x)
        assert np.array_equal(ha.view(np.uint64), hb.view(np.uint64))
    ta, tb = a.trace(), b.trace()
    for key in ("grad_norm", "f", "x_norm"):
        assert np.array_equal(np.asarray(ta[key]).view(np.uint64), np.asarray(tb[key]).view(np.uint64)), key
    assert np.array_equal(ta["chol_ok"], tb["chol_ok"])
    a.close()
    b.close()


def test_round_timestamps_follow_the_run():
    prob = datagen.config1()
    from ufednl_amd.engine import Engine
    eng = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False)
    eng.run(2)
    t0 = time.perf_counter()
    eng.run(300)
    wall = time.perf_counter() - t0
    sec = eng.trace()["seconds"]
    eng.close()
    assert len(sec) == 302
    assert np.all(np.diff(sec) > 0)
    span = sec[-1] - sec[1]
    # the device clock's span of the last 300 rounds is within the host's wall
    # time of them (and not far below it: the rounds run back to back)
    assert 0.5 * wall < span <= wall * 1.01 + 1e-3, (span, wall)
