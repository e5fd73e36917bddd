"""Batches of independent problems: every member of a batch must follow
exactly the trajectory it follows alone (same kernels, same per-problem
arithmetic), with the round launched directly or replayed from a captured
HIP graph. Members differ in d, n, m, compressor and Hessian mode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = [
    # (n, m, d_feat, density, compressor, k_mult, hessian, seed)
    (6, 57, 20, 0.3, "topk", 8.0, "fp64", 31),
    (9, 40, 44, 0.2, "toplek", 4.0, "fp64", 32),
    (5, 63, 70, 0.15, "randk", 2.0, "fp32", 33),
    (12, 35, 33, 0.25, "seqk", 3.0, "bf16x2", 34),
    (7, 90, 130, 0.08, "topk", 8.0, "fp32", 35),
    (4, 120, 9, 0.5, "topk", 2.0, "fp64", 36),
]


def _problems():
    from ufednl_amd import datagen
    return [datagen.binary_problem(f"b{i}", n, m, df, dens, seed=seed, target_pos=0.3)
            for i, (n, m, df, dens, _, _, _, seed) in enumerate(MEMBERS)]


def _engines(probs):
    from ufednl_amd.engine import Engine
    return [Engine.from_problem(p, compressor=c, k_mult=km, hessian=h, tracking=True, workers=2)
            for p, (_, _, _, _, c, km, h, _) in zip(probs, MEMBERS)]


def _solo(probs, rounds):
    out = []
    for eng in _engines(probs):
        eng.run(rounds, check=False)
        tr = eng.trace()
        out.append((tr["grad_norm"].copy(), tr["f"].copy(), eng.x().copy()))
        eng.close()
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_solo_runs(use_graph):
    from ufednl_amd.engine import Batch
    probs = _problems()
    rounds = 25
    solo = _solo(probs, rounds)
    engs = _engines(probs)
    b = Batch(engs, use_graph=use_graph)
    # several enqueues with host synchronisation between them (graph replays
    # interleaved with eager host work)
    b.run(10)
    b.run(1)
    b.enqueue(7)
    b.sync()
    b.run(rounds - 18)
    b.close()
    for i, eng in enumerate(engs):
        tr = eng.trace()
        g, f, x = solo[i]
        np.testing.assert_array_equal(tr["grad_norm"], g, err_msg=f"member {i} grad norms")
        np.testing.assert_array_equal(tr["f"], f, err_msg=f"member {i} f")
        np.testing.assert_array_equal(eng.x(), x, err_msg=f"member {i} x")
        eng.close()


def test_member_continues_alone_after_batch():
    from ufednl_amd.engine import Batch
    probs = _problems()[:3]
    solo = []
    from ufednl_amd.engine import Engine
    for p, (_, _, _, _, c, km, h, _) in zip(probs, MEMBERS[:3]):
        e = Engine.from_problem(p, compressor=c, k_mult=km, hessian=h, tracking=False, workers=2)
        e.run(12, check=False)
        solo.append(e.x())
        e.close()
    engs = [Engine.from_problem(p, compressor=c, k_mult=km, hessian=h, tracking=False, workers=2)
            for p, (_, _, _, _, c, km, h, _) in zip(probs, MEMBERS[:3])]
    b = Batch(engs)
    b.run(5)
    b.close()
    for e in engs:
        e.run(7, check=False)
    for e, x in zip(engs, solo):
        np.testing.assert_array_equal(e.x(), x)
        e.close()
