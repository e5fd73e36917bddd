"""The margin pass's tile shape (rows per wave) is picked at engine creation
by timing both candidates on the engine's data. Both shapes must compute the
same bits: margins, weights and the per-16-row -log sigma partials behind f,
with or without line search (whose trials run the same pass)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine(prob, rows, **kw):
    from ufednl_amd.engine import Engine
    os.environ["FEDNL_MARGIN_ROWS"] = str(rows)
    try:
        return Engine.from_problem(prob, **kw)
    finally:
        del os.environ["FEDNL_MARGIN_ROWS"]


@pytest.mark.parametrize("n,m,dfeat,ls", [(7, 45, 20, None), (5, 64, 40, (0.5, 0.5, 8)), (3, 96, 70, None)])
def test_margin_shapes_give_identical_rounds(n, m, dfeat, ls):
    from ufednl_amd import datagen
    prob = datagen.small_problem(n=n, m=m, d_feat=dfeat, seed=n + m)
    kw = dict(compressor="topk", k_mult=4.0, tracking=True, line_search=ls)
    e4 = _engine(prob, 4, **kw)
    e8 = _engine(prob, 8, **kw)
    assert e4.tuning()["margin_rows"] == 4 and e8.tuning()["margin_rows"] == 8
    e4.run(12)
    e8.run(12)
    t4, t8 = e4.trace(), e8.trace()
    for key in ("grad_norm", "f", "x_norm"):
        np.testing.assert_array_equal(t4[key], t8[key], err_msg=key)
    if ls:
        np.testing.assert_array_equal(t4["ls_trials"], t8["ls_trials"])
    np.testing.assert_array_equal(e4.x(), e8.x())
    e4.close()
    e8.close()


def test_tuner_picks_a_candidate():
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    os.environ.pop("FEDNL_MARGIN_ROWS", None)
    eng = Engine.from_problem(datagen.small_problem(n=9, m=64, d_feat=30, seed=4), compressor="topk", k_mult=2.0)
    assert eng.tuning()["margin_rows"] in (4, 8)
    eng.run(3)
    eng.close()


def _engine_chunk(prob, chunk, **kw):
    from ufednl_amd.engine import Engine
    os.environ["FEDNL_HESS_CHUNK"] = str(chunk)
    try:
        return Engine.from_problem(prob, **kw)
    finally:
        del os.environ["FEDNL_HESS_CHUNK"]


@pytest.mark.parametrize("n,m,dfeat", [(7, 45, 20), (5, 130, 70), (3, 96, 150)])
def test_hessian_chunks_give_identical_rounds(n, m, dfeat):
    """The FP64 A^T D A tile staged 16 or 32 samples at a time: the same MFMA
    sequence, so the same Delta, Lk, gradient and trajectory bits."""
    from ufednl_amd import datagen
    prob = datagen.small_problem(n=n, m=m, d_feat=dfeat, seed=n * m)
    kw = dict(compressor="topk", k_mult=4.0, tracking=True, hessian_form="dense")
    e16 = _engine_chunk(prob, 16, **kw)
    e32 = _engine_chunk(prob, 32, **kw)
    assert e16.tuning()["hessian_chunk"] == 16 and e32.tuning()["hessian_chunk"] == 32
    for r in range(6):
        e16.run(1)
        e32.run(1)
        d16, d32 = e16.last_delta(), e32.last_delta()
        np.testing.assert_array_equal(d16.view(np.uint64), d32.view(np.uint64), err_msg=f"round {r} Delta")
    t16, t32 = e16.trace(), e32.trace()
    for key in ("grad_norm", "f", "x_norm"):
        np.testing.assert_array_equal(t16[key], t32[key], err_msg=key)
    np.testing.assert_array_equal(e16.x(), e32.x())
    e16.close()
    e32.close()


def test_hessian_tuner_picks_a_candidate():
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    os.environ.pop("FEDNL_HESS_CHUNK", None)
    eng = Engine.from_problem(datagen.small_problem(n=9, m=64, d_feat=30, seed=4), compressor="topk", k_mult=2.0,
                              hessian_form="dense")
    assert eng.tuning()["hessian_chunk"] in (16, 32)
    eng.run(3)
    eng.close()
