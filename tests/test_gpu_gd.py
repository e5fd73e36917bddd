"""GD (--algorithm gd; LocalWorkerForGD.h, TrainLoop.h:993-1052) on the GPU:
the engine's rounds against the oracle's (orc_gd_run, pinned against the
reference build by tests/test_oracle_vs_ref.py::test_gd_matches_reference),
and bin_fednl_local --algorithm gd end to end: the tracking file's "GD"
header, step size and records, the theoretical step 2 / (mu + L)."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.test_gpu_cli import BIN, N_CLIENTS, N_SAMPLES, D_FEAT, _dataset, _read_tracking, _write_libsvm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tracking", [True, False])
@pytest.mark.parametrize("shape", ["small", "config2"])
def test_gd_rounds_match_oracle(oracle, tracking, shape):
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    if shape == "small":
        prob, step, rounds = datagen.small_problem(n=7, m=53, d_feat=20, density=0.3, seed=21), 2.0, 60
    else:
        prob, step, rounds = datagen.config2(), 5.0, 40
    print(f"gd {shape} tracking={tracking}: {rounds} rounds, step {step}", flush=True)
    eng = Engine.from_problem(prob, algorithm="gd", global_lr=step, tracking=tracking)
    eng.run(rounds)
    tr = eng.trace()
    x = eng.x()
    eng.close()
    ref = oracle.gd_run(prob, step, rounds, tracking)
    np.testing.assert_allclose(tr["grad_norm"], ref["grad_norm"], rtol=1e-11, atol=0.0)
    np.testing.assert_allclose(tr["x_norm"], ref["x_norm"], rtol=1e-11, atol=1e-300)
    if tracking:
        np.testing.assert_allclose(tr["f"], ref["f"], rtol=1e-13, atol=0.0)
    np.testing.assert_allclose(x, ref["x"], rtol=1e-10, atol=1e-14)
    assert (tr["participants"] == prob.features.shape[0]).all() and not tr["items"].any()


def test_gd_rejects_line_search_and_batches():
    from ufednl_amd import datagen
    from ufednl_amd.engine import Batch, Engine, FednlError
    prob = datagen.small_problem(n=3, m=20, d_feat=5, seed=2)
    with pytest.raises(FednlError):
        Engine.from_problem(prob, algorithm="gd", global_lr=1.0, line_search=(0.5, 0.5, 3))
    eng = Engine.from_problem(prob, algorithm="gd", global_lr=1.0)
    with pytest.raises(FednlError):
        Batch([eng])
    eng.close()


def _cli(tmp_path, extra):
    X, y = _dataset(seed=33)
    data = tmp_path / "gd.libsvm"
    _write_libsvm(data, X, y)
    out = tmp_path / "gd.bin"
    args = ["--train", "--train_dataset", str(data), "--clients", str(N_CLIENTS), "--add_intercept", "--lambda",
            "0.001", "--algorithm", "gd", "--rounds", "25", "--tracking", "--out", str(out), *extra]
    print(" ".join([BIN, *args]), flush=True)
    p = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return X, y, _read_tracking(str(out)), p.stdout


def _split(X, y):
    from ufednl_amd.datagen import Problem
    m = N_SAMPLES // N_CLIENTS
    Xi = np.hstack([X, np.ones((N_SAMPLES, 1))])[: m * N_CLIENTS]
    return Problem("cli", np.ascontiguousarray(Xi.reshape(N_CLIENTS, m, D_FEAT + 1)),
                   np.ascontiguousarray(y[: m * N_CLIENTS].reshape(N_CLIENTS, m)), 1e-3, np.zeros(D_FEAT + 1))


def test_cli_gd_tracking_file(oracle, tmp_path):
    X, y, (strs, ints, dbl, rec), out = _cli(tmp_path, ["--global_lr", "0.5", "--global_step_size_multiplier", "3"])
    d = D_FEAT + 1
    m = N_SAMPLES // N_CLIENTS
    # TrainLoop.h:753-792: the dataset, "GD", no compressor; k = 1 * d (the
    # multiplier's default); the global step with its multiplier; alpha 0
    assert strs == ["gd.libsvm", "GD", ""]
    assert ints == (m, N_SAMPLES, 25, N_CLIENTS, d, d)
    assert dbl == (1.0, 1.5, 0.0, 1e-3)
    ref = oracle.gd_run(_split(X, y), 1.5, 25, True)
    g = np.array([r[1] for r in rec])
    f = np.array([r[2] for r in rec])
    np.testing.assert_allclose(g, ref["grad_norm"], rtol=1e-11)
    np.testing.assert_allclose(f, ref["f"], rtol=1e-13)
    # bytes: every client's gradient, 8 d each, accumulated before the record
    np.testing.assert_array_equal([r[4] for r in rec], [8.0 * d * N_CLIENTS * r for r in range(25)])
    assert all(r[5] == 0.0 for r in rec)
    assert "Global step size (with multiplier): 1.5" in out
    assert "Main Loop preparation for: gd" in out


def test_cli_gd_theoretical_step(oracle, tmp_path):
    X, y, (strs, ints, dbl, rec), out = _cli(tmp_path, ["--theoretical_global_lr"])
    L = float(re.search(r"  L\(f\): ([0-9.eE+-]+)", out).group(1))
    step = dbl[1]
    assert abs(step - 2.0 / (1e-3 + L)) <= 1e-5 * step  # L printed with 6 digits
    assert "Is global step size theoretical: [YES]" in out
    if oracle.ref_available():
        prob = _split(X, y)
        L_ref, _ = oracle.ref_l_smoothness(prob.features.reshape(-1, D_FEAT + 1), prob.labels.reshape(-1), 1e-3)
        assert abs(step - 2.0 / (1e-3 + L_ref)) <= 1e-9 * step
    ref = oracle.gd_run(_split(X, y), step, 25, True)
    np.testing.assert_allclose(np.array([r[1] for r in rec]), ref["grad_norm"], rtol=1e-11)
