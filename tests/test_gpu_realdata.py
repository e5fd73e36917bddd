"""Real datasets through the bin_fednl_local drop-in on the GPU: phishing,
a9a and w8a (the datasets the reference ships, dopt/datasets/), as dopt's
loader reads them (tests/golden/realdata.npz; the loader itself is pinned by
tests/test_loader.py), written back as LIBSVM text, and run with
--add_intercept --reshuffle_train (dopt::shuffle, seed 123) at the client
counts whose shares match the measured workloads (50 x 221, 80 x 407,
142 x 350).

Each run's tracking file is compared with ||grad f(x_r)|| of the reference's
own building blocks on the same split (tests/golden/make_realdata_fixtures.py):
RandK for all 20 rounds within the shared trajectory tolerance; TopK for its
first two rounds within it (later rounds can differ by a selection made on a
tie within rounding) and within 5% for all 20. The same runs check
--check_split and the L(f) --compute-L-smooth prints against the reference's
computeLSmoothness (logistic_regression.h:707-740).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.test_gpu_cli import _read_tracking
from tests.test_loader import realdata, write_libsvm
from tests.tolerances import assert_grad_norms_close

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ufednl_amd", "bin_fednl_local")
FIX = np.load(os.path.join(ROOT, "tests", "golden", "realdata.npz"))
CLIENTS = {"phishing": 50, "a9a": 80, "w8a": 142}
ROUNDS = 20


def _run(tmp_path, name, comp, extra=()):
    X, y = realdata(name)
    data = tmp_path / f"{name}.libsvm"
    if not data.exists():
        write_libsvm(data, X, y)
    out = tmp_path / f"{name}_{comp}.bin"
    args = ["--train", "--train_dataset", str(data), "--clients", str(CLIENTS[name]), "--add_intercept",
            "--reshuffle_train", "--lambda", "0.001", "--algorithm", "fednl1", "--compressor", comp,
            "--k_compressor_as_d_mult", "8", "--theoretical_alpha", "--use_theoretical_alpha_option_2",
            "--fednl-option-b", "--rounds", str(ROUNDS), "--tracking", "--out", str(out), *extra]
    print(f"{name} {comp}: {BIN} {' '.join(args)}", flush=True)
    p = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    strs, ints, dbl, rec = _read_tracking(str(out))
    n = CLIENTS[name]
    assert ints[:5] == (X.shape[0] // n, X.shape[0], ROUNDS, n, X.shape[1] + 1)
    return np.array([r[1] for r in rec]), p.stdout


@pytest.mark.parametrize("name", sorted(CLIENTS))
def test_realdata_randk_matches_reference(tmp_path, name):
    g, out = _run(tmp_path, name, "randk", extra=("--check_split",))
    assert "Checking data split [OK]" in out
    assert_grad_norms_close(g, FIX[f"{name}_randk_grad_norm"])


@pytest.mark.parametrize("name", sorted(CLIENTS))
def test_realdata_topk_matches_reference(tmp_path, name):
    g, _ = _run(tmp_path, name, "topk")
    ref = FIX[f"{name}_topk_grad_norm"]
    assert_grad_norms_close(g[:2], ref[:2])
    rel = np.abs(g - ref) / ref
    print(f"{name} topk: max relative ||grad f|| difference {rel.max():.2e} "
          f"(first above 1e-10 at round {int(np.argmax(rel > 1e-10)) if (rel > 1e-10).any() else -1})", flush=True)
    assert rel.max() < 5e-2


def test_realdata_l_smoothness_matches_reference(tmp_path, oracle):
    name = "phishing"
    _, out = _run(tmp_path, name, "randk", extra=("--compute-L-smooth",))
    m = re.search(r"power iterations: (\d+)", out)
    L = re.search(r"  L\(f\): ([0-9.eE+-]+)", out)
    assert m and L, out[-2000:]
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built")
    X, y = realdata(name)
    n = CLIENTS[name]
    mpc = X.shape[0] // n
    order = np.zeros(X.shape[0], dtype=np.uint32)
    oracle.ref().ref_shuffle(123, X.shape[0], X.shape[0], oracle.up(order))
    idx = order[: mpc * n]
    Xi = np.hstack([X, np.ones((X.shape[0], 1))])[idx]
    L_ref, it_ref = oracle.ref_l_smoothness(Xi, y[idx], 1e-3)
    assert int(m.group(1)) == it_ref
    assert abs(float(L.group(1)) - L_ref) <= 1e-5 * L_ref  # printed with 6 digits
