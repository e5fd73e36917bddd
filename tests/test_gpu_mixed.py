"""Reduced-precision Hessian modes (FP32 MFMA, BF16 hi/lo split) against FP64.

Only the A^T D A contraction changes precision; the gradient, Delta, Lk, the
aggregation and the Newton solve stay FP64. What is checked:
  - lane layouts of the two MFMA instructions the kernels use;
  - per-client relative Frobenius error of H_c(x) against the oracle's FP64
    Hessian (logistic_regression.h:365-547 restated in oracle/fednl_oracle.c),
    with the bound of each mode written below;
  - the FP64 gradient path is unchanged (round 0 against the oracle);
  - a reduced-precision run still reaches ||grad f||^2 <= 1e-17 and ends at the
    FP64 run's optimum within 1e-8.
"""
import numpy as np
import pytest


pytestmark = pytest.mark.gpu

# per-Hessian relative Frobenius error bounds (workload: 80 x 407, d=123,
# binary features, so A is exact in both modes; the error comes from the fp32
# products D*A, the BF16 split of D*A and FP32 accumulation over 407 samples)
TOL_FP32 = 1e-5
TOL_BF16X2 = 5e-5


def _full_fro(packed, d):
    """Frobenius norm of the symmetric matrix held as packed upper (list order)."""
    diag_pos = np.array([j * (j + 3) // 2 for j in range(d)])
    diag = packed[diag_pos]
    return np.sqrt(2.0 * np.sum(packed ** 2) - np.sum(diag ** 2))


def test_mfma_f32_bf16_layouts():
    from ufednl_amd import engine as E
    rng = np.random.default_rng(5)
    a32 = rng.integers(-8, 9, 64).astype(np.float32)
    b32 = rng.integers(-8, 9, 64).astype(np.float32)
    abf = rng.integers(-8, 9, (64, 8)).astype(np.float32)
    bbf = rng.integers(-8, 9, (64, 8)).astype(np.float32)
    o1, o2 = E.probe_mfma_mixed(a32, b32, abf, bbf)
    # f32 16x16x4: A[m][k] from lane 16k + m, B[k][n] from lane 16k + n,
    # D[m][n] in lane 16 (m // 4) + n, item m % 4
    A = np.array([[a32[16 * k + m] for k in range(4)] for m in range(16)], dtype=np.float64)
    B = np.array([[b32[16 * k + n] for n in range(16)] for k in range(4)], dtype=np.float64)
    D = A @ B
    got = np.array([[o1[16 * (m // 4) + n, m % 4] for n in range(16)] for m in range(16)])
    np.testing.assert_array_equal(got, D)
    # bf16 16x16x32: A[m][k] from lane 16 (k // 8) + m, element k % 8; B likewise
    A = np.array([[abf[16 * (k // 8) + m, k % 8] for k in range(32)] for m in range(16)], dtype=np.float64)
    B = np.array([[bbf[16 * (k // 8) + n, k % 8] for n in range(16)] for k in range(32)], dtype=np.float64)
    D = A @ B
    got = np.array([[o2[16 * (m // 4) + n, m % 4] for n in range(16)] for m in range(16)])
    np.testing.assert_array_equal(got, D)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-13), ("fp32", TOL_FP32), ("bf16x2", TOL_BF16X2)])
def test_hessian_error_per_client(oracle, mode, tol):
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    prob = datagen.config3()
    n, m, d = prob.features.shape
    eng = Engine.from_problem(prob, compressor="seqk", k_mult=8.0, hessian=mode, tracking=False)
    rng = np.random.default_rng(11)
    for x in (np.zeros(d), rng.normal(scale=0.2, size=d)):
        H = eng.hessians_at(x)
        errs = []
        for c in range(0, n, 8):
            ref = oracle.client_eval(prob.features[c], prob.labels[c], prob.lam, x)["hess_up"]
            errs.append(_full_fro(H[c] - ref, d) / _full_fro(ref, d))
        assert max(errs) < tol, (mode, max(errs))
    eng.close()


@pytest.mark.parametrize("mode", ["fp32", "bf16x2"])
@pytest.mark.parametrize("shape", ["config3", "wide"])
def test_gradient_path_stays_fp64(oracle, mode, shape):
    """The client gradients come from the margin pass's 16-row partials, added
    up by the Hessian kernels' diagonal tiles: d = 124 keeps the rows in
    registers, d = 601 (d_pad 640) reads them again; three rounds, so that the
    gradient is also checked at x != 0."""
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    if shape == "config3":
        prob = datagen.config3()
    else:
        prob = datagen.small_problem(n=3, m=70, d_feat=600, density=0.05, seed=9)
    eng = Engine.from_problem(prob, compressor="seqk", k_mult=8.0, hessian=mode, tracking=False)
    n = prob.features.shape[0]
    for r in range(3):
        # the gradient at the engine's own iterate, from the oracle's FP64
        # client evaluations (logistic_regression.h:148-161), averaged
        x = eng.x()
        g_orc = sum(oracle.client_eval(prob.features[c], prob.labels[c], prob.lam, x)["grad"] for c in range(n)) / n
        eng.run(1)
        g_gpu = eng.last_round()["grad"]
        np.testing.assert_allclose(g_gpu, g_orc, rtol=1e-12, atol=1e-16, err_msg=f"round {r}")
    eng.close()


@pytest.mark.parametrize("comp,rounds", [("seqk", 150), ("topk", 250)])
def test_reduced_precision_converges_to_fp64_optimum(comp, rounds):
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    prob = datagen.config3()
    finals = {}
    for mode in ("fp64", "fp32", "bf16x2"):
        eng = Engine.from_problem(prob, compressor=comp, k_mult=8.0, hessian=mode, tracking=False)
        eng.run(rounds, check=False)
        tr = eng.trace()
        g2 = tr["grad_norm"] ** 2
        assert np.any(g2 <= 1e-17), (mode, comp, g2.min())
        finals[mode] = eng.x()
        eng.close()
    ref = finals["fp64"]
    for mode in ("fp32", "bf16x2"):
        assert np.max(np.abs(finals[mode] - ref)) <= 1e-8 * (1.0 + np.max(np.abs(ref))), mode
