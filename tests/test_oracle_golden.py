"""The oracle against the golden fixtures written from the reference's own
code (tests/golden/make_golden.py) and against the reference's known answers.
CPU only."""
import os

import numpy as np
import pytest

from tests.tolerances import assert_grad_norms_close

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _npz(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---------------------------------------------------------------- selectors

def test_topk_fixtures(oracle):
    z = _npz("selectors.npz")
    for q in range(int(z["topk_count"][0])):
        d, k = z[f"topk_{q}_d"]
        got = oracle.pos_to_flat(oracle.topk_packed(z[f"topk_{q}_in"], int(k)), int(d))
        np.testing.assert_array_equal(got, z[f"topk_{q}_out"], err_msg=f"case {q} d={d} k={k}")


def test_toplek_fixtures(oracle):
    z = _npz("selectors.npz")
    for q in range(int(z["toplek_count"][0])):
        d, k = (int(v) for v in z[f"toplek_{q}_d"])
        S = d * (d + 1) // 2
        st = 1234
        for c in range(6):
            pos, st = oracle.toplek_packed(z[f"toplek_{q}_in"][c], d, k, k * (1.0 / S), st)
            cnt = int(z[f"toplek_{q}_cnt"][c])
            assert len(pos) == cnt, (q, c)
            np.testing.assert_array_equal(oracle.pos_to_flat(pos, d), z[f"toplek_{q}_out"][c, :cnt])


def test_randk_seqk_fixtures(oracle):
    z = _npz("selectors.npz")
    for q in range(int(z["rand_count"][0])):
        d, k, seed = (int(v) for v in z[f"randk_{q}_dks"])
        S = d * (d + 1) // 2
        for name, fn in (("randk", oracle.randk_pos), ("seqk", oracle.randseqk_pos)):
            st = seed
            for r in range(4):
                pos, st = fn(st, S, k)
                np.testing.assert_array_equal(oracle.pos_to_flat(pos, d), z[f"{name}_{q}_out"][r])


@pytest.mark.parametrize("seed", [0, 1, 123, 1234, 4294967295])
def test_lcg_fixtures(oracle, seed):
    z = _npz("selectors.npz")
    L = oracle.lib()
    import ctypes
    s1 = ctypes.c_uint32(seed); s2 = ctypes.c_uint32(seed); s3 = ctypes.c_uint32(seed)
    ints = [L.orc_lcg_int_next(ctypes.byref(s1)) for _ in range(64)]
    reals = [L.orc_lcg_int_unit(ctypes.byref(s2)) for _ in range(64)]
    rr = [L.orc_lcg_real_unit(ctypes.byref(s3)) for _ in range(64)]
    np.testing.assert_array_equal(np.array(ints, dtype=np.uint32), z[f"lcg_{seed}_int"])
    np.testing.assert_array_equal(np.array(reals), z[f"lcg_{seed}_intreal"])
    np.testing.assert_array_equal(np.array(rr), z[f"lcg_{seed}_real"])


# ---------------------------------------------------------------- logistic oracle, Newton

def test_client_eval_fixture(oracle):
    z = _npz("logistic.npz")
    o = oracle.client_eval(z["feat"], z["labels"], float(z["lam"]), z["x"])
    # the reference computes sigmoid with vectorclass exp on 8-wide blocks and
    # libm on the tail; the oracle uses libm throughout: ~1 ulp
    np.testing.assert_allclose(o["margin"], z["margin"], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(o["sigmoid"], z["sigmoid"], rtol=1e-14)
    np.testing.assert_allclose(o["grad"], z["grad"], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(o["hess_up"], z["hess_up"], rtol=1e-13, atol=1e-17)
    assert abs(o["f"] - float(z["f"][0])) <= 1e-14 * abs(float(z["f"][0]))


def test_newton_fixture(oracle):
    z = _npz("logistic.npz")
    d = len(z["newton_b"])
    ok, dx, _ = oracle.cholesky_solve(z["newton_H_up"], d, float(z["newton_shift"]), z["newton_b"])
    assert ok and bool(z["newton_ok"])
    np.testing.assert_allclose(dx, z["newton_dx"], rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- FedNL trajectories

@pytest.mark.parametrize("comp", ["topk", "toplek", "randk", "seqk"])
def test_trajectory_fixtures(oracle, comp):
    z = _npz("trajectories.npz")
    from ufednl_amd.datagen import Problem
    feat = z[f"{comp}_features"].astype(np.float64)
    lab = z[f"{comp}_labels"].astype(np.float64)
    prob = Problem("fixture", feat, lab, 1e-3, None)
    R = int(z[f"{comp}_meta"][5])
    f = oracle.FedNL(prob, comp, k_mult=2.0, workers=2)
    outs = [f.round() for _ in range(R)]
    g = np.array([o["grad_norm"] for o in outs])
    assert_grad_norms_close(g, z[f"{comp}_grad_norm"])
    np.testing.assert_allclose(np.array([o["f"] for o in outs]), z[f"{comp}_f"], rtol=1e-12)
    np.testing.assert_allclose(np.array([o["x_norm"] for o in outs]), z[f"{comp}_x_norm"], rtol=1e-10,
                               atol=1e-14)
    np.testing.assert_allclose(f.get_state()["x"], z[f"{comp}_x"], rtol=1e-8, atol=1e-10)
    # lockstep dumps of the first rounds: selected indices exactly
    for r in range(3):
        np.testing.assert_allclose(outs[r]["grad"], z[f"{comp}_dump_grad"][r], rtol=1e-12, atol=1e-16)
        np.testing.assert_allclose(outs[r]["lk"], z[f"{comp}_dump_lk"][r], rtol=1e-10, atol=1e-16)
        np.testing.assert_array_equal(outs[r]["cnt"], z[f"{comp}_dump_cnt"][r])
        for c in range(prob.n):
            cnt = outs[r]["cnt"][c]
            flat = oracle.pos_to_flat(outs[r]["pos"][c, :cnt], prob.d)
            ref = z[f"{comp}_dump_idx"][r, c, :cnt]
            np.testing.assert_array_equal(flat, ref)


# ---------------------------------------------------------------- reference known answers

def test_topk_known_answers_5x5(oracle):
    """SimpleMathRoutinesGTest.cpp:16-146 (dopt/bin_tests/src/math_routines)."""
    rows = [[-5, 0, 1, -10, 2], [0, 12, 2, -14, 6], [1, 2, 1, 3, 17], [-10, -14, 3, -9, -10], [2, 6, 17, -10, 2]]
    d = 5
    lda = 8  # MatrixNMD(5,5) with 4-wide double batches pads rows to 8 (MatrixNMD.h:44-81)
    up = np.array([rows[i][j] for j in range(d) for i in range(j + 1)], dtype=np.float64)
    flat = lambda i, j: j * lda + i  # noqa: E731
    assert oracle.lda(d) == lda
    got2 = oracle.pos_to_flat(oracle.topk_packed(up, 2), d)
    # ignoreSign=true, k=2 -> {(2,4), (1,3)}
    assert got2.tolist() == sorted([flat(2, 4), flat(1, 3)])
    lst = [flat(i, j) for j in range(d) for i in range(j + 1)]
    got_k2 = sorted(lst[p] for p in oracle.topk_packed(up, 2))
    assert got_k2 == sorted([flat(2, 4), flat(1, 3)])
    mm = [[0, 5, 10, 15, 20], [1, 6, 11, 16, 21], [2, 7, 12, 17, 22], [3, 8, 13, 18, 23], [4, 9, 14, 19, 24]]
    up2 = np.array([mm[i][j] for j in range(d) for i in range(j + 1)], dtype=np.float64)
    got_k3 = sorted(lst[p] for p in oracle.topk_packed(up2, 3))
    assert got_k3 == sorted([flat(4, 4), flat(3, 4), flat(2, 4)])


def test_cholesky_known_factor(oracle):
    """CholeskyFactorizationGTest.cpp:11-35: chol of L L^T with L = [[1,0,0],[4,7,0],[1,2,5]]."""
    Lm = np.array([[1, 0, 0], [4, 7, 0], [1, 2, 5]], dtype=np.float64)
    A = Lm @ Lm.T
    np.testing.assert_array_equal(A, np.array([[1, 4, 1], [4, 65, 18], [1, 18, 30]], dtype=np.float64))
    up = np.array([A[i, j] for j in range(3) for i in range(j + 1)])
    ok, x, U = oracle.cholesky_solve(up, 3, 0.0, np.array([1.0, 2.0, 3.0]))
    assert ok
    Ufull = np.zeros((3, 3))
    q = 0
    for j in range(3):
        for i in range(j + 1):
            Ufull[i, j] = U[q]
            q += 1
    np.testing.assert_array_equal(Ufull.T, Lm)
    np.testing.assert_allclose(A @ x, [1.0, 2.0, 3.0], rtol=1e-14)
    # non positive definite forms are rejected (CholeskyFactorizationGTest.cpp:37-72)
    for bad in ([[1, 4], [4, 7]], [[1, 2], [2, 1]]):
        Bm = np.array(bad, dtype=np.float64)
        ok, _, _ = oracle.cholesky_solve(np.array([Bm[0, 0], Bm[0, 1], Bm[1, 1]]), 2, 0.0, np.ones(2))
        assert not ok


def test_reshuffle_fixture_matches_oracle_shuffle(oracle):
    """dopt's --reshuffle_train permutation (reference dopt::shuffle over all
    samples, fixture written from oracle/_ref) equals the oracle's partial
    Fisher-Yates run to k = n with the same LCG seed."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "reshuffle.npz"))
    for key in z.files:
        _, seed, n = key.split("_")
        seed, n = int(seed), int(n)
        v = np.arange(n, dtype=np.uint32)
        st = np.array([seed], dtype=np.uint32)
        oracle.lib().orc_shuffle_partial(oracle.up(v), n, n, oracle.up(st))
        np.testing.assert_array_equal(v, z[key], err_msg=key)
        assert sorted(z[key].tolist()) == list(range(n))


def test_natural_compressor_known_answers(oracle):
    """dopt's Natural compressor keeps sign and exponent and clears the
    mantissa (MatrixNMD.h:2009-2029): values round toward zero to a power of
    two; subnormals become zero of the same sign."""
    f = oracle.lib().orc_natural
    import ctypes
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    cases = [(3.0, 2.0), (-5.5, -4.0), (0.75, 0.5), (1.0, 1.0), (-1.9999, -1.0), (0.0, 0.0),
             (1e-310, 0.0), (-1e-310, -0.0), (np.inf, np.inf), (6.02e23, 2.0 ** 78)]
    for v, want in cases:
        got = f(v)
        assert got == want and np.signbit(got) == np.signbit(want), (v, got, want)
