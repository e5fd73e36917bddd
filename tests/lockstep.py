"""One FedNL round of the engine in lockstep with the oracle.

The oracle's state (x, learned Hessians, H, generator states) is loaded into
the engine, both run one round, and the engine's round is checked in two ways:

1. Exactly, against dopt's client and master arithmetic applied to the
   engine's own Hessian differences Delta_c (fednl_last_delta): the selected
   positions (the oracle's selectors run on the engine's Delta), the message
   values fl(item * fl(fl(1/n) * alpha)), the learned update
   learned + fl(alpha * item), the master's H += values in client order and
   the generator states (LocalWorkerForFedNL_TopK.h:70-129,
   LocalWorkerForFedNL_RandK.h:80-162, TrainLoop.h:562-564,1197-1274).
2. Within FP64 tolerance, against the oracle's own round: Delta itself (the
   Hessian is summed in another order), gradient, Lk and the values at common
   positions; the Newton step against a Cholesky solve of the engine's own
   post-round system; and, when both rounds selected the same positions, the
   Newton step and next iterate against the oracle's.
"""
import numpy as np

import oracle as O


def _items(comp, dv, mult):
    if comp in ("randk", "seqk"):
        return dv * mult
    if comp == "natural":
        nat = O.lib().orc_natural
        return np.array([nat(float(v)) for v in dv])
    return dv


def expected_selection(comp, D, pre, n, d, k, workers):
    """Positions (ascending) per client and the generator states after the
    round, from the oracle's selectors on the engine's Delta."""
    S = d * (d + 1) // 2
    pos = []
    rng = np.array(pre["rng"], dtype=np.uint32).copy()
    lek = np.array(pre["lek"], dtype=np.uint32).copy()
    for c in range(n):
        if comp == "topk":
            p = O.topk_packed(D[c], k)
        elif comp == "toplek":
            w = c % workers
            p, lek[w] = O.toplek_packed(D[c], d, k, k * (1.0 / S), int(lek[w]))
        elif comp == "randk":
            p, rng[c] = O.randk_pos(int(rng[c]), S, k)
        elif comp == "seqk":
            p, rng[c] = O.randseqk_pos(int(rng[c]), S, k)
        else:  # identical, natural: the whole upper triangle
            p = np.arange(S, dtype=np.uint32)
        pos.append(np.sort(np.asarray(p, dtype=np.uint32)))
    return pos, rng, lek


def lockstep_round(eng, orc, comp, workers=1, check_oracle=True, max_flip_frac=0.0):
    """Loads the oracle's state, runs one engine round and checks it
    (lockstep_check)."""
    pre = lockstep_load(eng, orc)
    eng.run(1)
    return lockstep_check(eng, orc, pre, comp, workers, check_oracle, max_flip_frac)


def lockstep_load(eng, orc):
    """Copies the oracle's state into the engine; returns that state."""
    pre = orc.get_state()
    eng.set_state(pre)
    return pre


def lockstep_check(eng, orc, pre, comp, workers=1, check_oracle=True, max_flip_frac=0.0, option_b=True):
    """After the engine ran one round from `pre`: runs the oracle's round and
    asserts both kinds of agreement.
    max_flip_frac: the share of selected positions allowed to differ from the
    oracle's own selection (its Delta differs from the engine's in the last
    bits, so entries tied to rounding at the selection boundary can swap).
    Returns a dict of diagnostics."""
    n, d, k, S = eng.n, eng.d, eng.k, eng.S
    out_g = eng.last_round()
    post_g = eng.get_state()
    D = eng.last_delta()
    out_o = orc.round()
    post_o = orc.get_state()

    alpha = eng.alpha
    prescale = (1.0 / float(n)) * alpha
    mult = float(S) / float(k)

    # ---- exact: dopt's arithmetic on the engine's Delta
    pos_e, rng_e, lek_e = expected_selection(comp, D, pre, n, d, k, workers)
    learned = np.array(pre["learned"], dtype=np.float64).reshape(n, S).copy()
    H = np.array(pre["H"], dtype=np.float64).copy()
    for c in range(n):
        cnt = len(pos_e[c])
        assert out_g["cnt"][c] == cnt, (c, out_g["cnt"][c], cnt)
        np.testing.assert_array_equal(out_g["pos"][c, :cnt], pos_e[c], err_msg=f"positions of client {c}")
        np.testing.assert_array_equal(out_g["idx"][c, :cnt], O.pos_to_flat(pos_e[c], d),
                                      err_msg=f"flat indices of client {c}")
        item = _items(comp, D[c][pos_e[c]], mult)
        vals = item * prescale
        np.testing.assert_array_equal(out_g["vals"][c, :cnt], vals, err_msg=f"message values of client {c}")
        learned[c, pos_e[c]] = learned[c, pos_e[c]] + alpha * item
        H[pos_e[c]] = H[pos_e[c]] + vals
    np.testing.assert_array_equal(post_g["learned"], learned, err_msg="learned Hessians after the round")
    np.testing.assert_array_equal(post_g["H"], H, err_msg="master H after the round")
    if comp in ("randk", "seqk"):
        np.testing.assert_array_equal(post_g["rng"], rng_e)
    if comp == "toplek":
        np.testing.assert_array_equal(post_g["lek"], lek_e)

    diag = dict(flips=0, selected=int(sum(len(p) for p in pos_e)))
    if not check_oracle:
        return diag

    # ---- tolerance: the oracle's own round
    scale = max(np.abs(out_o["grad"]).max(), 1e-300)
    np.testing.assert_allclose(out_g["grad"], out_o["grad"], rtol=1e-12, atol=1e-14 * scale)
    np.testing.assert_allclose(out_g["lk"], out_o["lk"], rtol=1e-10, atol=1e-15)
    # Delta against the oracle's client Hessians at the same x
    H_o = np.array([O.client_eval(f_c, y_c, orc.prob.lam, pre["x"])["hess_up"]
                    for f_c, y_c in zip(orc.prob.features, orc.prob.labels)])
    D_o = H_o - np.asarray(pre["learned"]).reshape(n, S)
    hs = np.abs(H_o).max()
    diag["delta_err"] = float(np.abs(D - D_o).max() / hs)
    assert diag["delta_err"] <= 1e-13, diag["delta_err"]
    flips = 0
    for c in range(n):
        cg = out_g["cnt"][c]
        co = out_o["cnt"][c]
        a = set(out_g["pos"][c, :cg].tolist())
        b = set(out_o["pos"][c, :co].tolist())
        flips += len(a ^ b)
        common = np.array(sorted(a & b), dtype=np.int64)
        if len(common):
            vg = dict(zip(out_g["pos"][c, :cg].tolist(), out_g["vals"][c, :cg].tolist()))
            vo = dict(zip(out_o["pos"][c, :co].tolist(), out_o["vals"][c, :co].tolist()))
            g_ = np.array([vg[p] for p in common.tolist()])
            o_ = np.array([vo[p] for p in common.tolist()])
            # the values carry Delta's error, bounded above by 1e-13 of |H|
            vs = 1e-13 * hs * prescale * (mult if comp in ("randk", "seqk") else 1.0)
            assert np.abs(g_ - o_).max() <= vs, (c, np.abs(g_ - o_).max(), vs)
    diag["flips"] = flips
    assert flips <= max_flip_frac * diag["selected"], (flips, diag["selected"])
    dxs = max(np.abs(out_o["dx"]).max(), 1e-300)
    # the Newton step solves the engine's own post-round system
    # (H + shift I) dx = g, shift = mean Lk (option b)
    lk_sum = 0.0
    for v in out_g["lk"]:
        lk_sum = lk_sum + v
    shift = lk_sum * (1.0 / float(n)) if option_b else orc.prob.lam
    ok, dx_sys, _ = O.cholesky_solve(post_g["H"], d, shift, out_g["grad"])
    assert ok
    np.testing.assert_allclose(out_g["dx"], dx_sys, rtol=1e-9, atol=1e-11 * np.abs(dx_sys).max())
    if flips == 0:
        np.testing.assert_allclose(out_g["dx"], out_o["dx"], rtol=1e-9, atol=1e-12 * dxs)
        np.testing.assert_allclose(post_g["x"], post_o["x"], rtol=1e-12, atol=1e-12 * dxs)
    # with flips the oracle's H differs by whole entries at the swapped
    # positions (each system is checked on its own above and in the oracle)
    if comp in ("randk", "seqk"):
        np.testing.assert_array_equal(post_g["rng"], post_o["rng"])
    return diag
