"""Checker-only bindings.

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg import this package, and only to check or time the
restatement. ``liboracle.so`` is the C restatement of dopt's FedNL round
(oracle/fednl_oracle.c); ``_ref/libdopt_ref.so`` is the reference's own
building blocks compiled by oracle/Makefile.ref (present only where the
reference sources were available at build time).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdopt_ref.so")

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_up = ctypes.POINTER(ctypes.c_uint32)
_c_ip = ctypes.POINTER(ctypes.c_int)

COMPRESSORS = {"identical": 0, "randk": 1, "topk": 2, "seqk": 3, "toplek": 4, "natural": 5}


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def dp(a):
    return _p(a, _c_dp)


def up(a):
    return _p(a, _c_up)


def ip(a):
    return _p(a, _c_ip)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.orc_lda.restype = ctypes.c_int
        L.orc_lcg_int_next.restype = ctypes.c_uint32
        L.orc_lcg_int_unit.restype = ctypes.c_double
        L.orc_lcg_real_unit.restype = ctypes.c_double
        L.orc_topk_packed.restype = ctypes.c_int
        L.orc_topk_packed.argtypes = [_c_dp, ctypes.c_size_t, ctypes.c_size_t, _c_up]
        L.orc_toplek_packed.restype = ctypes.c_int
        L.orc_toplek_packed.argtypes = [_c_dp, ctypes.c_int, ctypes.c_size_t, ctypes.c_double, _c_up, _c_up]
        L.orc_randk_pos.argtypes = [_c_up, ctypes.c_size_t, ctypes.c_size_t, _c_up]
        L.orc_randseqk_pos.argtypes = [_c_up, ctypes.c_size_t, ctypes.c_size_t, _c_up]
        L.orc_topk.restype = ctypes.c_int
        L.orc_frob_upper_sqr_packed.restype = ctypes.c_double
        L.orc_frob_upper_packed.restype = ctypes.c_double
        L.orc_flat_of_pos.restype = ctypes.c_uint32
        L.orc_flat_of_pos.argtypes = [ctypes.c_size_t, ctypes.c_int]
        L.orc_client_eval.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp,
                                      _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp]
        L.orc_cholesky_packed.restype = ctypes.c_int
        L.orc_cholesky_packed.argtypes = [_c_dp, ctypes.c_int, ctypes.c_double, _c_dp]
        L.orc_tri_solve_packed.argtypes = [_c_dp, ctypes.c_int, _c_dp, _c_dp]
        L.orc_fednl_create.restype = ctypes.c_void_p
        L.orc_fednl_create.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_fednl_destroy.argtypes = [ctypes.c_void_p]
        L.orc_fednl_round.restype = ctypes.c_int
        L.orc_fednl_round.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_dp, _c_dp, _c_ip, _c_up, _c_dp, _c_dp, _c_dp]
        L.orc_fednl_get_state.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_dp, _c_up, _c_up]
        L.orc_fednl_set_state.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_dp, _c_up, _c_up]
        L.orc_fednl_set_line_search.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int]
        L.orc_fednl_line_search_count.restype = ctypes.c_int
        L.orc_fednl_line_search_count.argtypes = [ctypes.c_void_p]
        L.orc_gauss_solve.argtypes = [_c_dp, _c_dp, ctypes.c_int, _c_dp]
        L.orc_fednl_set_pp.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
        L.orc_fednl_pp_round.restype = ctypes.c_int
        L.orc_fednl_pp_round.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_ip, _c_ip, _c_dp, _c_dp]
        L.orc_natural.restype = ctypes.c_double
        L.orc_natural.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.ref_lda.restype = ctypes.c_int
        R.ref_topk.restype = ctypes.c_int
        R.ref_topk_signed.restype = ctypes.c_int
        R.ref_frob_upper.restype = ctypes.c_double
        R.ref_frob_upper_sqr.restype = ctypes.c_double
        R.ref_newton_solve.restype = ctypes.c_int
        R.ref_toplek_stream.argtypes = [ctypes.c_int, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_up, _c_ip]
        R.ref_client_eval.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp,
                                      _c_dp, _c_dp, _c_dp, _c_dp, _c_dp]
        R.ref_newton_solve.argtypes = [_c_dp, ctypes.c_int, ctypes.c_double, _c_dp, _c_dp, _c_dp]
        R.ref_fednl_run.restype = ctypes.c_int
        R.ref_fednl_run.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, _c_dp, _c_dp, _c_dp, _c_dp, ctypes.c_int,
                                    _c_dp, _c_dp, _c_ip, _c_up, _c_dp, _c_dp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, _c_ip, _c_ip]
        R.ref_load_libsvm.restype = ctypes.c_int
        R.ref_load_libsvm.argtypes = [ctypes.c_char_p, ctypes.c_int, _c_ip, _c_ip, _c_dp, _c_dp, _c_up, _c_up]
        R.ref_crc32.restype = ctypes.c_uint32
        R.ref_crc32.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32]
        R.ref_crc32_seed.restype = ctypes.c_uint32
        R.ref_l_smoothness.restype = ctypes.c_double
        R.ref_l_smoothness.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_ip]
        R.ref_fednl_pp_run.restype = ctypes.c_int
        R.ref_fednl_pp_run.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _c_dp, _c_dp,
                                       _c_dp, _c_dp, _c_ip, _c_ip]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers


def ref_load_libsvm(path, add_intercept):
    """dopt's own LIBSVM loading as train() runs it (TrainLoop.h:123-170),
    from oracle/_ref: returns (X (N x d), y, crc_samples, crc_labels) with the
    two digests Reports.h:25-33 prints, or raises with the loader's code."""
    R = ref()
    N, d = ctypes.c_int(), ctypes.c_int()
    cx, cy = ctypes.c_uint32(), ctypes.c_uint32()
    rc = R.ref_load_libsvm(path.encode(), int(add_intercept), ctypes.byref(N), ctypes.byref(d), None, None,
                           ctypes.byref(cx), ctypes.byref(cy))
    if rc != 0:
        raise RuntimeError(f"reference loader failed ({rc}) on {path}")
    X = np.zeros((N.value, d.value))
    y = np.zeros(N.value)
    R.ref_load_libsvm(path.encode(), int(add_intercept), ctypes.byref(N), ctypes.byref(d), X.ctypes.data_as(_c_dp),
                      y.ctypes.data_as(_c_dp), None, None)
    return X, y, int(cx.value), int(cy.value)


def ref_l_smoothness(features, labels, lam):
    """(L(f), power iterations) of the reference's computeLSmoothness on the
    samples (m x d) as train() calls it (TrainLoop.h:306-316)."""
    f = np.ascontiguousarray(features, dtype=np.float64)
    b = np.ascontiguousarray(labels, dtype=np.float64)
    it = np.zeros(1, dtype=np.int32)
    L = ref().ref_l_smoothness(dp(f), dp(b), f.shape[1], f.shape[0], lam, ip(it))
    return float(L), int(it[0])


def gd_run(prob, step, rounds, tracking=True):
    """The oracle's GD (orc_gd_run: LocalWorkerForGD.h, TrainLoop.h:993-1052)
    from x = 0: per-round ||g||, recorded f, ||x||, and the last iterate."""
    n, m, d = prob.features.shape
    f = np.ascontiguousarray(prob.features, dtype=np.float64)
    b = np.ascontiguousarray(prob.labels, dtype=np.float64)
    g = np.zeros(rounds); fv = np.zeros(rounds); xn = np.zeros(rounds); x = np.zeros(d)
    L = lib()
    L.orc_gd_run.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _c_dp, _c_dp, _c_dp, _c_dp]
    L.orc_gd_run(dp(f), dp(b), n, m, d, prob.lam, step, rounds, 1 if tracking else 0, dp(g), dp(fv), dp(xn), dp(x))
    return dict(grad_norm=g, f=fv, x_norm=xn, x=x)


def ref_gd_run(prob, step, rounds, tracking=True):
    """GD from the reference's own client oracle and vector arithmetic
    (oracle/_ref ref_gd_run); same outputs as gd_run."""
    n, m, d = prob.features.shape
    f = np.ascontiguousarray(prob.features, dtype=np.float64)
    b = np.ascontiguousarray(prob.labels, dtype=np.float64)
    g = np.zeros(rounds); fv = np.zeros(rounds); xn = np.zeros(rounds); x = np.zeros(d)
    R = ref()
    R.ref_gd_run.argtypes = [_c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _c_dp, _c_dp, _c_dp, _c_dp]
    R.ref_gd_run(dp(f), dp(b), n, m, d, prob.lam, step, rounds, 1 if tracking else 0, dp(g), dp(fv), dp(xn), dp(x))
    return dict(grad_norm=g, f=fv, x_norm=xn, x=x)


def ref_crc32(b):
    """dopt::crc32(buf, size, crc32Seed()) from oracle/_ref (Crc.cpp:74-90)."""
    R = ref()
    buf = ctypes.create_string_buffer(bytes(b), len(b)) if len(b) else ctypes.create_string_buffer(1)
    return int(R.ref_crc32(buf, len(b), R.ref_crc32_seed()))

def lda(d: int) -> int:
    return (d + 3) // 4 * 4


def upper_list(d: int, lda_: int | None = None) -> np.ndarray:
    L = lda(d) if lda_ is None else lda_
    return np.array([j * L + i for j in range(d) for i in range(j + 1)], dtype=np.uint32)


def pack_upper(full: np.ndarray, d: int) -> np.ndarray:
    """full: flat d*LDA column-major -> packed upper (list order)."""
    return np.ascontiguousarray(full[upper_list(d)])


def unpack_full(packed: np.ndarray, d: int) -> np.ndarray:
    """packed upper -> flat d*LDA symmetric column-major (padding rows 0)."""
    L = lda(d)
    out = np.zeros(d * L)
    idx = 0
    for j in range(d):
        for i in range(j + 1):
            out[j * L + i] = packed[idx]
            out[i * L + j] = packed[idx]
            idx += 1
    return out


def pos_to_flat(pos: np.ndarray, d: int) -> np.ndarray:
    return upper_list(d)[np.asarray(pos, dtype=np.int64)]


def topk_packed(vals: np.ndarray, k: int) -> np.ndarray:
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    S = vals.shape[0]
    out = np.zeros(max(S, k), dtype=np.uint32)
    c = lib().orc_topk_packed(dp(vals), S, k, up(out))
    return out[:c].copy()


def toplek_packed(vals: np.ndarray, d: int, k: int, delta: float, state: int):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    S = vals.shape[0]
    out = np.zeros(max(S, k), dtype=np.uint32)
    st = np.array([state], dtype=np.uint32)
    c = lib().orc_toplek_packed(dp(vals), d, k, delta, up(st), up(out))
    return out[:c].copy(), int(st[0])


def randk_pos(state: int, S: int, k: int):
    st = np.array([state], dtype=np.uint32)
    out = np.zeros(k, dtype=np.uint32)
    lib().orc_randk_pos(up(st), S, k, up(out))
    return out, int(st[0])


def randseqk_pos(state: int, S: int, k: int):
    st = np.array([state], dtype=np.uint32)
    out = np.zeros(k, dtype=np.uint32)
    lib().orc_randseqk_pos(up(st), S, k, up(out))
    return out, int(st[0])


def client_eval(feat, labels, lam, x):
    feat = np.ascontiguousarray(feat, dtype=np.float64)
    labels = np.ascontiguousarray(labels, dtype=np.float64)
    m, d = feat.shape
    S = d * (d + 1) // 2
    t = np.zeros(m); s = np.zeros(m); g = np.zeros(d); h = np.zeros(S); f = np.zeros(1); D = np.zeros(m)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_client_eval(dp(feat), dp(labels), d, m, lam, dp(x), dp(t), dp(s), dp(g), dp(h), dp(f), dp(D))
    return dict(margin=t, sigmoid=s, grad=g, hess_up=h, f=float(f[0]), D=D)


def cholesky_solve(H_up: np.ndarray, d: int, shift: float, b: np.ndarray):
    H_up = np.ascontiguousarray(H_up, dtype=np.float64)
    U = np.zeros_like(H_up)
    ok = lib().orc_cholesky_packed(dp(H_up), d, shift, dp(U))
    x = np.zeros(d)
    if ok:
        lib().orc_tri_solve_packed(dp(U), d, dp(np.ascontiguousarray(b, dtype=np.float64)), dp(x))
    return bool(ok), x, U


class FedNL:
    """The oracle's FedNL run (forcesequpdate semantics)."""

    def __init__(self, prob, compressor="topk", k_mult=8.0, alpha=None, option_b=True, workers=1,
                 line_search=None, pp=None):
        self.prob = prob
        n, m, d = prob.features.shape
        self.n, self.m, self.d = n, m, d
        self.S = d * (d + 1) // 2
        # Identical and Natural send the whole upper triangle (TrainLoop.h:354-377)
        self.k = self.S if compressor in ("identical", "natural") else int(k_mult * d)
        self.comp = COMPRESSORS[compressor]
        if alpha is None:
            alpha = theoretical_alpha(compressor, self.k, d, option=2)
        self.alpha = alpha
        self.workers = workers
        self._feat = np.ascontiguousarray(prob.features, dtype=np.float64)
        self._lab = np.ascontiguousarray(prob.labels, dtype=np.float64)
        self.h = lib().orc_fednl_create(dp(self._feat), dp(self._lab), n, m, d, prob.lam, self.comp,
                                        self.k, alpha, 1 if option_b else 0, workers)
        if line_search:
            c, gamma, steps = line_search
            lib().orc_fednl_set_line_search(self.h, 1, c, gamma, int(steps))
        if pp:
            # FedNL-PP: (clients per round, client sampler seed)
            tau, seed = pp
            lib().orc_fednl_set_pp(self.h, int(tau), int(seed))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_fednl_destroy(self.h)
            self.h = None

    def round(self, tracking=True):
        n, d, k = self.n, self.d, self.k
        kk = max(k, self.S) if self.comp in (0, 5) else k
        g = np.zeros(d); lk = np.zeros(n); cnt = np.zeros(n, dtype=np.int32)
        pos = np.zeros(n * kk, dtype=np.uint32); vals = np.zeros(n * kk); dx = np.zeros(d); st = np.zeros(3)
        ok = lib().orc_fednl_round(self.h, 1 if tracking else 0, dp(g), dp(lk), ip(cnt), up(pos), dp(vals),
                                   dp(dx), dp(st))
        return dict(ok=bool(ok), grad=g, lk=lk, cnt=cnt, pos=pos.reshape(n, kk), vals=vals.reshape(n, kk),
                    dx=dx, grad_norm=st[0], f=st[1], x_norm=st[2],
                    ls_count=lib().orc_fednl_line_search_count(self.h))

    def pp_round(self, tracking=True):
        """One FedNL-PP round: selected clients, full gradient at x_r, stats."""
        sel = np.zeros(self.n, dtype=np.int32); cnt = np.zeros(self.n, dtype=np.int32)
        g = np.zeros(self.d); st = np.zeros(3)
        ok = lib().orc_fednl_pp_round(self.h, 1 if tracking else 0, ip(sel), ip(cnt), dp(g), dp(st))
        return dict(ok=bool(ok), sel=sel, cnt=cnt, grad=g, grad_norm=st[0], f=st[1], x_norm=st[2])

    def get_state(self):
        x = np.zeros(self.d); L = np.zeros(self.n * self.S); H = np.zeros(self.S)
        rng = np.zeros(self.n, dtype=np.uint32); lek = np.zeros(max(1, min(self.workers, self.n)), dtype=np.uint32)
        lib().orc_fednl_get_state(self.h, dp(x), dp(L), dp(H), up(rng), up(lek))
        return dict(x=x, learned=L.reshape(self.n, self.S), H=H, rng=rng, lek=lek)

    def set_state(self, st):
        lib().orc_fednl_set_state(self.h, dp(np.ascontiguousarray(st["x"])),
                                  dp(np.ascontiguousarray(st["learned"]).reshape(-1)),
                                  dp(np.ascontiguousarray(st["H"])), up(np.ascontiguousarray(st["rng"])),
                                  up(np.ascontiguousarray(st["lek"])))


def theoretical_alpha(compressor: str, k: int, d: int, option: int = 2) -> float:
    """TrainLoop.h:409-518 with SpecialMathRoutinesForMatrix.h:195-218."""
    S = d * (d + 1) // 2
    if compressor in ("randk", "seqk"):
        w = S / k - 1.0
        return 1.0 / (w + 1.0)
    if compressor in ("topk", "toplek"):
        if option == 2:
            return 1.0
        delta = float(k) / (float(d) * float(d))
        return 1.0 - np.sqrt(1.0 - delta)
    if compressor == "natural":
        return 1.0 / (1.0 / 8.0 + 1.0)
    return 1.0


def ref_fednl_run(prob, compressor="topk", k_mult=8.0, alpha=None, option_b=True, rounds=10, workers=1,
                  tracking=True, dump_rounds=0, line_search=None):
    """Run the reference building blocks (oracle/_ref) for `rounds` rounds.
    line_search: None, or (c, gamma, max_steps) for FedNL-LS."""
    n, m, d = prob.features.shape
    k = d * (d + 1) // 2 if compressor in ("identical", "natural") else int(k_mult * d)
    if alpha is None:
        alpha = theoretical_alpha(compressor, k, d, option=2)
    feat = np.ascontiguousarray(prob.features, dtype=np.float64)
    lab = np.ascontiguousarray(prob.labels, dtype=np.float64)
    gn = np.zeros(rounds); fv = np.zeros(rounds); xn = np.zeros(rounds); x = np.zeros(d); sec = np.zeros(rounds)
    dr = max(dump_rounds, 0)
    S = d * (d + 1) // 2
    kk = max(k, S)
    dg = np.zeros(max(dr, 1) * d); dl = np.zeros(max(dr, 1) * n); dc = np.zeros(max(dr, 1) * n, dtype=np.int32)
    di = np.zeros(max(dr, 1) * n * k, dtype=np.uint32); ddx = np.zeros(max(dr, 1) * d)
    ls = line_search if line_search else (0.5, 0.5, 0)
    lsc = np.zeros(rounds, dtype=np.int32)
    cok = np.zeros(rounds, dtype=np.int32)
    st = ref().ref_fednl_run(dp(feat), dp(lab), n, m, d, prob.lam, COMPRESSORS[compressor], k, alpha,
                             1 if option_b else 0, rounds, workers, 1 if tracking else 0, dp(gn), dp(fv), dp(xn),
                             dp(x), dr, dp(dg), dp(dl), ip(dc), up(di), dp(ddx), dp(sec),
                             ctypes.c_int(ls[2]), ctypes.c_double(ls[0]), ctypes.c_double(ls[1]), ip(lsc), ip(cok))
    return dict(status=st, grad_norm=gn, f=fv, x_norm=xn, x=x, seconds=sec, ls_counts=lsc, chol_ok=cok,
                dump=dict(grad=dg.reshape(-1, d)[:dr], lk=dl.reshape(-1, n)[:dr], cnt=dc.reshape(-1, n)[:dr],
                          idx=di.reshape(-1, n, k)[:dr], dx=ddx.reshape(-1, d)[:dr]))


def ref_fednl_pp_run(prob, compressor="topk", k_mult=8.0, tau=1, seed=1234, alpha=None, option_b=True,
                     rounds=10, workers=1, tracking=True):
    """FedNL-PP with the reference's own building blocks (oracle/_ref): tau
    clients per round, client sampler seeded `seed`."""
    n, m, d = prob.features.shape
    k = d * (d + 1) // 2 if compressor in ("identical", "natural") else int(k_mult * d)
    if alpha is None:
        alpha = theoretical_alpha(compressor, k, d, option=2)
    feat = np.ascontiguousarray(prob.features, dtype=np.float64)
    lab = np.ascontiguousarray(prob.labels, dtype=np.float64)
    gn = np.zeros(rounds); fv = np.zeros(rounds); xn = np.zeros(rounds); x = np.zeros(d)
    cok = np.zeros(rounds, dtype=np.int32); sel = np.zeros(rounds * n, dtype=np.int32)
    st = ref().ref_fednl_pp_run(dp(feat), dp(lab), n, m, d, prob.lam, COMPRESSORS[compressor], k, alpha,
                                1 if option_b else 0, rounds, workers, 1 if tracking else 0, int(tau), int(seed),
                                dp(gn), dp(fv), dp(xn), dp(x), ip(cok), ip(sel))
    return dict(status=st, grad_norm=gn, f=fv, x_norm=xn, x=x, chol_ok=cok, sel=sel.reshape(rounds, n))
