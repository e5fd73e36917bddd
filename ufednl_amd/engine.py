"""ctypes binding of libfednl_hip.so (C ABI in include/fednl_hip.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfednl_hip.so")

COMPRESSORS = {"identical": 0, "randk": 1, "topk": 2, "seqk": 3, "toplek": 4, "natural": 5}
HESSIAN_MODES = {"fp64": 0, "fp32": 1, "bf16x2": 2}
KERNEL_NAMES_MAX = 16

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_up = ctypes.POINTER(ctypes.c_uint32)
_c_ip = ctypes.POINTER(ctypes.c_int)


class FednlOpts(ctypes.Structure):
    _fields_ = [
        ("lambda_", ctypes.c_double),
        ("compressor", ctypes.c_int),
        ("k_mult", ctypes.c_double),
        ("alpha_mode", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("option_b", ctypes.c_int),
        ("workers", ctypes.c_int),
        ("tracking", ctypes.c_int),
        ("hessian_mode", ctypes.c_int),
        ("line_search", ctypes.c_int),
        ("max_line_search_steps", ctypes.c_int),
        ("c_line_search", ctypes.c_double),
        ("gamma_line_search", ctypes.c_double),
        ("clients_per_round", ctypes.c_int),
        ("client_selection_seed", ctypes.c_uint),
        ("algorithm", ctypes.c_int),
        ("hessian_form", ctypes.c_int),
        ("global_lr", ctypes.c_double),
    ]


class FednlRequest(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("n", ctypes.c_int), ("m", ctypes.c_int), ("d", ctypes.c_int),
        ("features", ctypes.POINTER(ctypes.c_double)),
        ("labels", ctypes.POINTER(ctypes.c_double)),
        ("opts", FednlOpts),
        ("max_rounds", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("x_out", ctypes.POINTER(ctypes.c_double)),
    ]


class FednlResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("rounds", ctypes.c_int), ("rounds_to_tol", ctypes.c_int),
        ("admitted_s", ctypes.c_double), ("finished_s", ctypes.c_double), ("grad_norm", ctypes.c_double),
    ]


class FednlError(RuntimeError):
    pass


_lib = None


def lib_path() -> str:
    return LIB_PATH


def lib():
    """Loads libfednl_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FednlError(f"{LIB_PATH} is missing: build it with `python -m ufednl_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.fednl_last_error.restype = ctypes.c_char_p
        L.fednl_default_opts.argtypes = [ctypes.POINTER(FednlOpts)]
        L.fednl_create.argtypes = [ctypes.POINTER(FednlOpts), ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dp,
                                   _c_dp, ctypes.POINTER(ctypes.c_void_p)]
        L.fednl_destroy.argtypes = [ctypes.c_void_p]
        L.fednl_sizes.argtypes = [ctypes.c_void_p, _c_ip, ctypes.POINTER(ctypes.c_int64), _c_ip]
        L.fednl_alpha.restype = ctypes.c_double
        L.fednl_alpha.argtypes = [ctypes.c_void_p]
        L.fednl_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.fednl_enqueue.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.fednl_sync.argtypes = [ctypes.c_void_p]
        L.fednl_trace.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_dp, _c_dp, _c_ip, _c_ip]
        L.fednl_get_x.argtypes = [ctypes.c_void_p, _c_dp]
        L.fednl_trace_line_search.argtypes = [ctypes.c_void_p, _c_ip, _c_ip]
        L.fednl_trace_clients.argtypes = [ctypes.c_void_p, _c_ip, _c_dp, _c_ip]
        L.fednl_pp_selection.argtypes = [ctypes.c_void_p, _c_ip]
        L.fednl_last_round.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_ip, _c_up, _c_dp, _c_up, _c_dp]
        L.fednl_get_state.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_dp, _c_up, _c_up]
        L.fednl_last_delta.argtypes = [ctypes.c_void_p, _c_dp]
        L.fednl_set_state.argtypes = [ctypes.c_void_p, _c_dp, _c_dp, _c_dp, _c_up, _c_up]
        L.fednl_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.fednl_kernel_timing_mask.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.fednl_tuning.argtypes = [ctypes.c_void_p, _c_ip, _c_ip]
        L.fednl_hessian_form.argtypes = [ctypes.c_void_p]
        L.fednl_kernel_times.argtypes = [ctypes.c_void_p, _c_dp, _c_ip, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_char_p)]
        L.fednl_select.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dp, _c_up, _c_up,
                                   ctypes.c_int, ctypes.c_double, _c_up, _c_ip]
        L.fednl_newton_solve.argtypes = [ctypes.c_int, _c_dp, ctypes.c_double, _c_dp, _c_dp]
        vp = ctypes.c_void_p
        L.fednl_newton_solve_noalloc.argtypes = [ctypes.c_int, vp, ctypes.c_double, vp, vp, vp, vp]
        L.fednl_hessians_at.argtypes = [ctypes.c_void_p, _c_dp, _c_dp]
        L.fednl_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.fednl_batch_destroy.argtypes = [ctypes.c_void_p]
        L.fednl_batch_enqueue.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.fednl_batch_sync.argtypes = [ctypes.c_void_p]
        L.fednl_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.fednl_footprint.restype = ctypes.c_size_t
        L.fednl_footprint.argtypes = [ctypes.POINTER(FednlOpts), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.fednl_serve.argtypes = [ctypes.POINTER(FednlRequest), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(FednlResult)]
        fp = ctypes.POINTER(ctypes.c_float)
        L.fednl_probe_mfma_mixed.argtypes = [fp, fp, fp, fp, fp, fp]
        _lib = L
    return _lib


def _err(rc, what):
    if rc < 0:
        msg = lib().fednl_last_error().decode(errors="replace")
        raise FednlError(f"{what} failed ({rc}): {msg}")
    return rc


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def dp(a):
    return _p(a, _c_dp)


def up(a):
    return _p(a, _c_up)


def ip(a):
    return _p(a, _c_ip)


class Engine:
    """One FedNL problem resident on the GPU."""

    def __init__(self, features, labels, lam=1e-3, compressor="topk", k_mult=8.0, alpha_mode=2, alpha=1.0,
                 option_b=True, workers=1, tracking=True, hessian="fp64", line_search=None, pp=None,
                 algorithm="fednl1", global_lr=0.0, hessian_form="auto"):
        L = lib()
        feats = np.ascontiguousarray(features, dtype=np.float64)
        labs = np.ascontiguousarray(labels, dtype=np.float64)
        n, m, d = feats.shape
        o = make_opts(lam, compressor, k_mult, alpha_mode, alpha, option_b, workers, tracking, hessian, line_search,
                      pp, algorithm, global_lr, hessian_form)
        h = ctypes.c_void_p()
        _err(L.fednl_create(ctypes.byref(o), n, m, d, dp(feats), dp(labs), ctypes.byref(h)), "fednl_create")
        self.h = h
        self.n, self.m, self.d = n, m, d
        k = ctypes.c_int(); S = ctypes.c_int64(); lda = ctypes.c_int()
        L.fednl_sizes(self.h, ctypes.byref(k), ctypes.byref(S), ctypes.byref(lda))
        self.k, self.S, self.lda = k.value, S.value, lda.value
        self.workers = max(1, min(workers, n))
        self.alpha = L.fednl_alpha(self.h)

    def hessians_at(self, x):
        """Client Hessians H_c(x) (n x S packed upper) as the engine's
        Hessian mode computes them."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.n * self.S)
        _err(lib().fednl_hessians_at(self.h, dp(x), dp(out)), "fednl_hessians_at")
        return out.reshape(self.n, self.S)

    @classmethod
    def from_problem(cls, prob, **kw):
        return cls(prob.features, prob.labels, lam=prob.lam, **kw)

    def close(self):
        if getattr(self, "h", None):
            lib().fednl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, rounds: int, check=True):
        rc = lib().fednl_run(self.h, rounds)
        if check:
            _err(rc, "fednl_run")
        return rc

    def enqueue(self, rounds: int):
        _err(lib().fednl_enqueue(self.h, rounds), "fednl_enqueue")

    def sync(self):
        _err(lib().fednl_sync(self.h), "fednl_sync")

    def trace(self):
        cnt = ctypes.c_int()
        _err(lib().fednl_trace(self.h, None, None, None, None, None, ctypes.byref(cnt)), "fednl_trace")
        R = cnt.value
        g = np.zeros(R); f = np.zeros(R); xn = np.zeros(R); sec = np.zeros(R); ok = np.zeros(R, dtype=np.int32)
        _err(lib().fednl_trace(self.h, dp(g), dp(f), dp(xn), dp(sec), ip(ok), ctypes.byref(cnt)), "fednl_trace")
        ls = np.zeros(R, dtype=np.int32)
        _err(lib().fednl_trace_line_search(self.h, ip(ls), ctypes.byref(cnt)), "fednl_trace_line_search")
        parts = np.zeros(R, dtype=np.int32); items = np.zeros(R)
        _err(lib().fednl_trace_clients(self.h, ip(parts), dp(items), ctypes.byref(cnt)), "fednl_trace_clients")
        return dict(grad_norm=g, f=f, x_norm=xn, seconds=sec, chol_ok=ok, ls_trials=ls, participants=parts,
                    items=items)

    def pp_selection(self):
        sel = np.zeros(self.n, dtype=np.int32)
        _err(lib().fednl_pp_selection(self.h, ip(sel)), "fednl_pp_selection")
        return sel

    def x(self):
        out = np.zeros(self.d)
        _err(lib().fednl_get_x(self.h, dp(out)), "fednl_get_x")
        return out

    def last_round(self):
        n, k, d = self.n, self.k, self.d
        g = np.zeros(d); lk = np.zeros(n); cnt = np.zeros(n, dtype=np.int32); pos = np.zeros(n * k, dtype=np.uint32)
        vals = np.zeros(n * k); idx = np.zeros(n * k, dtype=np.uint32); dx = np.zeros(d)
        _err(lib().fednl_last_round(self.h, dp(g), dp(lk), ip(cnt), up(pos), dp(vals), up(idx), dp(dx)),
             "fednl_last_round")
        return dict(grad=g, lk=lk, cnt=cnt, pos=pos.reshape(n, k), vals=vals.reshape(n, k), idx=idx.reshape(n, k),
                    dx=dx)

    def last_delta(self):
        """Delta_c = H_c(x_r) - learned_c of the last round (n x S packed upper)."""
        out = np.zeros(self.n * self.S)
        _err(lib().fednl_last_delta(self.h, dp(out)), "fednl_last_delta")
        return out.reshape(self.n, self.S)

    def get_state(self):
        x = np.zeros(self.d); L = np.zeros(self.n * self.S); H = np.zeros(self.S)
        rng = np.zeros(self.n, dtype=np.uint32); lek = np.zeros(self.workers, dtype=np.uint32)
        _err(lib().fednl_get_state(self.h, dp(x), dp(L), dp(H), up(rng), up(lek)), "fednl_get_state")
        return dict(x=x, learned=L.reshape(self.n, self.S), H=H, rng=rng, lek=lek)

    def set_state(self, st):
        def c(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt).reshape(-1)
        _err(lib().fednl_set_state(self.h, dp(c(st.get("x"), np.float64)), dp(c(st.get("learned"), np.float64)),
                                   dp(c(st.get("H"), np.float64)), up(c(st.get("rng"), np.uint32)),
                                   up(c(st.get("lek"), np.uint32))), "fednl_set_state")

    def tuning(self):
        """Tile shapes picked at creation: {"margin_rows": 4 or 8,
        "hessian_chunk": 16 or 32 (samples per LDS stage of the FP64 A^T D A)}."""
        r = np.zeros(2, dtype=np.int32)
        _err(lib().fednl_tuning(self.h, ip(r[:1]), ip(r[1:])), "fednl_tuning")
        return {"margin_rows": int(r[0]), "hessian_chunk": int(r[1])}

    @property
    def hessian_form(self):
        """The FP64 Hessian form picked at creation: "dense" (MFMA over the
        padded design) or "sparse" (binary designs: sums over the samples
        holding both features)."""
        v = lib().fednl_hessian_form(self.h)
        _err(min(v, 0), "fednl_hessian_form")
        return {v_: k_ for k_, v_ in HESSIAN_FORMS.items()}[v]

    def kernel_timing(self, enable=True, only=None):
        """Per-kernel device timing; only: a kernel name to time alone."""
        if only is None:
            _err(lib().fednl_kernel_timing(self.h, 1 if enable else 0), "fednl_kernel_timing")
            return
        names = self.kernel_names()
        mask = (1 << names.index(only)) if enable else 0
        _err(lib().fednl_kernel_timing_mask(self.h, mask), "fednl_kernel_timing_mask")

    def kernel_names(self):
        names = (ctypes.c_char_p * KERNEL_NAMES_MAX)()
        nk = _err(lib().fednl_kernel_times(self.h, None, None, KERNEL_NAMES_MAX, names), "fednl_kernel_times")
        return [names[q].decode() for q in range(nk)]

    def kernel_times(self):
        ms = np.zeros(KERNEL_NAMES_MAX); lc = np.zeros(KERNEL_NAMES_MAX, dtype=np.int32)
        names = (ctypes.c_char_p * KERNEL_NAMES_MAX)()
        nk = _err(lib().fednl_kernel_times(self.h, dp(ms), ip(lc), KERNEL_NAMES_MAX, names), "fednl_kernel_times")
        # kernels that ran: (total ms, launches)
        return {names[q].decode(): (float(ms[q]), int(lc[q])) for q in range(nk) if lc[q] > 0}


def select(compressor, delta, d, k, rng=None, toplek_rng=None, workers=1):
    """Standalone selection on identical inputs. delta: (nclients, S) packed upper."""
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    nc, S = delta.shape
    assert S == d * (d + 1) // 2
    rng_a = None if rng is None else np.ascontiguousarray(rng, dtype=np.uint32).copy()
    lek_a = None if toplek_rng is None else np.ascontiguousarray(toplek_rng, dtype=np.uint32).copy()
    out = np.zeros(nc * k, dtype=np.uint32)
    cnt = np.zeros(nc, dtype=np.int32)
    comp = COMPRESSORS[compressor]
    _err(lib().fednl_select(comp, nc, d, k, dp(delta), up(rng_a), up(lek_a), workers, k * (1.0 / S), up(out),
                            ip(cnt)), "fednl_select")
    return out.reshape(nc, k), cnt, rng_a, lek_a


def newton_solve(H_up, d, shift, b):
    H_up = np.ascontiguousarray(H_up, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    dx = np.zeros(d)
    ok = _err(lib().fednl_newton_solve(d, dp(H_up), shift, dp(b), dp(dx)), "fednl_newton_solve")
    return bool(ok), dx


def newton_solve_noalloc(d, H_up_ptr, shift, b_ptr, dx_ptr, ok_ptr, stream=None):
    """fednl_newton_solve_noalloc on device pointers (ints), enqueued on
    `stream` (a hipStream_t handle, None for the null stream)."""
    _err(lib().fednl_newton_solve_noalloc(d, H_up_ptr, shift, b_ptr, dx_ptr, ok_ptr, stream),
         "fednl_newton_solve_noalloc")


class Batch:
    """Independent engines stepped together: each round kernel is launched
    once for every member (optionally one captured HIP graph per round)."""

    def __init__(self, engines, use_graph=True):
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        h = ctypes.c_void_p()
        _err(lib().fednl_batch_create(arr, len(self.engines), 1 if use_graph else 0, ctypes.byref(h)),
             "fednl_batch_create")
        self.h = h

    def run(self, rounds: int):
        _err(lib().fednl_batch_run(self.h, rounds), "fednl_batch_run")

    def enqueue(self, rounds: int):
        _err(lib().fednl_batch_enqueue(self.h, rounds), "fednl_batch_enqueue")

    def sync(self):
        _err(lib().fednl_batch_sync(self.h), "fednl_batch_sync")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().fednl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ALGORITHMS = {"fednl1": 0, "gd": 1}
HESSIAN_FORMS = {"auto": 0, "dense": 1, "sparse": 2}


def make_opts(lam=1e-3, compressor="topk", k_mult=8.0, alpha_mode=2, alpha=1.0, option_b=True, workers=1,
              tracking=False, hessian="fp64", line_search=None, pp=None, algorithm="fednl1", global_lr=0.0,
              hessian_form="auto"):
    """fednl_opts; line_search: None (FedNL) or (c, gamma, max_steps) for
    FedNL-LS; pp: None or (clients per round, sampler seed) for FedNL-PP;
    algorithm "gd" with its step size global_lr for gradient descent;
    hessian_form "auto", "dense" or "sparse" (the FP64 A^T D A)."""
    o = FednlOpts()
    lib().fednl_default_opts(ctypes.byref(o))
    o.lambda_ = lam
    o.compressor = COMPRESSORS[compressor]
    o.k_mult = k_mult
    o.alpha_mode = alpha_mode
    o.alpha = alpha
    o.option_b = 1 if option_b else 0
    o.workers = workers
    o.tracking = 1 if tracking else 0
    o.hessian_mode = HESSIAN_MODES[hessian]
    if line_search is not None:
        c, gamma, steps = line_search
        o.line_search = 1
        o.c_line_search = c
        o.gamma_line_search = gamma
        o.max_line_search_steps = int(steps)
    if pp is not None:
        tau, seed = pp
        o.clients_per_round = int(tau)
        o.client_selection_seed = int(seed)
    o.algorithm = ALGORITHMS[algorithm]
    o.global_lr = global_lr
    o.hessian_form = HESSIAN_FORMS[hessian_form]
    return o


def serve(requests, hbm_budget_bytes, rounds_per_step=4, use_graph=True):
    """Runs fednl_serve. requests: list of dicts with arrival_s, problem
    (features n x m x d, labels n x m, lam), max_rounds, tol and the engine
    options (compressor, k_mult, hessian ...). Returns (results, final x
    per request)."""
    L = lib()
    keep = []
    arr = (FednlRequest * len(requests))()
    xs = []
    for i, rq in enumerate(requests):
        prob = rq["problem"]
        feats = np.ascontiguousarray(prob.features, dtype=np.float64)
        labs = np.ascontiguousarray(prob.labels, dtype=np.float64)
        n, m, d = feats.shape
        x = np.zeros(d)
        keep += [feats, labs]
        xs.append(x)
        r = arr[i]
        r.arrival_s = rq.get("arrival_s", 0.0)
        r.n, r.m, r.d = n, m, d
        r.features = dp(feats)
        r.labels = dp(labs)
        r.opts = make_opts(lam=prob.lam, **rq.get("opts", {}))
        r.max_rounds = rq.get("max_rounds", 500)
        r.tol = rq.get("tol", 1e-17)
        r.x_out = dp(x)
    res = (FednlResult * len(requests))()
    _err(L.fednl_serve(arr, len(requests), float(hbm_budget_bytes), rounds_per_step, 1 if use_graph else 0, res),
         "fednl_serve")
    out = [dict(status=r.status, rounds=r.rounds, rounds_to_tol=r.rounds_to_tol, admitted_s=r.admitted_s,
                finished_s=r.finished_s, grad_norm=r.grad_norm) for r in res]
    return out, xs


def footprint(n, m, d, **opts):
    o = make_opts(**opts)
    return lib().fednl_footprint(ctypes.byref(o), n, m, d)


def probe_mfma_mixed(a32, b32, abf, bbf):
    """Raw outputs of one v_mfma_f32_16x16x4f32 (a32, b32: 64 lane values) and
    one v_mfma_f32_16x16x32_bf16 (abf, bbf: 64 x 8 lane values); 64 x 4 each."""
    f = ctypes.POINTER(ctypes.c_float)
    arrs = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in (a32, b32, abf, bbf)]
    o1 = np.zeros(256, dtype=np.float32)
    o2 = np.zeros(256, dtype=np.float32)
    a32_, b32_, abf_, bbf_ = (a.ctypes.data_as(f) for a in arrs)
    _err(lib().fednl_probe_mfma_mixed(a32_, b32_, o1.ctypes.data_as(f), abf_, bbf_, o2.ctypes.data_as(f)),
         "fednl_probe_mfma_mixed")
    return o1.reshape(64, 4), o2.reshape(64, 4)
