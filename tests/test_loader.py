"""The CLI's LIBSVM reader (fednl_cli_load_libsvm) and CRC-32 (fednl_cli_crc32)
against dopt's own: the reference loader and dopt::crc32 compiled into
oracle/_ref (dataset_loader_fs.h:381-883 as TrainLoop.h:123-170 runs it;
Crc.cpp:74-90), the fixture tests/golden/datasets.json (the digests
bin_fednl_local prints under --debug crc32fordata, Reports.h:25-33, for every
dataset the reference ships) and zlib's CRC-32. CPU only: the library is
loaded, no GPU call is made.
"""
import ctypes
import json
import os
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = "/root/reference/dopt/datasets"
DIGESTS = os.path.join(ROOT, "tests", "golden", "datasets.json")
REALDATA = os.path.join(ROOT, "tests", "golden", "realdata.npz")
NAMES = sorted(json.load(open(DIGESTS)))


def _lib():
    from ufednl_amd import engine as E
    L = E.lib()
    L.fednl_cli_load_libsvm.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_void_p]
    L.fednl_cli_crc32.restype = ctypes.c_uint32
    L.fednl_cli_crc32.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    return L


def load_ours(path, intercept):
    L = _lib()
    N, d = ctypes.c_int(), ctypes.c_int()
    rc = L.fednl_cli_load_libsvm(str(path).encode(), int(intercept), ctypes.byref(N), ctypes.byref(d), None, None)
    if rc != 0:
        return rc, None, None
    X = np.zeros((N.value, d.value))
    y = np.zeros(N.value)
    rc = L.fednl_cli_load_libsvm(str(path).encode(), int(intercept), ctypes.byref(N), ctypes.byref(d),
                                 X.ctypes.data, y.ctypes.data)
    return rc, X, y


def crc_ours(b):
    b = bytes(b)
    buf = ctypes.create_string_buffer(b, max(len(b), 1))
    return int(_lib().fednl_cli_crc32(buf, len(b)))


def digests(X, y):
    """Reports.h:25-33: the d x N design by columns, each padded to
    LDA = roundUp(d, 4), then the labels."""
    N, d = X.shape
    lda = (d + 3) // 4 * 4
    col = np.zeros((N, lda))
    col[:, :d] = X
    return crc_ours(col.tobytes()), crc_ours(y.tobytes())


def _file(name):
    p = os.path.join(DATA, name)
    if not os.path.exists(p):
        pytest.skip("the reference datasets are not present")
    if zlib.crc32(open(p, "rb").read()) & 0xFFFFFFFF != json.load(open(DIGESTS))[name]["file_crc32"]:
        pytest.skip(f"{name} is not the file the fixture was made from")
    return p


@pytest.mark.parametrize("name", NAMES)
def test_loader_digests_match_the_reference_fixture(name):
    p = _file(name)
    fx = json.load(open(DIGESTS))[name]
    for ic in (0, 1):
        rc, X, y = load_ours(p, ic)
        assert rc == 0
        ref = fx[f"intercept{ic}"]
        assert X.shape == (ref["N"], ref["d"])
        assert digests(X, y) == (ref["crc_samples"], ref["crc_labels"])


@pytest.mark.parametrize("name", ["a1a", "a9a", "mushrooms", "phishing", "australian", "madelon.t", "my_gen_small.txt"])
def test_loader_matches_the_reference_loader(oracle, name):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built")
    p = _file(name)
    for ic in (0, 1):
        Xr, yr, cx, cy = oracle.ref_load_libsvm(p, ic)
        rc, X, y = load_ours(p, ic)
        assert rc == 0 and X.shape == Xr.shape
        assert np.array_equal(X.view(np.uint64), Xr.view(np.uint64))
        assert np.array_equal(y, yr)
        assert digests(X, y) == (cx, cy)
        assert oracle.ref().ref_lda(X.shape[1]) == (X.shape[1] + 3) // 4 * 4


def realdata(name):
    """A dataset of tests/golden/realdata.npz as (X without intercept, y)."""
    z = np.load(REALDATA)
    shape = tuple(z[f"{name}_shape"])
    vals = z[f"{name}_vals"]
    if f"{name}_bits" in z:
        codes = np.unpackbits(z[f"{name}_bits"], axis=1, count=shape[1])
    else:
        codes = z[f"{name}_codes"]
    return vals[codes].reshape(shape), np.where(z[f"{name}_pos"], 1.0, -1.0)


def write_libsvm(path, X, y):
    """Labels +1 / -1, 1-based indices of the nonzero attributes; the first
    line names the last attribute explicitly (so that both ends of the index
    range are seen, whatever dopt's first-line rule makes of them)."""
    d = X.shape[1]
    with open(path, "w") as f:
        for s in range(X.shape[0]):
            toks = ["+1" if y[s] > 0 else "-1"]
            nz = np.nonzero(X[s])[0]
            if s == 0:
                nz = np.unique(np.concatenate([[0], nz, [d - 1]]))
            toks += [f"{i + 1}:{repr(float(X[s, i]))}" for i in nz]
            f.write(" ".join(toks) + "\n")


@pytest.mark.parametrize("name", ["phishing", "a9a", "w8a"])
def test_realdata_fixture_roundtrips_through_the_loader(tmp_path, name):
    X, y = realdata(name)
    p = tmp_path / f"{name}.libsvm"
    write_libsvm(p, X, y)
    rc, X2, y2 = load_ours(p, 0)
    assert rc == 0
    assert np.array_equal(X2.view(np.uint64), X.view(np.uint64)) and np.array_equal(y2, y)
    if os.path.isdir(DATA):
        rc, X3, y3 = load_ours(_file(name), 0)
        assert np.array_equal(X3.view(np.uint64), X.view(np.uint64)) and np.array_equal(y3, y)


CASES = {
    # dataset_loader_fs.h:526-541: a label equal to neither end of the range
    # is a new class, so 5, 1, 3 are three classes
    "three_classes": ("5 1:1\n1 1:2\n3 2:1\n", None),
    "zero_one": ("1 1:1 2:0.5\n0 2:1\n1 1:2\n", ((3, 2), [1, -1, 1])),
    "signed": ("+1 1:1\n-1 2:3\n", ((2, 2), [1, -1])),
    # int32 labels only (StringUtils.h:171-268): 0.2 and 0.8 are no class
    "fractional_labels": ("0.2 1:1\n0.8 1:2\n", None),
    # the first line resets the index range per feature: it ends at 3, later
    # lines widen it to [3, 9]
    "first_line_rule": ("1 7:1 3:1\n-1 5:2 9:1\n", ((2, 7), [1, -1])),
    # dopt's decimal parser: integer + fraction * 10^-len, exponents by powers
    "number_forms": ("1 1:1e-3 2:-2.5E2 3:.5 4:1. 5:+0.3\n-1 1:0.1 2:123456.789 3:-7e+2 5:0.30000000000000004\n",
                     ((2, 5), [1, -1])),
    "blank_lines_and_tabs": ("\n1\t1:1\r\n\n-1 2:1\t3:2\n\n", ((2, 3), [1, -1])),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_loader_rules(oracle, tmp_path, case):
    text, expect = CASES[case]
    p = tmp_path / "d.libsvm"
    p.write_text(text)
    rc, X, y = load_ours(p, 0)
    if expect is None:
        assert rc == -4
    else:
        assert rc == 0 and X.shape == expect[0]
        assert list(y) == expect[1]
    if oracle.ref_available():
        if expect is None:
            with pytest.raises(RuntimeError):
                oracle.ref_load_libsvm(str(p), 0)
        else:
            Xr, yr, _, _ = oracle.ref_load_libsvm(str(p), 0)
            assert np.array_equal(X.view(np.uint64), Xr.view(np.uint64)) and np.array_equal(y, yr)


def test_crc32_is_dopts(oracle):
    rng = np.random.default_rng(5)
    for n in (0, 1, 7, 64, 1000, 4097):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        c = crc_ours(b)
        # seed 0xFFFFFFFF and no final inversion: zlib's CRC-32 inverted
        assert c == (zlib.crc32(b) ^ 0xFFFFFFFF)
        if oracle.ref_available():
            assert c == oracle.ref_crc32(b)
