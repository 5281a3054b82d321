"""Host-side checks of libqsx.so that need no GPU: the C-ABI exports every
entry point declared in include/qsx.h, ctr-idx parsing, and the exact SU(2)
coefficient engine (CG and X-symbols) against sympy fixtures and the oracle."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import _capi
from qspace_amd.qspace import parse_ctridx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "qsx.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qsx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    L = _capi.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), f"libqsx.so does not export {n}"
    assert set(names) == set(_capi.exported_symbols())
    assert L.qsx_version().decode().startswith("qsx")


@pytest.mark.parametrize("s,idx,conj", [("13", [0, 2], 0), ("1,3", [0, 2], 0), ("[1 3]*", [0, 2], 1),
                                         ("12**", [0, 1], 0), ("2", [1], 0), ("1, 2 ;3", [0, 1, 2], 0),
                                         ("10,11", [9, 10], 0), ("A", [9], 0)])
def test_parse_ctridx(s, idx, conj):
    L = _capi.lib()
    out = (C.c_int32 * 16)()
    cj = C.c_int32()
    n = L.qsx_parse_ctridx(s.encode(), out, 16, C.byref(cj))
    assert n == len(idx)
    assert list(out[:n]) == idx and cj.value == conj
    assert parse_ctridx(s) == (idx, conj)


@pytest.mark.parametrize("s", ["", "11", "0", "1,x"])
def test_parse_ctridx_errors(s):
    out = (C.c_int32 * 16)()
    assert _capi.lib().qsx_parse_ctridx(s.encode(), out, 16, None) < 0
    with pytest.raises(ValueError):
        parse_ctridx(s)


def test_cg_exact_against_sympy_fixture():
    with open(os.path.join(ROOT, "tests", "golden", "su2_cg.json")) as f:
        rows = json.load(f)["rows"]
    err = max(abs(Q.su2_cg(*r[:6]) - r[6]) for r in rows)
    assert err < 2e-16


def test_cg_exact_large_spin_identities():
    # <j m; j -m | 0 0> = (-1)^{j-m} / sqrt(2j+1) at j = 100
    tj = 200
    for tm in (200, 0, -198):
        v = Q.su2_cg(tj, tm, tj, -tm, 0, 0)
        ref = (-1) ** ((tj - tm) // 2) / np.sqrt(tj + 1)
        assert abs(v - ref) < 1e-16
    # stretched coupling <j j; 1/2 1/2 | j+1/2 j+1/2> = 1
    assert Q.su2_cg(291, 291, 1, 1, 292, 292) == 1.0


def _random_valid_x3(rng):
    """random OM=1 SU(2) contraction patterns with consistent labels."""
    kind = rng.randint(4)
    tjs = int(rng.randint(0, 9))
    if kind == 0:   # rank2 (in,out) . rank3 (in,in,out) over leg 0
        ts = int(rng.randint(1, 3))
        tc = tjs + ts - 2 * int(rng.randint(0, min(tjs, ts) + 1))
        return [tjs, tjs], [1, -1], [tjs, ts, tc], [1, 1, -1], [1], [0]
    if kind == 1:   # rank3 . conj rank3 over two legs -> rank 2
        ts = int(rng.randint(1, 3))
        tc = tjs + ts - 2 * int(rng.randint(0, min(tjs, ts) + 1))
        return [tjs, ts, tc], [1, 1, -1], [tjs, ts, tc], [-1, -1, 1], [0, 1], [0, 1]
    if kind == 2:   # H*psi stage 1: HL(a, a'*, w) . psi(a', b) over a'
        tw = 2 * int(rng.randint(0, 2))
        ta = tjs + tw - 2 * int(rng.randint(0, min(tjs, tw) + 1))
        return [ta, tjs, tw], [1, -1, 1], [tjs, tjs], [1, 1], [1], [0]
    # H*psi stage 2: T(a, w, b') . HR(b', b*, w*)... contracted over (w, b')
    tw = 2 * int(rng.randint(0, 2))
    tb = tjs + tw - 2 * int(rng.randint(0, min(tjs, tw) + 1))
    return [tjs, tw, tb], [1, 1, 1], [tjs, tb, tw], [1, -1, -1], [1, 2], [2, 1]


def test_x3_product_vs_oracle():
    rng = np.random.RandomState(11)
    n = 0
    for _ in range(300):
        case = _random_valid_x3(rng)
        try:
            ref = O.x3(*case)
        except O.OracleError:
            with pytest.raises(Q.QsxError):
                Q.su2_x3(*case)
            continue
        got = Q.su2_x3(*case)
        assert abs(got - ref) < 1e-13 * max(1.0, abs(ref)), case
        n += 1
    assert n > 150


def test_x3_large_spin_vs_oracle():
    cases = [([140, 2, 142], [1, 1, -1], [142, 140, 2], [1, -1, -1], [0, 1], [1, 2]),
             ([100, 2, 100], [1, 1, -1], [100, 100], [-1, 1], [0], [0]),
             ([290, 2, 290], [1, 1, -1], [290, 290], [1, 1], [2], [0])]
    for c in cases:
        assert abs(Q.su2_x3(*c) - O.x3(*c)) < 1e-13


def test_x3_rejects_same_direction():
    with pytest.raises(Q.QsxError) as e:
        Q.su2_x3([2, 2], [1, -1], [2, 2], [1, -1], [1], [1])  # both contracted legs outgoing
    assert e.value.code in (2, 8)


def test_no_device_is_a_clean_error():
    """on a host without a GPU the device entry points return QSX_ERR_DEVICE."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("GPU device present")
    err = C.create_string_buffer(256)
    rc = _capi.lib().qsx_device_init(0, err, 256)
    assert rc == 6 and b"device" in err.value


# ---- cell-form options and itag matching (host logic, no GPU) -------------


def _mk(tags):
    return Q.QSpace([np.array([[0]])] * len(tags), [np.ones((1,) * len(tags))], "A", tags)


# the options and the matching run in the library (qsx_itag_match, host only)
def _legs(opt, tags=("L", "s", "R*")):
    """legs of a rank-3 operand that a fully matching partner contracts
    under option opt, and the conj flag the step takes"""
    A = _mk(list(tags))
    B = _mk([t[:-1] if t.endswith("*") else t + "*" for t in tags])
    ia, ib, cA, cB = Q.itag_match(A, opt if isinstance(opt, list) and opt and isinstance(opt[0], str) else
                                  ([opt] if opt is not None else []), B, [])
    return ia, cA


@pytest.mark.parametrize("opt,legs,conj", [(None, [0, 1, 2], 0), ("23", [1, 2], 0), ("!2", [0, 2], 0),
                                           ("!13", [1], 0), ([1, 3], [0, 2], 0)])
def test_cell_options(opt, legs, conj):
    ia, cA = _legs(opt)
    assert (ia, cA) == (legs, conj)


def test_cell_option_conj():
    """'*' on A flips the matching relation: A with itself matches all legs
    with conj(A) taken (no missing-conj retry needed); '!2*' also excludes
    leg 2"""
    A = _mk(["L", "s", "R*"])
    assert Q.itag_match(A, ["*"], A, []) == ([0, 1, 2], [0, 1, 2], 1, 0)
    assert Q.itag_match(A, ["!2*"], A, []) == ([0, 2], [0, 2], 1, 0)
    assert Q.itag_match(A, [], A, ["*"]) == ([0, 1, 2], [0, 1, 2], 0, 1)


@pytest.mark.parametrize("bad,code", [("", 1), ("2x", 1), ("**", 1), ("-op:S", 5), ("--itag:s/a/b/", 5),
                                      ([0], 1), ([10], 1), ("4", 1), ("11", 1)])
def test_cell_option_errors(bad, code):
    A = _mk(["L", "s", "R*"])
    opt = ["*", bad] if bad == "*" else [bad]
    if bad == "**":
        opt = ["*", "*"]
    with pytest.raises(Q.QsxError) as e:
        Q.itag_match(A, opt, A, [])
    assert e.value.code == code


def test_itag_match_rules():
    A = _mk(["L", "s", "R*"])
    B = _mk(["L*", "s*", "R"])
    assert Q.itag_match(A, [], B, [])[:2] == ([0, 1, 2], [0, 1, 2])
    assert Q.itag_match(A, ["13"], B, [])[:2] == ([0, 2], [0, 2])
    # same directions: no match at first, retried with conj(A) (the
    # reference's "missing conj flag" rule)
    assert Q.itag_match(A, [], A, []) == ([0, 1, 2], [0, 1, 2], 1, 0)
    # a leg that matches twice is an error (iTags::getCtr)
    with pytest.raises(Q.QsxError) as e:
        Q.itag_match(_mk(["a", "a"]), [], _mk(["a*", "b"]), [])
    assert e.value.code == 2
    # two empty names count as equal
    assert Q.itag_match(_mk(["", "x"]), [], _mk(["*", "y"]), [])[:2] == ([0], [0])
    # nothing to match at all
    with pytest.raises(Q.QsxError) as e:
        Q.itag_match(_mk(["a"]), [], _mk(["b"]), [])
    assert e.value.code == 2


def test_complex_rejected_where_unsupported():
    """H*psi plans take real tensors: a complex operand is a clear error
    before any device use"""
    A = Q.QSpace([np.array([[0]]), np.array([[0]])], [np.ones((2, 2)) * (1 + 1j)], "A", ["a", "a*"])
    for f in (lambda: Q.HPsi(A, A, A),):
        with pytest.raises(Q.QsxError) as e:
            f()
        assert e.value.code == 5


def test_ortho_argument_errors():
    """orthoQS rejects what it does not restate before any device use"""
    from qspace_amd import gen
    P = gen.u1_two_site_psi(seed=1, Lq=(-2, 0, 2), Rq=(-2, 0, 2), dmin=1, dmax=2)
    A = Q.QSpace([np.array([[0], [1]], dtype=np.int32) for _ in range(4)], [np.ones((2, 1, 1, 2))] * 2, "SU2",
                 ["a", "b", "c*", "d*"])  # SU(2) of rank 4 split other than K = 2: not restated
    cases = [(lambda: Q.orthoQS(A, 1), 5), (lambda: Q.orthoQS(P, 0), 1), (lambda: Q.orthoQS(P, 4), 1),
             (lambda: Q.orthoQS(P, 3), 2),
             (lambda: Q.orthoQS(Q.QSpace(P.Q, P.data, "A", ["L", "s1", "s2*", "R"]), 2), 2),
             (lambda: Q.orthoQS(Q.QSpace(P.Q, [d * (1 + 1j) for d in P.data], "A", P.itags), 2), 5)]
    for f, code in cases:
        with pytest.raises(Q.QsxError) as e:
            f()
        assert e.value.code == code


def test_plain_c_caller(tmp_path):
    """include/qsx.h is a C99 header and a C program links against
    libqsx.so: tests/c/capi_example.c (host-only entry points: ctr-idx
    parsing, CG coefficients, outer-multiplicity count, packed offsets)
    built with -std=c99 -pedantic -Werror and run"""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "qspace_amd")
    exe = str(tmp_path / "capi_example")
    subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c", "capi_example.c"), "-L", lib, "-lqsx", f"-Wl,-rpath,{lib}", "-lm",
                    "-o", exe], check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "capi_example: ok" in out.stdout


def test_kernel_arguments_in_device_memory():
    """loading the library asks the HIP runtime for kernel arguments in
    device memory (HIP_FORCE_DEV_KERNARG=1) unless the host program set the
    variable itself"""
    _capi.lib()
    libc = C.CDLL(None)
    libc.getenv.restype = C.c_char_p
    user = os.environ.get("HIP_FORCE_DEV_KERNARG")
    assert libc.getenv(b"HIP_FORCE_DEV_KERNARG") == (user.encode() if user is not None else b"1")
