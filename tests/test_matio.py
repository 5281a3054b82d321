"""QSpace .mat struct I/O (host only): the struct layout {Q, data, info} of
the reference's MATLAB interface, round trips, MATLAB's dropped trailing
singleton dimensions, and a hand-built struct in the reference's layout."""
import os

import numpy as np
import pytest

import oracle as O
from qspace_amd import QSpace, gen, matio
from qspace_amd.matio import _norm_ext, load_qspace, qspace_from_mx, qspace_to_mx, save_qspace


def _same(A, B):
    assert A.rank == B.rank and A.nblk == B.nblk
    assert np.array_equal(A.QIDX, B.QIDX)
    assert A.itags == B.itags and A.qtype == B.qtype
    for a, b in zip(A.data, B.data):
        assert a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_round_trip(tmp_path, kind):
    if kind == "u1":
        A = gen.u1_mps(seed=3, Lq=tuple(range(-4, 5, 2)), Rq=tuple(range(-5, 6, 2)), dmin=2, dmax=5)
    else:
        A, _ = gen.su2_mps(seed=4, tS=tuple(range(5)), dmin=2, dmax=5)
    p = os.path.join(tmp_path, "t.mat")
    save_qspace(p, {"A": A})
    B = load_qspace(p, "A")
    _same(A, B)
    if kind == "su2":   # weights written as the external norms, read back
        dirs = [(-1 if t.endswith("*") else 1) for t in A.itags]
        ref = [[O.norm_ext([int(A.Q[k][i, 0]) for k in range(3)], dirs)] for i in range(A.nblk)]
        assert np.allclose(B.info["cgw"], ref, rtol=0, atol=1e-15)


def test_norm_ext_matches_oracle():
    for tj, d in [([1, 1, 0], [1, 1, -1]), ([2, 3, 5], [1, 1, -1]), ([4, 2, 4], [1, -1, 1]), ([3, 3], [1, -1]),
                  ([5, 1, 4], [-1, -1, 1]), ([2, 2, 2], [1, 1, 1]), ([0, 0], [1, -1])]:
        assert _norm_ext(tj, d) == pytest.approx(O.norm_ext(tj, d), abs=1e-15)


def test_reference_layout_struct():
    """a struct as MATLAB writes it: Q a 1 x r cell, data an n x 1 cell whose
    blocks lost their trailing singleton dims, info.itags a cell, info.cgr an
    n x 1 struct array of CRef records"""
    q0 = np.array([[0.0], [2.0]]); q1 = np.array([[1.0], [1.0]]); q2 = np.array([[1.0], [1.0]])
    Q = np.empty((1, 3), dtype=object); Q[0, 0], Q[0, 1], Q[0, 2] = q0, q1, q2
    D = np.empty((2, 1), dtype=object)
    D[0, 0] = np.arange(6.0).reshape(3, 2, order="F")      # dims (3,2,1) stored as 3 x 2
    D[1, 0] = np.arange(4.0).reshape(2, 2, order="F")
    it = np.empty((1, 3), dtype=object); it[0, 0], it[0, 1], it[0, 2] = "L", "s", "R*"
    cgr = np.zeros((2, 1), dtype=[("type", object), ("qset", object), ("qdir", object), ("cid", object),
                                  ("size", object), ("nnz", object), ("cgw", object), ("cgt", object)])
    cgr[0, 0]["cgw"] = np.array([[1.5]]); cgr[1, 0]["cgw"] = np.array([[2.0]])
    for i, qs in enumerate(([0.0, 1.0, 1.0], [2.0, 1.0, 1.0])):
        cgr[i, 0]["qset"] = np.array([qs]); cgr[i, 0]["qdir"] = np.array(["++-"])
        cgr[i, 0]["type"] = np.array(["SU2"])
    S = {"Q": Q, "data": D, "info": {"qtype": "SU2", "itags": it, "cgr": cgr}}
    T = qspace_from_mx(S)
    assert T.rank == 3 and T.nblk == 2 and T.qtype == "SU2" and T.itags == ["L", "s", "R*"]
    assert T.data[0].shape == (3, 2, 1) and T.data[0][2, 1, 0] == 5.0
    assert np.array_equal(T.info["cgw"], [[1.5], [2.0]])
    M = qspace_to_mx(T)
    _same(T, qspace_from_mx(M))


def test_cgr_record_in_another_gauge_is_refused():
    """a cgr record whose qset is not the block's labels in leg order (a CGT
    permutation the reference would have brought to the standard form before
    writing, Source/clebsch_io.cc:1748-1756), or whose qdir is not the legs'
    directions, is refused instead of losing the gauge's sign"""
    A, _ = gen.su2_mps(seed=3, tS=(0, 1, 2, 3, 4), dmin=2, dmax=3)
    M = qspace_to_mx(A)
    qspace_from_mx(M)  # the writer's own records pass the check
    i = next(i for i in range(A.nblk) if len({int(A.Q[k][i, 0]) for k in range(3)}) == 3)
    bad = M["info"]["cgr"].copy()
    bad[i, 0]["qset"] = bad[i, 0]["qset"][:, [1, 0, 2]]
    with pytest.raises(ValueError, match="non-standard gauge"):
        qspace_from_mx({"Q": M["Q"], "data": M["data"], "info": dict(M["info"], cgr=bad)})
    bad = M["info"]["cgr"].copy()
    d = bad[i, 0]["qdir"]
    bad[i, 0]["qdir"] = d[::-1] if d[::-1] != d else d.replace("+", "x")
    with pytest.raises(ValueError, match="qdir"):
        qspace_from_mx({"Q": M["Q"], "data": M["data"], "info": dict(M["info"], cgr=bad)})


def test_invalid_structs():
    Q = np.empty((1, 2), dtype=object); Q[0, 0] = np.array([[0.0], [1.0]]); Q[0, 1] = np.array([[0.0]])
    D = np.empty((2, 1), dtype=object); D[0, 0] = np.ones((1, 1)); D[1, 0] = np.ones((1, 1))
    with pytest.raises(ValueError):
        qspace_from_mx({"Q": Q, "data": D})                    # Q cells of different lengths
    Q[0, 1] = np.array([[0.5], [1.0]])
    with pytest.raises(ValueError):
        qspace_from_mx({"Q": Q, "data": D})                    # non-integer labels
    with pytest.raises(ValueError):
        qspace_from_mx({"Q": Q})                               # no data


def test_complex_round_trip(tmp_path):
    """complex blocks keep their imaginary parts through a .mat file"""
    rng = np.random.RandomState(7)
    T = QSpace([np.array([[0], [1]]), np.array([[0], [1]])],
               [rng.standard_normal((2, 3)) + 1j * rng.standard_normal((2, 3)), rng.standard_normal((1, 2)) + 0j],
               "A", ["a", "a*"])
    path = str(tmp_path / "c.mat")
    save_qspace(path, {"T": T})
    R = load_qspace(path, "T")
    assert R.is_complex()
    for a, b in zip(R.data, T.data):
        assert np.array_equal(a, b)


def test_outer_multiplicity_round_trip_and_foreign_files(tmp_path):
    """blocks with an outer-multiplicity dimension round-trip through files
    marked as this library's basis; the same data without the mark (the
    reference's RC_STORE basis) are refused"""
    from qspace_amd import gen
    T = gen.su2_om_tensor(3, [[(0, 2), (1, 1), (2, 2)]] * 4, ["a", "b", "c*", "d*"])
    assert max(T.om()) > 1
    p = str(tmp_path / "om.mat")
    matio.save_qspace(p, {"T": T})
    U = matio.load_qspace(p, "T")
    assert all(np.array_equal(a, b) for a, b in zip(T.data, U.data))
    assert np.array_equal(U.om(), T.om())
    S = matio.qspace_to_mx(T)
    S["info"]["otype"] = ""
    with pytest.raises(ValueError, match="outer-multiplicity"):
        matio.qspace_from_mx(S)
