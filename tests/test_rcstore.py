"""RC_STORE CData files (CPU only): file names by the store's rule
(system/RCStore.jl qset2fname, QSet::QStrS), the CData layout (header
struct + sparse CGT, Source/clebsch_io.cc:313-441, Source/wbsparray.hh:1144-1154)
written and read back for rank 2, rank 3 and rank-4 bases with outer
multiplicity, and the map from a stored basis to this library's.  No store
written by the reference exists here: the files are written in its layout
(parity with real stores unpinned)."""
import os

import numpy as np
import pytest

import qspace_amd as Q
from qspace_amd import rcstore as RC


def test_file_names():
    assert RC.label_string([1, 1, 2], [1, 1, -1]) == "(1,1;2)"
    assert RC.label_string([2, 1, 1], [-1, 1, 1]) == "(2;1,1)"
    assert RC.label_string([12, 1, 11], [1, 1, -1]) == "(C,1;B)"
    assert RC.label_string([40, 1, 39], [1, 1, -1]) == "(40,1;39)"
    assert RC.cdata_path("/s", "SU2", [1, 1, 2], [1, 1, -1]) == "/s/SU2/CStore/++-/(1,1;2).cgd"
    assert RC.cdata_path("/s", "SU2", [2, 2, 2, 2], [1, 1, -1, -1]) == "/s/SU2/CStore/++--/(2,2;2,2).cgd"


def _sorted_basis(tj, dirs):
    """this library's basis of a label set with the legs in the stored order"""
    o = RC.canonical_legs(tj, dirs)
    B = Q.su2_om_basis(tj, dirs)
    return [tj[k] for k in o], [dirs[k] for k in o], np.transpose(B, [0] + [k + 1 for k in o])


@pytest.mark.parametrize("tj,dirs", [([3, 3], [1, -1]), ([2, 1, 3], [1, 1, -1]), ([2, 2, 2, 2], [1, 1, -1, -1]),
                                     ([1, 1, 1, 1, 2], [1, 1, 1, -1, -1])])
def test_write_read_round_trip(tmp_path, tj, dirs):
    qs, qd, B = _sorted_basis(tj, dirs)
    cgt = np.moveaxis(B, 0, -1) if len(B) > 1 else B[0]
    cd = RC.CData("SU2", np.array(qs), "".join("+" if d > 0 else "-" for d in qd), np.arange(6.0), cgt)
    p = RC.cdata_path(str(tmp_path), "SU2", qs, qd)
    RC.write_cdata(p, cd)
    assert os.path.exists(p)
    got = RC.find_cdata(str(tmp_path), "SU2", qs, qd)
    assert got.qtype == "SU2" and list(got.qset) == qs and got.qdir == cd.qdir
    assert np.array_equal(got.cid, np.arange(6.0))
    assert got.om == len(B)
    assert np.array_equal(got.cgt, cgt)
    # the library's own basis maps to itself
    T = RC.basis_overlap(got, B)
    assert np.allclose(T, np.eye(len(B)), atol=1e-14)
    assert RC.find_cdata(str(tmp_path), "SU2", [9, 9], [1, -1]) is None


def test_foreign_basis_coefficients(tmp_path):
    """a store with another basis of the same invariant space (a mixing with
    norms other than 1): coefficients on it, mapped by basis_overlap, give the
    same tensor on this library's basis"""
    tj, dirs = [2, 2, 2, 2], [1, 1, -1, -1]
    qs, qd, B = _sorted_basis(tj, dirs)
    rng = np.random.RandomState(3)
    M = rng.standard_normal((len(B), len(B)))
    S = np.tensordot(M, B, axes=(1, 0))           # stored basis tensors
    p = RC.cdata_path(str(tmp_path), "SU2", qs, qd)
    RC.write_cdata(p, RC.CData("SU2", np.array(qs), "++--", np.zeros(6), np.moveaxis(S, 0, -1)))
    got = RC.read_cdata(p)
    x = rng.standard_normal(len(B))               # coefficients on the stored basis
    full = np.tensordot(x, got.basis(), axes=(0, 0))
    y = RC.basis_overlap(got, B).T @ x
    assert np.allclose(y, [np.sum(b * full) for b in B], atol=1e-13)
    assert np.allclose(np.tensordot(y, B, axes=(0, 0)), full, atol=1e-13)


def test_not_a_cdata_file(tmp_path):
    from scipy.io import savemat
    p = str(tmp_path / "x.cgd")
    savemat(p, {"other": np.ones(3)})
    with pytest.raises(ValueError):
        RC.read_cdata(p)
