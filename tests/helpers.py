"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_err(A, B):
    """relative Frobenius error over all blocks (same structure assumed)."""
    num = sum(float(np.sum(np.abs(a - b) ** 2)) for a, b in zip(A.data, B.data))
    den = sum(float(np.sum(np.abs(b) ** 2)) for b in B.data)
    return np.sqrt(num / max(den, 1e-300))


def assert_same_structure(C, R):
    """bit-exact Q records, block order, dims, itags, cgw of two QSpaces."""
    assert C.rank == R.rank, (C.rank, R.rank)
    assert C.nblk == R.nblk, (C.nblk, R.nblk)
    if C.rank:
        assert np.array_equal(C.QIDX, R.QIDX)
        assert np.array_equal(C.dims(), R.dims())
    assert C.itags == R.itags, (C.itags, R.itags)
    if R.info["cgw"] is not None or C.info["cgw"] is not None:
        assert np.array_equal(np.asarray(C.info["cgw"]), np.asarray(R.info["cgw"]))


def oracle_ortho(P, K, **kw):
    """the oracle's orthoQS core called with the Nkeep / stol the front end
    resolves (qspace_amd.ortho_args, orthoQS.cc:137-170), so that it sees the
    same arguments as the device call with the same keywords"""
    import oracle as O
    import qspace_amd as Q
    kw = dict(kw)
    kw["Nkeep"], kw["stol"] = Q.ortho_args(P, K, kw.get("Nkeep"), kw.get("stol"))
    return O.ortho(P, K, **kw)
