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


def pseudo_hpsi_oracle(HKL, psi, HKR, OPL, OPR, hconj):
    """the oracle's pseudo-MPO H*psi (get_HPsi.m:62-83) as a psi-shaped
    QSpace: HK_L.psi + psi.HK_R + sum_j {OP_L(j),'*',psi}.OP_R(j)
    [+ {OP_L(j),psi}.OP_R(j)' where hconj[j]], blocks summed by label"""
    import numpy as np
    import oracle as O
    terms = []
    if HKL is not None:
        terms.append(O.contract(HKL, [1], psi, [0], skip_zero=0))
    if HKR is not None:
        terms.append(O.contract(psi, [1], HKR, [1], skip_zero=0))
    for L, R, h in zip(OPL, OPR, hconj):
        T1 = O.contract(L, [0], psi, [0], conjA=1, skip_zero=0)
        terms.append(O.contract(T1, [1, 2], R, [2, 1], skip_zero=0))
        if h:
            T2 = O.contract(L, [1], psi, [0], skip_zero=0)
            terms.append(O.contract(T2, [1, 2], R, [2, 0], conjB=1, skip_zero=0))
    key = {tuple(r): i for i, r in enumerate(psi.QIDX)}
    out = [np.zeros_like(d) for d in psi.data]
    for Ht in terms:
        assert Ht.itags == psi.itags, (Ht.itags, psi.itags)
        for r, d in zip(Ht.QIDX, Ht.data):
            out[key[tuple(r)]] += d
    return psi.with_packed(np.concatenate([o.ravel(order="F") for o in out]))
