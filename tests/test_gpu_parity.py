"""HIP path vs the CPU oracle on the same seeded inputs (GPU tests).

Bar: Q records, block order, dims, itags and cgw bit-exact; data within
1e-12 relative Frobenius error (FP64 paths).
"""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from helpers import assert_same_structure, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _both(A, ia, B, ib, perm=None, conjA=0, conjB=0, mode="batched"):
    C = Q.api._contract_pair(A, ia, conjA, B, ib, conjB, perm, mode=mode)
    R = O.contract(A, ia, B, ib, perm=perm, conjA=conjA, conjB=conjB)
    return C, R


def test_u1_mps_overlap(gpu):
    """workload u1_mps at full size: C = contractQS(A,'12*',A,'12')."""
    A = gen.u1_mps()
    C = Q.contractQS(A, "12*", A, "12")
    R = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL
    assert C.itags == ["R", "R*"]


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_u1_modes_and_perm(gpu, mode):
    A = gen.u1_mps(seed=7)
    C, R = _both(A, [2], A, [2], perm=[2, 0, 1, 3], conjB=1, mode=mode)
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL


def test_su2_env_small(gpu):
    """SU(2) L.A then X.conj(A) at reduced size (2S in 0..6, dims 4..12)."""
    A, L = gen.su2_mps(seed=11, tS=tuple(range(7)), dmin=4, dmax=12)
    X, XR = _both(L, [1], A, [0])
    assert_same_structure(X, XR)
    assert rel_err(X, XR) < TOL
    C, CR = _both(XR, [0, 1], A, [0, 1], conjB=1)
    assert_same_structure(C, CR)
    assert rel_err(C, CR) < TOL


def test_su2_noncontiguous_legs(gpu):
    """an operator on the middle (local) leg forces an operand permute."""
    A, L = gen.su2_mps(seed=12, tS=tuple(range(6)), dmin=3, dmax=9)
    rng = np.random.RandomState(4)
    W = Q.QSpace([np.array([[1]]), np.array([[1]])], [rng.standard_normal((1, 1))], "SU2", ["s", "s*"])
    C, R = _both(A, [1], W, [1])
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL
    C2, R2 = _both(W, [1], A, [1])
    assert_same_structure(C2, R2)
    assert rel_err(C2, R2) < TOL


def test_single_block_gemm_layouts(gpu):
    """one block per operand, all four matricisations vs numpy (FP64)."""
    rng = np.random.RandomState(9)
    M, K, N = 37, 29, 41
    for sa in ("mk", "km"):
        for sb in ("kn", "nk"):
            a = rng.standard_normal((M, K) if sa == "mk" else (K, M))
            b = rng.standard_normal((K, N) if sb == "kn" else (N, K))
            A = Q.QSpace([[[0]], [[0]]], [a], "A")
            B = Q.QSpace([[[0]], [[0]]], [b], "A")
            ia = [1] if sa == "mk" else [0]
            ib = [0] if sb == "kn" else [1]
            C = Q.api._contract_pair(A, ia, 0, B, ib, 0)
            ref = (a if sa == "mk" else a.T) @ (b if sb == "kn" else b.T)
            assert np.abs(C.data[0] - ref).max() / np.abs(ref).max() < 1e-14


def _hpsi_oracle(HL, psi, HR, hconj=1, fermA=None):
    T = O.contract(HL, [1], psi, [0], fermA=fermA, skip_zero=0)
    H = O.contract(T, [1, 2], HR, [2, 1], skip_zero=0)
    terms = [H]
    if hconj:
        T2 = O.contract(HL, [0], psi, [0], conjA=1, fermA=fermA, skip_zero=0)
        H2 = O.contract(T2, [1, 2], HR, [2, 0], conjB=1, skip_zero=0)
        terms.append(H2)
    # map onto psi's block order
    key = {tuple(r): i for i, r in enumerate(psi.QIDX)}
    out = [np.zeros_like(d) for d in psi.data]
    for Ht in terms:
        for r, d in zip(Ht.QIDX, Ht.data):
            out[key[tuple(r)]] += d
    return psi.with_packed(np.concatenate([o.ravel(order="F") for o in out]))


@pytest.mark.parametrize("hconj", [0, 1])
def test_hpsi_heisenberg_small(gpu, hconj):
    HL, psi, HR = gen.heis_hpsi(seed=21, Dstar=40, mu=1.5, sigma=0.5, dmax=12)
    H = Q.HPsi(HL, psi, HR, hconj=hconj)
    got = H.apply(psi.packed())
    ref = _hpsi_oracle(HL, psi, HR, hconj).packed()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < TOL


def test_hpsi_plain_equals_batched(gpu):
    HL, psi, HR = gen.heis_hpsi(seed=22, Dstar=60, mu=2.0, sigma=0.6, dmax=40)
    v = psi.packed()
    a = Q.HPsi(HL, psi, HR, hconj=1, mode="batched").apply(v)
    b = Q.HPsi(HL, psi, HR, hconj=1, mode="plain").apply(v)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < TOL


def test_hpsi_hermitian_full_size(gpu):
    """size-independent property at the D*=2000 workload: <x,Hy>_w = <Hx,y>_w
    in the full-space (CGT-weighted) inner product."""
    HL, psi, HR = gen.heis_hpsi(seed=1003, Dstar=2000)
    H = Q.HPsi(HL, psi, HR, hconj=1)
    w = H.weights()
    rng = np.random.RandomState(5)
    x = rng.standard_normal(H.N)
    y = rng.standard_normal(H.N)
    hx = H.apply(x)
    hy = H.apply(y)
    nrm = np.sqrt((w * hx) @ hx) * np.sqrt((w * y) @ y)
    assert abs((w * x) @ hy - (w * hx) @ y) / nrm < 1e-12
    # linearity
    z = H.apply(2.0 * x - 3.0 * y)
    assert np.linalg.norm(z - (2.0 * hx - 3.0 * hy)) / np.linalg.norm(z) < 1e-12


def test_hpsi_deterministic_graph_replays(gpu):
    """split-k tiles reduce in a fixed order: graph replays and direct launches
    give bitwise identical H*psi."""
    HL, psi, HR = gen.heis_hpsi(seed=1003, Dstar=2000)
    H = Q.HPsi(HL, psi, HR, hconj=1)
    x = np.random.RandomState(3).standard_normal(H.N)
    dx = Q.DeviceBuffer.from_array(x)
    dy = Q.DeviceBuffer(H.N)
    outs = []
    for graph in (True, True, True, False, False):
        H.use_graph(graph)
        for _ in range(3):
            H.apply_dev(dx, dy)
        outs.append(dy.download())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


# every work decomposition (engine, tile size, workgroup sharing, k-splits)
# is the same contraction: each must meet the oracle bar
POLICIES = [
    {"engine": "grouped", "split_steps": 1 << 28},
    {"engine": "grouped", "split_steps": 2, "tile_overhead": 0},
    {"engine": "wave", "wave_max": 4},
    {"engine": "wave", "wave_max": 1, "group_mfma": 1},
    {"engine": "wave", "wave_max": 2, "group_mfma": 8, "split_mfma": 16},
    {"engine": "wave", "wave_max": 2, "split_mfma": 8},
    {"engine": "wave", "wave_max": 3, "group_mfma": 4, "split_mfma": 64},
    {"engine": "wave", "wave_max": 1, "group_mfma": 16, "xcd_bins": 0},
    {"engine": "wave", "wave_max": 2, "group_mfma": 16, "split_mfma": 32, "xcd_bins": 3},
    {"engine": "wave", "wave_max": 1, "group_mfma": 8, "group_waves": 16},
    {"engine": "wave", "wave_max": 1, "group_mfma": 8, "group_waves": 8, "split_mfma": 16, "xcd_bins": 5},
]


@pytest.mark.parametrize("pi", range(len(POLICIES)))
def test_hpsi_policies_match_oracle(gpu, pi):
    """blocks up to 90 wide (tiles of several fragments, long k-ranges) under
    each policy vs the oracle; two applies bitwise equal (fixed sum order)."""
    HL, psi, HR = gen.heis_hpsi(seed=23, Dstar=400, mu=3.0, sigma=0.7, dmax=90)
    H = Q.HPsi(HL, psi, HR, hconj=1)
    H.set_policy(POLICIES[pi], POLICIES[pi])
    v = psi.packed()
    got = H.apply(v)
    ref = _hpsi_oracle(HL, psi, HR, 1).packed()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < TOL
    assert np.array_equal(H.apply(v), got)


@pytest.mark.parametrize("pi", range(len(POLICIES)))
def test_contract_policies_match_oracle(gpu, pi):
    """SU(2) environment step L.A (long k: the L bond) under each policy."""
    A, L = gen.su2_mps(seed=13, tS=tuple(range(9)), dmin=20, dmax=70)
    P = Q.CtrPlan(L, "2", A, "1")
    P.set_policy(**POLICIES[pi])
    dA = Q.DeviceBuffer.from_array(L.packed())
    dB = Q.DeviceBuffer.from_array(A.packed())
    dC = Q.DeviceBuffer(P.nout)
    P.execute(dA, dB, dC)
    got = dC.download()
    R = O.contract(L, [1], A, [0], skip_zero=0)
    ref = R.packed()
    assert got.shape == ref.shape
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < TOL


def test_hpsi_hubbard_small(gpu):
    """SU(2)xU(1) Hubbard H*psi with exact +-1 parity signs on HL vs the oracle."""
    HL, psi, HR = gen.hubbard_hpsi(seed=24, Dstar=120, mu=1.5, sigma=0.6, dmax=12, qmax=4)
    ferm = gen.fermion_parity_signs(HL)
    assert (ferm < 0).any() and (ferm > 0).any()
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    got = H.apply(psi.packed())
    ref = _hpsi_oracle(HL, psi, HR, 1, fermA=ferm).packed()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < TOL


def test_split_precision_hubbard_signs(gpu):
    """config 5 in split precision: the SU(2)xU(1) Hubbard operator with
    exact +-1 parity signs on HL (D* = 1500) against the oracle, BF16 with
    the default 7 slices to 1e-12 and FP32 x 3 to its stated 5e-10; the
    signs enter through the pairs' afac like the X-symbols"""
    HL, psi, HR = gen.hubbard_hpsi(seed=28, Dstar=1500, mu=3.0, sigma=0.8, dmax=200)
    ferm = gen.fermion_parity_signs(HL)
    assert (ferm < 0).any() and (ferm > 0).any()
    x = np.random.RandomState(12).standard_normal(psi.numel())
    ref = _hpsi_oracle(HL, psi.with_packed(x), HR, 1, fermA=ferm).packed()
    errs = {}
    for mode, bound in (("split_bf16", TOL), ("split_f32", 5e-10)):
        got = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm, mode=mode).apply(x)
        errs[mode] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        assert errs[mode] < bound, (mode, errs[mode])
    print("split precision, Hubbard D*=1500: rel err vs oracle", errs, flush=True)


def _dense_weighted(H):
    """H as a dense matrix in the orthonormal basis sqrt(w) x (symmetric)."""
    N = H.N
    cols = []
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        cols.append(H.apply(e))
    M = np.column_stack(cols)
    sw = np.sqrt(H.weights())
    S = (sw[:, None] * M) / sw[None, :]
    return S


@pytest.mark.parametrize("kind", ["heis", "hubbard"])
def test_eigensolvers_vs_dense(gpu, kind):
    """Lanczos and Davidson with a Krylov space of the full dimension give the
    lowest eigenvalue of the dense operator; their ground states are
    eigenvectors (weighted residual)."""
    if kind == "heis":
        HL, psi, HR = gen.heis_hpsi(seed=25, Dstar=40, mu=1.2, sigma=0.4, dmax=6)
        ferm = None
    else:
        HL, psi, HR = gen.hubbard_hpsi(seed=26, Dstar=30, mu=1.0, sigma=0.4, dmax=4, qmax=3)
        ferm = gen.fermion_parity_signs(HL)
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    S = _dense_weighted(H)
    assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
    ev = np.linalg.eigvalsh(0.5 * (S + S.T))
    e_ref = ev[0]
    w = H.weights()
    x0 = np.random.RandomState(1).standard_normal(H.N)
    for solver in (H.lanczos, H.davidson):
        e0, ritz, gs = solver(x0, H.N)
        assert abs(e0 - e_ref) < 1e-10 * max(1.0, abs(e_ref)), (solver.__name__, e0, e_ref)
        r = H.apply(gs) - e0 * gs
        assert np.sqrt((w * r) @ r) / np.sqrt((w * gs) @ gs) < 1e-8 * max(1.0, abs(e_ref))


def test_lanczos_invariant_start(gpu):
    """Lanczos started on an exact eigenvector (the highest, positive
    eigenvalue of a small Heisenberg operator): the Krylov space is
    invariant after one vector, beta_0 is at rounding level, and the
    tridiagonal matrix is cut there -- the result is that eigenvalue and
    that vector, not a value of the rounding-noise directions or 0."""
    HL, psi, HR = gen.heis_hpsi(seed=25, Dstar=40, mu=1.2, sigma=0.4, dmax=6)
    H = Q.HPsi(HL, psi, HR, hconj=1)
    S = _dense_weighted(H)
    ev, U = np.linalg.eigh(0.5 * (S + S.T))
    assert ev[-1] > 0 and ev[-1] - ev[-2] > 1e-6
    sw = np.sqrt(H.weights())
    x0 = U[:, -1] / sw  # back from the orthonormal basis sqrt(w) x
    e0, ritz, gs = H.lanczos(x0, 12)
    assert abs(e0 - ev[-1]) < 1e-10 * abs(ev[-1]), (e0, ev[-1])
    assert np.allclose(ritz, e0, rtol=1e-10, atol=0)
    g = gs * sw
    assert abs(abs(g @ U[:, -1]) / np.linalg.norm(g) - 1.0) < 1e-10


@pytest.mark.parametrize("niter", [3, 8, 25])
def test_lanczos_invariant_start_interior(gpu, niter):
    """Lanczos started on an exact eigenvector of an interior eigenvalue much
    smaller than |H| (the Hubbard operator's eigenvalue nearest 0, and one
    of magnitude ~|H|/60): the rounding noise of H v is ~eps |H|, far above
    eps |lambda|, and the breakdown is still recognised (scale from the whole
    tridiagonal matrix), so the result is that eigenvalue, not the ground
    energy reached through normalised noise directions."""
    HL, psi, HR = gen.hubbard_hpsi(seed=26, Dstar=33, mu=1.0, sigma=0.4, dmax=4, qmax=3)
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=gen.fermion_parity_signs(HL))
    S = _dense_weighted(H)
    ev, U = np.linalg.eigh(0.5 * (S + S.T))
    hn = np.abs(ev).max()
    sw = np.sqrt(H.weights())
    k_small = int(np.argmin(np.abs(ev)))
    k_mid = int(np.argmin(np.abs(np.abs(ev) - hn / 60)))
    for k in (k_small, k_mid):
        assert abs(ev[k]) < hn / 30 and ev[k] - ev[0] > 1.0
        e0, ritz, gs = H.lanczos(U[:, k] / sw, niter)
        assert abs(e0 - ev[k]) < 1e-10 * hn, (k, e0, ev[k])
        g = gs * sw
        assert abs(abs(g @ U[:, k]) / np.linalg.norm(g) - 1.0) < 1e-10


def test_split_precision_slices(gpu):
    """BF16-slice H*psi (exact FP32 accumulation, FP64 recombination) against
    the oracle at D* = 1000: the error falls by ~2^-7 per slice and 7
    slices meet the 1e-12 bar."""
    HL, psi, HR = gen.heis_hpsi(seed=27, Dstar=1000)
    v = np.random.RandomState(2).standard_normal(psi.numel())
    ref = _hpsi_oracle(HL, psi.with_packed(v), HR, 1).packed()
    errs = {}
    for S in (3, 4, 5, 6, 7, 8):
        H = Q.HPsi(HL, psi, HR, hconj=1, mode="split_bf16", nslices=S)
        got = H.apply(v)
        errs[S] = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("split errors vs oracle", errs, flush=True)
    for S in (3, 4, 5, 6):
        assert errs[S + 1] < errs[S] * 0.1, errs
    assert errs[7] < TOL and errs[8] < TOL, errs
    assert errs[3] > 1e-9, errs  # the slices really are narrow


def test_split_precision_config4_vs_oracle(gpu, heis4000):
    """config 4 (the headline operator at D* = 4000) in split precision with
    the default slice count: within 1e-12 of the oracle; at full size the
    split H*psi is symmetric in the weighted inner product (<x,Hy>_w =
    <Hx,y>_w), and 20 Lanczos steps give the FP64 ground energy to 1e-13."""
    HL, psi, HR, x, ref = heis4000
    H = Q.HPsi(HL, psi, HR, hconj=1, mode="split_bf16")
    got = H.apply(x)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"split_bf16 default slices, D*=4000: rel err vs oracle {err:.3e}", flush=True)
    assert err < TOL
    w = H.weights()
    y = np.random.RandomState(8).standard_normal(H.N)
    hy = H.apply(y)
    lhs, rhs = (w * x) @ hy, (w * got) @ y
    assert abs(lhs - rhs) < 1e-12 * np.sqrt(((w * x) @ x) * ((w * hy) @ hy))
    Hd = Q.HPsi(HL, psi, HR, hconj=1)
    e_ref = Hd.lanczos(psi.packed(), 20)[0]
    e_split = H.lanczos(psi.packed(), 20)[0]
    assert abs(e_split - e_ref) < 1e-13 * abs(e_ref), (e_split, e_ref)


def test_hpsi_hubbard_replays_match_oracle(gpu):
    """SU(2)xU(1) Hubbard H*psi (D* = 1500, parity signs on HL) under the
    autotuned policy: bitwise stable over 20 graph replays and 20 direct
    launches, and within 1e-12 of the oracle."""
    HL, psi, HR = gen.hubbard_hpsi(seed=28, Dstar=1500, mu=3.0, sigma=0.8, dmax=200)
    ferm = gen.fermion_parity_signs(HL)
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    H.autotune(reps=3)
    x = np.random.RandomState(9).standard_normal(H.N)
    dx = Q.DeviceBuffer.from_array(x)
    dy = Q.DeviceBuffer(H.N)
    outs = []
    for graph in (True, False):
        H.use_graph(graph)
        for _ in range(20):
            H.apply_dev(dx, dy)
            outs.append(dy.download())
    for o in outs:
        assert np.array_equal(o, outs[0])
    ref = _hpsi_oracle(HL, psi.with_packed(x), HR, 1, fermA=ferm).packed()
    assert np.linalg.norm(outs[0] - ref) / np.linalg.norm(ref) < TOL


def test_cell_form_options(gpu):
    """cell-form leg options (icFlags, contractQS.cc:159-293): {A,'!2*',A},
    {A,'13*',A} and {A,[1 3],'*',A} contract legs 1 and 3 like usage #1; the
    nested {A,'!2*',{L,A}} (the environment-update shape of updateHK.m)
    equals the explicit chain; all against the oracle."""
    A = gen.u1_mps(seed=31, Lq=tuple(range(-6, 7, 2)), Rq=tuple(range(-7, 8, 2)), dmin=3, dmax=9)
    R = O.contract(A, [0, 2], A, [0, 2], conjA=1)
    for form in ([A, "!2*", A], [A, "13*", A], [A, [1, 3], "*", A]):
        C = Q.contractQS(form)
        assert_same_structure(C, R)
        assert rel_err(C, R) < TOL
    rng = np.random.RandomState(3)
    Lq = sorted({int(q) for q in A.Q[0][:, 0]})
    dims = {int(q): A.data[i].shape[0] for i, q in enumerate(A.Q[0][:, 0])}
    Lenv = Q.QSpace([np.array([[q] for q in Lq]), np.array([[q] for q in Lq])],
                    [rng.standard_normal((dims[q], dims[q])) for q in Lq], "A", ["L", "L*"])
    C = Q.contractQS([A, "!2*", [Lenv, A]])
    Y = O.contract(Lenv, [1], A, [0])
    Z = O.contract(A, [0, 2], Y, [0, 2], conjA=1)
    assert_same_structure(C, Z)
    assert rel_err(C, Z) < TOL


def test_cell_form_single_operand_and_errors(gpu):
    """one top-level operand (contractQS_itags at level 0 with one entry,
    contractQS.cc:676-704): its conj flag and index option give a conj /
    permutation of it, bit for bit as permuteQS; a nested cell as that
    operand is contracted first; malformed cells are errors"""
    A, L = gen.su2_mps(seed=11, tS=tuple(range(5)), dmin=3, dmax=8)
    for got, ref in ((Q.contractQS(A, "*"), Q.permuteQS(A, [], conj=True)),
                     (Q.contractQS(A, "312"), Q.permuteQS(A, [2, 0, 1])),
                     (Q.contractQS(A, "312*"), Q.permuteQS(Q.permuteQS(A, [], conj=True), [2, 0, 1]))):
        assert_same_structure(got, ref)
        assert all(np.array_equal(a, b) for a, b in zip(got.data, ref.data))
    X = Q.contractQS([L, A], "*")
    Y = Q.permuteQS(Q.contractQS(L, "2", A, "1"), [], conj=True)
    assert_same_structure(X, Y)
    assert rel_err(X, Y) < TOL
    bad = {"one-entry cell": [[A, "*"]], "one-entry nested cell": [A, [L]], "index beyond the rank": [A, A, "5"],
           "one-entry first cell": [[A], L], "one-entry cell with a cell": [[[L, A], "*"]]}
    for name, args in bad.items():
        with pytest.raises(Q.QsxError) as e:
            print("malformed cell form:", name, flush=True)
            Q.contractQS(*args)
        assert e.value.code in (1, 2), name


def test_env_update_u1_cell_form(gpu):
    """environment update L' = A'.(HK.W.A) in the cell form of updateHK.m
    ({A,'!2*',{{A,HK},W}}), contracted on device-resident operands, against
    the oracle's explicit chain."""
    A, HK, W = gen.u1_env_update(seed=41, dw=2)
    C = Q.contractQS([A, "!2*", [[A, HK], W]])
    X1 = O.contract(A, [0], HK, [0])            # (R*, s, L, w*)
    X2 = O.contract(X1, [1, 3], W, [2, 0])      # (R*, L, w2*, s)
    R = O.contract(A, [0, 2], X2, [1, 3], conjA=1)
    assert C.itags == ["R", "R*", "w2*"]
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL


def _complexify(T, seed, real_blocks=False):
    rng = np.random.RandomState(seed)
    data = [np.asarray(d, dtype=np.complex128) if real_blocks and i % 2 else d + 1j * rng.standard_normal(d.shape)
            for i, d in enumerate(T.data)]
    return Q.QSpace(T.Q, [np.asfortranarray(d) for d in data], T.qtype, T.itags, T.info.get("cgw"))


def test_complex_u1_mps_overlap(gpu):
    """complex workload u1_mps: contractQS(A,'12*',A,'12') conjugates the
    data of the first operand (device vs oracle by bilinearity)"""
    A = _complexify(gen.u1_mps(seed=21), 1)
    C = Q.contractQS(A, "12*", A, "12")
    R = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert C.is_complex()
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL


@pytest.mark.parametrize("which", ["both", "A", "B"])
def test_complex_u1_perm_conjB(gpu, which):
    A0 = gen.u1_mps(seed=22)
    A = _complexify(A0, 2, True) if which in ("both", "A") else A0
    B = _complexify(A0, 3) if which in ("both", "B") else A0
    C, R = _both(A, [2], B, [2], perm=[2, 0, 1, 3], conjB=1)
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL


def test_complex_su2_env(gpu):
    """complex SU(2) L.A then X.conj(A), and the missing-conj rule"""
    A, L = gen.su2_mps(seed=23, tS=tuple(range(7)), dmin=4, dmax=12)
    A = _complexify(A, 4)
    L = _complexify(L, 5, True)
    X, XR = _both(L, [1], A, [0])
    assert_same_structure(X, XR)
    assert rel_err(X, XR) < TOL
    C, CR = _both(XR, [0, 1], A, [0, 1], conjB=1)
    assert_same_structure(C, CR)
    assert rel_err(C, CR) < TOL
    N, NR = _both(A, [0, 1], A, [0, 1])
    assert_same_structure(N, NR)
    assert rel_err(N, NR) < TOL


def test_complex_env_update_cell_form(gpu):
    """the cell form with complex leaves (a complex MPS tensor, real HK and
    W): {A,'!2*',{{A,HK},W}} against the oracle's explicit complex chain;
    the '*' option conjugates A's data"""
    A0, HK, W = gen.u1_env_update(seed=42, dw=2)
    A = _complexify(A0, 6)
    C = Q.contractQS([A, "!2*", [[A, HK], W]])
    X1 = O.contract(A, [0], HK, [0])
    X2 = O.contract(X1, [1, 3], W, [2, 0])
    R = O.contract(A, [0, 2], X2, [1, 3], conjA=1)
    assert C.is_complex() and C.itags == ["R", "R*", "w2*"]
    assert_same_structure(C, R)
    assert rel_err(C, R) < TOL


def test_split_f32x3_vs_oracle(gpu, heis4000):
    """the FP32 x 3 variant of config 4 (three FP32 slices on the FP32
    matrix cores, FP32 accumulation, FP64 recombination) at D* = 1000 and at
    the headline D* = 4000 against the oracle, to its stated bound of 5e-10
    (the u = 1 products accumulate in FP32: about 2^-35 of |row| |col| per
    term); weighted symmetry and the 20-step Lanczos energy (relative 1e-10)"""
    BOUND = 5e-10
    HL, psi, HR = gen.heis_hpsi(seed=27, Dstar=1000)
    v = np.random.RandomState(2).standard_normal(psi.numel())
    ref = _hpsi_oracle(HL, psi.with_packed(v), HR, 1).packed()
    got = Q.HPsi(HL, psi, HR, hconj=1, mode="split_f32").apply(v)
    err1 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    HL, psi, HR, x, ref = heis4000
    H = Q.HPsi(HL, psi, HR, hconj=1, mode="split_f32")
    got = H.apply(x)
    err4 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"split_f32x3: rel err vs oracle D*=1000 {err1:.3e}, D*=4000 {err4:.3e}", flush=True)
    assert err1 < BOUND and err4 < BOUND
    assert err4 > 1e-13  # FP32 accumulation really is in the path
    w = H.weights()
    y = np.random.RandomState(8).standard_normal(H.N)
    hy = H.apply(y)
    assert abs((w * x) @ hy - (w * got) @ y) < 10 * BOUND * np.sqrt(((w * x) @ x) * ((w * hy) @ hy))
    e_ref = Q.HPsi(HL, psi, HR, hconj=1).lanczos(psi.packed(), 20)[0]
    e_f32 = H.lanczos(psi.packed(), 20)[0]
    assert abs(e_f32 - e_ref) < 1e-10 * abs(e_ref), (e_f32, e_ref)
    with pytest.raises(Q.QsxError):
        Q.HPsi(HL, psi, HR, hconj=1, mode="split_f32", nslices=4)
