"""Pseudo-MPO two-site H*psi (get_HPsi.m:62-83) on the device against the
oracle's contraction chain (GPU tests): HK_L.psi + psi.HK_R + sum_j
{OP_L(j),'*',psi}.OP_R(j) [+ {OP_L(j),psi}.OP_R(j)' where hconj(j)]."""
import numpy as np
import pytest

import qspace_amd as Q
from qspace_amd import gen
from helpers import pseudo_hpsi_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _err(got, ref):
    return np.linalg.norm(got - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("case", ["full", "no_local", "left_only", "right_only"])
def test_pseudo_hpsi_matches_oracle(gpu, case):
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=41, Dstar=700, nop=2)
    hconj = [1, 0]
    if case == "no_local":
        HKL = HKR = None
    elif case == "left_only":
        HKR, OPL, OPR, hconj = None, [], [], []
    elif case == "right_only":
        HKL, OPL, OPR, hconj = None, OPL[:1], OPR[:1], [0]
    H = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, hconj)
    x = np.random.RandomState(2).standard_normal(H.N)
    ref = pseudo_hpsi_oracle(HKL, psi.with_packed(x), HKR, OPL, OPR, hconj).packed()
    got = H.apply(x)
    print(f"pseudo H*psi {case}: N={H.N} rel err {_err(got, ref):.2e}", flush=True)
    assert _err(got, ref) < TOL
    H.autotune(reps=2)
    assert _err(H.apply(x), ref) < TOL


@pytest.mark.parametrize("policy", [{"engine": "grouped", "split_steps": 2},
                                    {"engine": "wave", "wave_max": 2, "group_mfma": 16, "split_mfma": 32},
                                    {"engine": "wave", "wave_max": 1, "group_mfma": 8, "xcd_bins": 0}])
def test_pseudo_policies_and_plain(gpu, policy):
    """every engine reads psi through the local terms' operand flag; the
    per-pair plain path gives the same H*psi"""
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=42, Dstar=300, mu=2.5, sigma=0.7, dmax=70, nop=1)
    x = np.random.RandomState(5).standard_normal(psi.numel())
    ref = pseudo_hpsi_oracle(HKL, psi.with_packed(x), HKR, OPL, OPR, [1]).packed()
    H = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, [1])
    H.set_policy(policy, policy)
    assert _err(H.apply(x), ref) < TOL
    Hp = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, [1], mode="plain")
    assert _err(Hp.apply(x), ref) < TOL


def test_pseudo_hermitian_and_eigensolvers(gpu):
    """with hconj on every operator (and symmetric block Hamiltonians) the
    operator is symmetric in the weighted inner product; Lanczos and
    Davidson with a Krylov space of the full dimension give the lowest
    eigenvalue of the dense operator"""
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=43, Dstar=40, mu=1.2, sigma=0.4, dmax=6, nop=2)
    H = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, [1, 1])
    w = H.weights()
    rng = np.random.RandomState(6)
    x, y = rng.standard_normal(H.N), rng.standard_normal(H.N)
    hx, hy = H.apply(x), H.apply(y)
    assert abs((w * x) @ hy - (w * hx) @ y) < 1e-12 * np.sqrt(((w * hx) @ hx) * ((w * y) @ y))
    cols = [H.apply(e) for e in np.eye(H.N)]
    M = np.column_stack(cols)
    sw = np.sqrt(w)
    S = (sw[:, None] * M) / sw[None, :]
    assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
    e_ref = np.linalg.eigvalsh(0.5 * (S + S.T))[0]
    for solver in (H.lanczos, H.davidson):
        e0, _, gs = solver(rng.standard_normal(H.N), H.N)
        assert abs(e0 - e_ref) < 1e-10 * max(1.0, abs(e_ref)), (solver.__name__, e0, e_ref)


@pytest.mark.parametrize("seed", range(10))
def test_random_pseudo_operators(gpu, seed):
    """random pseudo-MPO operators (sizes, 1-3 operator pairs, random hconj
    per pair) under the default decomposition: the oracle bar, and a second
    apply bitwise equal"""
    rng = np.random.RandomState(500 + seed)
    Dstar = int(rng.choice([20, 60, 150, 400]))
    nop = int(rng.randint(1, 4))
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=500 + seed, Dstar=Dstar, mu=float(rng.uniform(1.0, 3.0)),
                                                   sigma=float(rng.uniform(0.3, 0.9)),
                                                   dmax=int(rng.choice([6, 30, 90])), nop=nop)
    hc = [int(rng.randint(0, 2)) for _ in range(nop)]
    print(f"seed {seed}: D*={Dstar} nop={nop} hconj={hc}", flush=True)
    H = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, hc)
    x = rng.standard_normal(H.N)
    ref = pseudo_hpsi_oracle(HKL, psi.with_packed(x), HKR, OPL, OPR, hc).packed()
    got = H.apply(x)
    assert _err(got, ref) < TOL
    assert np.array_equal(H.apply(x), got)
