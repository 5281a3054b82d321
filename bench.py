"""Benchmark of the SU(2) two-site H*psi (the contractQS hot path) on one MI355X.

Default workload: spin-1/2 Heisenberg two-site H*psi in SU(2) at D* = 4000
kept multiplets (enlarged-bond psi, rank-3 environments, hconj term), FP64.
A "step" is one H*psi on device-resident inputs (graph replay of the two
batched-GEMM launches).  Prints ONE JSON line.

Before the timed region the run also makes one device copy of 64 MiB
(k_copy, 64 MiB read + 64 MiB written): counter runs of this same command
(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) calibrate the per-dispatch HBM
byte counts of the GEMM kernels against it.

    python bench.py [--steps K] [--warmup W] [--workload heis4000|heis2000|hubbard4000]
                    [--no-cpu] [--lite] [--policy FILE]

Besides the headline H*psi the JSON line carries "configs": the benchmark
configurations 1 (U(1) MPS overlap, CPU oracle and GPU), 2 (SU(2)
environment step), 3 (Heisenberg D*=2000, 20 Lanczos iterations), 4 (split
precision: BF16-slice and FP32x3 H*psi error/time on the headline operator),
5 (SU(2)xU(1) Hubbard D*=4000 with parity signs, 30 Davidson iterations, the
plain per-pair path and both split modes) and the pseudo-MPO H*psi of
get_HPsi.m:62-83 at D*=4000; and "block_ops": permuteQS / plusQS bandwidth
on the SU(2) environments, the U(1) and SU(2) environment updates in the
cell form of updateHK.m and the orthoQS truncations.

    python bench.py --split-profile bf16|f32    # only the split H*psi (profiler runs)
    python bench.py --eig-profile [--workload W] # only the eigensolver of an untuned plan (profiler runs)
    python bench.py --save-policy FILE          # also write the autotuned stage policies
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "useful GFLOP/s + H·ψ ms at D*; rel err vs FP64 oracle (1 MI355X)"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "heis4000": dict(kind="heis", seed=1003, Dstar=4000, mu=3.0, sigma=0.8, dmax=256,
                     name="su2_heisenberg_two_site_hpsi_Dstar4000"),
    "heis2000": dict(kind="heis", seed=1003, Dstar=2000, mu=3.0, sigma=0.8, dmax=256,
                     name="su2_heisenberg_two_site_hpsi_Dstar2000"),
    "heis8000": dict(kind="heis", seed=1003, Dstar=8000, mu=3.0, sigma=0.8, dmax=256,
                     name="su2_heisenberg_two_site_hpsi_Dstar8000"),
    "heis16000": dict(kind="heis", seed=1003, Dstar=16000, mu=3.0, sigma=0.8, dmax=256,
                      name="su2_heisenberg_two_site_hpsi_Dstar16000"),
    "hubbard4000": dict(kind="hubbard", seed=1005, Dstar=4000, mu=3.2, sigma=0.9, dmax=384,
                        name="su2xu1_hubbard_two_site_hpsi_Dstar4000"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(w):
    from qspace_amd import gen
    if w["kind"] == "heis":
        HL, psi, HR = gen.heis_hpsi(seed=w["seed"], Dstar=w["Dstar"], mu=w["mu"], sigma=w["sigma"], dmax=w["dmax"])
        ferm = None
    else:
        HL, psi, HR = gen.hubbard_hpsi(seed=w["seed"], Dstar=w["Dstar"], mu=w["mu"], sigma=w["sigma"], dmax=w["dmax"])
        ferm = gen.fermion_parity_signs(HL)
    return HL, psi, HR, ferm


def restrict(T, keep_fn):
    """sub-QSpace of the blocks whose label records satisfy keep_fn."""
    from qspace_amd import QSpace
    idx = [i for i, r in enumerate(T.QIDX) if keep_fn(r)]
    return QSpace([q[idx] for q in T.Q], [T.data[i] for i in idx], T.qtype, T.itags)


CALIB_MIB = 64


def calibration_copy():
    """one k_copy of CALIB_MIB MiB (known HBM traffic) for counter calibration"""
    import qspace_amd as Q
    n = CALIB_MIB * (1 << 20) // 8
    a = Q.DeviceBuffer(n)
    b = Q.DeviceBuffer(n)
    Q._capi.lib().qsx_memset_device(a.ptr, 0, n * 8, None)
    Q._capi.lib().qsx_copy_device(b.ptr, a.ptr, n, None)
    Q._capi.lib().qsx_synchronize()
    return {"kernel": "k_copy", "read_bytes": n * 8, "write_bytes": n * 8}


def time_hpsi(H, dpsi, dout, steps, warmup, graph=True):
    import qspace_amd as Q
    H.use_graph(graph)
    for _ in range(warmup):
        H.apply_dev(dpsi, dout)
    Q._capi.lib().qsx_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        H.apply_dev(dpsi, dout)
    Q._capi.lib().qsx_synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def stage_timing(H, dpsi, dout, reps):
    """per-launch device time of the two grouped-GEMM kernels of one H*psi
    (HIP events on the stream they run on), direct launches, mean over reps"""
    H.use_graph(False)
    t = []
    for _ in range(reps):
        H.apply_dev(dpsi, dout)
        t.append(H.last_timing())
    H.use_graph(True)
    return [float(x) for x in np.mean(np.array(t), axis=0)]


def oracle_hpsi(HL, psi, HR, ferm, x):
    """the oracle's H*psi of the packed vector x (both terms, parity signs on
    HL): the explicit contraction chain of get_HPsi.m:54-57"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xs = psi.with_packed(x)
    T1 = O.contract(HL, [1], xs, [0], fermA=ferm, skip_zero=0)
    T2 = O.contract(HL, [0], xs, [0], conjA=1, fermA=ferm, skip_zero=0)
    key = {tuple(r): i for i, r in enumerate(psi.QIDX)}
    acc = [np.zeros_like(d) for d in psi.data]
    for Ht in (O.contract(T1, [1, 2], HR, [2, 1], skip_zero=0),
               O.contract(T2, [1, 2], HR, [2, 0], conjB=1, skip_zero=0)):
        for r, d in zip(Ht.QIDX, Ht.data):
            acc[key[tuple(r)]] += d
    return np.concatenate([a.ravel(order="F") for a in acc])


def oracle_pseudo_hpsi(HKL, psi, HKR, OPL, OPR, hconj, x):
    """the oracle's pseudo-MPO H*psi of the packed vector x: the explicit
    contraction chain of get_HPsi.m:62-83 (HK_L.psi + psi.HK_R + per
    operator pair {OP_L,'*',psi}.OP_R [+ {OP_L,psi}.OP_R' where hconj])"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xs = psi.with_packed(x)
    terms = [O.contract(HKL, [1], xs, [0], skip_zero=0), O.contract(xs, [1], HKR, [1], skip_zero=0)]
    for L, R, h in zip(OPL, OPR, hconj):
        T1 = O.contract(L, [0], xs, [0], conjA=1, skip_zero=0)
        terms.append(O.contract(T1, [1, 2], R, [2, 1], skip_zero=0))
        if h:
            T2 = O.contract(L, [1], xs, [0], skip_zero=0)
            terms.append(O.contract(T2, [1, 2], R, [2, 0], conjB=1, skip_zero=0))
    key = {tuple(r): i for i, r in enumerate(psi.QIDX)}
    acc = [np.zeros_like(d) for d in psi.data]
    for Ht in terms:
        for r, d in zip(Ht.QIDX, Ht.data):
            acc[key[tuple(r)]] += d
    return np.concatenate([a.ravel(order="F") for a in acc])


def pseudo_mpo():
    """pseudo-MPO two-site H*psi (get_HPsi.m:62-83, the form update_psi_2site
    uses without a full MPO): SU(2) Heisenberg block Hamiltonians HK_L, HK_R
    and one spin operator pair with its Hermitian-conjugate term, D* = 4000;
    one multi-term plan, hipGraph replay, against the oracle's chain and the
    per-pair plain path"""
    import qspace_amd as Q
    from qspace_amd import gen
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=1008, Dstar=4000, nop=1)
    hc = [1]
    H = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, hc)
    H.autotune(5)
    x = Q.DeviceBuffer.from_array(psi.packed())
    y = Q.DeviceBuffer(H.N)
    ms = time_hpsi(H, x, y, 50, 5)
    got = y.download()
    ref = oracle_pseudo_hpsi(HKL, psi, HKR, OPL, OPR, hc, psi.packed())
    Hp = Q.HPsi.pseudo(HKL, psi, HKR, OPL, OPR, hc, mode="plain")
    ms_plain = time_hpsi(Hp, x, y, 3, 1, graph=False)
    out = {"workload": "su2_heisenberg_pseudo_mpo_two_site_hpsi_Dstar4000", "psi_elements": H.N,
           "useful_gflop_per_hpsi": H.flops / 1e9, "hpsi_ms": ms, "gflops": H.flops / ms / 1e6,
           "rel_err_hpsi_vs_oracle": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
           "plain_fp64_per_pair_ms": ms_plain, "speedup_vs_plain_fp64": ms_plain / ms}
    log(f"[bench] pseudo-MPO H*psi: {out}")
    return out


def host_cpus():
    """CPUs this process may use (its affinity mask, capped by
    OMP_NUM_THREADS when the host sets it), with nproc and the CPU model"""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": omp or None, "cpu_model": model}


def cpu_baseline(w, sample_smax=None, target_s=10.0):
    """Reference-structure CPU path (oracle/, OpenMP over output blocks, per
    pair OpenBLAS dgemm + axpy): the full workload (or, with sample_smax, the
    same generated operands restricted to sectors with 2S <= sample_smax),
    repeated for about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    HL, psi, HR, ferm = make_workload(w)
    spin_col = 0 if w["kind"] == "heis" else 1
    qd = psi.qdim

    def keep(r):
        return sample_smax is None or all(r[k * qd + spin_col] <= sample_smax for k in range(len(r) // qd))
    HLs, psis, HRs = restrict(HL, keep), restrict(psi, keep), restrict(HR, keep)
    if ferm is not None:  # the signs of the sample's own HL blocks (signs do not change the cost)
        from qspace_amd import gen
        ferm = gen.fermion_parity_signs(HLs)
    cores, host = host_cpus()
    # the four contractions of H*psi (two stages x hconj term)
    T1 = O.contract(HLs, [1], psis, [0], fermA=ferm, skip_zero=0)
    T2 = O.contract(HLs, [0], psis, [0], conjA=1, fermA=ferm, skip_zero=0)
    jobs = [O.CpuBaseline(HLs, [1], psis, [0]), O.CpuBaseline(HLs, [0], psis, [0], conjA=1),
            O.CpuBaseline(T1, [1, 2], HRs, [2, 1]), O.CpuBaseline(T2, [1, 2], HRs, [2, 0], conjB=1)]
    flops = sum(j.flops for j in jobs)

    def timed(nthreads, budget):
        for j in jobs:
            j.run(nthreads)  # warm
        reps = 0
        t0 = time.perf_counter()
        while True:
            for j in jobs:
                j.run(nthreads)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget:
                return el / reps, reps, el

    sec, reps, el = timed(cores, target_s)
    sec8, reps8, _ = timed(8, 0.5 * target_s)
    # GPU vs CPU on the same sample: rel err of H*psi vs the FP64 oracle
    import qspace_amd as Q
    Hg = Q.HPsi(HLs, psis, HRs, hconj=1, fermA=ferm)
    got = Hg.apply(psis.packed())
    ref = oracle_hpsi(HLs, psis, HRs, ferm, psis.packed())
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return {
        "value": flops / sec / 1e9, "unit": "GFLOP/s", "cores": cores, "kind": "port",
        "sample": ((f"the full workload ({psi.nblk} psi blocks" if sample_smax is None else
                    f"same seeded operands restricted to sectors with 2S<={sample_smax} ({psis.nblk} of {psi.nblk} "
                    "psi blocks") + f", {flops / 1e9:.3f} GFLOP per H*psi), {reps} H*psi in {el:.1f} s on all "
                   f"{cores} CPUs this process may use; OpenMP over output blocks, per-pair dgemm + axpy "
                   f"({os.path.basename(jobs[0].blas)}, 1 BLAS thread)"),
        "host": host,
        "value_8_threads": flops / sec8 / 1e9,
        "ms_per_hpsi_8_threads": sec8 * 1e3,
        "reps_8_threads": reps8,
        "ms_per_hpsi_sample": sec * 1e3,
        "rel_err_gpu_vs_fp64_oracle_sample": err,
    }


def _rel(C, R):
    """relative Frobenius error of C vs R; structure (records, dims) must agree"""
    if not (np.array_equal(C.QIDX, R.QIDX) and [d.shape for d in C.data] == [d.shape for d in R.data]):
        return float("inf")
    c, r = C.packed(), R.packed()
    return float(np.linalg.norm(c - r) / max(np.linalg.norm(r), 1e-300))


def _time_eig(fn, x0, niter, reps=3):
    """(first call ms, steady-state ms, E0): the steady state is the best of
    `reps` further calls (a single call is exposed to host scheduling noise)"""
    t0 = time.perf_counter()
    fn(x0, niter)
    first = (time.perf_counter() - t0) * 1e3
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        e0 = fn(x0, niter)[0]
        t = (time.perf_counter() - t0) * 1e3
        best = t if best is None else min(best, t)
    return first, best, e0


def _time_plan(P, A, B, reps=20):
    """device time of one batched execution of a contraction plan (inputs resident)"""
    import qspace_amd as Q
    dA = Q.DeviceBuffer.from_array(A.packed())
    dB = Q.DeviceBuffer.from_array(B.packed())
    dC = Q.DeviceBuffer(P.nout)
    for _ in range(3):
        P.execute(dA, dB, dC)
    Q._capi.lib().qsx_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        P.execute(dA, dB, dC)
    Q._capi.lib().qsx_synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def extra_configs():
    """the five benchmark configurations (config 4 = the headline workload)"""
    import qspace_amd as Q
    from qspace_amd import gen
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {}
    # 1: U(1) MPS overlap contractQS(A,'12*',A,'12'): CPU oracle and GPU
    A = gen.u1_mps()
    t0 = time.perf_counter()
    R = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    C = Q.contractQS(A, "12*", A, "12")
    P = Q.CtrPlan(A, "12*", A, "12")
    out["1_u1_mps_overlap"] = {"blocks": A.nblk, "useful_flops": C.flops, "oracle_cpu_ms": cpu_ms,
                               "gpu_plan_execute_ms": _time_plan(P, A, A), "rel_err_vs_oracle": _rel(C, R)}
    log(f"[bench] config 1: {out['1_u1_mps_overlap']}")
    # 2: SU(2) environment step X = L.A, C = X.conj(A)
    A, L = gen.su2_mps()
    X = Q.contractQS(L, "2", A, "1")
    XR = O.contract(L, [1], A, [0])
    Cg = Q.contractQS(X, "12", A, "12*")
    CR = O.contract(XR, [0, 1], A, [0, 1], conjB=1)
    P1 = Q.CtrPlan(L, "2", A, "1")
    P2 = Q.CtrPlan(X, "12", A, "12*")
    out["2_su2_env"] = {"multiplets_per_bond": int(sum(d.shape[0] for d in L.data)),
                        "useful_flops": X.flops + Cg.flops,
                        "gpu_plan_execute_ms": _time_plan(P1, L, A) + _time_plan(P2, X, A),
                        "rel_err_vs_oracle": max(_rel(X, XR), _rel(Cg, CR))}
    log(f"[bench] config 2: {out['2_su2_env']}")
    # 3: Heisenberg H*psi at D*=2000, 20 Lanczos iterations
    w = WORKLOADS["heis2000"]
    HL, psi, HR, ferm = make_workload(w)
    H = Q.HPsi(HL, psi, HR, hconj=1)
    H.autotune(5)
    x = Q.DeviceBuffer.from_array(psi.packed())
    y = Q.DeviceBuffer(H.N)
    ms = time_hpsi(H, x, y, 50, 5)
    lan_first, lan, e0 = _time_eig(H.lanczos, psi.packed(), 20)
    ritz = H.lanczos(psi.packed(), 20)[1]
    out["3_heis2000_lanczos20"] = {"psi_elements": H.N, "useful_gflop_per_hpsi": H.flops / 1e9, "hpsi_ms": ms,
                                   "gflops": H.flops / ms / 1e6, "lanczos20_ms": lan,
                                   "lanczos20_first_call_ms": lan_first, "E0": e0,
                                   "ritz_last3": [float(r) for r in ritz[-3:]]}
    log(f"[bench] config 3: {out['3_heis2000_lanczos20']}")
    del H
    # 4: split precision (BF16 slices, exact FP32 accumulation, FP64
    # recombination) on the headline operator: error table per slice count
    # against the FP64 oracle (and the GPU's own FP64 result)
    w = WORKLOADS["heis4000"]
    HL, psi, HR, ferm = make_workload(w)
    Hd = Q.HPsi(HL, psi, HR, hconj=1)
    Hd.autotune(5)
    xv = psi.packed()
    ref = Hd.apply(xv)
    oref = oracle_hpsi(HL, psi, HR, None, xv)
    e_ref, _, _ = Hd.lanczos(xv, 20)
    x = Q.DeviceBuffer.from_array(xv)
    y = Q.DeviceBuffer(Hd.N)
    ms_fp64 = time_hpsi(Hd, x, y, 50, 5)
    fp64_err = float(np.linalg.norm(ref - oref) / np.linalg.norm(oref))
    del Hd
    table = {}
    for S in (5, 6, 7, 8):
        Hs = Q.HPsi(HL, psi, HR, hconj=1, mode="split_bf16", nslices=S)
        got = Hs.apply(xv)
        ms = time_hpsi(Hs, x, y, 20, 3)
        es, _, _ = Hs.lanczos(xv, 20)
        inf = Hs.stage_info()
        table[str(S)] = {"rel_err_hpsi_vs_oracle": float(np.linalg.norm(got - oref) / np.linalg.norm(oref)),
                         "rel_err_hpsi_vs_gpu_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                         "E0_rel_err_lanczos20": float(abs(es - e_ref) / abs(e_ref)), "hpsi_ms": ms,
                         "executed_slice_gflop": float(sum(inf["split_exec_flops"]) / 1e9)}
        del Hs
    Hs = Q.HPsi(HL, psi, HR, hconj=1, mode="split_bf16")
    default_err = float(np.linalg.norm(Hs.apply(xv) - oref) / np.linalg.norm(oref))
    del Hs
    # FP32 x 3 slices (FP32 matrix cores, FP32 accumulation, FP64 recombination)
    Hf = Q.HPsi(HL, psi, HR, hconj=1, mode="split_f32")
    got = Hf.apply(xv)
    ms_f32 = time_hpsi(Hf, x, y, 20, 3)
    ef, _, _ = Hf.lanczos(xv, 20)
    inf = Hf.stage_info()
    f32 = {"rel_err_hpsi_vs_oracle": float(np.linalg.norm(got - oref) / np.linalg.norm(oref)),
           "rel_err_hpsi_vs_gpu_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
           "E0_rel_err_lanczos20": float(abs(ef - e_ref) / abs(e_ref)), "hpsi_ms": ms_f32,
           "executed_slice_gflop": float(sum(inf["split_exec_flops"]) / 1e9),
           "stated_bound_rel_err": 5e-10}
    del Hf
    out["4_split_bf16_heis4000"] = {"fp64_hpsi_ms": ms_fp64, "fp64_rel_err_vs_oracle": fp64_err,
                                    "E0_fp64_lanczos20": e_ref, "default_slices_rel_err_vs_oracle": default_err,
                                    "slices": table, "fp32x3": f32}
    log(f"[bench] config 4: {out['4_split_bf16_heis4000']}")
    # 5: SU(2)xU(1) Hubbard H*psi at D*=4000 with parity signs, 30 Davidson iterations
    w = WORKLOADS["hubbard4000"]
    HL, psi, HR, ferm = make_workload(w)
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    H.autotune(5)
    x = Q.DeviceBuffer.from_array(psi.packed())
    y = Q.DeviceBuffer(H.N)
    ms = time_hpsi(H, x, y, 50, 5)
    got = y.download()
    ref = oracle_hpsi(HL, psi, HR, ferm, psi.packed())
    dav_first, dav, e0 = _time_eig(H.davidson, psi.packed(), 30)
    # the same H*psi on the plain per-pair FP64 path and in both split modes
    Hp = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm, mode="plain")
    ms_plain = time_hpsi(Hp, x, y, 3, 1, graph=False)
    del Hp
    split = {}
    for mode in ("split_bf16", "split_f32"):
        Hs = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm, mode=mode)
        ms_s = time_hpsi(Hs, x, y, 20, 3)
        gs = y.download()
        split[mode] = {"hpsi_ms": ms_s, "rel_err_hpsi_vs_oracle": float(np.linalg.norm(gs - ref) / np.linalg.norm(ref)),
                       "executed_slice_gflop": float(sum(Hs.stage_info()["split_exec_flops"]) / 1e9)}
        del Hs
    out["5_hubbard4000_davidson30"] = {"psi_elements": H.N, "useful_gflop_per_hpsi": H.flops / 1e9, "hpsi_ms": ms,
                                       "gflops": H.flops / ms / 1e6, "davidson30_ms": dav,
                                       "davidson30_first_call_ms": dav_first, "E0": e0,
                                       "rel_err_hpsi_vs_oracle": float(np.linalg.norm(got - ref) /
                                                                       np.linalg.norm(ref)),
                                       "plain_fp64_per_pair_ms": ms_plain, "speedup_vs_plain_fp64": ms_plain / ms,
                                       "split": split}
    log(f"[bench] config 5: {out['5_hubbard4000_davidson30']}")
    return out


def block_ops():
    """permuteQS / plusQS on the device over the SU(2) Heisenberg environment
    HL (rank 3, all blocks in one launch), at D* = 4000 and 32000; device time
    of the one launch (events), best of 5.  Algorithmic bytes: permute 16 B
    per element (read + write), plus 24 B (two reads + write)."""
    import qspace_amd as Q
    from qspace_amd import gen
    rng = np.random.RandomState(11)
    out = {}
    for D in (4000, 32000):
        HL, _, _ = gen.heis_hpsi(seed=1003, Dstar=D)
        n = HL.numel()
        row = {"blocks": HL.nblk, "elements": n, "max_block": [int(x) for x in max(d.shape for d in HL.data)]}
        for perm in ([1, 0, 2], [2, 1, 0]):
            ms = min(Q.permuteQS(HL, perm).gpu_ms for _ in range(5))
            gbs = 16.0 * n / (ms * 1e-3) / 1e9
            row[f"permute_{''.join(str(k + 1) for k in perm)}"] = {"ms": ms, "GB/s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS}
        B = HL.with_packed(rng.standard_normal(n))
        ms = min(Q.plusQS(HL, B, 0.5).gpu_ms for _ in range(5))
        gbs = 24.0 * n / (ms * 1e-3) / 1e9
        row["plus"] = {"ms": ms, "GB/s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS}
        out[f"su2_heisenberg_env_Dstar{D}"] = row
        log(f"[bench] block ops D*={D}: {row}")
    return out


def env_update():
    """U(1) environment update L' = A'.(HK.W.A) in the cell form of
    updateHK.m, {A,'!2*',{{A,HK},W}}: bond of 21 charge sectors with
    40..120 states each, MPO bond 3 states per charge.  GPU: the whole
    contractQS call (upload of the leaves, three plans built and executed on
    device buffers, download), mean of 3 after one warm-up; CPU: the
    oracle's explicit chain, one thread (the checker and the CPU baseline)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import qspace_amd as Q
    from qspace_amd import gen
    A, HK, W = gen.u1_env_update(seed=1006, Lq=tuple(range(-20, 21, 2)), Rq=tuple(range(-21, 22, 2)), dmin=40,
                                 dmax=120, dw=3)
    Q.contractQS([A, "!2*", [[A, HK], W]])
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        C = Q.contractQS([A, "!2*", [[A, HK], W]])
        ts.append((time.perf_counter() - t0) * 1e3)
    t0 = time.perf_counter()
    X1 = O.contract(A, [0], HK, [0])
    X2 = O.contract(X1, [1, 3], W, [2, 0])
    R = O.contract(A, [0, 2], X2, [1, 3], conjA=1)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    out = {"bond_states": int(sum(A.data[i].shape[0] for i in range(A.nblk) if int(A.Q[2][i, 0]) == 1)),
           "blocks": [A.nblk, HK.nblk, W.nblk], "useful_gflop": (X1.flops + X2.flops + R.flops) / 1e9,
           "gpu_wall_ms": sum(ts) / len(ts), "oracle_cpu_ms_1thread": cpu_ms, "rel_err_vs_oracle": _rel(C, R)}
    log(f"[bench] env update: {out}")
    return out


def env_update_su2():
    """SU(2) environment update L' = A'.(HK.W.A), {A,'!2*',{{A,HK},W}}: bond
    labels 2S in 0..16 with 8..64 states per multiplet, a rank-4 MPO (operator
    legs of spin 0 and 1, 4 states each) and rank-4 intermediates with outer
    multiplicity.  GPU: the whole contractQS call as above; the oracle's
    chain (numpy restatement of the outer-multiplicity contraction) for the
    error, compared block by block (record-keyed)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import qspace_amd as Q
    from qspace_amd import gen
    A, HK, W = gen.su2_env_update(seed=1010, tS=tuple(range(17)), dmin=8, dmax=64, dw=4)
    Q.contractQS([A, "!2*", [[A, HK], W]])
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        C = Q.contractQS([A, "!2*", [[A, HK], W]])
        ts.append((time.perf_counter() - t0) * 1e3)
    t0 = time.perf_counter()
    X1 = O.contract(A, [0], HK, [0])
    X2 = O.contract(X1, [1, 3], W, [2, 0])
    R = O.contract(A, [0, 2], X2, [1, 3], conjA=1)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    key = lambda T: [tuple(tuple(int(x) for x in T.Q[k][i]) for k in range(T.rank)) for i in range(T.nblk)]
    dr = dict(zip(key(R), R.data))
    num = sum(float(np.sum((d - dr[k]) ** 2)) for k, d in zip(key(C), C.data) if k in dr)
    den = sum(float(np.sum(d ** 2)) for d in R.data)
    out = {"bond_multiplets": int(sum(A.data[i].shape[0] for i in range(A.nblk))),
           "blocks": [A.nblk, HK.nblk, W.nblk], "max_om_intermediate": int(max(X2.om())),
           "same_records": sorted(key(C)) == sorted(key(R)), "gpu_wall_ms": sum(ts) / len(ts),
           "oracle_numpy_ms": cpu_ms, "rel_err_vs_oracle": float(np.sqrt(num / den))}
    log(f"[bench] env update SU(2): {out}")
    return out


def truncation():
    """orthoQS (block SVD truncation after the eigensolve, mpsOrthoQS) of a
    U(1) two-site wave function psi(L,s1,s2*,R*) split between (L,s1) and
    (s2,R): 11 charge sectors per bond, "lds" with U[4,32] states per sector
    (group matrices up to 64 x 64, decomposed in LDS), "bond1000" with
    U[40,140] (about 1000 states per bond, group matrices about 180 x 180,
    decomposed in place in HBM/L2); Nkeep = 512.  And an SU(2) MPS tensor
    A(L,s,R*) (17 multiplets per bond, U[8,64] per multiplet, group matrices
    up to 128 x 64) orthonormalised to the left (K = 2) and to the right
    (K = 1), Nkeep = 64 multiplets.  GPU: device time of the call (events:
    assembly, Jacobi SVD, result blocks) and its wall time, best of 3; CPU:
    the oracle (numpy LAPACK gesdd), wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import qspace_amd as Q
    from qspace_amd import gen
    cases = [(f"u1_two_site_psi_{name}", gen.u1_two_site_psi(seed=1009, dmin=dmin, dmax=dmax), 2, 512)
             for name, dmin, dmax in (("lds", 4, 32), ("bond1000", 40, 140))]
    A = gen.su2_mps(seed=1002)[0]
    cases += [("su2_mps_left", A, 2, 64), ("su2_mps_right", A, 1, 64)]
    out = {}
    for key, P, K, nkeep in cases:
        Q.orthoQS(P, K, Nkeep=nkeep)
        best, wall = None, []
        for _ in range(3):
            t0 = time.perf_counter()
            A1, A2, info = Q.orthoQS(P, K, Nkeep=nkeep)
            wall.append((time.perf_counter() - t0) * 1e3)
            best = info if best is None or info["gpu_ms"] < best["gpu_ms"] else best
        t0 = time.perf_counter()
        R1, R2, rinfo = O.ortho(P, K, Nkeep=nkeep)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        smax = float(rinfo["svd"][0])
        X = O.contract(A1, [K], A2, [P.rank - K])
        XR = O.contract(R1, [K], R2, [P.rank - K])
        out[key] = {
            "blocks": P.nblk, "elements": P.numel(), "K": K, "groups": best["ngroups"], "sweeps": best["sweeps"],
            "Ntot": best["Ntot"], "Nkeep": best["Nkeep"], "gpu_ms": best["gpu_ms"], "gpu_wall_ms": min(wall),
            "oracle_cpu_ms": cpu_ms,
            "svd_max_err_rel": float(np.abs(best["svd"] - rinfo["svd"]).max() / smax),
            "product_rel_err_vs_oracle": _rel(X, XR)}
        log(f"[bench] truncation {key}: {out[key]}")
    return out


def split_profile(mode, steps, warmup):
    """only the split-precision H*psi of the headline operator (for profiler
    runs of its kernels): time per apply and executed slice flops"""
    import qspace_amd as Q
    w = WORKLOADS["heis4000"]
    HL, psi, HR, ferm = make_workload(w)
    H = Q.HPsi(HL, psi, HR, hconj=1, mode=f"split_{mode}")
    calib = calibration_copy()
    x = Q.DeviceBuffer.from_array(psi.packed())
    y = Q.DeviceBuffer(H.N)
    ms = time_hpsi(H, x, y, steps, warmup, graph=False)
    inf = H.stage_info()
    fl = float(sum(inf["split_exec_flops"]))
    peak = 2500.0 if mode == "bf16" else 157.3
    return {"split_mode": mode, "workload": w["name"], "hpsi_ms": ms, "executed_slice_gflop": fl / 1e9,
            "slice_tflops": fl / (ms * 1e-3) / 1e12, "peak_tflops": peak,
            "frac_of_peak": fl / (ms * 1e-3) / 1e12 / peak, "useful_gflop": H.flops / 1e9, "calibration": calib}


def eig_profile(workload):
    """only the eigensolver of an untuned plan (for profiler runs of its
    loop): Lanczos 20 on the Heisenberg workloads, Davidson 30 on the Hubbard
    one; first call and best of three further calls"""
    import qspace_amd as Q
    w = WORKLOADS[workload]
    HL, psi, HR, ferm = make_workload(w)
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    solver, niter = (H.davidson, 30) if w["kind"] == "hubbard" else (H.lanczos, 20)
    first, best, e0 = _time_eig(solver, psi.packed(), niter)
    return {"eig_profile": w["name"], "solver": solver.__name__, "niter": niter, "first_call_ms": first,
            "ms": best, "E0": e0, "psi_elements": H.N, "policy": "untuned plan (default decomposition)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="heis4000", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lite", action="store_true",
                    help="timed H*psi only (no CPU baseline, plain path, Lanczos): for profiler runs")
    ap.add_argument("--block-ops", action="store_true",
                    help="only the block operations (permuteQS / plusQS, environment update): one JSON line")
    ap.add_argument("--policy", default=None,
                    help="JSON file {stage1: {...}, stage2: {...}} fixing the work decomposition (no autotune)")
    ap.add_argument("--save-policy", default=None, help="write the autotuned policy to this JSON file")
    ap.add_argument("--split-profile", default=None, choices=["bf16", "f32"],
                    help="only the split-precision H*psi (for profiler runs): one JSON line")
    ap.add_argument("--eig-profile", action="store_true",
                    help="only the eigensolver of an untuned plan of --workload (for profiler runs): one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding of the output blocks is a design note, not built.", file=sys.stderr)
        sys.exit(2)
    import ctypes
    import qspace_amd as Q
    err = ctypes.create_string_buffer(256)
    rc = Q._capi.lib().qsx_device_init(0, err, 256)
    if rc:
        print(f"bench.py: no GPU: {err.value.decode()}", file=sys.stderr)
        sys.exit(1)
    if args.split_profile:
        print(json.dumps(split_profile(args.split_profile, args.steps, args.warmup)), flush=True)
        return
    if args.eig_profile:
        print(json.dumps(eig_profile(args.workload)), flush=True)
        return
    if args.block_ops:
        bops = block_ops()
        bops["u1_env_update"] = env_update()
        bops["su2_env_update"] = env_update_su2()
        bops["truncation"] = truncation()
        print(json.dumps({"block_ops": bops}), flush=True)
        return
    w = WORKLOADS[args.workload]
    log(f"[bench] generating {w['name']}")
    HL, psi, HR, ferm = make_workload(w)
    t0 = time.perf_counter()
    H = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    plan_s = time.perf_counter() - t0
    # a second build of the same plan: the X-symbol cache and the device code
    # are warm (the state of every later site of a sweep)
    t0 = time.perf_counter()
    H2 = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm)
    plan_warm_s = time.perf_counter() - t0
    del H2
    t0 = time.perf_counter()
    if args.policy:
        with open(args.policy) as f:
            pj = json.load(f)
        H.set_policy(pj["stage1"], pj["stage2"])
        tune = {"stage_ms": (-1.0, -1.0), "policy": (pj["stage1"], pj["stage2"])}
    else:
        tune = H.autotune(5)
    tune_s = time.perf_counter() - t0
    log(f"[bench] autotune {tune_s:.2f} s: {tune}")
    if args.save_policy:
        with open(args.save_policy, "w") as f:
            json.dump({"stage1": tune["policy"][0], "stage2": tune["policy"][1]}, f)
    info = H.stage_info()
    log(f"[bench] plan {plan_s:.2f} s; N={H.N} flops/H*psi={H.flops / 1e9:.3f} GFLOP pairs={info['pairs']} "
        f"items={info['tiles']} kernels={info['kernel']} grid={info['grid']} reduce={info['reduce_grid']}")
    calib = calibration_copy()
    rng = np.random.RandomState(7)
    x = rng.standard_normal(H.N)
    dpsi = Q.DeviceBuffer.from_array(x)
    dout = Q.DeviceBuffer(H.N)
    ms = time_hpsi(H, dpsi, dout, args.steps, args.warmup, graph=True)
    log(f"[bench] batched+graph H*psi {ms:.4f} ms -> {H.flops / ms / 1e6:.1f} GFLOP/s")
    ms_nograph = time_hpsi(H, dpsi, dout, max(5, args.steps // 5), 2, graph=False)
    kt_ev = stage_timing(H, dpsi, dout, 20)
    # kernel durations: 200 back-to-back launches of one stage between two
    # events on its stream (dispatch overlapped with the previous launch)
    kt = [H.stage_ms(1, 200, dpsi, dout), H.stage_ms(2, 200, dpsi, dout)]
    t1, t2 = kt[0], kt[1]
    log(f"[bench] stage kernels {info['kernel'][0]} {kt[0]:.4f} ms, {info['kernel'][1]} {kt[1]:.4f} ms "
        f"(per-launch events {kt_ev[0]:.4f}, {kt_ev[1]:.4f} ms)")
    ms_plain = err_plain = herm = lan_ms = lan_first = e0 = None
    if not args.lite:
        # plain FP64 per-pair GPU path (one launch per pair, no graph)
        Hp = Q.HPsi(HL, psi, HR, hconj=1, fermA=ferm, mode="plain")
        ms_plain = time_hpsi(Hp, dpsi, dout, 3, 1, graph=False)
        hp = dout.download()
        H.apply_dev(dpsi, dout)
        hb = dout.download()
        err_plain = float(np.linalg.norm(hb - hp) / np.linalg.norm(hp))
        log(f"[bench] plain per-pair FP64 H*psi {ms_plain:.3f} ms (batched/plain rel diff {err_plain:.2e})")
        del Hp
        # Hermiticity in the full-space inner product (size-independent check)
        wts = H.weights()
        y = rng.standard_normal(H.N)
        hx = H.apply(x)
        hy = H.apply(y)
        herm = float(abs((wts * x) @ hy - (wts * hx) @ y) / (np.sqrt((wts * hx) @ hx) * np.sqrt((wts * y) @ y)))
        # Lanczos, 20 iterations in one graph (second call: first-call setup excluded)
        lan_first, lan_ms, e0 = _time_eig(H.lanczos, psi.packed(), 20)
        log(f"[bench] lanczos 20 it: {lan_ms:.2f} ms (first call {lan_first:.2f} ms), E0={e0:.10f}")
    # dominant kernel roofline
    dom = 0 if t1 >= t2 else 1
    tk = (t1, t2)[dom]
    fl = info["flops"][dom]
    by = info["bytes"][dom]
    ai = fl / by
    ridge = PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"traffic_{w['name']}.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        traffic = tj.get(("stage1", "stage2")[dom])
    if ai >= ridge:
        roof = {"bound": "mfma", "achieved": fl / (tk * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / (tk * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = traffic
    roof["kernel"] = f"{info['kernel'][dom]} (H*psi stage {dom + 1})"
    roof["kernel_grid"] = info["grid"][dom]
    roof["reduce_grid"] = info["reduce_grid"][dom]
    roof["kernel_ms"] = {"stage1": kt[0], "stage2": kt[1]}
    roof["kernel_ms_timing"] = "mean of 200 back-to-back launches of the stage between two HIP events"
    roof["kernel_ms_per_launch_events"] = {"stage1": kt_ev[0], "stage2": kt_ev[1]}
    roof["algorithmic_flops"] = fl
    roof["algorithmic_bytes"] = by
    roof["avg_launch_ms"] = tk
    roof["fp64_mfma_frac"] = fl / (tk * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
    roof["hbm_frac"] = by / (tk * 1e-3) / 1e9 / PEAK_HBM_GBS
    # FP64 MFMAs the kernel issues (tile and k padding included)
    ex = info.get("mfma_exec_flops", (0.0, 0.0))[dom]
    if ex:
        roof["mfma_issued_flops"] = ex
        roof["mfma_issued_frac"] = ex / (tk * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
    configs = bops = None
    if not args.lite:
        configs = extra_configs()
        configs["pseudo_mpo_heis4000"] = pseudo_mpo()
        bops = block_ops()
        bops["u1_env_update"] = env_update()
        bops["su2_env_update"] = env_update_su2()
        bops["truncation"] = truncation()
    cpu = None
    if not args.no_cpu and not args.lite:
        log("[bench] cpu baseline (sample) ...")
        cpu = cpu_baseline(w)
        log(f"[bench] cpu {cpu['value']:.2f} GFLOP/s on {cpu['cores']} cores")
    value = H.flops / ms / 1e6
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal sector dims, N(0,1) block entries)",
        "config": {"workload": w["name"], "Dstar": w["Dstar"], "psi_elements": H.N, "hconj": 1,
                   "useful_gflop_per_hpsi": H.flops / 1e9, "pairs": list(info["pairs"]),
                   "items": list(info["tiles"]), "kernels": list(info["kernel"]), "grid": list(info["grid"]),
                   "stage_flops": list(info["flops"]), "stage_bytes": list(info["bytes"]),
                   "path": "batched FP64 MFMA (wave-tile or workgroup-tile engine, autotuned) + hipGraph"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "hpsi_ms": ms,
        "hpsi_ms_no_graph": ms_nograph,
        "plain_fp64_per_pair_ms": ms_plain,
        "speedup_vs_plain_fp64": ms_plain / ms if ms_plain else None,
        "rel_diff_batched_vs_plain": err_plain,
        "hermiticity_err_full_size": herm,
        "lanczos20_ms": lan_ms,
        "lanczos20_first_call_ms": lan_first,
        "lanczos20_E0": e0,
        "plan_create_s": plan_s,
        "plan_create_warm_s": plan_warm_s,
        "autotune": {"seconds": tune_s, **{k: list(v) for k, v in tune.items()}},
        "calibration": calib,
        "configs": configs,
        "block_ops": bops,
    }
    if cpu is not None:
        out["rel_err_vs_fp64_oracle"] = cpu["rel_err_gpu_vs_fp64_oracle_sample"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
