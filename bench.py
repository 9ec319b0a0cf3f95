"""Benchmark: two-vertex H_eff matvec on one MI355X (ChainAMPS hot path).

Workload: the d=64 Z2 coupled-chain effective Hamiltonian (levels E_n
sorted uniform[0,8], parity n mod 2, Gaussian inter-chain operator between
opposite parities, J_perp = 0.1, MPO bond dim 3) with environments of bond
dimension chi split into two parity sectors (default chi = 500, the largest
chi of the d=64 configurations whose FP64 matvec, Lanczos basis and
environments fit one 288 GB device; chi = 2000 needs 131 GB per FP64
vector).  Environments and psi are complex normal on every QN-allowed entry.

A "step" is one H_eff . psi matvec (TwoVertexMPOMPSMultiply,
DMRG_routines.cpp:898-974) on a device-resident vector in block layout.
`value` is the FP64 matvec throughput in TFLOP/s of algorithmic flops
(sum of 8*m*n*k over the QN-allowed block triples of the four-stage order
L, W, W, R).  The same line carries the FP32 split-complex throughput, the
roofline of the dominant kernel (HIP events on the launching stream) and
the CPU baseline (the oracle's reference-order matvec, timed on a bounded
sample on the host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 500]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

FP64_MFMA_PEAK = 78.6      # TFLOP/s, MI355X dense FP64 matrix
FP32_MFMA_PEAK = 157.3     # TFLOP/s, MI355X dense FP32 matrix
HERE = os.path.dirname(os.path.abspath(__file__))
# HBM traffic of the dominant stage kernel, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE and calibrated on a plain device copy; the summary
# (with the commands that produced it) is committed next to the kernel trace
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "heff_matvec_chi500_d64.md")


def committed_traffic(chi, d):
    """Bytes per launch of stage L from the committed profile summary (the
    line `traffic_bytes_per_launch: <value>`), or None for another workload
    or when the summary is absent."""
    if (chi, d) != (500, 64) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    import re
    with open(TRAFFIC_PROFILE) as f:
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def copy_calibration(gib=1.0, reps=3):
    """A plain device-to-device copy of `gib` GiB, `reps` times: run under
    rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE), the counters per dispatch
    against the known byte count give the calibration factors."""
    import torch
    n = int(gib * (1 << 30))
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "reps": reps, "read_bytes_per_copy": n, "write_bytes_per_copy": n}))


def cpu_baseline(budget_s=20.0):
    """Oracle restatement of the reference-order matvec (LeftPart . V, then
    . RightPart via sparse products; DMRG_routines.cpp:779-974) on a reduced
    chi of the same d=64 model, single host thread.  Reported as
    algorithmic (optimal-order) TFLOP/s of the sample."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import models, dmrg
    rules, phys, W, E, sig = models.coupled_chain_model(d=64, seed=0, Jperp=0.1)
    rng = np.random.default_rng(1)
    chi = 24
    aq = (np.arange(chi) % 2).reshape(-1, 1)
    bq = W.indices[1].charges
    LB = models.random_left_block(rules, aq, bq, rng)
    RB = models.random_right_block(rules, aq, bq, rng)
    comp = dmrg.TwoVertexComponents(LB, W, RB)
    v = rng.standard_normal(comp.n) + 1j * rng.standard_normal(comp.n)
    comp.matvec(v)                         # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        comp.matvec(v)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= 50:
            break
    flops = optimal_order_flops_dense(LB, W, RB, comp)
    return dict(value=flops * reps / el / 1e12, unit="TFLOP/s", cores=1, kind="port",
                sample=f"oracle reference-order matvec (scipy CSC SpGEMM), d=64 Z2 coupled chain, "
                       f"chi={chi}, n={comp.n}, {reps} matvecs in {el:.1f}s; flops counted in the "
                       f"optimal four-stage order ({flops:.3e}/matvec)")


def optimal_order_flops_dense(LB, W, RB, comp):
    """Sum of 8*m*n*k over nonzero QN block triples of the four stages,
    counted from the dense tensors (independent of the device planner)."""
    L = LB.to_dense_tensor()[0, :, 0, :, 0, :]
    R = RB.to_dense_tensor()[:, 0, :, 0, :, 0]
    Wd = W.to_dense_tensor()
    aq = LB.indices[5].charges[:, 0]
    a2q = RB.indices[4].charges[:, 0]
    sq = W.indices[2].charges[:, 0]
    mod = W.rules.rules[0] if W.rules.rules[0] > 0 else None
    add = (lambda a, b: (a + b) % mod) if mod else (lambda a, b: a + b)
    secs = lambda q: {c: int(np.sum(q == c)) for c in np.unique(q)}
    A, A2, S = secs(aq), secs(a2q), secs(sq)
    w = Wd.shape[1]
    nzL = {(b, c0): c0p for b in range(w) for c0 in A for c0p in A
           if np.any(L[np.ix_(aq == c0p, [b], aq == c0)])}
    nzR = {(b, c3): c3p for b in range(w) for c3 in A2 for c3p in A2
           if np.any(R[np.ix_(a2q == c3p, [b], a2q == c3)])}
    nzW = {(bl, br, sp, s_) for bl in range(w) for br in range(w) for sp in S for s_ in S
           if np.any(Wd[np.ix_(sq == sp, [bl], sq == s_, [br])])}
    flops = 0.0
    psi = [(s1, s0, s2, add(add(s1, s0), s2)) for s1 in S for s0 in A for s2 in S if add(add(s1, s0), s2) in A2]
    T1 = set()
    for (s1, s0, s2, s3) in psi:
        for b in range(w):
            if (b, s0) in nzL:
                s0p = nzL[(b, s0)]
                flops += 8 * A[s0p] * A[s0] * A2[s3] * S[s1] * S[s2]
                T1.add((b, s0p, s3, s1, s2))
    T2 = set()
    for (b, s0p, s3, s1, s2) in T1:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s1:
                flops += 8 * A[s0p] * A2[s3] * S[sp] * S[s1] * S[s2]
                T2.add((br, s0p, s3, sp, s2))
    T3 = set()
    for (b, s0p, s3, s1p, s2) in T2:
        for (bl, br, sp, s_) in nzW:
            if bl == b and s_ == s2:
                flops += 8 * A[s0p] * A2[s3] * S[s1p] * S[sp] * S[s2]
                T3.add((br, s0p, s3, s1p, sp))
    for (b, s0p, s3, s1p, s2p) in T3:
        if (b, s3) in nzR:
            flops += 8 * A[s0p] * A2[nzR[(b, s3)]] * A2[s3] * S[s1p] * S[s2p]
    return flops


def orthogonalise_step(dm, W):
    """The iDMRG driver's measurement after a growth step (iDMRG_DRV.cpp:
    64-82): the central unit cell, its left transfer-matrix fixed point alone
    (matvec rate), then Orthogonalise and iTwoVertexEnergy."""
    cell = dm.unitcell()
    t0 = time.perf_counter()
    val, tinfo = cell.transfer_leading(0)
    t_lead = time.perf_counter() - t0
    t0 = time.perf_counter()
    ocell, info = cell.orthogonalise()
    t_ortho = time.perf_counter() - t0
    t0 = time.perf_counter()
    e, _ = ocell.energy(W)
    t_energy = time.perf_counter() - t0
    lam = ocell.lam(0)
    out = {"left_value": round(val.real, 12), "left_matvecs": tinfo["matvecs"], "left_restarts": tinfo["restarts"],
           "left_seconds": round(t_lead, 4), "ms_per_transfer_matvec": round(1e3 * t_lead / max(tinfo["matvecs"], 1), 3),
           "orthogonalise_seconds": round(t_ortho, 4), "energy_seconds": round(t_energy, 4),
           "matvecs": [int(info["left_matvecs"]), int(info["right_matvecs"])], "new_chi": int(info["chi"]),
           "entropy": round(info["entropy"], 10), "energy_per_vertex": round(e.real, 12),
           "boundary_lambda_max": round(float(lam.max()), 12)}
    print(f"# orthogonalise chi={out['new_chi']} {t_ortho:.2f} s, energy {t_energy:.2f} s, "
          f"transfer matvec {out['ms_per_transfer_matvec']} ms", file=sys.stderr, flush=True)
    ocell.close()
    cell.close()
    return out


def idmrg_steps(ctx, cam, workloads, a):
    """Times cam.IDMRG growth steps on the d-level coupled-chain MPO with
    bond dimension cap chi, FP64 and FP32-mixed; the last step is the one at
    full size (n = chi^2 d^2 / 2 for two parity sectors)."""
    W, _, _, _ = workloads.coupled_chain_mpo(ctx, d=a.d, seed=0, Jperp=0.1)
    out = {}
    for mixed, name in ((False, "fp64"), (True, "mixed")):
        dm = cam.IDMRG(W, converge=False, mixed=mixed)
        steps = []
        for _ in range(a.idmrg_steps):
            r = dm.grow(a.chi)
            steps.append({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                          if k in ("N", "chi", "n", "energy", "truncation", "seconds", "eig_seconds",
                                   "svd_seconds", "block_seconds", "matvecs_fp64", "matvecs_fp32")})
            print(f"# idmrg {name} step N={r['N']} chi={r['chi']} n={r['n']} {r['seconds']:.2f} s", file=sys.stderr,
                  flush=True)
        if not mixed:
            out["orthogonalise"] = orthogonalise_step(dm, W)
        dm.close()
        out[name] = steps
    out["last_step_seconds"] = {k: out[k][-1]["seconds"] for k in ("fp64", "mixed")}
    out["speedup_mixed_vs_fp64"] = round(out["fp64"][-1]["seconds"] / out["mixed"][-1]["seconds"], 3)
    out["note"] = ("growth steps with the reference's growth-step eigensolver setting (tol 0.1, ncv 20, "
                   "predicted start vector); seconds = eigensolve + SVD + block updates")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chi", type=int, default=500)
    ap.add_argument("--d", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eig", action="store_true", help="skip the growth-step eigensolve timing")
    ap.add_argument("--no-idmrg", action="store_true", help="skip the iDMRG step timing")
    ap.add_argument("--idmrg-steps", type=int, default=4,
                    help="growth steps per precision (the bond reaches chi at step 3 for d=64)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy used to calibrate the HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "sharding of this path is a design, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return

    import torch
    import chainamps_amd as cam
    from chainamps_amd import workloads

    ctx = cam.Context(0)
    t0 = time.perf_counter()
    L, W, R = workloads.coupled_chain_heff_inputs(ctx, chi=a.chi, d=a.d, seed=0, Jperp=0.1, hermitian=True)
    H = cam.Heff(L, W, R)
    setup_s = time.perf_counter() - t0
    n = H.n
    flops = H.flops
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.randn(n, dtype=torch.complex128, device=dev, generator=g)
    y = torch.empty_like(x)
    torch.cuda.synchronize()

    # FP32 vectors (split complex) for the FP32 matvec as the mixed solver
    # runs it; the FP64-vector variant (conversion in and out) is kept too
    x32 = x.to(torch.complex64)
    y32 = torch.empty_like(x32)

    def run(prec, k):
        for _ in range(k):
            if prec == "fp32":
                H.apply_device32(x32.data_ptr(), y32.data_ptr())
            else:
                H.apply_device(x.data_ptr(), y.data_ptr(), prec)
        ctx.sync()

    res = {}
    for prec, name in ((cam.CAM_FP64, "fp64"), ("fp32", "fp32"), (cam.CAM_FP32, "fp32_fp64_vectors")):
        run(prec, a.warmup)
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(prec, a.steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        res[name] = dict(ms=el / a.steps * 1e3, tflops=flops * a.steps / el / 1e12)
    del x32, y32
    # roofline of the dominant kernel, from HIP events on the context stream
    ms64, fl = H.profile(cam.CAM_FP64, reps=max(3, min(a.steps, 10)))
    ms32, _ = H.profile(cam.CAM_FP32, reps=max(3, min(a.steps, 10)))
    dom = int(np.argmax(ms64))
    achieved = fl[dom] / (ms64[dom] * 1e-3) / 1e12
    stage_names = ["L", "W1", "W2", "R"]
    stage_bytes = H.stage_bytes()
    # eigensolve of one iDMRG growth step (the reference's setting for it:
    # converge=0, i.e. tol 0.1 and one Krylov cycle of ncv=20 vectors,
    # arpack_interface.cpp:28-39) from the same start vector, FP64 against
    # FP32-mixed (FP64 first step and reorthogonalisation, FP32 matvecs)
    eig = {}
    free, _ = torch.cuda.mem_get_info()
    need = 23 * 16 * n          # Krylov basis (ncv+1 vectors) and work vector, FP64
    if free < 1.05 * need:
        eig = {"skipped": f"needs {need / 1e9:.0f} GB of free device memory, {free / 1e9:.0f} GB free"}
    elif not a.no_eig:
        for mixed, name in ((False, "fp64"), (True, "mixed")):
            y.copy_(x)
            torch.cuda.synchronize()
            t = time.perf_counter()
            E, info = H.lowest_device(y.data_ptr(), have_start=True, converge=False, ncv=20, mixed=mixed)
            el = time.perf_counter() - t
            eig[name] = dict(seconds=round(el, 3), energy=E, matvecs_fp64=info["matvecs_fp64"],
                             matvecs_fp32=info["matvecs_fp32"], residual=info["residual"])
        eig["speedup_mixed_vs_fp64"] = round(eig["fp64"]["seconds"] / eig["mixed"]["seconds"], 3)
        eig["note"] = ("Lanczos eigensolve with the growth-step setting; the SVD and the block updates of "
                       "the step are not included")
    launch = {"fp64": H.get_launch(cam.CAM_FP64), "fp32": H.get_launch(cam.CAM_FP32)}
    # whole iDMRG growth steps (SuperBlock::grow_two_vertex, DMRG_routines.cpp:309-374)
    # of the same model MPO: block updates, H_eff build, eigensolve from the
    # predicted guess (growth setting), SVD truncation to chi and prediction,
    # FP64 against FP32-mixed; the matvec buffers above are released first
    del x, y
    H.close()
    torch.cuda.empty_cache()
    idmrg = None if a.no_idmrg else idmrg_steps(ctx, cam, workloads, a)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_budget)
    line = {
        "metric": "H_eff matvec TFLOP/s + iDMRG step time, chi=2000 d=64, FP32-mixed vs FP64",
        "value": round(res["fp64"]["tflops"], 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(res["fp64"]["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 (complex128)",
        "data": "synthetic (seeded QN-allowed complex normal environments and psi; model MPO)",
        "config": {"workload": f"two-vertex H_eff matvec, d={a.d} Z2 coupled chain, w=3, "
                               f"chi={a.chi} (2 parity sectors)",
                   "chi": a.chi, "d": a.d, "mpo_bond": 3, "n": n, "flops_per_matvec": flops,
                   "parallelism": "single GPU"},
        "fp32_split": {"tflops": round(res["fp32"]["tflops"], 4), "ms_per_step": round(res["fp32"]["ms"], 4),
                       "speedup_vs_fp64": round(res["fp64"]["ms"] / res["fp32"]["ms"], 3),
                       "vectors": "FP32 split complex in the block layout, as the mixed-precision eigensolver "
                                  "keeps its Krylov vectors",
                       "ms_with_fp64_vectors": round(res["fp32_fp64_vectors"]["ms"], 4)},
        "stages": {stage_names[i]: {"fp64_ms": round(ms64[i], 4), "fp32_ms": round(ms32[i], 4),
                                    "gflop": round(fl[i] / 1e9, 3)} for i in range(4)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_MFMA_PEAK,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK, 4),
                     "traffic": committed_traffic(a.chi, a.d) if stage_names[dom] == "L" else None,
                     "traffic_source": "profiles/heff_matvec_chi500_d64.md (stage L, rocprofv3 --pmc)",
                     "algorithmic_bytes": stage_bytes[dom],
                     "kernel": f"grouped_gemm_kernel<double> stage {stage_names[dom]}"},
        "cpu_baseline": cpu,
        "growth_step_eigensolve": eig or None,
        "idmrg_step": idmrg,
        "autotuned_launch": {"fields": "(algo 0=4M/1=3M, tile_mode, wg_per_cu, kchunk) per stage L, W1, W2, R",
                             **launch},
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
