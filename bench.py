#!/usr/bin/env python3
"""Throughput of the MI355X Biot-Savart FMM: B-field points/s on one GPU.

Headline workload (`value`): 4,194,304 sources in 4 coaxial cylindrical shells (radii 0.5 / 0.8 /
1.1 / 1.4, thickness 0.1, |z| <= 0.5, uniform in volume, one random current sign per shell,
J = s_k phi-hat with w = s_k / N), 4,194,304 targets uniform in [-2, 2]^3, expansion order
p = 16, leaf capacity 64, shared-frame trees, reference MAC.  One step = one full evaluation
from device-resident inputs to device-resident B in input order: both octrees, interaction
lists, FP64 operator tables, P2M, M2M, M2L, L2L, L2P and near-field P2P.

The mixed-precision run gives `value` (F16X3: the M2L on FP16 split products, hi + lo operands
equilibrated by powers of 2, FP32 accumulate; M2M and L2L on FP32 MFMA); the plain FP64 FMM
(FP64 MFMA translations, FP64 P2P) is timed on the same inputs (`speedup_vs_fp64`).  rel-L2 of
each is measured against an FP64 direct sum on the first 16,384 targets.  A second workload
(1,048,576 shell sources r in [0.8, 1], 1,048,576 uniform targets, p = 12, leaf capacity 64) is
reported under `secondary`.  The CPU baseline is the oracle's restatement of the reference CPU
FMM path on the reference's own case (32,768 shell sources, 4,096 uniform targets, p = 10, leaf
capacity 32), timed on all host cores and on one.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f16x3|fp32|bf16x3]
                       [--workload shells4m|shell1m] [--no-secondary] [--no-fp64] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "FMM B-field points/s at rel-L2 ≤1e-6 vs FP64 oracle, mixed vs plain FP64"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0,   # dense MFMA; bf16x3 = 3 bf16 GEMMs,
               "f16x3": 2500.0 / 3.0}                                 # f16x3 = 3 fp16 GEMMs
HBM_GBS = 8000.0
TRANSLATIONS = {"f16x3": "M2L on FP16 split MFMA products (F16X3), M2M/L2L on FP32 MFMA",
                "fp32": "FP32 MFMA translations", "bf16x3": "BF16 split MFMA translations (BF16X3)",
                "fp64": "FP64 MFMA translations"}
SEED = 20261202


def shell_sources(rng, n, r0=0.8, r1=1.0, zh=0.5):
    u = rng.random((n, 3))
    r = np.sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * u[:, 0])
    ph = 2 * np.pi * u[:, 1]
    z = (2 * u[:, 2] - 1) * zh
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def coaxial_shells(rng, n_per, radii=(0.5, 0.8, 1.1, 1.4), thickness=0.1, zh=0.5):
    signs = rng.choice([-1.0, 1.0], size=len(radii))
    n = n_per * len(radii)
    P, W = [], []
    for R, s in zip(radii, signs):
        P.append(shell_sources(rng, n_per, R - thickness / 2, R + thickness / 2, zh))
        W.append(np.full(n_per, s / n))
    return np.concatenate(P), np.concatenate(W)


WORKLOADS = {
    # name: (description, sources, targets, p, leaf capacity, seed offset)
    "shells4m": ("4 coaxial shells (r = 0.5/0.8/1.1/1.4 +- 0.05, |z|<=0.5, random sign per shell, J=s phi-hat, "
                 "w=s/N)", 1 << 22, 1 << 22, 16, 64, 3),
    "shell1m": ("cylindrical shell (r 0.8-1.0, |z|<=0.5, J=phi-hat, w=1/N)", 1 << 20, 1 << 20, 12, 64, 2),
}


def make_workload(name, n_src=None, n_tgt=None):
    desc, ns, nt, p, cap, off = WORKLOADS[name]
    ns = n_src or ns
    nt = n_tgt or nt
    rng = np.random.default_rng(SEED + off)
    if name == "shells4m":
        src, w = coaxial_shells(rng, ns // 4)
    else:
        src = shell_sources(rng, ns)
        w = np.full(ns, 1.0 / ns)
    tgt = rng.uniform(-2.0, 2.0, (nt, 3))
    return desc, src, w, tgt, p, cap


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def time_fmm(fm, torch, prec, d_src, d_w, d_tgt, p, cap, steps, warmup, subset, B_ref, staged_runs=3):
    F = fm.FMM(order=p, source_leaf_capacity=cap, target_leaf_capacity=cap, precision=prec, frame=fm.FRAME_SHARED)
    F.set_sources_dev(d_src.shape[0], vp(d_src), vp(d_w))
    F.set_targets_dev(d_tgt.shape[0], vp(d_tgt))
    B = torch.empty_like(d_tgt)
    for _ in range(warmup):
        F.run_B_dev(1.0, vp(B))
    F.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        F.run_B_dev(1.0, vp(B))
    F.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    # per-stage HIP-event times of the last timed evaluation (the library records them on its
    # streams in every run; the lists and the near field overlap other stages there)
    st = F.stats()
    Bs = B[:subset].cpu().numpy()
    err = float(np.linalg.norm(Bs - B_ref) / np.linalg.norm(B_ref))
    finite = bool(torch.isfinite(B).all().item())
    # untimed stage-by-stage evaluations (no overlap) for the stage costs in isolation: the
    # minimum of each stage over the runs
    runs = []
    for _ in range(staged_runs):
        F.build_trees()
        F.leaf_expansions()
        F.upward()
        F.downward()
        F.evaluate_B_dev(1.0, vp(B))
        F.synchronize()
        runs.append(F.stats())
    staged = {k: (min(r[k] for r in runs) if k.startswith("t_") else runs[-1][k]) for k in runs[-1]} if runs else {}
    F.close()
    return ms, st, staged, err, finite


def m2l_roofline(st, L, prec, traffic):
    """The dominant kernel, the M2L class GEMM, on the MFMA roofline: algorithmic flops 2 L^2 x 3
    components per far pair over the launch's duration (HIP events around its single launch)."""
    flops = 2.0 * L * L * 3 * st["far_pairs"]
    s = st["t_m2l_gemm"] * 1e-3
    achieved = flops / s / 1e12 if s > 0 else 0.0
    peak = PEAK_TFLOPS[prec]
    # unique operands of the launch: each source multipole, each class operator and each target
    # local once (FP32 coefficients; FP16 hi + lo operator tables)
    uniq = (st["source_nodes"] + st["target_nodes"]) * 3 * L * 4 + st["m2l_classes"] * L * L * 4
    return {"bound": "mfma", "kernel": "M2L class-batched GEMM (%s)" % prec, "achieved": achieved, "peak": peak,
            "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
            "flops_per_launch": flops, "launch_ms": st["t_m2l_gemm"],
            "hbm_unique_operands": {"bytes": uniq, "achieved_GBs": uniq / s / 1e9 if s > 0 else 0.0,
                                    "peak": HBM_GBS}}


def other_stage_rooflines(st, L, n_src, n_tgt, prec):
    """Achieved rates of the other stages from the stage-by-stage (no overlap) event times, with the
    algorithmic units of each: P2M and L2P bytes (positions, strengths, FP32 coefficients, FP64 B),
    P2P pair interactions at 20 flop, the M2L slab reduction's bytes (one FP32 slab per far pair
    read, the local expansions read and written)."""
    Lp = (L + 63) // 64 * 64
    b_coef = 4 if prec != "fp64" else 8
    rows = {}
    p2m_bytes = n_src * 56 + st["source_leaves"] * 3 * L * b_coef
    rows["p2m"] = {"ms": st["t_p2m"], "bytes": p2m_bytes, "bound": "hbm",
                   "achieved_GBs": p2m_bytes / (st["t_p2m"] * 1e-3) / 1e9 if st["t_p2m"] > 0 else 0.0}
    l2p_bytes = n_tgt * 24 + st["target_leaves"] * 3 * L * b_coef + n_tgt * 24
    rows["l2p"] = {"ms": st["t_l2p"], "bytes": l2p_bytes, "bound": "hbm",
                   "achieved_GBs": l2p_bytes / (st["t_l2p"] * 1e-3) / 1e9 if st["t_l2p"] > 0 else 0.0}
    flops = 20.0 * st["near_interactions"]
    rows["p2p"] = {"ms": st["t_p2p"], "interactions": st["near_interactions"], "bound": "valu_fp32",
                   "achieved_TFs_at_20_flop": flops / (st["t_p2p"] * 1e-3) / 1e12 if st["t_p2p"] > 0 else 0.0,
                   "peak_TFs": PEAK_TFLOPS["fp32"]}
    red_ms = st["t_m2l"] - st["t_m2l_gemm"]
    R = Lp if prec == "bf16x3" else min(Lp, (L + 15) // 16 * 16)
    red_bytes = st["far_pairs"] * 3 * R * b_coef + 2 * st["target_nodes"] * 3 * Lp * b_coef
    rows["m2l_reduction"] = {"ms": red_ms, "bytes": red_bytes, "bound": "hbm",
                             "achieved_GBs": red_bytes / (red_ms * 1e-3) / 1e9 if red_ms > 0 else 0.0}
    for r in rows.values():
        if r["bound"] == "hbm":
            r["frac"] = r["achieved_GBs"] / HBM_GBS
        else:
            r["frac"] = r["achieved_TFs_at_20_flop"] / r["peak_TFs"]
    return rows


def committed_traffic(workload, prec, p):
    """HBM bytes per M2L GEMM launch (calibrated FETCH_SIZE + WRITE_SIZE) from the committed
    rocprofv3 counter summary of this exact workload and precision; None when there is none."""
    path = os.path.join(HERE, "profiles", "m2l_gemm_hbm_traffic_%s_p%d_%s.md" % (workload, p, prec))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return {"bytes": float(line.split(":")[1]), "source": "committed rocprofv3 counters: profiles/" +
                    os.path.basename(path)}
    return None


def direct_reference(fm, torch, d_src, d_w, d_tgt, sub, cap):
    """FP64 direct Biot-Savart sum (integration.cpp:101-120) on the error subset, on the GPU."""
    Fd = fm.FMM(order=2, source_leaf_capacity=cap, target_leaf_capacity=cap)
    Fd.set_sources_dev(d_src.shape[0], vp(d_src), vp(d_w))
    Fd.set_targets_dev(sub, vp(d_tgt))
    Fd.build_trees()
    d_ref = torch.empty((sub, 3), dtype=torch.float64, device=d_tgt.device)
    Fd.direct_B_dev(1.0, sub, vp(d_tgt), vp(d_ref))
    out = d_ref.cpu().numpy()
    Fd.close()
    return out


def run_workload(fm, torch, name, args, with_fp64):
    desc, src, w, tgt, p, cap = make_workload(name)
    dev = torch.device("cuda:0")
    d_src = torch.from_numpy(src).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    d_tgt = torch.from_numpy(tgt).to(dev)
    ns, nt = len(src), len(tgt)
    sub = min(args.subset, nt)
    B_ref = direct_reference(fm, torch, d_src, d_w, d_tgt, sub, cap)
    print("bench: %s: %d sources, %d targets, p=%d: timing %s" % (name, ns, nt, p, args.precision),
          file=sys.stderr, flush=True)
    ms, st, staged, err, finite = time_fmm(fm, torch, args.precision, d_src, d_w, d_tgt, p, cap, args.steps,
                                           args.warmup, sub, B_ref)
    L = (p + 1) ** 2
    out = {
        "workload": "Biot-Savart FMM: %d sources, %s; %d uniform targets in [-2,2]^3; p=%d, leaf capacity %d, "
                    "shared frame, %s, FP64 tables/harmonics/accumulation"
                    % (ns, desc, nt, p, cap, TRANSLATIONS[args.precision]),
        "points_per_s": nt / (ms * 1e-3), "ms_per_step": ms,
        "rel_l2_vs_fp64_direct": err, "error_subset": sub, "finite": finite,
        "stage_ms": {k[2:]: round(v, 4) for k, v in st.items() if k.startswith("t_")},
        "stage_ms_staged": {k[2:]: round(v, 4) for k, v in staged.items() if k.startswith("t_")},
        "counts": {k: st[k] for k in ("source_nodes", "target_nodes", "far_pairs", "close_pairs", "m2l_classes",
                                      "near_interactions", "source_leaves", "target_leaves")},
        "roofline": m2l_roofline(st, L, args.precision, committed_traffic(name, args.precision, p)),
        "other_stages": other_stage_rooflines(staged, L, ns, nt, args.precision),
    }
    if staged.get("t_m2l_gemm", 0) > 0:
        # the same GEMM launch in the stage-by-stage runs, with nothing beside it (in the timed
        # fused evaluation the near field may share the chip with it)
        alone = m2l_roofline(staged, L, args.precision, None)
        out["roofline"]["alone"] = {"launch_ms": alone["launch_ms"], "achieved": alone["achieved"],
                                    "frac": alone["frac"]}
    if with_fp64:
        print("bench: %s: timing fp64" % name, file=sys.stderr, flush=True)
        ms64, st64, _, err64, fin64 = time_fmm(fm, torch, "fp64", d_src, d_w, d_tgt, p, cap, args.steps,
                                               args.warmup, sub, B_ref, staged_runs=0)
        out["fp64"] = {"points_per_s": nt / (ms64 * 1e-3), "ms_per_step": ms64, "rel_l2_vs_fp64_direct": err64,
                       "finite": fin64, "stage_ms": {k[2:]: round(v, 4) for k, v in st64.items() if k.startswith("t_")}}
        out["speedup_vs_fp64"] = ms64 / ms
    del d_src, d_w, d_tgt
    torch.cuda.empty_cache()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The oracle's restatement of the reference CPU FMM path (FMMGPU's Device::CPU path: P2M, M2M
    matrices, M2L complex per pair with the reference's l + lambda <= p truncation, complex L2L,
    finite-difference curl, treecode close list) on the reference's own case: 32,768 sources in
    the hollow cylinder, 4,096 uniform targets, p = 10, leaf capacity 32.  Timed on all the host
    cores this process may use (OpenMP over independent outputs: same results as one thread) and
    on one; the all-fixes restatement (near field, analytic curl, full M2L) is timed beside it."""
    from oracle import oracle as O
    rng = np.random.default_rng(SEED + 1)
    ns, nt = 32768, 4096
    src = shell_sources(rng, ns)
    w = np.full(ns, 1.0 / ns)
    tgt = rng.uniform(-2.0, 2.0, (nt, 3))
    nproc = len(os.sched_getaffinity(0))
    cores = min(nproc, int(os.environ.get("OMP_NUM_THREADS", nproc)))
    Bd = O.direct_B(src, w, tgt)

    def timed(flags, threads):
        O.set_threads(threads)
        t0 = time.perf_counter()
        R = O.FMMRun(src, w, tgt, p=10, cap_s=32, cap_t=32, flags=flags)
        dt = time.perf_counter() - t0
        B = R.B_input_order(nt)
        ok = np.isfinite(B[:, 0])
        return dt, float(np.linalg.norm(B[ok] - Bd[ok]) / np.linalg.norm(Bd[ok]))

    t_all, e_faithful = timed(O.FAITHFUL, cores)
    t_one, _ = timed(O.FAITHFUL, 1)
    t_fix, e_fixed = timed(O.CORRECTED, cores)
    O.set_threads(cores)
    return {"value": nt / t_all, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": "oracle restatement of the reference CPU FMM path (M2M matrices, M2L complex per pair with "
                      "l+lambda<=p, L2L complex, FD curl eps=1e-3, treecode close list) on the reference's own case: "
                      "32768 shell sources x 4096 uniform targets, p=10, leaf capacity 32, FP64; %.3f s on %d threads"
                      % (t_all, cores),
            "single_thread": {"value": nt / t_one, "seconds": t_one},
            "nproc": nproc, "cpu_model": cpu_model(),
            "rel_l2_vs_fp64_direct": e_faithful,
            "corrected": {"value": nt / t_fix, "seconds": t_fix, "rel_l2_vs_fp64_direct": e_fixed,
                          "what": "same restatement with every fix: direct near field, analytic curl, full 2p M2L"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="f16x3", choices=["f16x3", "fp32", "bf16x3"])
    ap.add_argument("--workload", default="shells4m", choices=sorted(WORKLOADS))
    ap.add_argument("--subset", type=int, default=16384)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="first copy 1 GiB device-to-device (the calibration kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (sharding by target Morton "
              "range is a design note); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import bachelors_amd as fm
    fm.lib()
    if args.traffic_probe:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    head = run_workload(fm, torch, args.workload, args, not args.no_fp64)
    out = {
        "metric": METRIC,
        "value": head["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded source and target clouds)",
        "config": {"workload": head["workload"]},
        "rel_l2_vs_fp64_direct": head["rel_l2_vs_fp64_direct"],
        "roofline": head["roofline"],
    }
    if "speedup_vs_fp64" in head:
        out["speedup_vs_fp64"] = head["speedup_vs_fp64"]
    out["detail"] = head
    if not args.no_secondary:
        other = "shell1m" if args.workload == "shells4m" else "shells4m"
        out["secondary"] = run_workload(fm, torch, other, args, not args.no_fp64)
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
