#!/usr/bin/env python3
"""Throughput of the MI355X Biot-Savart FMM: B-field points/s on one GPU.

Workload (the metric's configuration): 1,048,576 sources uniform in volume in the hollow
cylinder r in [0.8, 1.0], |z| <= 0.5, current along phi-hat with equal weights 1/N, and
1,048,576 targets uniform in [-2, 2]^3; expansion order p = 12, leaf capacity 64,
shared-frame trees, reference MAC.  One step = one full evaluation from device-resident
inputs to device-resident B in input order: both octrees, interaction lists, FP64
operator tables, P2M, M2M, M2L, L2L, L2P and near-field P2P.

The mixed-precision run gives `value`: by default F16X3 (the M2L on FP16 split products, hi + lo
operands equilibrated by powers of 2, FP32 accumulate; M2M and L2L on FP32 MFMA); the plain
FP64 FMM (FP64 MFMA translations, FP64 P2P) is timed on the same inputs.  rel-L2 of
each is measured against an FP64 direct sum on the first 16,384 targets.  The CPU
baseline is the oracle's restatement of the reference CPU FMM path, single thread, on
a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f16x3|fp32|bf16x3]
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


def shell_sources(rng, n, r0=0.8, r1=1.0, zh=0.5):
    u = rng.random((n, 3))
    r = np.sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * u[:, 0])
    ph = 2 * np.pi * u[:, 1]
    z = (2 * u[:, 2] - 1) * zh
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def workload(n_src, n_tgt, seed):
    rng = np.random.default_rng(seed)
    src = shell_sources(rng, n_src)
    w = np.full(n_src, 1.0 / n_src)
    tgt = rng.uniform(-2.0, 2.0, (n_tgt, 3))
    return src, w, tgt


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def time_fmm(fm, torch, prec, d_src, d_w, d_tgt, p, cap, steps, warmup, subset, B_ref):
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
    finite = bool(np.isfinite(B.cpu().numpy()).all())
    # untimed stage-by-stage evaluations (no overlap) for the stage costs in isolation: the
    # minimum of each stage over three runs
    runs = []
    for _ in range(3):
        F.build_trees()
        F.leaf_expansions()
        F.upward()
        F.downward()
        F.evaluate_B_dev(1.0, vp(B))
        F.synchronize()
        runs.append(F.stats())
    staged = {k: (min(r[k] for r in runs) if k.startswith("t_") else runs[-1][k]) for k in runs[-1]}
    F.close()
    return ms, st, staged, err, finite


def other_stage_rooflines(st, L, args):
    """Achieved rates of the other stages from the stage-by-stage (no overlap) event times, with the
    algorithmic units of each: P2M and L2P bytes (positions, strengths, FP32 coefficients, FP64 B),
    P2P pair interactions at 20 flop, the M2L slab reduction's bytes (one FP32 slab per far pair
    read, the local expansions read and written)."""
    Lp = (L + 63) // 64 * 64
    b_coef = 4 if args.precision != "fp64" else 8
    rows = {}
    p2m_bytes = args.sources * 56 + st["source_leaves"] * 3 * L * b_coef
    rows["p2m"] = {"ms": st["t_p2m"], "bytes": p2m_bytes, "bound": "hbm",
                   "achieved_GBs": p2m_bytes / (st["t_p2m"] * 1e-3) / 1e9 if st["t_p2m"] > 0 else 0.0}
    l2p_bytes = args.targets * 24 + st["target_leaves"] * 3 * L * b_coef + args.targets * 24
    rows["l2p"] = {"ms": st["t_l2p"], "bytes": l2p_bytes, "bound": "hbm",
                   "achieved_GBs": l2p_bytes / (st["t_l2p"] * 1e-3) / 1e9 if st["t_l2p"] > 0 else 0.0}
    flops = 20.0 * st["near_interactions"]
    rows["p2p"] = {"ms": st["t_p2p"], "interactions": st["near_interactions"], "bound": "valu_fp32",
                   "achieved_TFs_at_20_flop": flops / (st["t_p2p"] * 1e-3) / 1e12 if st["t_p2p"] > 0 else 0.0,
                   "peak_TFs": PEAK_TFLOPS["fp32"]}
    red_ms = st["t_m2l"] - st["t_m2l_gemm"]
    # slab rows: L padded to 16 for FP32/FP64 (the library's M2L kernel instances cover every
    # order whose 16-padded length is below Lp), Lp for BF16x3
    R = Lp if args.precision == "bf16x3" else min(Lp, (L + 15) // 16 * 16)
    red_bytes = st["far_pairs"] * 3 * R * b_coef + 2 * st["target_nodes"] * 3 * Lp * b_coef
    rows["m2l_reduction"] = {"ms": red_ms, "bytes": red_bytes, "bound": "hbm",
                             "achieved_GBs": red_bytes / (red_ms * 1e-3) / 1e9 if red_ms > 0 else 0.0}
    for r in rows.values():
        if r["bound"] == "hbm":
            r["frac"] = r["achieved_GBs"] / 8000.0
        else:
            r["frac"] = r["achieved_TFs_at_20_flop"] / r["peak_TFs"]
    return rows


def measured_traffic(args):
    """HBM bytes per M2L GEMM launch from the committed rocprofv3 counter summary of the default
    workload and this precision (FETCH_SIZE + WRITE_SIZE, calibrated by a 1 GiB copy); None for
    other workloads."""
    default = (args.sources, args.targets, args.order, args.cap) == (1 << 20, 1 << 20, 12, 64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "m2l_gemm_hbm_traffic_1M_shell_p12_%s.md" % args.precision)
    if not default or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(seed, n, p, cap):
    from oracle import oracle as O
    src, w, tgt = workload(n, n, seed + 100)
    t0 = time.perf_counter()
    R = O.FMMRun(src, w, tgt, p=p, cap_s=cap, cap_t=cap, flags=O.FACT_FIX)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": "oracle restatement of the reference CPU FMM path (M2M matrices, M2L complex per pair, "
                      "L2L complex, FD curl, treecode close list; factorial fix for p=%d), single thread, "
                      "%d shell sources x %d uniform targets, p=%d, leaf capacity %d: %.2f s" % (p, n, n, p, cap, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="f16x3", choices=["f16x3", "fp32", "bf16x3"])
    ap.add_argument("--sources", type=int, default=1 << 20)
    ap.add_argument("--targets", type=int, default=1 << 20)
    ap.add_argument("--order", type=int, default=12)
    ap.add_argument("--cap", type=int, default=64)
    ap.add_argument("--subset", type=int, default=16384)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=32768)
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
    seed = 20261202 + 2
    src, w, tgt = workload(args.sources, args.targets, seed)
    dev = torch.device("cuda:0")
    if args.traffic_probe:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    d_src = torch.from_numpy(src).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    d_tgt = torch.from_numpy(tgt).to(dev)
    sub = min(args.subset, args.targets)
    # FP64 direct Biot-Savart sum on the error subset (integration.cpp:101-120), on the GPU
    Fd = fm.FMM(order=2, source_leaf_capacity=args.cap, target_leaf_capacity=args.cap)
    Fd.set_sources_dev(d_src.shape[0], vp(d_src), vp(d_w))
    Fd.set_targets_dev(sub, vp(d_tgt))
    Fd.build_trees()
    d_ref = torch.empty((sub, 3), dtype=torch.float64, device=dev)
    Fd.direct_B_dev(1.0, sub, vp(d_tgt), vp(d_ref))
    B_ref = d_ref.cpu().numpy()
    Fd.close()

    ms, st, staged, err, finite = time_fmm(fm, torch, args.precision, d_src, d_w, d_tgt, args.order, args.cap,
                                           args.steps, args.warmup, sub, B_ref)
    L = (args.order + 1) ** 2
    m2l_flops = 2.0 * L * L * 3 * st["far_pairs"]
    m2l_s = st["t_m2l_gemm"] * 1e-3      # the translation GEMM alone (HIP events around its launch)
    achieved = m2l_flops / m2l_s / 1e12 if m2l_s > 0 else 0.0
    peak = PEAK_TFLOPS[args.precision]
    # algorithmic bytes of the launch: each far pair reads its source multipole and writes its
    # translated local once (3 x L FP32 values each way), plus one operator per class (L x L
    # entries of 4 bytes: FP32, or FP16 hi + lo)
    m2l_bytes = st["far_pairs"] * 2 * 3 * L * 4 + st["m2l_classes"] * L * L * 4
    gbs = m2l_bytes / m2l_s / 1e9 if m2l_s > 0 else 0.0
    if gbs / HBM_GBS > achieved / peak:   # the binding resource: HBM bytes or MFMA flops
        roof = {"bound": "hbm", "kernel": "M2L class-batched GEMM", "achieved": gbs, "peak": HBM_GBS, "unit": "GB/s",
                "frac": gbs / HBM_GBS, "traffic": measured_traffic(args),
                "mfma": {"achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak}}
    else:
        roof = {"bound": "mfma", "kernel": "M2L class-batched GEMM", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": measured_traffic(args),
                "hbm": {"achieved": gbs, "peak": HBM_GBS, "unit": "GB/s", "frac": gbs / HBM_GBS}}
    out = {
        "metric": METRIC,
        "value": args.targets / (ms * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded cylindrical-shell sources, uniform targets)",
        "config": {"workload": "Biot-Savart FMM: %d shell sources (r 0.8-1.0, |z|<=0.5, J=phi-hat, w=1/N), "
                               "%d uniform targets in [-2,2]^3, p=%d, leaf capacity %d, shared frame, %s, "
                               "FP64 tables/harmonics/accumulation"
                               % (args.sources, args.targets, args.order, args.cap, TRANSLATIONS[args.precision])},
        "rel_l2_vs_fp64_direct": err,
        "error_subset": sub,
        "finite": finite,
        "stage_ms": {k[2:]: round(v, 4) for k, v in st.items() if k.startswith("t_")},
        "stage_ms_staged": {k[2:]: round(v, 4) for k, v in staged.items() if k.startswith("t_")},
        "counts": {k: st[k] for k in ("source_nodes", "target_nodes", "far_pairs", "close_pairs", "m2l_classes")},
        "roofline": roof,
        "other_stages": other_stage_rooflines(staged, L, args),
    }
    if not args.no_fp64:
        ms64, st64, _, err64, fin64 = time_fmm(fm, torch, "fp64", d_src, d_w, d_tgt, args.order, args.cap, args.steps,
                                            args.warmup, sub, B_ref)
        out["fp64"] = {"points_per_s": args.targets / (ms64 * 1e-3), "ms_per_step": ms64,
                       "rel_l2_vs_fp64_direct": err64, "finite": fin64,
                       "stage_ms": {k[2:]: round(v, 4) for k, v in st64.items() if k.startswith("t_")}}
        out["speedup_vs_fp64"] = ms64 / ms
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(seed, args.cpu_sample, args.order, args.cap)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
