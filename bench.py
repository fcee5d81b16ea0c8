"""Benchmark of the ijkl AO->MO transform on one MI355X.

Workload (the headline configuration): N_AO = N_MO = 240, synthetic PSD AO
integrals (ij|kl) = sum_P B^P_ij B^P_kl with rank 720 and decay length 20, C the
Q factor of a seeded Gaussian 240 x 240 matrix.  A step is one whole transform,
packed AO block -> packed (pq|rs) MO stream, with the inputs resident in HBM.
The metric is MO integrals per second: P_M (P_M + 1) / 2 = 418,197,660 per step.

Also reported:
  roofline      the dominant kernel (half 1, Q1 + Q2 on FP64 MFMA) from HIP events on
                the library stream: algorithmic flops 2 (M N^2 + N P_M) per AO pair
                x P_A pairs / its mean launch time, against 78.6 TFLOP/s FP64 matrix.
  cpu_baseline  the oracle (C restatement of tpqkld / tijkld) on the host cores, on a
                bounded sample of AO pairs and MO pairs of the same workload,
                extrapolated to one whole transform.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|split_fp16|both]
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

METRIC = "MO integrals/s at N_AO=240: FP64 MFMA vs split-FP16 at <=1e-10 abs err"
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this workload (fetched = 2.0 x FETCH_SIZE, written = 1.0 x
# WRITE_SIZE, factors from a 1 GiB copy in the same process); keyed by
# (n_ao, n_mo, precision).  Summaries: profiles/fp64_n240_kernels.md.
MEASURED_TRAFFIC = {(240, 240, "fp64"): (2.0 * 3924450.4 + 6780965.6) * 1024}
FP16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def flops_half1(n, m):
    pa, pm = n * (n + 1) // 2, m * (m + 1) // 2
    return 2.0 * pa * (m * n * n + n * pm)


def flops_half2(n, m):
    pm = m * (m + 1) // 2
    return 2.0 * (n * n * m * (m + 1) * (2 * m + 1) / 6.0 + n * pm * (pm + 1) / 2.0)


def make_inputs(n, m, rank, decay, seed, torch):
    from cipsi_amd import synth
    B = synth.eri_factors(n, rank, decay, seed)
    ao = synth.eri_packed_torch(B, "cuda")
    c = synth.orthonormal(n, seed + 1)[:, :m]
    return B, ao, torch.as_tensor(np.ascontiguousarray(c), device="cuda"), c


def cpu_baseline(ao_host, c, h_gpu, out_gpu, n, m, nthreads=1, n_pq=1200, n_kl=1600):
    """Time the oracle on sampled AO pairs (half 1) and MO pairs (half 2) of the same
    workload and extrapolate to a whole transform.  Cost per AO pair is constant;
    MO pairs are taken with a fixed stride so the sample is unbiased over ilk."""
    import oracle
    pa, pm = n * (n + 1) // 2, m * (m + 1) // 2
    pq_list = np.linspace(0, pa - 1, n_pq).astype(np.int64)
    t0 = time.perf_counter()
    h_cpu = oracle.tpqkld_rows(ao_host, c, pq_list, nthreads=nthreads)
    t_h1 = time.perf_counter() - t0
    kl_list = np.arange(pm // (2 * n_kl), pm, pm // n_kl)[:n_kl].astype(np.int64)
    t0 = time.perf_counter()
    out_cpu = oracle.tijkld_rows(h_gpu, c, kl_list, nthreads=nthreads)
    t_h2 = time.perf_counter() - t0
    est = t_h1 * pa / len(pq_list) + t_h2 * pm / len(kl_list)
    # cross-check the sampled rows against the GPU results
    d1 = float(np.max(np.abs(h_cpu[pq_list] - h_gpu[pq_list])))
    lo = kl_list * (kl_list + 1) // 2
    idx = np.concatenate([np.arange(a, a + k + 1) for a, k in zip(lo, kl_list)])
    d2 = float(np.max(np.abs(out_cpu[idx] - out_gpu[idx])))
    return est, t_h1 + t_h2, d1, d2, len(pq_list), len(kl_list)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--n-ao", type=int, default=240)
    ap.add_argument("--n-mo", type=int, default=240)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baseline")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from cipsi_amd import ijkl

    n, m = args.n_ao, args.n_mo
    rank, decay, seed = 3 * n, 20.0, 2026
    pa, pm = n * (n + 1) // 2, m * (m + 1) // 2
    n_out = pm * (pm + 1) // 2
    print(f"[bench] generating N_AO={n} N_MO={m} rank={rank}", file=sys.stderr, flush=True)
    B, ao, c_dev, c_host = make_inputs(n, m, rank, decay, seed, torch)
    torch.cuda.synchronize()

    if args.calib_copy_mib:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE scale the counters
        src = torch.ones(args.calib_copy_mib * 2**17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ctx = ijkl.Context(0)
    prec = args.precision
    for _ in range(args.warmup):
        out = ctx.transform(ao, c_dev, precision=prec)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    ctx.enable_timing(True)
    ctx.reset_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctx.transform(ao, c_dev, precision=prec)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.enable_timing(False)
    ms_step = 1e3 * dt / args.steps
    value = n_out / (dt / args.steps)
    stages = {}
    for st in ("pad_c", "expand", "half1", "transpose", "half2"):
        tot, cnt = ctx.read_timing(st)
        stages[st] = tot / cnt if cnt else 0.0
    print(f"[bench] {ms_step:.3f} ms/step, stages(ms): "
          + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()), file=sys.stderr, flush=True)

    f1 = flops_half1(n, m)
    achieved = f1 / (stages["half1"] * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "k_half_fp64 (half 1: Q1+Q2)",
                "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                "traffic": MEASURED_TRAFFIC.get((n, m, prec)),
                "stage_ms": {k: round(v, 4) for k, v in stages.items()},
                "half2_tflops": round(flops_half2(n, m) / (stages["half2"] * 1e-3) / 1e12, 3)
                if stages["half2"] else None}

    # accuracy against the closed form on the GPU (independent rounding path)
    result = {
        "metric": METRIC, "value": value, "unit": "MO-integrals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded low-rank PSD AO integrals, orthonormal C)",
        "config": {"workload": f"ijkl C1 transform N_AO={n} N_MO={m} rank={rank} decay={decay}",
                   "n_ao": n, "n_mo": m, "precision": prec, "mo_integrals_per_step": n_out},
        "roofline": roofline,
    }

    if not args.no_cpu:
        print("[bench] cpu baseline sample", file=sys.stderr, flush=True)
        h_gpu = ctx.tpqkld(ao, c_dev).cpu().numpy()
        out_host = out.cpu().numpy()
        ao_host = ao.cpu().numpy()
        est, spent, d1, d2, s1, s2 = cpu_baseline(ao_host, c_host, h_gpu, out_host, n, m,
                                                  nthreads=args.cpu_threads)
        result["cpu_baseline"] = {
            "value": n_out / est, "unit": "MO-integrals/s", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"oracle tpqkld on {s1} AO pairs + tijkld on {s2} MO pairs (stride-sampled) "
                      f"of the same N={n} workload, {spent:.1f} s measured, extrapolated to "
                      f"{est:.1f} s per transform; no disk I/O",
            "est_seconds_per_transform": est,
            "gpu_vs_cpu_sampled_maxabs": {"half1": d1, "half2": d2},
        }
        result["speedup_vs_cpu"] = value / (n_out / est)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
