#!/usr/bin/env python3
"""All-mode MM-CSF MTTKRP benchmark on one MI355X.

Workload (the configuration the headline metric is quoted on): a 3-mode
500000 x 200000 x 50000 tensor with 100M distinct nonzeros, mode-0 and mode-1
coordinates Zipf(1.2), mode 2 uniform, values U(0,1) float32, generated on the
GPU from a fixed seed; MM-CSF built with -f 128; factors R=32 U(0,1).

A step is one all-mode MTTKRP (every mode's output from the same factors).
The timed region runs K steps with float32 factors, then K steps with
bfloat16 factor storage (float32 accumulation).  Inputs are resident in HBM
before timing starts.

Reported: all-mode ms, algorithmic HBM GB/s (value), GFLOP/s, per-mode kernel
times from HIP events, error of every output against a float64 contraction,
the roofline of the dominant kernel, and the reference's sequential COO CPU
MTTKRP (-t 1) restated in C, timed on a bounded sample on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mm-csf_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "All-mode MTTKRP ms & HBM GB/s, FP32 vs BF16-factor at <=1e-3 rel err"
HBM_PEAK_GBS = 8000.0
SEED = 0x4D4D43534600 + 1


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(info, distinct, R, sf):
    """Compulsory HBM bytes of each mode's MTTKRP (4-byte indices/values):
    nonzeros (value + leaf index), fiber levels (pointer + index), the slice
    coordinate of each level-1 entry, each distinct input factor row read
    once (sf bytes per element) and each distinct output row written once."""
    N = info["ndims"]
    nnz = info["totNnz"]
    lv = info["nLevel"]
    struct = nnz * 8 + sum(lv[l] * 8 for l in range(1, N - 1)) + lv[1] * 4
    out = []
    for n in range(N):
        fac = sum(distinct[m] * R * sf for m in range(N) if m != n)
        out.append(struct + fac + distinct[n] * R * 4)
    return out


def fp64_reference(inds, vals, U, mode, dims, chunk=4_000_000):
    """float64 MTTKRP on the GPU: contributions sorted by output row are
    reduced per row segment with prefix sums along the nonzero axis (no
    atomics on hot rows)."""
    import torch
    R = U[0].shape[1]
    ref = torch.zeros((dims[mode], R), dtype=torch.float64, device=vals.device)
    UT = [u.double().t().contiguous() for u in U]          # R x dims[m]
    order = torch.argsort(inds[mode], stable=True)
    n = vals.numel()
    for s in range(0, n, chunk):
        sel = order[s:min(s + chunk, n)]
        rows = inds[mode][sel].long()
        c = vals[sel].double()[None, :].expand(R, -1).clone()   # R x chunk
        for m in range(len(dims)):
            if m != mode:
                c *= UT[m][:, inds[m][sel].long()]
        cs = torch.cumsum(c, dim=1)
        last = torch.nonzero(rows[1:] != rows[:-1]).flatten()
        last = torch.cat([last, torch.tensor([rows.numel() - 1], device=rows.device)])
        seg = cs[:, last]
        seg[:, 1:] -= cs[:, last[:-1]]
        ref.index_add_(0, rows[last], seg.t())
    return ref


def rel_err(out, ref):
    import torch
    d = out.double() - ref
    return float(torch.linalg.norm(d) / torch.linalg.norm(ref)), float(d.abs().max() / ref.abs().max())


def cpu_baseline(inds, vals, U, dims, sample, bytes_f32):
    """The reference's -t 1 path (sequential float32 COO MTTKRP) restated in
    C (oracle/), one core, on the first `sample` nonzeros sorted like
    sort_COOtensor; time scaled to the full nonzero count."""
    import numpy as np

    import oracle as O

    nnz = vals.numel()
    S = min(sample, nnz)
    hi = [a[:S].cpu().numpy().astype(np.int64) for a in inds]
    hv = vals[:S].cpu().numpy()
    hU = [u.cpu().numpy() for u in U]
    per_mode = []
    for n in range(len(dims)):
        order = O.order_tensormode(dims, n)
        perm = np.lexsort((hi[order[1]], hi[order[0]]))   # first two keys (util.h:389-397)
        si = [a[perm] for a in hi]
        out = np.zeros((dims[n], hU[0].shape[1]), dtype=np.float32)
        t0 = time.perf_counter()
        O.mttkrp_coo_f32(si, hv[perm], n, hU, out)
        per_mode.append(time.perf_counter() - t0)
        log(f"cpu baseline mode {n}: {per_mode[-1]:.2f} s on {S} nonzeros")
    scale = nnz / S
    full_s = [t * scale for t in per_mode]
    return {
        "value": round(sum(bytes_f32) / sum(full_s) / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"first {S} of {nnz} nonzeros per mode, sorted by modeOrder[0..1]; "
                   f"single-thread float32 COO MTTKRP (mttkrp_cpu.cpp:60-113 restated), "
                   f"time x{scale:.2f} to the full tensor"),
        "all_mode_ms_extrapolated": round(sum(full_s) * 1e3, 1),
        "sample_seconds": [round(t, 3) for t in per_mode],
        "host_cores_available": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nnz", type=int, default=100_000_000)
    ap.add_argument("--cpu-sample", type=int, default=50_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one GPU and the multi-GPU "
              "path is not built (see DESIGN.md, sharding)", file=sys.stderr)
        sys.exit(2)

    import torch

    import mmcsf

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    dims = [500000, 200000, 50000]
    R = 32
    log("generating tensor")
    t0 = time.perf_counter()
    X = mmcsf.Tensor.generate(dims, args.nnz, [1, 1, 0], [1.2, 1.2, 0.0], SEED)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    log(f"generated {args.nnz} nonzeros in {gen_s:.2f} s; building MM-CSF")
    t0 = time.perf_counter()
    fmt = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = fmt.info()
    log(f"built in {build_s:.2f} s: {info}")
    inds, vals = X.to_torch(dev)
    distinct = [int((torch.bincount(inds[m].long(), minlength=dims[m]) > 0).sum()) for m in range(3)]
    g = torch.Generator(device=dev).manual_seed(SEED)
    U = [torch.rand((d, R), generator=g, device=dev) for d in dims]
    Ub = [u.to(torch.bfloat16) for u in U]
    outs = [torch.empty((d, R), dtype=torch.float32, device=dev) for d in dims]
    bytes_f32 = algorithmic_bytes(info, distinct, R, 4)
    bytes_bf16 = algorithmic_bytes(info, distinct, R, 2)

    def timed(factors):
        for _ in range(args.warmup):
            mmcsf.mttkrp_allmode(fmt, factors, outs)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            mmcsf.mttkrp_allmode(fmt, factors, outs)
        e1.record()
        torch.cuda.synchronize()
        step_ms = e0.elapsed_time(e1) / args.steps
        # per-kernel durations (HIP events on the launch stream, kernel only)
        km = [0.0] * 3
        for _ in range(args.steps):
            _, ms = mmcsf.mttkrp_allmode(fmt, factors, outs, timed=True)
            km = [a + b for a, b in zip(km, ms)]
        return step_ms, [k / args.steps for k in km]

    log("timing float32 factors")
    f32_ms, f32_k = timed(U)
    f32_out = [o.clone() for o in outs]
    log(f"float32: {f32_ms:.3f} ms/step, kernels {f32_k}")
    log("timing bfloat16 factors")
    b16_ms, b16_k = timed(Ub)
    b16_out = [o.clone() for o in outs]
    log(f"bfloat16: {b16_ms:.3f} ms/step, kernels {b16_k}")

    # the reference's layout: one MM-CSF tree, non-root modes scatter (LDS
    # privatisation of hot rows, atomics for the rest)
    log("timing the single-tree path")
    fmt1 = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128, mode_trees=False)
    single_steps = []
    for _ in range(args.warmup):
        mmcsf.mttkrp_allmode(fmt1, U, outs)
    for _ in range(max(3, args.steps // 4)):
        _, ms = mmcsf.mttkrp_allmode(fmt1, U, outs, timed=True)
        single_steps.append(ms)
    single_k = [sum(s[m] for s in single_steps) / len(single_steps) for m in range(3)]
    del fmt1
    torch.cuda.empty_cache()
    log(f"single tree: kernels {single_k}")

    log("float64 reference")
    err32, err16 = [], []
    for m in range(3):
        ref = fp64_reference(inds, vals, U, m, dims)
        err32.append(rel_err(f32_out[m], ref))
        err16.append(rel_err(b16_out[m], ref))
        log(f"mode {m}: rel err fp32 {err32[-1][0]:.3e}, bf16 {err16[-1][0]:.3e}")
        del ref
    torch.cuda.empty_cache()

    dom = max(range(3), key=lambda m: f32_k[m])
    achieved = bytes_f32[dom] / (f32_k[dom] * 1e-3) / 1e9
    flops = 3 * args.nnz * R * 3
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic_c2.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tr = json.load(fh)
        traffic = tr.get("modes", {}).get(str(dom), {}).get("hbm_bytes")

    result = {
        "metric": METRIC,
        "value": round(sum(bytes_f32) / (f32_ms * 1e-3) / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU generator, seed 0x4D4D43534601)",
        "config": {
            "workload": ("all-mode MM-CSF MTTKRP (-t 12 -f 128), 500000x200000x50000, "
                         f"{args.nnz} distinct nonzeros, Zipf1.2 x Zipf1.2 x uniform, R=32"),
            "R": R, "nnz": args.nnz, "dims": dims, "fbr_threshold": 128,
            "mode_order": info["modeOrder"], "n_slices": info["nLevel"][0],
            "n_fibers": info["nLevel"][1], "distinct_rows": distinct,
        },
        "gflops": round(flops / (f32_ms * 1e-3) / 1e9, 1),
        "modes_f32": [{"mode": m, "kernel_ms": round(f32_k[m], 4), "alg_bytes": bytes_f32[m],
                       "gbs": round(bytes_f32[m] / (f32_k[m] * 1e-3) / 1e9, 1),
                       "rel_err_fp64": err32[m][0], "max_rel_err_fp64": err32[m][1]} for m in range(3)],
        "bf16": {
            "ms_per_step": round(b16_ms, 4),
            "gbs": round(sum(bytes_bf16) / (b16_ms * 1e-3) / 1e9, 2),
            "speedup_vs_f32": round(f32_ms / b16_ms, 3),
            "modes": [{"mode": m, "kernel_ms": round(b16_k[m], 4), "alg_bytes": bytes_bf16[m],
                       "rel_err_fp64": err16[m][0], "max_rel_err_fp64": err16[m][1]} for m in range(3)],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"mode {dom} (level {info['modeOrder'].index(dom)}) MM-CSF MTTKRP, float32 factors",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "single_tree": {
            "note": "the reference's MM-CSF layout: one tree, modes other than its root scatter",
            "kernel_ms": [round(k, 4) for k in single_k],
            "all_mode_kernel_ms": round(sum(single_k), 4),
            "gbs": round(sum(bytes_f32) / (sum(single_k) * 1e-3) / 1e9, 2),
        },
        "generate_s": round(gen_s, 2),
        "build_s": round(build_s, 2),
    }
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(inds, vals, U, dims, args.cpu_sample, bytes_f32)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
