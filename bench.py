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
GATHER_TOOL = os.path.join(ROOT, "mm-csf_amd", "bin", "gather_ceiling")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(info, distinct, R, sf, single_tree=False):
    """Compulsory HBM bytes of each mode's MTTKRP (4-byte indices/values):
    nonzeros (value + leaf index), fiber levels (pointer + index) and the
    slice coordinate of each level-1 entry of the tree that mode traverses,
    each distinct input factor row read once (sf bytes per element) and each
    distinct output row written once."""
    N = info["ndims"]
    nnz = info["totNnz"]
    out = []
    for n in range(N):
        lv = info["nLevel"] if single_tree else info["modeLevel"][n]
        struct = nnz * 8 + sum(lv[l] * 8 for l in range(1, N - 1)) + lv[1] * 4
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


def cpu_model():
    """CPU model name of this host (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sorted_for_t1(torch, inds, vals, dims, mode):
    """The nonzeros ordered as sort_COOtensor leaves them before -t 1
    (mttkrp.cu:79): by modeOrder[0], then modeOrder[1]; sorted on the GPU,
    returned on the host."""
    import numpy as np

    import oracle as O
    order = O.order_tensormode(dims, mode)
    key = inds[order[0]].long() * dims[order[1]] + inds[order[1]].long()
    perm = torch.sort(key, stable=True).indices
    hi = [a[perm].cpu().numpy().astype(np.int64) for a in inds]
    return hi, vals[perm].cpu().numpy()


def cpu_baseline(torch, inds, vals, U, dims, bytes_f32):
    """The reference's -t 1 path (sequential float32 COO MTTKRP,
    mttkrp_cpu.cpp:60-113) restated in C (oracle/), one core, on every
    nonzero of the tensor, each mode timed over its loop only
    (mttkrp.cu:106-108) with the nonzeros sorted like sort_COOtensor."""
    import numpy as np

    import oracle as O

    nnz = vals.numel()
    hU = [u.cpu().numpy() for u in U]
    per_mode = []
    for n in range(len(dims)):
        si, sv = sorted_for_t1(torch, inds, vals, dims, n)
        out = np.zeros((dims[n], hU[0].shape[1]), dtype=np.float32)
        t0 = time.perf_counter()
        O.mttkrp_coo_f32(si, sv, n, hU, out)
        per_mode.append(time.perf_counter() - t0)
        log(f"cpu baseline mode {n}: {per_mode[-1]:.2f} s on {nnz} nonzeros")
        del si, sv, out
    return {
        "value": round(sum(bytes_f32) / sum(per_mode) / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"all {nnz} nonzeros, every mode, sorted by modeOrder[0..1]; single-thread float32 "
                   f"COO MTTKRP (mttkrp_cpu.cpp:60-113 restated in C, gcc -O2), loop only"),
        "all_mode_ms": round(sum(per_mode) * 1e3, 1),
        "mode_seconds": [round(t, 3) for t in per_mode],
        "cpu_model": cpu_model(),
        "host_cores_available": os.cpu_count(),
    }


def gather_ceiling(table_mb):
    """Random row-gather rates of this GPU from the gather-only kernels of
    mm-csf_amd/tools/gather_ceiling.hip: 128- and 64-byte rows (float32 and
    bfloat16 rows at R=32) uniformly at random from 1 MB and 3.2 MB tables
    (resident in every XCD's 4 MB L2; the better of the two is the rate no
    gather of such rows exceeds) and from a
    table of table_mb (uniform rows over the whole factor table: the rate
    without the locality a skewed tensor gives), and the coalesced
    stream-read rate."""
    import subprocess
    if not os.path.exists(GATHER_TOOL):
        raise RuntimeError(f"{GATHER_TOOL} not built (run make -C mm-csf_amd)")
    r = subprocess.run([GATHER_TOOL, "-b", "128,64", "-t", f"1,3.2,{table_mb:.3f}"], capture_output=True,
                       text=True, timeout=120, check=True)
    lines = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    out = {"leaf_table_mb": round(table_mb, 3)}
    leaf = round(table_mb, 3)
    for d in lines:
        if d["kind"] == "stream":
            out["stream_gb_per_s"] = d["gb_per_s"]
        if d["kind"] != "gather":
            continue
        rb, t, v = d["row_bytes"], d["table_mb"], d["g_rows_per_s"]
        if t in (1.0, 3.2):  # L2-resident: the better of the two
            out[f"g_rows_per_s_{rb}B_l2"] = max(out.get(f"g_rows_per_s_{rb}B_l2", 0.0), v)
        if abs(t - leaf) < 1e-6:
            out[f"g_rows_per_s_{rb}B_leaf_table"] = v
    return out


def gathered_rows(info, single_tree=False):
    """Factor rows a fiber-level-factored traversal must gather per mode: one
    leaf row per nonzero and one row per entry of every level between the
    root and the leaf (the fiber-level rows, each read once per fiber)."""
    N = info["ndims"]
    out = []
    for n in range(N):
        lv = info["nLevel"] if single_tree else info["modeLevel"][n]
        out.append(info["totNnz"] + sum(lv[l] for l in range(1, N - 1)))
    return out


def run_c2(args):
    import torch

    import mmcsf

    dev = torch.device("cuda:0")
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
    rows = gathered_rows(info)

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
        # per-kernel durations (HIP events recorded by the library on the
        # stream the kernel is launched on, kernel only)
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

    if args.profile:
        print(json.dumps({"metric": METRIC, "profile": True, "ms_per_step": round(f32_ms, 4),
                          "kernel_ms_f32": f32_k, "kernel_ms_bf16": b16_k}), flush=True)
        return

    # the reference's layout: one MM-CSF tree, non-root modes scatter (LDS
    # privatisation of hot rows, atomics for the rest)
    log("timing the single-tree path")
    fmt1 = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128, mode_trees=False)
    info1 = fmt1.info()
    single_steps = []
    for _ in range(args.warmup):
        mmcsf.mttkrp_allmode(fmt1, U, outs)
    for _ in range(max(3, args.steps // 4)):
        _, ms = mmcsf.mttkrp_allmode(fmt1, U, outs, timed=True)
        single_steps.append(ms)
    single_k = [sum(s[m] for s in single_steps) / len(single_steps) for m in range(3)]
    bytes_single = algorithmic_bytes(info1, distinct, R, 4, single_tree=True)
    del fmt1
    torch.cuda.empty_cache()
    log(f"single tree: kernels {single_k}")

    # B-CSF (-t 8) and HB-CSF (-t 10) rooted at each mode, as the reference's
    # run_all.sh runs them (one -t 8 / -t 10 run per -m)
    per_type = {}
    o1 = [torch.empty((d, R), dtype=torch.float32, device=dev) for d in dims]
    for typ, name in ((8, "B-CSF"), (10, "HB-CSF")):
        ms32, ms16 = [], []
        for m in range(3):
            f = mmcsf.Format(X, typ, mode=m, fbr_threshold=128)
            ms32.append(_event_ms(torch, lambda: mmcsf.mttkrp(f, m, U, o1[m]), args.steps, args.warmup))
            ms16.append(_event_ms(torch, lambda: mmcsf.mttkrp(f, m, Ub, o1[m]), args.steps, args.warmup))
            del f
            torch.cuda.empty_cache()
        per_type[name] = {"type": typ, "f32_ms_per_mode": [round(v, 4) for v in ms32],
                          "f32_all_mode_ms": round(sum(ms32), 4),
                          "bf16_ms_per_mode": [round(v, 4) for v in ms16],
                          "bf16_all_mode_ms": round(sum(ms16), 4)}
        log(f"-t {typ} per mode: f32 {ms32}, bf16 {ms16}")
    del o1

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
    leaf = info["modeTreeOrder"][dom][-1]
    log("gather ceiling")
    ceil = gather_ceiling(dims[leaf] * R * 4 / 1e6)
    log(f"gather ceiling: {ceil}")
    flops = 3 * args.nnz * R * 3
    gbytes = lambda b: round(b / 1e9, 3)  # noqa: E731

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
            "workload": ("all-mode MM-CSF MTTKRP, -t 12 -f 128 -a 1 (the reference's MM-CSF tree plus a root "
                         "tree per other mode, so every mode is a root traversal: "
                         f"{gbytes(info['deviceBytes'])} GB of structure; the reference's single tree, -a 0, "
                         f"holds {gbytes(info1['deviceBytes'])} GB and is timed under single_tree), "
                         f"500000x200000x50000, {args.nnz} distinct nonzeros, Zipf1.2 x Zipf1.2 x uniform, R=32; "
                         "-w 1 and -w 4 give the same launch (the walk splits work by chunk cost, not per slice)"),
            "R": R, "nnz": args.nnz, "dims": dims, "fbr_threshold": 128, "layout": "-a 1",
            "structure_bytes": {"-a 1": info["deviceBytes"], "-a 0": info1["deviceBytes"]},
            "mode_order": info["modeOrder"], "tree_orders": info["modeTreeOrder"], "n_slices": info["nLevel"][0],
            "n_fibers": info["nLevel"][1], "mode_tree_levels": info["modeLevel"], "distinct_rows": distinct,
        },
        "gflops": round(flops / (f32_ms * 1e-3) / 1e9, 1),
        "modes_f32": [{"mode": m, "kernel_ms": round(f32_k[m], 4), "alg_bytes": bytes_f32[m],
                       "gbs": round(bytes_f32[m] / (f32_k[m] * 1e-3) / 1e9, 1), "gathered_rows": rows[m],
                       "g_rows_per_s": round(rows[m] / (f32_k[m] * 1e-3) / 1e9, 1),
                       "rel_err_fp64": err32[m][0], "max_rel_err_fp64": err32[m][1]} for m in range(3)],
        "bf16": {
            "ms_per_step": round(b16_ms, 4),
            "gbs": round(sum(bytes_bf16) / (b16_ms * 1e-3) / 1e9, 2),
            "speedup_vs_f32": round(f32_ms / b16_ms, 3),
            "gather_ceiling_64B_over_128B": round(ceil["g_rows_per_s_64B_l2"] / ceil["g_rows_per_s_128B_l2"], 3),
            "modes": [{"mode": m, "kernel_ms": round(b16_k[m], 4), "alg_bytes": bytes_bf16[m],
                       "g_rows_per_s": round(rows[m] / (b16_k[m] * 1e-3) / 1e9, 1),
                       "rel_err_fp64": err16[m][0], "max_rel_err_fp64": err16[m][1]} for m in range(3)],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"mode {dom} MTTKRP (chunked root walk k_walk over the mode's root tree), float32 factors",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # HBM bytes per launch are not measurable with the L2 counters
            # (FETCH_SIZE counts L2-to-fabric requests, most of which the
            # 256 MB MALL serves); the per-launch L2-fabric request bytes are
            # in the committed rocprofv3 summary of this kernel
            "traffic": None,
            "gather": {
                "rows_per_launch": rows[dom],
                "achieved_g_rows_per_s": round(rows[dom] / (f32_k[dom] * 1e-3) / 1e9, 1),
                "ceiling_g_rows_per_s": ceil["g_rows_per_s_128B_l2"],
                "frac": round(rows[dom] / (f32_k[dom] * 1e-3) / 1e9 / ceil["g_rows_per_s_128B_l2"], 4),
                "ceiling": ceil,
                "note": ("ceiling: 128-byte rows gathered at random from an L2-resident table by a kernel that "
                         "does nothing else (mm-csf_amd/bin/gather_ceiling, measured in this run); uniform rows "
                         "over a table of the leaf factor table's size gather at the *_leaf_table rate"),
            },
        },
        "single_tree": {
            "note": "the reference's MM-CSF layout (-a 0): one tree, modes other than its root scatter",
            "kernel_ms": [round(k, 4) for k in single_k],
            "all_mode_kernel_ms": round(sum(single_k), 4),
            "gbs": round(sum(bytes_single) / (sum(single_k) * 1e-3) / 1e9, 2),
        },
        "per_mode_types": per_type,
        "generate_s": round(gen_s, 2),
        "build_s": round(build_s, 2),
    }
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(torch, inds, vals, U, dims, bytes_f32)
    print(json.dumps(result), flush=True)


# ---------------------------------------------------------------------------
# further workloads (parity cases of the path, each printing its own line)
# ---------------------------------------------------------------------------
def _event_ms(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _line(metric, value, unit, ms, args, dtype, config, higher=True, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": higher, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype, "data": "synthetic (GPU generator)", "config": config}
    d.update(extra)
    print(json.dumps(d), flush=True)


def run_c1(args):
    """Mode-0 MTTKRP on a uniform 2000^3 tensor with 1M nonzeros, R=32, -f 128:
    COO (-t 3), B-CSF (-t 8), HB-CSF (-t 10), MM-CSF (-t 12) against the
    float64 contraction and the sequential float32 COO CPU MTTKRP (-t 1)."""
    import numpy as np
    import torch

    import mmcsf
    import oracle as O  # the -t 1 CPU leg

    dev = torch.device("cuda:0")
    dims, nnz, R = [2000, 2000, 2000], 1_000_000, 32
    X = mmcsf.Tensor.generate(dims, nnz, [0, 0, 0], [0.0, 0.0, 0.0], 0x4D4D43534600)
    inds, vals = X.to_torch(dev)
    g = torch.Generator(device=dev).manual_seed(0x4D4D43534600)
    U = [torch.rand((d, R), generator=g, device=dev) for d in dims]
    ref = fp64_reference(inds, vals, U, 0, dims)
    hi = [a.cpu().numpy().astype(np.int64) for a in inds]
    hv = vals.cpu().numpy()
    order = O.order_tensormode(dims, 0)
    perm = np.lexsort((hi[order[1]], hi[order[0]]))
    t1 = np.zeros((dims[0], R), dtype=np.float32)
    c0 = time.perf_counter()
    O.mttkrp_coo_f32([a[perm] for a in hi], hv[perm], 0, [u.cpu().numpy() for u in U], t1)
    t1_s = time.perf_counter() - c0
    t1d = torch.from_numpy(t1).to(dev).double()
    t1_err = rel_err(t1d.float(), ref)[0]
    types = {}
    out = torch.empty((dims[0], R), device=dev)
    for typ, name in ((3, "COO"), (8, "B-CSF"), (10, "HB-CSF"), (12, "MM-CSF")):
        fmt = mmcsf.Format(X, typ, mode=0, fbr_threshold=128)
        ms = _event_ms(torch, lambda: mmcsf.mttkrp(fmt, 0, U, out), args.steps, args.warmup)
        e64 = rel_err(out, ref)[0]
        et1 = float(torch.linalg.norm(out.double() - t1d) / torch.linalg.norm(t1d))
        types[name] = {"type": typ, "ms": round(ms, 4), "gflops": round(3 * nnz * R / (ms * 1e-3) / 1e9, 1),
                       "rel_err_fp64": e64, "rel_err_vs_t1": et1}
        log(f"c1 {name}: {ms:.4f} ms, err fp64 {e64:.2e}, vs -t 1 {et1:.2e}")
        del fmt
    best = min(types.values(), key=lambda v: v["ms"])
    _line("Mode-0 MTTKRP ms, uniform 2000^3, 1M nonzeros, R=32", best["ms"], "ms", best["ms"], args, "f32",
          {"workload": "mode-0 MTTKRP, 2000x2000x2000 uniform, 1000000 distinct nonzeros, R=32, -f 128",
           "R": R, "nnz": nnz, "dims": dims}, higher=False, types=types,
          t1_cpu={"ms": round(t1_s * 1e3, 2), "cores": 1, "rel_err_fp64": t1_err})


def run_next(args):
    """The rows beyond the five MTTKRP types, each timed with its parity check:
    (1) the parallel .tns reader and the binary COO cache against the
    single-threaded reader of the oracle (load_tensor, util.h:313-370);
    (2) the one-tree all-mode types -t 9 / -t 13 / -t 14 and the MM-CSF variants
    (-a 0: the reference's one tree; -M 1: the mixed-mode partition) against
    float64, beside the default -t 12 (a root tree per mode), with build times,
    and the binary format cache (save, load, bit-identical outputs); (3)
    leaf-mode TTM (ttm.cu) against float64 and the sequential host TTM
    (TTM_CPU, ttm_cpu.cpp:60-86)."""
    import tempfile

    import numpy as np
    import torch

    import mmcsf
    import oracle as O  # the single-threaded reader restated

    dev = torch.device("cuda:0")
    # (1) text reader and binary cache: 2000^3 uniform, 2M nonzeros
    dims1 = [2000, 2000, 2000]
    X1 = mmcsf.Tensor.generate(dims1, 2_000_000, [0, 0, 0], [0.0, 0.0, 0.0], 0x4D4D43534600)
    _, gi, gv = X1.to_numpy()
    with tempfile.TemporaryDirectory() as d:
        p_tns, p_bin = os.path.join(d, "x.tns"), os.path.join(d, "x.bin")
        X1.save_tns(p_tns)
        X1.save_bin(p_bin)
        tns_mb = os.path.getsize(p_tns) / 1e6
        t0 = time.perf_counter()
        Y = mmcsf.Tensor.load_tns(p_tns)
        par_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref = O.load_tns(p_tns)
        ora_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        Z = mmcsf.Tensor.load_bin(p_bin)
        bin_s = time.perf_counter() - t0
    _, yi, yv = Y.to_numpy()
    _, zi, zv = Z.to_numpy()
    same = all(np.array_equal(yi[m], ref[1][m]) and np.array_equal(zi[m], gi[m]) for m in range(3)) and \
        np.array_equal(yv, ref[2]) and np.array_equal(zv.view(np.uint32), gv.view(np.uint32))
    io = {"tns_MB": round(tns_mb, 1), "reader_s": round(par_s, 3), "reader_MB_s": round(tns_mb / par_s, 1),
          "oracle_reader_s": round(ora_s, 3), "speedup": round(ora_s / par_s, 1), "bin_load_s": round(bin_s, 3),
          "identical": bool(same)}
    log(f"next io: {io}")
    del X1, Y, Z, gi, gv, yi, yv, zi, zv, ref

    # (2) one-tree types on the c2 shape with fewer nonzeros
    dims, R = [500000, 200000, 50000], 32
    X = mmcsf.Tensor.generate(dims, args.nnz_next, [1, 1, 0], [1.2, 1.2, 0.0], SEED)
    inds, vals = X.to_torch(dev)
    g = torch.Generator(device=dev).manual_seed(SEED)
    U = [torch.rand((d, R), generator=g, device=dev) for d in dims]
    refs = [fp64_reference(inds, vals, U, m, dims) for m in range(3)]
    outs = [torch.empty((d, R), dtype=torch.float32, device=dev) for d in dims]
    types = {}
    cache = None
    for key, typ, name, kw in (("12", 12, "MM-CSF, a root tree per mode (-a 1)", {}),
                               ("12 -a 0", 12, "MM-CSF, the reference's one tree (-a 0)", {"mode_trees": False}),
                               ("12 -M 1", 12, "MM-CSF, mixed-mode partition (-M 1)", {"mixed": True}),
                               ("9", 9, "ONE-B-CSF", {}), ("13", 13, "single CSF, no bins", {}),
                               ("14", 14, "single CSF", {})):
        t0 = time.perf_counter()
        fmt = mmcsf.Format(X, typ, fbr_threshold=128, **kw)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        ms = _event_ms(torch, lambda: mmcsf.mttkrp_allmode(fmt, U, outs), args.steps, args.warmup)
        errs = [rel_err(outs[m], refs[m])[0] for m in range(3)]
        info = fmt.info()
        types[key] = {"name": name, "type": typ, "all_mode_ms": round(ms, 4), "build_s": round(build_s, 3),
                      "structure_bytes": info["deviceBytes"], "rel_err_fp64": errs}
        if kw.get("mixed"):
            types[key]["partition_nnz"] = info["partNnz"][:3]
        log(f"next -t {key}: {types[key]}")
        if key == "12":
            # binary cache of the built format: save, reload, same outputs bit for bit
            want = [o.clone() for o in outs]
            with tempfile.TemporaryDirectory() as d:
                p_fmt = os.path.join(d, "x.fmt")
                t0 = time.perf_counter()
                fmt.save(p_fmt)
                save_s = time.perf_counter() - t0
                size = os.path.getsize(p_fmt)
                t0 = time.perf_counter()
                f2 = mmcsf.Format.load(p_fmt)
                torch.cuda.synchronize()
                load_s = time.perf_counter() - t0
            mmcsf.mttkrp_allmode(f2, U, outs)
            torch.cuda.synchronize()
            # slices spanning waves are added with float atomics, so two runs
            # of one format agree to rounding, not bit for bit (the cache's
            # arrays round-trip bit for bit; the GPU tests check that)
            cache = {"file_MB": round(size / 1e6, 1), "save_s": round(save_s, 3), "load_s": round(load_s, 3),
                     "build_s": round(build_s, 3), "load_vs_build": round(build_s / load_s, 2),
                     "rel_diff_vs_built": max(rel_err(outs[m], want[m].double())[0] for m in range(3)),
                     "rel_err_fp64": [rel_err(outs[m], refs[m])[0] for m in range(3)]}
            log(f"next format cache: {cache}")
            del f2, want
        del fmt
        torch.cuda.empty_cache()
    del refs

    # (3) TTM on the leaf mode of the B-CSF tree rooted at mode 0
    fmt = mmcsf.Format(X, 8, mode=0, fbr_threshold=128)
    info = fmt.info()
    leaf_mode = info["modeOrder"][-1]
    ptr = torch.from_numpy(fmt.export("fbrPtr", 1).astype(np.int64)).to(dev)
    leaf = torch.from_numpy(fmt.export("inds", leaf_mode).astype(np.int64)).to(dev)
    v = torch.from_numpy(fmt.export("vals")).to(dev).double()
    nf = ptr.numel() - 1
    fid = torch.repeat_interleave(torch.arange(nf, device=dev), ptr[1:] - ptr[:-1])
    Y64 = torch.zeros((nf, R), dtype=torch.float64, device=dev)
    Ul = U[leaf_mode]
    for s0 in range(0, v.numel(), 4_000_000):
        sl = slice(s0, s0 + 4_000_000)
        Y64.index_add_(0, fid[sl], v[sl, None] * Ul[leaf[sl]].double())
    Yd = torch.empty((nf, R), dtype=torch.float32, device=dev)
    ms = _event_ms(torch, lambda: mmcsf.ttm(fmt, Ul, Yd), args.steps, args.warmup)
    err = float(torch.linalg.norm(Yd.double() - Y64) / torch.linalg.norm(Y64))
    t0 = time.perf_counter()
    Yh = mmcsf.ttm_host(fmt, Ul.cpu().numpy())
    host_s = time.perf_counter() - t0
    errh = float(torch.linalg.norm(torch.from_numpy(Yh).to(dev).double() - Y64) / torch.linalg.norm(Y64))
    nnz = int(v.numel())
    ttm = {"leaf_fibers": nf, "gpu_ms": round(ms, 4), "rel_err_fp64": err, "host_ttm_ms": round(host_s * 1e3, 1),
           "host_cores": 1, "host_rel_err_fp64": errh,
           "gbs": round((nnz * 8 + nf * 8 + nf * R * 4) / (ms * 1e-3) / 1e9, 1)}
    log(f"next ttm: {ttm}")
    _line("Rows beyond the MTTKRP types: .tns reader, one-tree and MM-CSF variants, format cache, TTM",
          types["9"]["all_mode_ms"], "ms", types["9"]["all_mode_ms"], args, "f32",
          {"workload": f"reader: 2000^3 uniform 2M nonzeros; all-mode variants, format cache and TTM: "
                       f"500000x200000x50000, {args.nnz_next} nonzeros, Zipf1.2 x Zipf1.2 x uniform, R=32, -f 128",
           "R": R, "nnz": args.nnz_next}, higher=False, io=io, types=types, format_cache=cache, ttm=ttm)


def run_c3(args):
    """4-mode precision sweep: 100000x50000x20000x1000, 50M nonzeros, Zipf1.1 on
    modes 0 and 2; all-mode MTTKRP at R=16 and R=64, float32 vs bfloat16
    factor storage, errors against float64."""
    import torch

    import mmcsf

    dev = torch.device("cuda:0")
    dims, nnz = [100000, 50000, 20000, 1000], args.nnz_c3
    X = mmcsf.Tensor.generate(dims, nnz, [1, 0, 1, 0], [1.1, 0.0, 1.1, 0.0], 0x4D4D43534602)
    t0 = time.perf_counter()
    fmt = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = fmt.info()
    inds, vals = X.to_torch(dev)
    distinct = [int((torch.bincount(inds[m].long(), minlength=dims[m]) > 0).sum()) for m in range(4)]
    runs = []
    for R in (16, 64):
        g = torch.Generator(device=dev).manual_seed(0x4D4D43534602 + R)
        U = [torch.rand((d, R), generator=g, device=dev) for d in dims]
        Ub = [u.to(torch.bfloat16) for u in U]
        outs = [torch.empty((d, R), device=dev) for d in dims]
        f32 = _event_ms(torch, lambda: mmcsf.mttkrp_allmode(fmt, U, outs), args.steps, args.warmup)
        o32 = [o.clone() for o in outs]
        _, k32 = mmcsf.mttkrp_allmode(fmt, U, outs, timed=True)
        b16 = _event_ms(torch, lambda: mmcsf.mttkrp_allmode(fmt, Ub, outs), args.steps, args.warmup)
        o16 = [o.clone() for o in outs]
        _, k16 = mmcsf.mttkrp_allmode(fmt, Ub, outs, timed=True)
        e32, e16 = [], []
        for m in range(4):
            ref = fp64_reference(inds, vals, U, m, dims, chunk=2_000_000)
            e32.append(rel_err(o32[m], ref)[0])
            e16.append(rel_err(o16[m], ref)[0])
            del ref
        torch.cuda.empty_cache()
        b32 = algorithmic_bytes(info, distinct, R, 4)
        bb = algorithmic_bytes(info, distinct, R, 2)
        runs.append({"R": R, "f32_ms": round(f32, 4), "bf16_ms": round(b16, 4),
                     "f32_gbs": round(sum(b32) / (f32 * 1e-3) / 1e9, 1),
                     "bf16_gbs": round(sum(bb) / (b16 * 1e-3) / 1e9, 1),
                     "bf16_speedup": round(f32 / b16, 3),
                     "f32_kernel_ms": [round(k, 4) for k in k32], "bf16_kernel_ms": [round(k, 4) for k in k16],
                     "rel_err_fp64_f32": e32, "rel_err_fp64_bf16": e16})
        log(f"c3 R={R}: f32 {f32:.3f} ms, bf16 {b16:.3f} ms, err f32 {max(e32):.2e} bf16 {max(e16):.2e}")
    r = runs[0]
    _line("All-mode MTTKRP ms & HBM GB/s, 4-mode precision sweep", r["f32_gbs"], "GB/s", r["f32_ms"], args,
          "f32", {"workload": f"all-mode MM-CSF MTTKRP, 100000x50000x20000x1000, {nnz} distinct nonzeros, "
                              "Zipf1.1 x U x Zipf1.1 x U, R=16 and 64, f32 vs bf16 factors",
                  "dims": dims, "nnz": nnz}, runs=runs, build_s=round(build_s, 2))


def _planted_values(torch, inds, A, noise, seed, chunk=4_000_000):
    """sum_r prod_m A_m[i_m][r] + noise * N(0,1), float64 math, float32 values."""
    n = inds[0].numel()
    g = torch.Generator(device=inds[0].device).manual_seed(seed)
    out = torch.empty(n, dtype=torch.float32, device=inds[0].device)
    for s in range(0, n, chunk):
        e = min(s + chunk, n)
        p = torch.ones((e - s, A[0].shape[1]), dtype=torch.float64, device=inds[0].device)
        for m, a in enumerate(A):
            p *= a[inds[m][s:e].long()]
        z = torch.randn(e - s, generator=g, device=inds[0].device, dtype=torch.float64)
        out[s:e] = (p.sum(1) + noise * z).float()
    return out


def _als64(torch, inds, vals, dims, U0, iters):
    """float64 CP-ALS with the update, normalisation and fit rules of cpd()
    (cpd_cpu.cu:484-691), on the GPU in torch, as the numerical reference."""
    N = len(dims)
    U = [u.double().clone() for u in U0]
    R = U[0].shape[1]
    ATA = [u.t() @ u for u in U]
    xn = float((vals.double() ** 2).sum())
    fits = []
    lam = torch.ones(R, dtype=torch.float64, device=vals.device)
    for it in range(iters):
        for m in range(N):
            M = fp64_reference(inds, vals, U, m, dims)
            if m == N - 1:
                Mlast = M
            G = torch.ones((R, R), dtype=torch.float64, device=vals.device)
            for n in range(N):
                if n != m:
                    G = G * ATA[n]
            U[m] = torch.linalg.solve(G, M.t()).t()
            lam = U[m].norm(dim=0) if it == 0 else torch.clamp(U[m].max(dim=0).values, min=1.0)
            U[m] = U[m] / lam
            ATA[m] = U[m].t() @ U[m]
        H = torch.ones((R, R), dtype=torch.float64, device=vals.device)
        for n in range(N):
            H = H * ATA[n]
        nm = abs(float(lam @ H @ lam))
        inner = float(((U[N - 1] * Mlast) @ lam).sum())
        res = xn + nm - 2 * inner
        fits.append(1 - (res ** 0.5 if res > 0 else res) / xn ** 0.5)
    return fits


def run_c4(args):
    """CP-ALS, 20 iterations at R=32 on a 200000x100000x20000 tensor with 20M
    nonzeros sampled from a planted rank-32 |N(0,1)| model + N(0, 0.01)
    noise (sigma 0.01); fits against a float64 ALS."""
    import torch

    import mmcsf

    dev = torch.device("cuda:0")
    dims, nnz, R, iters = [200000, 100000, 20000], args.nnz_c4, 32, 20
    seed = 0x4D4D43534603
    X0 = mmcsf.Tensor.generate(dims, nnz, [1, 0, 0], [1.2, 0.0, 0.0], seed)
    inds, _ = X0.to_torch(dev)
    del X0
    g = torch.Generator(device=dev).manual_seed(seed)
    A = [torch.randn((d, R), generator=g, device=dev, dtype=torch.float64).abs() for d in dims]
    vals = _planted_values(torch, inds, A, 0.01, seed + 1)
    X = mmcsf.Tensor.from_torch(dims, [a.int().contiguous() for a in inds], vals)
    fmt = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U0 = [torch.rand((d, R), generator=g, device=dev) for d in dims]
    U = [u.clone() for u in U0]
    lam, st = mmcsf.cpd_als(fmt, U, iters=iters, tol=0.0)
    Ub = [u.clone() for u in U0]
    _, stb = mmcsf.cpd_als(fmt, Ub, iters=iters, tol=0.0, mttkrp_bf16=True)
    fits64 = _als64(torch, inds, vals, dims, U0, iters)
    diff = max(abs(a - b) for a, b in zip(st["fit_history"], fits64))
    log(f"c4 fits gpu {st['fit_history'][-1]:.6f} fp64 {fits64[-1]:.6f} max diff {diff:.2e}")
    ms_it = st["ms_total"] / st["iterations"]
    _line("CP-ALS ms per iteration", round(ms_it, 3), "ms", ms_it, args, "f32",
          {"workload": f"CP-ALS R=32, 20 iterations, 200000x100000x20000, {nnz} nonzeros, planted rank-32 "
                       "|N(0,1)| + N(0,0.01) noise, Zipf1.2 x U x U", "dims": dims, "nnz": nnz, "R": R},
          higher=False, fit=st["fit"], fit_history=st["fit_history"], fit_fp64=fits64,
          max_abs_fit_diff_vs_fp64=diff, ms_mttkrp_total=st["ms_mttkrp"], lu_fallbacks=st["lu_fallbacks"],
          bf16_mttkrp={"fit": stb["fit"], "max_abs_fit_diff_vs_fp64":
                       max(abs(a - b) for a, b in zip(stb["fit_history"], fits64)),
                       "ms_per_iteration": stb["ms_total"] / stb["iterations"]})


def run_c5(args):
    """Differentiable CP completion: 1000000x500000x100000, 50M observed
    entries of a planted rank-16 N(0,1)/4 model + N(0, 0.01) noise, Zipf1.1 on
    modes 0-1; 100 Adam steps (lr 1e-2) at R=16 and R=64 with float32 and
    bfloat16 factors; gradient check against float64 central differences on a
    10k-nonzero subsample."""
    import torch

    import mmcsf

    dev = torch.device("cuda:0")
    dims, nnz = [1000000, 500000, 100000], args.nnz_c5
    seed = 0x4D4D43534604
    X0 = mmcsf.Tensor.generate(dims, nnz, [1, 1, 0], [1.1, 1.1, 0.0], seed)
    inds, _ = X0.to_torch(dev)
    del X0
    g = torch.Generator(device=dev).manual_seed(seed)
    A = [torch.randn((d, 16), generator=g, device=dev, dtype=torch.float64) / 4.0 for d in dims]
    vals = _planted_values(torch, inds, A, 0.01, seed + 1)
    X = mmcsf.Tensor.from_torch(dims, [a.int().contiguous() for a in inds], vals)
    fmt = mmcsf.Format(X, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    runs = []
    for R in (16, 64):
        g2 = torch.Generator(device=dev).manual_seed(seed + R)
        init = [0.3 * torch.randn((d, R), generator=g2, device=dev) for d in dims]
        for bf16 in (False, True):
            P = [u.clone().requires_grad_(True) for u in init]
            opt = torch.optim.Adam(P, lr=1e-2, fused=True)  # one pass per parameter tensor
            losses = []

            def step():
                opt.zero_grad(set_to_none=False)
                F = [p.to(torch.bfloat16) for p in P] if bf16 else P
                loss = mmcsf.cp_completion_loss(fmt, F)
                loss.backward()
                opt.step()
                return loss

            for _ in range(3):
                losses.append(float(step()))
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ls = [step() for _ in range(97)]
            e1.record()
            torch.cuda.synchronize()
            losses += [float(v) for v in ls]
            ms = e0.elapsed_time(e1) / 97
            # kernel-only loss + all-mode gradient at the final factors
            F = [p.detach().to(torch.bfloat16) for p in P] if bf16 else [p.detach() for p in P]
            kms = _event_ms(torch, lambda: mmcsf.cp_loss_grad(fmt, F), 5, 1)
            runs.append({"R": R, "factors": "bf16" if bf16 else "f32", "ms_per_adam_step": round(ms, 4),
                         "loss_grad_ms": round(kms, 4), "loss_first": losses[0], "loss_last": losses[-1],
                         "rmse_last": (2 * losses[-1] / nnz) ** 0.5})
            log(f"c5 R={R} {'bf16' if bf16 else 'f32'}: {ms:.3f} ms/step, loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    # gradient check on a 10k-nonzero subsample: <grad, D> vs float64 central differences
    sel = torch.randperm(nnz, generator=torch.Generator(device=dev).manual_seed(seed + 9), device=dev)[:10000]
    sub = [a[sel].contiguous() for a in inds]
    sv = vals[sel].contiguous()
    Xs = mmcsf.Tensor.from_torch(dims, [a.int() for a in sub], sv)
    fs = mmcsf.Format(Xs, mmcsf.TYPE_MMCSF, fbr_threshold=128)
    U = [0.3 * torch.randn((d, 16), generator=g, device=dev) for d in dims]
    _, grads = mmcsf.cp_loss_grad(fs, U)
    D = [torch.randn(u.shape, generator=g, device=dev, dtype=torch.float64) for u in U]

    def loss64(W):
        p = torch.ones((10000, 16), dtype=torch.float64, device=dev)
        for m, w in enumerate(W):
            p *= w[sub[m].long()]
        return 0.5 * float(((p.sum(1) - sv.double()) ** 2).sum())

    h = 1e-4
    U64 = [u.double() for u in U]
    fd = (loss64([u + h * d for u, d in zip(U64, D)]) - loss64([u - h * d for u, d in zip(U64, D)])) / (2 * h)
    an = sum(float((gr.double() * d).sum()) for gr, d in zip(grads, D))
    fd_rel = abs(an - fd) / abs(fd)
    log(f"c5 FD check: analytic {an:.6e} fd {fd:.6e} rel {fd_rel:.2e}")
    r = runs[0]
    _line("CP completion ms per Adam step", r["ms_per_adam_step"], "ms", r["ms_per_adam_step"], args, "f32",
          {"workload": f"CP completion, 1000000x500000x100000, {nnz} observed entries, planted rank-16 "
                       "N(0,1)/4 + N(0,0.01), Zipf1.1 x Zipf1.1 x U, 100 Adam steps lr 1e-2",
           "dims": dims, "nnz": nnz}, higher=False, runs=runs,
          fd_check={"sample_nnz": 10000, "R": 16, "analytic": an, "central_difference_fp64": fd,
                    "rel_err": fd_rel})


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nnz", type=int, default=100_000_000)
    ap.add_argument("--nnz-c3", type=int, default=50_000_000)
    ap.add_argument("--nnz-c4", type=int, default=20_000_000)
    ap.add_argument("--nnz-c5", type=int, default=50_000_000)
    ap.add_argument("--nnz-next", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile", action="store_true",
                    help="c2: only the timed MTTKRP launches (for rocprofv3 runs)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "next"],
                    help="c2 (default): the headline all-mode MTTKRP line; c1/c3/c4/c5: the other "
                         "configurations; next: the .tns reader, one-tree types and TTM; one JSON line each")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the multi-GPU path (slice-range shards per root tree) "
              "is not built, so only --gpus 1 runs", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5,
     "next": run_next}[args.workload](args)


if __name__ == "__main__":
    main()
