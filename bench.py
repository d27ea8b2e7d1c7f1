"""Benchmark: Tanimoto pairs/s of the LINGO similarity path on one MI355X.

Workload (one step): the full 32768 x 32768 float32 Tanimoto matrix of a
seeded synthetic set (strings U[20,120] over 'CNOSPFcnos()=#[]@+-123456',
seed 2) against itself, rows = columns, via getMultipleRows(0, 32768).  The
packed set is resident in HBM and the 4 GiB result stays in HBM; no host
copies are timed.

Prints ONE JSON line.  Besides the contract fields it carries
  roofline      the dominant kernel's algorithmic integer ops per launch
                (4 * (La + Lb) per pair) over its mean event-timed duration,
                against the VALU peak 256 CU * 64 lanes * 2.4 GHz
  cpu_baseline  the CPU oracle (C restatement of _CPULingo.c, OpenMP over
                rows) timed on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P_VALU = 256 * 64 * 2.4e9          # int32 lane-ops/s
B_HBM = 8.0e12                     # bytes/s
N_ALLPAIRS = 32768


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(s, budget_s=10.0):
    """Oracle pairs/s on R rows x all columns, R grown until ~budget_s."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    rows = 8
    while True:
        sub = tuple(x[:rows] for x in s)
        t0 = time.perf_counter()
        oracle.tanimoto_block(sub, s, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= budget_s or rows >= len(s[0]):
            break
        rows = min(len(s[0]), max(rows * 2, int(rows * budget_s / max(dt, 1e-3) * 1.05)))
    pairs = rows * len(s[0])
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{rows} rows x {len(s[0])} columns of the same workload, {dt:.1f} s",
            "seconds": round(dt, 3)}


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (hipMemcpy), the known-size
    reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_ALLPAIRS, help="molecules (rows = columns)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--policy", choices=["auto", "hash8", "merge"], default="auto",
                    help="auto: bit-sliced 28-row kernels; hash8: 8-row hashed kernels; "
                         "merge: one-row-per-workgroup merge kernel")
    ap.add_argument("--layout", choices=["block", "natural"], default="block",
                    help="query storage order: length-sorted 256-column blocks, or the caller's order")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_calibration(args.copy_bytes)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    from siml_amd import synth
    from siml_amd.GPULingo import GPULingo

    n = args.n
    log(f"packing {n} synthetic strings (U[20,120], seed 2)")
    s = synth.packed_padded(n, 20, 120, seed=2)
    lens = s[2].ravel().astype(np.int64)
    from siml_amd import _lib
    g = GPULingo(0)
    h = g._h
    h.call("siml_set_layout", _lib.LAYOUT_NATURAL if args.layout == "natural" else _lib.LAYOUT_BLOCK)
    h.call("siml_set_policy", {"auto": _lib.POLICY_AUTO, "hash8": _lib.POLICY_HASH8,
                               "merge": _lib.POLICY_MERGE}[args.policy])
    g.set_refsmiles(*s)
    g.set_qsmiles(*s)

    for i in range(args.warmup):
        h.call("siml_multirow_async", 0, n)
        h.call("siml_synchronize")
        log(f"warmup step {i + 1}/{args.warmup} done")
    h.call("siml_set_timing", 1)
    h.kernel_stats(reset=True)
    h.call("siml_synchronize")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.call("siml_multirow_async", 0, n)
    h.call("siml_synchronize")
    wall = time.perf_counter() - t0
    kms, launches, kpairs = h.kernel_stats(reset=True)
    kernel = h.last_kernel()
    log(f"timed {args.steps} steps: {wall * 1e3:.1f} ms wall, kernel {kms:.1f} ms over {launches} launches")

    pairs_per_step = n * n
    value = pairs_per_step * args.steps / wall
    kernel_s = kms / 1e3 / max(launches, 1)
    ops = 4.0 * (2 * n * lens.sum())            # sum over pairs of 4*(La+Lb), rows == columns
    bytes_alg = 4.0 * n * n + 8.0 * 2 * lens.sum() + 16.0 * n
    achieved = ops / kernel_s
    rec = {
        "metric": "Tanimoto pairs/s (1 GPU, bit-exact)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 merge, float32 division",
        "data": "synthetic (seeded SMILES-like strings, packed with the native packer)",
        "config": {"workload": f"all-pairs {n} x {n}, strings U[20,120], full float32 matrix in HBM",
                   "rows": n, "columns": n, "seed": 2, "policy": args.policy,
                   "layout": args.layout},
        "kernel": kernel,
        "kernel_ms_mean": kernel_s * 1e3,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": P_VALU / 1e12,
                     "unit": "Tops/s (int32 lane-ops)", "frac": achieved / P_VALU, "traffic": None,
                     "alg_ops_per_launch": ops, "alg_bytes_per_launch": bytes_alg,
                     "hbm_frac": bytes_alg / kernel_s / B_HBM},
    }
    if not args.no_cpu:
        log("timing the CPU oracle on a bounded sample")
        rec["cpu_baseline"] = cpu_baseline(s, args.cpu_budget)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
