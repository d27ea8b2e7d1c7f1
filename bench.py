"""Benchmark: Tanimoto pairs/s of the LINGO similarity path on one MI355X.

Default run (one JSON line):
  value       the all-pairs workload: the full 32768 x 32768 float32
              Tanimoto matrix of seeded strings (U[20,120] over
              'CNOSPFcnos()=#[]@+-123456', seed 2) against themselves via
              getMultipleRows(0, 32768); packed set and 4 GiB result resident
              in HBM, no host copies timed.  One step = one matrix.
  roofline    the all-pairs kernel against HBM (algorithmic bytes per
              launch over its mean event-timed duration, peak 8 TB/s) and,
              beside it, against its limiter, the LDS table probes (bytes
              the probes read over the same duration, against the LDS rate
              siml_selftest_lds measures on this device)
  batched     the batched-uneven-sets workload: 1M-molecule resident
              database (lognormal lengths, median 45, seed 3), 2000 query
              sets of geometric(1/128) sizes clipped to [1, 4096], top-100 per
              query molecule; "plain" = one call per set, "batched" = all sets
              in one call; pairs/s of each, their ratio, and the similarity
              kernel's roofline
  serving     the serving loop on a 10M-molecule resident database: Poisson
              arrivals of query batches (geometric(1/32) sizes clipped to
              [1, 512]), half top-50, half threshold 0.8, continuous batching
              over 2 slots; p50/p99 latency and pairs/s per arrival rate
  cpu_baseline  the CPU oracle (C restatement of _CPULingo.c, OpenMP over
              rows) on a bounded sample of the all-pairs workload, on the
              host threads allotted to this process and on one core

--quick leaves out the batched and serving parts.  Other workloads
(--workload, each one JSON line): zinc (the reference's demo4096.smi cycled
to 32768), cpu4096 (the reference's CPU case through CPULingo), batched,
threshold (10M database x 1024 near-duplicate queries: CSR for T >= 0.7 and
the 1000-bin int64 histogram), serve.
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_HBM = 8.0e12                     # bytes/s, MI355X HBM3E
# int32 VALU lane-operations/s: 256 CUs x 4 SIMDs x 16 lanes, two VALU
# instructions issued per SIMD per cycle (simple VOP2 forms), 2.4 GHz
VALU_PEAK = 256 * 4 * 16 * 2 * 2.4e9
N_CU, N_SIMD, N_SE = 256, 1024, 32   # shader engines: 8 XCDs x 4 (SQ_CYCLES is counted per engine)
N_ALLPAIRS = 32768
MAX_LEVEL = 15                     # levels per Lingo in the bit-sliced kernels


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Host threads allotted to this process: OMP_NUM_THREADS when set (the
    per-GPU allotment of a shared machine), else every visible core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def sources_sha():
    """Hash of the library's sources: recorded counter figures apply to the
    build whose sources hash the same."""
    h = hashlib.sha256()
    for p in sorted(glob.glob(os.path.join(ROOT, "siml_amd", "csrc", "*"))):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def recorded_counters(name):
    """The `counters:` JSON line of profiles/<name>.md (figures from
    rocprofv3 counter runs), or None."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("counters: "):
                    rec = json.loads(line[len("counters: "):])
                    rec["file"] = os.path.relpath(path, ROOT)
                    return rec
    except OSError:
        pass
    return None


def postings_met_self(s):
    """Postings the sparse path's rows meet when a padded set is compared
    with itself: sum over level keys (Lingo, t <= 15) of the number of
    molecules holding the level, squared."""
    lingos, counts, lengths = s[0], s[1], s[2].ravel()
    mask = np.arange(lingos.shape[1])[None, :] < lengths[:, None]
    ids = lingos[mask].astype(np.int64)
    cts = np.minimum(counts[mask].astype(np.int64), 15)
    keys = np.repeat(ids << 4, cts) + (np.arange(int(cts.sum())) - np.repeat(np.cumsum(cts) - cts, cts) + 1)
    _, n_k = np.unique(keys, return_counts=True)
    return int((n_k.astype(np.int64) ** 2).sum())


def level_keys_ragged(ids, counts, offsets):
    """Level keys (Lingo << 4 | t for t <= min(count, 15)) of a ragged set,
    ASCII Lingos only, and the molecule of each."""
    n = len(offsets) - 1
    ids, counts = ids[:offsets[-1]].astype(np.int64), counts[:offsets[-1]].astype(np.int64)
    cts = np.where((ids & 0x80808080) == 0, np.clip(counts, 0, MAX_LEVEL), 0)
    keys = np.repeat(ids << 4, cts) + (np.arange(int(cts.sum())) - np.repeat(np.cumsum(cts) - cts, cts) + 1)
    mol = np.repeat(np.repeat(np.arange(n), np.diff(offsets)), cts)
    return keys, mol


def postings_met_ragged(q, db):
    """Postings the sparse top-k's rows meet: over the query rows' level
    keys, the number of database molecules holding the key."""
    dk, _ = level_keys_ragged(*db[:3])
    uk, nk = np.unique(dk, return_counts=True)
    qk, _ = level_keys_ragged(*q[:3])
    pos = np.minimum(np.searchsorted(uk, qk), len(uk) - 1)
    return int(nk[pos][uk[pos] == qk].sum())


def levels_padded(lingos, counts, lengths):
    """Levels per molecule of a padded set: sum over its ASCII Lingos of
    min(count, 15), the column levels the bit-sliced kernels probe."""
    w = lingos.shape[1]
    live = np.arange(w)[None, :] < lengths.reshape(-1, 1)
    ascii_ = (lingos.astype(np.int64) & 0x80808080) == 0
    c = np.where(live & ascii_ & (counts > 0), np.minimum(counts, MAX_LEVEL), 0)
    return c.sum(axis=1).astype(np.int64)


def levels_ragged(ids, counts, offsets):
    n = len(offsets) - 1
    ids, counts = ids[:offsets[-1]], counts[:offsets[-1]]
    ascii_ = (ids.astype(np.int64) & 0x80808080) == 0
    c = np.where(ascii_ & (counts > 0), np.minimum(counts, MAX_LEVEL), 0).astype(np.int64)
    cs = np.zeros(len(c) + 1, np.int64)
    np.cumsum(c, out=cs[1:])
    return cs[offsets[1:]] - cs[offsets[:-1]] if n else np.zeros(0, np.int64)


def device_peaks():
    """Measured LDS read rates (bytes/s) of this device: 8-byte reads from a
    32 KiB table without bank conflicts and at random slots."""
    import ctypes
    from siml_amd import _lib
    lib = _lib.load()
    out = {}
    for op, name in ((0, "lds_conflict_free"), (1, "lds_random_slots")):
        v = ctypes.c_double()
        _lib.check(lib.siml_selftest_lds(0, op, ctypes.byref(v)))
        out[name] = v.value
    return out


def _oracle_rate(s, threads, budget_s):
    """Oracle pairs/s on R rows x all columns, R grown until ~budget_s."""
    import oracle
    rows = 8
    while True:
        sub = tuple(x[:rows] for x in s)
        t0 = time.perf_counter()
        oracle.tanimoto_block(sub, s, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= budget_s or rows >= len(s[0]):
            break
        rows = min(len(s[0]), max(rows * 2, int(rows * budget_s / max(dt, 1e-3) * 1.05)))
    return rows, dt


def cpu_baseline(s, budget_s=10.0):
    """The oracle's C restatement of _CPULingo.c timed on this host: on the
    host threads allotted to this process (the reported value), and on one
    core."""
    threads = host_threads()
    rows, dt = _oracle_rate(s, threads, budget_s)
    rows1, dt1 = _oracle_rate(s, 1, budget_s / 4)
    n = len(s[0])
    return {"value": rows * n / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{rows} rows x {n} columns of the same workload, {dt:.1f} s",
            "seconds": round(dt, 3), "value_1core": rows1 * n / dt1,
            "sample_1core": f"{rows1} rows x {n} columns, {dt1:.1f} s", "cpu_model": cpu_model(),
            "cores_visible": os.cpu_count(),
            "thread_policy": "OMP_NUM_THREADS when set (the host-thread allotment of this GPU), else all cores"}


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


def configure(g, args):
    from siml_amd import _lib
    h = g._h
    h.call("siml_set_layout", {"natural": _lib.LAYOUT_NATURAL, "block": _lib.LAYOUT_BLOCK,
                               "sorted": _lib.LAYOUT_SORTED}[args.layout])
    h.call("siml_set_policy", {"auto": _lib.POLICY_AUTO, "merge": _lib.POLICY_MERGE, "bits": _lib.POLICY_BITS,
                               "sparse": _lib.POLICY_SPARSE}[args.policy])
    g.set_group_tables(args.group_tables)


def base_record(args, metric, value, unit, wall, config, dtype="int32 intersections, float32 division", steps=None,
                warmup=None):
    """`wall` is the timed region in seconds over `steps` steps (default
    --steps)."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": wall * 1e3 / max(steps, 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded SMILES-like strings, packed with the native packer)", "config": config}


def roofline(alg_bytes, probe_bytes, kernel_s, peaks, counters=None):
    """Roofline object of one kernel: HBM (algorithmic bytes per launch over
    its duration, peak 8 TB/s; `traffic` = measured HBM bytes per launch
    from counter runs of the same sources) and the LDS table probes (bytes
    read per launch over the same duration, against the measured rates)."""
    sha = sources_sha()
    traffic, note = None, None
    if counters is not None:
        if counters.get("sources_sha") == sha:
            traffic = counters["fetch_kib"] * 1024 * counters["fetch_factor"] + \
                counters["write_kib"] * 1024 * counters["write_factor"]
        else:
            note = f"{counters['file']} was measured on sources {counters.get('sources_sha')}, not {sha}"
    achieved = alg_bytes / kernel_s
    rec = {"bound": "hbm", "achieved": achieved / 1e9, "peak": B_HBM / 1e9, "unit": "GB/s",
           "frac": achieved / B_HBM, "traffic": traffic,
           "alg_bytes_per_launch": alg_bytes, "kernel_ms": kernel_s * 1e3,
           "hbm_measured_frac": (traffic / kernel_s / B_HBM) if traffic else None,
           "lds": {}, "sources_sha": sha}
    if probe_bytes is not None:
        # table probes of the bit-sliced kernels
        lds = probe_bytes / kernel_s
        rec["lds"] = {"achieved": lds / 1e9, "unit": "GB/s", "probe_bytes_per_launch": probe_bytes,
                      "peak_conflict_free": peaks["lds_conflict_free"] / 1e9,
                      "frac": lds / peaks["lds_conflict_free"],
                      "peak_random_slots": peaks["lds_random_slots"] / 1e9,
                      "frac_of_random_slot_rate": lds / peaks["lds_random_slots"]}
    if counters is not None and counters.get("sources_sha") == sha:
        rec["counters"] = {k: v for k, v in counters.items() if k not in ("fetch_kib", "write_kib")}
        c = counters
        if "sq_cycles" in c:
            # utilisation over the counter run's own launch: SQ_CYCLES per
            # shader engine; VALU quad-cycles with one or two instructions
            # issued (SQ_ACTIVE_INST_VALU counts each instruction once, one
            # quad-cycle each; SQ_ACTIVE_INST_VALU2 the quad-cycles that
            # issued two) over every SIMD's cycles
            cyc = c["sq_cycles"] / N_SE
            rec["valu"] = {
                "busy_frac": (c["valu_active"] - c["valu_active2"]) * 4.0 / (N_SIMD * cyc),
                "dual_issue_frac": 2.0 * c["valu_active2"] / c["valu_active"],
                "insts_per_launch": c["valu_insts"],
                "lane_ops_per_s": c["valu_insts"] * 64.0 / kernel_s,
                "peak_lane_ops_per_s": VALU_PEAK,
                "lane_ops_frac": c["valu_insts"] * 64.0 / kernel_s / VALU_PEAK}
            rec["lds"]["busy_frac"] = c["lds_idx_active"] / (N_CU * cyc)
            rec["lds"]["bank_conflict_frac"] = c["lds_bank_conflict"] / (N_CU * cyc)
            rec["waves_per_simd"] = c["wave_cycles"] * 4.0 / (N_SIMD * cyc)
    if note:
        rec["traffic_note"] = note
    return rec


# ---------------------------------------------------------------------------
# all-pairs 32768 x 32768 (the value of the default line)
# ---------------------------------------------------------------------------
def workload_allpairs(args):
    from siml_amd import _lib, synth
    from siml_amd.GPULingo import GPULingo

    n = args.n
    log(f"packing {n} synthetic strings (U[20,120], seed 2)")
    s = synth.packed_padded(n, 20, 120, seed=2)
    g = GPULingo(0)
    h = g._h
    configure(g, args)
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
    kms, launches, _ = h.kernel_stats(reset=True)
    kernel = h.last_kernel()
    log(f"timed {args.steps} steps: {wall * 1e3:.1f} ms wall, kernel {kms:.1f} ms over {launches} launches")

    value = n * n * args.steps / wall
    kernel_s = kms / 1e3 / max(launches, 1)
    lv = levels_padded(s[0], s[1], s[2])
    groups, merged = h.row_groups()
    peaks = device_peaks()
    if kernel.startswith("rows_sparse"):
        # sparse path: the float32 matrix written once, the postings read
        # once (4 B per query level), the rows' elements (8 B per Lingo) and
        # 4 B of column magnitudes; the rows' levels meet sum_k n_k^2
        # postings (n_k: query molecules holding level key k; rows ==
        # columns here), each an LDS add
        alg_bytes = 4.0 * n * n + 4.0 * lv.sum() + 8.0 * int(s[2].sum()) + 4.0 * n
        probe_bytes = None
        postings_met = postings_met_self(s)
        profile = "k_sparse_f32_allpairs32768"
    else:
        # bit-sliced: the float32 matrix written once, the column level
        # stream read once (8 B per level), the row tables (16 B per table
        # entry: at most one per row level) and 8 B of per-column vectors
        alg_bytes = 4.0 * n * n + 8.0 * lv.sum() + 16.0 * lv.sum() + 8.0 * n
        probe_bytes = 16.0 * groups * lv.sum()
        postings_met = None
        profile = "k_bits_f32block_allpairs32768"
    counters = recorded_counters(profile) if (n == N_ALLPAIRS and args.layout == "block") else None
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact)", value, "pairs/s", wall,
                      {"workload": f"all-pairs {n} x {n}, strings U[20,120], full float32 matrix in HBM",
                       "rows": n, "columns": n, "seed": 2, "policy": args.policy, "layout": args.layout})
    rec.update({"kernel": kernel, "kernel_ms_mean": kernel_s * 1e3, "row_groups": groups, "merge_rows": merged,
                "roofline": roofline(alg_bytes, probe_bytes, kernel_s, peaks, counters)})
    if postings_met is not None:
        rec["postings_met_per_launch"] = postings_met
    if not args.no_cpu:
        import oracle
        rows = [0, 1, n // 3, n // 2, n - 2, n - 1]
        peek = np.empty((len(rows), n), np.float32)
        for i, r in enumerate(rows):
            h.call("siml_peek_rows_f32", r, r + 1, _lib.f32p(peek[i:i + 1]), n)
        want = oracle.tanimoto_block(tuple(x[rows] for x in s), s, nthreads=host_threads())
        rec["parity"] = {"rows_checked": rows, "bit_exact": bool(np.array_equal(peek.view(np.int32),
                                                                                 want.view(np.int32)))}
        log("timing the CPU oracle on a bounded sample")
        rec["cpu_baseline"] = cpu_baseline(s, args.cpu_budget)
    del g
    return rec


# ---------------------------------------------------------------------------
# real SMILES: the reference's 4096 ZINC molecules, cycled to 32768
# ---------------------------------------------------------------------------
def workload_zinc(args):
    import oracle
    from siml_amd import _lib, compiler
    from siml_amd.GPULingo import GPULingo

    path = os.path.join(ROOT, "tests", "golden", "demo4096.smi")
    with open(path, "rb") as f:
        lines = f.readlines()
    n = args.n
    lines = [lines[i % len(lines)] for i in range(n)]
    s = compiler.cSMILEStoMatrices(lines)[:4]
    lens = s[2].ravel().astype(np.int64)
    g = GPULingo(0)
    configure(g, args)
    g.set_refsmiles(*s)
    g.set_qsmiles(*s)
    h = g._h
    if args.autotune:
        g.set_autotune(True)
        g.getTopK(0, n, 100)            # tunes the top-k shape before it is timed
    for _ in range(args.warmup):
        h.call("siml_multirow_async", 0, n)
        h.call("siml_synchronize")
    h.call("siml_set_timing", 1)
    h.kernel_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.call("siml_multirow_async", 0, n)
    h.call("siml_synchronize")
    wall = time.perf_counter() - t0
    kms, launches, _ = h.kernel_stats(reset=True)
    rows = [0, 1, n // 2, n - 1]
    peek = np.empty((len(rows), n), np.float32)
    for i, r in enumerate(rows):
        h.call("siml_peek_rows_f32", r, r + 1, _lib.f32p(peek[i:i + 1]), n)
    want = oracle.tanimoto_block(tuple(x[rows] for x in s), s, nthreads=host_threads())
    t0 = time.perf_counter()
    ti, tt = g.getTopK(0, n, 100)
    t_topk = time.perf_counter() - t0
    oi, ot = oracle.topk(tuple(x[rows] for x in s), s, 100, nthreads=host_threads())
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact), real SMILES", n * n * args.steps / wall,
                      "pairs/s", wall, {"workload": f"all-pairs {n} x {n} of the reference's demo4096.smi ZINC "
                                                    f"molecules cycled, full float32 matrix in HBM", "rows": n,
                                        "columns": n})
    rec["data"] = "examples/demo4096.smi of the reference (4096 ZINC SMILES), cycled"
    rec.update({"kernel": h.last_kernel(), "kernel_ms_mean": kms / max(launches, 1),
                "mean_unique_lingos": float(lens.mean()),
                "topk100_s": t_topk, "topk100_pairs_per_s": n * n / t_topk,
                "parity": {"rows_checked": rows,
                           "bit_exact": bool(np.array_equal(peek.view(np.int32), want.view(np.int32))),
                           "topk_equal": bool(np.array_equal(ti[rows], oi) and
                                              np.array_equal(tt[rows].view(np.int32), ot.view(np.int32)))}})
    return rec


# ---------------------------------------------------------------------------
# CPULingo all-pairs 4096 x 4096 (the reference's own CPU-runnable case)
# ---------------------------------------------------------------------------
def workload_cpu4096(args):
    import oracle
    from siml_amd import synth
    from siml_amd.CPULingo import CPULingo

    n = 4096
    log("packing 4096 synthetic strings (U[20,80], seed 1)")
    L, C, lens, mags = synth.packed_padded(n, 20, 80, seed=1)
    cpu = CPULingo()
    cpu.set_refsmiles(L, C, lens)      # default magnitudes: the reference's lengths.sum(1)
    cpu.set_qsmiles(L, C, lens)

    def run():
        # examples/benchmark.py:123-167: the N^2 matrix in 256-row blocks
        return np.concatenate([cpu.getMultipleRows(r, min(n, r + 256)) for r in range(0, n, 256)])

    for _ in range(args.warmup):
        run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got = run()
    wall = (time.perf_counter() - t0) / args.steps
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact), CPULingo interface", n * n / wall, "pairs/s",
                      wall * args.steps, {"workload": "CPULingo all-pairs 4096 x 4096, strings U[20,80], 256-row "
                                                      "blocks, default (reference) magnitudes", "rows": n,
                                          "columns": n, "seed": 1})
    if not args.no_cpu:
        q = (L, C, lens, np.int32(lens.sum(1)))
        threads = host_threads()
        t0 = time.perf_counter()
        want = oracle.tanimoto_block(q, q, nthreads=threads)
        dt = time.perf_counter() - t0
        rows1 = 512
        t0 = time.perf_counter()
        oracle.tanimoto_block(tuple(x[:rows1] for x in q), q, nthreads=1)
        dt1 = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": n * n / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
                               "sample": f"the full {n} x {n} matrix, {dt:.2f} s", "seconds": round(dt, 3),
                               "value_1core": rows1 * n / dt1, "sample_1core": f"{rows1} rows x {n} columns, {dt1:.2f} s",
                               "cpu_model": cpu_model(), "cores_visible": os.cpu_count()}
        rec["bit_exact"] = bool(np.array_equal(got.view(np.int32), want.view(np.int32)))
    return rec


# ---------------------------------------------------------------------------
# batched uneven query sets, top-k
# ---------------------------------------------------------------------------
def lognormal_set(n, seed):
    from siml_amd import synth
    return synth.packed_ragged(n, 8, 256, seed=seed, dist="lognormal", median=45.0, sigma=0.35)


def oracle_topk_ragged(q, db, rows, k, chunk=1 << 20):
    """Oracle top-k of query rows against a ragged database, column chunks
    merged with the (T desc, index asc) order."""
    import oracle
    from siml_amd import synth
    Q = synth.padded_columns(*q, 0, len(q[2]) - 1)
    Q = tuple(x[rows] for x in Q)
    n = len(db[2]) - 1
    best_i = np.zeros((len(rows), 0), np.int32)
    best_t = np.zeros((len(rows), 0), np.float32)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        D = synth.padded_columns(*db, c0, c1)
        oi, ot = oracle.topk(Q, D, min(k, c1 - c0), nthreads=host_threads())
        best_i = np.concatenate([best_i, oi + c0], axis=1)
        best_t = np.concatenate([best_t, ot], axis=1)
        order = np.lexsort((best_i, -best_t.astype(np.float64)), axis=1)[:, :k]
        best_i = np.take_along_axis(best_i, order, 1)
        best_t = np.take_along_axis(best_t, order, 1)
    return best_i, best_t


def batched_part(args, check_oracle=False, peaks=None):
    """Plain (one call per set) and batched (one call) top-100 over the 1M
    database; both end with the results on the host."""
    from siml_amd import synth
    from siml_amd.GPULingo import GPULingo, LingoDB

    ndb = args.db or 1_000_000
    nsets, k = args.sets, 100
    log(f"batched: packing the {ndb}-molecule database (lognormal, seed 3)")
    db_set = lognormal_set(ndb, 3)
    sizes = synth.geometric_sizes(nsets, 1.0 / 128, 1, 4096, seed=3)
    nq = int(sizes.sum())
    q = lognormal_set(nq, 33)
    bounds = np.zeros(nsets + 1, np.int64)
    np.cumsum(sizes, out=bounds[1:])
    sets = [synth.ragged_slice(*q, int(bounds[i]), int(bounds[i + 1])) for i in range(nsets)]
    db = LingoDB(*db_set)
    g = GPULingo(0)
    configure(g, args)
    g.attach_db(db)
    h = g._h
    if args.launch_shape:
        g.set_launch_shape(*[int(x) for x in args.launch_shape.split(",")])
    elif args.autotune:
        g.set_autotune(True)
    # warm-up (and, with the autotuner, shape tuning): every set once and
    # the batched call once
    log("batched: warm-up pass over the sets and the batched call")
    if not args.no_plain:
        for s in sets:
            g.set_refsmiles_ragged(*s)
            g.getTopK(0, len(s[2]) - 1, k)
    g.set_refsmiles_ragged(*q)
    g.getTopK(0, nq, k)
    pairs = nq * ndb

    h.call("siml_set_timing", 1)
    h.kernel_stats(reset=True)
    plain = []
    t_plain = sim_plain = sel_plain = kms_plain = float("nan")
    if not args.no_plain:
        log("batched: plain, one getTopK per query set")
        t0 = time.perf_counter()
        for s in sets:
            g.set_refsmiles_ragged(*s)
            plain.append(g.getTopK(0, len(s[2]) - 1, k))
        t_plain = time.perf_counter() - t0
        sim_plain, sel_plain = h.kernel_split()
        kms_plain, n_plain, _ = h.kernel_stats(reset=True)

    log("batched: all sets as one launch sequence")
    t0 = time.perf_counter()
    g.set_refsmiles_ragged(*q)
    bi, bt = g.getTopK(0, nq, k)
    t_batched = time.perf_counter() - t0
    sim_b, sel_b = h.kernel_split()
    kms_b, n_b, _ = h.kernel_stats(reset=True)
    kernel = h.last_kernel()
    shape = g.launch_shape()
    groups, merged = h.row_groups()
    log(f"batched: plain {t_plain:.2f} s ({kms_plain:.0f} ms kernels), batched {t_batched:.2f} s "
        f"({kms_b:.0f} ms kernels: {sim_b:.0f} similarity + {sel_b:.0f} selection)")

    same = None
    if plain:
        pi = np.concatenate([p[0] for p in plain])
        pt = np.concatenate([p[1] for p in plain])
        same = bool(np.array_equal(pi, bi) and np.array_equal(pt.view(np.int32), bt.view(np.int32)))
    graphs = graph_calls(args, g, sets, q, k, pairs, (bi, bt)) if args.graphs else None
    persistent = persistent_call(g, q, k, pairs, (bi, bt)) if args.graphs else None
    out = {"workload": f"{nsets} query sets (geometric 1/128 sizes clipped [1, 4096], {nq} molecules) x "
                       f"{ndb}-molecule resident database (lognormal lengths, median 45), top-{k} per query "
                       f"molecule", "sets": nsets, "queries": nq, "database": ndb, "k": k, "seed": 3,
           "pairs": pairs,
           "plain": {"wall_s": t_plain, "pairs_per_s": pairs / t_plain, "kernel_ms": kms_plain,
                     "similarity_ms": sim_plain, "selection_ms": sel_plain, "calls": nsets},
           "batched": {"wall_s": t_batched, "pairs_per_s": pairs / t_batched, "kernel_ms": kms_b,
                       "similarity_ms": sim_b, "selection_ms": sel_b, "calls": 1},
           "speedup_wall": t_plain / t_batched, "speedup_kernel": kms_plain / max(kms_b, 1e-9),
           "speedup_similarity_kernel": sim_plain / max(sim_b, 1e-9),
           "plain_equals_batched": same, "graphs": graphs, "persistent_grid": persistent, "kernel": kernel,
           "autotune": bool(args.autotune),
           "batched_launch_shape": {"target_workgroups": shape[0], "topk_wave_cap": shape[1],
                                    "topk_seed": shape[2]}}
    if peaks is not None:
        lv_db = levels_ragged(*db_set[:3])
        lv_q = levels_ragged(*q[:3])
        out["row_groups"] = groups
        out["merge_rows"] = merged
        if kernel == "topk_sparse":
            # sparse top-k: the postings the query rows' levels meet (4 B
            # each), a 16-byte index slot per query level and the top-k of
            # each row (8 B per entry) written once
            met = postings_met_ragged(q, db_set)
            alg = 4.0 * met + 16.0 * lv_q.sum() + 8.0 * k * nq
            out["postings_met_per_launch"] = met
            out["roofline"] = roofline(alg, None, kms_b / 1e3, peaks, recorded_counters("k_sparse_topk_batched1M"))
        else:
            # algorithmic HBM bytes of the batched similarity kernel: the
            # database level stream once (8 B per level), the query tables
            # (16 B per entry) and 8 B per candidate key for k keys per row
            alg = 8.0 * lv_db.sum() + 16.0 * lv_q.sum() + 8.0 * k * nq
            out["roofline"] = roofline(alg, 16.0 * groups * lv_db.sum(), sim_b / 1e3, peaks,
                                       recorded_counters("k_bits_topk_batched1M"))
    if check_oracle:
        rng = np.random.default_rng(7)
        rows = np.sort(rng.choice(nq, size=min(6, nq), replace=False))
        log(f"batched: oracle top-{k} of query rows {rows.tolist()} over the whole database")
        oi, ot = oracle_topk_ragged(q, db_set, rows, k)
        out["oracle_rows_match"] = bool(np.array_equal(oi, bi[rows]) and
                                        np.array_equal(ot.view(np.int32), bt[rows].view(np.int32)))
    del g
    db.close()
    return out


def graph_calls(args, g, sets, q, k, pairs, eager):
    """The plain and batched calls again with graph mode on (each call's
    kernel sequence launched as one HIP graph): times, and whether the
    batched result equals the eager one."""
    h = g._h
    log("batched: the same calls in graph mode")
    g.set_graphs(True)
    h.call("siml_set_timing", 0)
    # warm-up: instantiates the graphs of every kernel sequence the calls use
    if not args.no_plain:
        for s in sets:
            g.set_refsmiles_ragged(*s)
            g.getTopK(0, len(s[2]) - 1, k)
    g.set_refsmiles_ragged(*q)
    g.getTopK(0, len(q[2]) - 1, k)
    h.call("siml_set_timing", 1)
    h.kernel_stats(reset=True)
    out = {}
    if not args.no_plain:
        t0 = time.perf_counter()
        for s in sets:
            g.set_refsmiles_ragged(*s)
            g.getTopK(0, len(s[2]) - 1, k)
        wall = time.perf_counter() - t0
        kms, _, _ = h.kernel_stats(reset=True)
        out["plain"] = {"wall_s": wall, "pairs_per_s": pairs / wall, "kernel_ms": kms, "calls": len(sets)}
    t0 = time.perf_counter()
    g.set_refsmiles_ragged(*q)
    gi, gt = g.getTopK(0, len(q[2]) - 1, k)
    wall = time.perf_counter() - t0
    kms, _, _ = h.kernel_stats(reset=True)
    out["batched"] = {"wall_s": wall, "pairs_per_s": pairs / wall, "kernel_ms": kms, "calls": 1}
    out["batched_equals_eager"] = bool(np.array_equal(gi, eager[0]) and
                                       np.array_equal(gt.view(np.int32), eager[1].view(np.int32)))
    out["graph_launches"], out["graph_instantiations"] = g.graph_stats()
    g.set_graphs(False)
    log(f"batched, graph mode: plain {out.get('plain', {}).get('wall_s', float('nan')):.2f} s, "
        f"batched {wall:.2f} s ({kms:.0f} ms kernels), equal {out['batched_equals_eager']}")
    return out


def persistent_call(g, q, k, pairs, eager):
    """The batched call at the same launch shape with the grid mode
    switched (persistent grid with per-XCD work queues <-> ordinary grid):
    time, and whether the result equals the first one."""
    h = g._h
    shape = g.launch_shape()
    out = {}
    for mode in (True, False):
        g.set_launch_shape(*shape)
        g.set_persistent(mode)
        g.set_refsmiles_ragged(*q)
        g.getTopK(0, len(q[2]) - 1, k)          # warm-up at this mode
        h.call("siml_set_timing", 1)
        h.kernel_stats(reset=True)
        t0 = time.perf_counter()
        g.set_refsmiles_ragged(*q)
        pi, pt = g.getTopK(0, len(q[2]) - 1, k)
        wall = time.perf_counter() - t0
        sim, sel = h.kernel_split()
        kms, _, _ = h.kernel_stats(reset=True)
        key = "persistent" if mode else "ordinary"
        out[key] = {"wall_s": wall, "pairs_per_s": pairs / wall, "kernel_ms": kms, "similarity_ms": sim,
                    "selection_ms": sel,
                    "equals_first": bool(np.array_equal(pi, eager[0]) and
                                         np.array_equal(pt.view(np.int32), eager[1].view(np.int32)))}
        log(f"batched, {key} grid at shape {shape}: {wall:.3f} s ({kms:.0f} ms kernels, {sim:.0f} similarity)")
    g.set_persistent(False)
    return out


def workload_batched(args):
    part = batched_part(args, check_oracle=not args.no_cpu, peaks=device_peaks())
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact), batched top-k", part["batched"]["pairs_per_s"],
                      "pairs/s", part["batched"]["wall_s"],
                      {"workload": "batched uneven query sets: " + part["workload"]}, steps=1, warmup=1)
    rec.update(part)
    return rec


# ---------------------------------------------------------------------------
# threshold neighbours + global histogram
# ---------------------------------------------------------------------------
def near_duplicate_queries(nq, db_buf, db_offs, seed):
    from siml_amd import compiler, synth
    rng = np.random.default_rng(seed)
    picks = rng.choice(len(db_offs) - 1, size=nq, replace=False)
    buf, offs = synth.mutated_strings(db_buf, db_offs, picks, 2, seed)
    return compiler.pack_ragged_buffer(buf, offs)


def workload_threshold(args):
    from siml_amd import compiler, synth
    from siml_amd.GPULingo import GPULingo, LingoDB
    import oracle

    ndb, nq, lo = args.db or 10_000_000, 1024, 0.7
    log(f"packing the {ndb}-molecule database (lognormal, seed 4)")
    buf, offs = synth.random_strings(ndb, 8, 256, seed=4, dist="lognormal", median=45.0, sigma=0.35)
    db_set = compiler.pack_ragged_buffer(buf, offs)
    q = near_duplicate_queries(nq, buf, offs, 44)
    del buf
    db = LingoDB(*db_set)
    g = GPULingo(0)
    configure(g, args)
    g.attach_db(db)
    g.set_refsmiles_ragged(*q)
    h = g._h
    g.getNeighborsCSR(0, 32, lo)
    g.getGlobalHistogram(0, 32)
    pairs = nq * ndb
    peaks = device_peaks()
    h.call("siml_set_timing", 1)
    h.kernel_stats(reset=True)
    h.work_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        roff, cols = g.getNeighborsCSR(0, nq, lo)
    t_csr = (time.perf_counter() - t0) / args.steps
    kms_csr, n_csr, _ = h.kernel_stats(reset=True)
    k_csr = h.last_kernel()
    # pairs the magnitude band did not prune, counted on the device (summed
    # over the count pass and, when one runs, the fill pass)
    ev_pairs, ev_len = h.work_stats(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hist = g.getGlobalHistogram(0, nq)
    t_hist = (time.perf_counter() - t0) / args.steps
    kms_hist, n_hist, _ = h.kernel_stats(reset=True)
    k_hist = h.last_kernel()
    groups, merged = h.row_groups()
    log(f"CSR: {t_csr * 1e3:.1f} ms ({int(roff[-1])} neighbours); hist: {t_hist * 1e3:.1f} ms")
    lv_db = levels_ragged(*db_set[:3])
    lv_q = levels_ragged(*q[:3])
    hist_s = kms_hist / 1e3 / args.steps
    checks = {"hist_total": int(hist.sum()), "hist_total_ok": int(hist.sum()) == pairs, "nnz": int(roff[-1])}
    if not args.no_cpu:
        rows = [0, 1, nq // 2, nq - 1]
        log(f"oracle neighbours and histogram of query rows {rows} over the whole database")
        Q = tuple(x[rows] for x in synth.padded_columns(*q, 0, nq))
        nb = [[] for _ in rows]
        oh = np.zeros(1000, np.int64)
        for c0 in range(0, ndb, 1 << 20):
            c1 = min(ndb, c0 + (1 << 20))
            D = synth.padded_columns(*db_set, c0, c1)
            lst, _ = oracle.neighbors(Q, D, lo, nthreads=host_threads())
            for i, l in enumerate(lst):
                nb[i].append(l + c0)
            oh += oracle.hist1000(Q, D)
        ok_nb = all(np.array_equal(np.concatenate(nb[i]), cols[roff[r]:roff[r + 1]]) for i, r in enumerate(rows))
        gh = np.zeros(1000, np.int64)
        for r in rows:
            gh += g.getGlobalHistogram(r, r + 1)
        checks.update({"oracle_rows": rows, "neighbours_match": bool(ok_nb), "hist_rows_match": bool(np.array_equal(gh, oh))})
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact), threshold CSR", pairs / t_csr, "pairs/s", t_csr * args.steps,
                      {"workload": f"threshold neighbours T >= {lo} + 1000-bin histogram: {nq} near-duplicate "
                                   f"queries x {ndb}-molecule resident database", "queries": nq, "database": ndb,
                       "seed": 4})
    if k_hist == "hist1000_sparse":
        # sparse path: the postings the query rows' levels meet (4 B each;
        # estimated from the first tenth of the database, whose molecules
        # are drawn like the rest), a 16-byte index slot per query level and
        # the bins
        tenth = max(ndb // 10, 1)
        met = postings_met_ragged(q, synth.ragged_slice(*db_set, 0, tenth)) * (ndb / tenth)
        hist_alg = 4.0 * met + 16.0 * lv_q.sum() + 8000.0
        hist_roof = roofline(hist_alg, None, hist_s, peaks, recorded_counters("k_sparse_hist1000_db10M"))
        hist_roof["postings_met_estimate"] = met
    else:
        # the histogram evaluates every pair: its probes are every group
        # against every database level
        hist_alg = 8.0 * lv_db.sum() + 16.0 * lv_q.sum() + 8000.0
        hist_roof = roofline(hist_alg, 16.0 * groups * lv_db.sum(), hist_s, peaks,
                             recorded_counters("k_bits_hist1000_db10M"))
    rec.update({"csr": {"wall_ms": t_csr * 1e3, "pairs_per_s": pairs / t_csr, "kernel_ms": kms_csr / args.steps,
                        "kernel": k_csr, "evaluated_pairs_per_call": ev_pairs / args.steps,
                        "evaluated_frac": ev_pairs / args.steps / pairs},
                "hist1000": {"wall_ms": t_hist * 1e3, "pairs_per_s": pairs / t_hist,
                             "kernel_ms": kms_hist / args.steps, "kernel": k_hist, "roofline": hist_roof},
                "checks": checks})
    rec["roofline"] = rec["hist1000"]["roofline"]
    db.close()
    return rec


# ---------------------------------------------------------------------------
# serving loop
# ---------------------------------------------------------------------------
def serving_part(args, rates, requests_for):
    """The serving loop at each arrival rate: p50/p99 latency and pairs/s
    (and again in graph mode)."""
    from siml_amd import synth
    from siml_amd.GPULingo import LingoDB

    ndb = args.db_serve or 10_000_000
    log(f"serving: packing the {ndb}-molecule database (lognormal, seed 5)")
    db_set = lognormal_set(ndb, 5)
    db = LingoDB(*db_set)
    nmax = max(requests_for(r) for r in rates)
    sizes = synth.geometric_sizes(nmax, 1.0 / 32, 1, 512, seed=5)
    pool = lognormal_set(int(sizes.sum()), 55)
    bounds = np.zeros(nmax + 1, np.int64)
    np.cumsum(sizes, out=bounds[1:])
    modes = {}
    for graphs in ((False, True) if args.graphs else (False,)):
        modes[graphs] = serve_rates(args, db, pool, sizes, bounds, rates, requests_for, ndb, graphs)
    results = modes[False]
    ok = [r for r in results if r["p99_ms"] < 50.0]
    best = max(ok or results, key=lambda r: r["pairs_per_s"])
    db.close()
    return {"workload": f"{ndb}-molecule resident database, Poisson arrivals of query batches (geometric 1/32 "
                        f"sizes clipped [1, 512], lognormal lengths), half top-50, half threshold 0.8",
            "database": ndb, "slots": args.slots, "max_rows": args.max_rows, "seed": 5, "rates": results,
            "graphs": ({"what": "the same arrivals with graph mode on (top-k batches launched as HIP graphs)",
                        "rates": modes[True]} if True in modes else None),
            "best_rate_with_p99_below_50ms": best["rate_per_s"] if ok else None,
            "best_pairs_per_s_with_p99_below_50ms": best["pairs_per_s"] if ok else None}


def serve_rates(args, db, pool, sizes, bounds, rates, requests_for, ndb, graphs):
    """One server over `db` (graph mode or not): p50/p99 per arrival rate."""
    from siml_amd import synth
    from siml_amd.serving import Request, Server, run_poisson

    results = []
    tag = "serving (graph mode)" if graphs else "serving"
    with Server(db, slots=args.slots, max_rows=args.max_rows, mem_budget=args.mem_budget << 30,
                autotune=args.autotune, group_tables=args.group_tables, graphs=graphs) as srv:
        # warm-up: one request of each kind per power-of-two size up to the
        # batch limit (with the autotuner this also fixes every shape bucket
        # before the timed arrivals)
        log(f"{tag}: warm-up requests" + (" (tuning launch shapes)" if args.autotune else ""))
        n = 1
        while n <= args.max_rows:
            for op, prm in (("topk", 50), ("threshold", 0.8)):
                s = synth.ragged_slice(*pool, 0, min(n, len(pool[2]) - 1))
                r = srv.submit(Request(-1, *s, op, prm))
                r.event.wait()
                if r.error is not None:
                    raise r.error
            n *= 2
        for rate in rates:
            nreq = requests_for(rate)
            reqs = []
            for i in range(nreq):
                s = synth.ragged_slice(*pool, int(bounds[i]), int(bounds[i + 1]))
                op, prm = ("topk", 50) if i % 2 == 0 else ("threshold", 0.8)
                reqs.append(Request(i, *s, op, prm))
            b0 = srv.batches
            wall, lat = run_poisson(srv, reqs, rate, seed=int(rate * 1000))
            errs = [r.error for r in reqs if r.error is not None]
            if errs:
                raise errs[0]
            pairs = int(sizes[:nreq].sum()) * ndb
            results.append({"rate_per_s": rate, "requests": nreq, "wall_s": wall, "batches": srv.batches - b0,
                            "p50_ms": float(np.percentile(lat, 50) * 1e3),
                            "p99_ms": float(np.percentile(lat, 99) * 1e3),
                            "max_ms": float(lat.max() * 1e3), "pairs_per_s": pairs / wall,
                            "requests_per_s": nreq / wall})
            log(f"{tag}: rate {rate}/s: p50 {results[-1]['p50_ms']:.2f} ms, p99 {results[-1]['p99_ms']:.2f} ms, "
                f"{results[-1]['pairs_per_s']:.3g} pairs/s")
    return results


def workload_request_split(args):
    """Where a serving request's time goes, per step of GpuExecutor.run, on
    the serving database: building and uploading the request's row groups
    (set_refsmiles_ragged), enqueueing the call, waiting for its kernels,
    copying the result; kernel time from HIP events.  Medians of 10 calls
    per request size and operation."""
    from siml_amd import synth
    from siml_amd.GPULingo import GPULingo, LingoDB

    ndb = args.db_serve or 10_000_000
    log(f"request split: packing the {ndb}-molecule database (lognormal, seed 5)")
    db = LingoDB(*lognormal_set(ndb, 5))
    pool = lognormal_set(4096, 55)
    g = GPULingo(0)
    g.attach_db(db)
    g.set_group_tables(args.group_tables)
    h = g._h
    out = []
    ops = ("topk", "topk_graph", "threshold") if args.graphs else ("topk", "threshold")
    for n in (1, 4, 16, 32, 64, 128, 512):
        for op in ops:
            rec = []
            g.set_graphs(op == "topk_graph")
            for it in range(12):
                ids, counts, offs, mags = synth.ragged_slice(*pool, it * 3, it * 3 + n)
                a = time.perf_counter()
                g.set_refsmiles_ragged(ids, counts, offs, mags)
                b = time.perf_counter()
                h.call("siml_set_timing", 1)
                h.kernel_stats(reset=True)
                if op.startswith("topk"):
                    h.call("siml_topk_async", 0, n, 50)
                else:
                    h.call("siml_neighbors_async", 0, n, 0.8, 1.1)
                c = time.perf_counter()
                h.call("siml_synchronize")
                d = time.perf_counter()
                h.retrieve_topk() if op.startswith("topk") else h.retrieve_csr()
                e = time.perf_counter()
                kms, _, _ = h.kernel_stats()
                h.call("siml_set_timing", 0)
                rec.append(((b - a) * 1e3, (c - b) * 1e3, (d - c) * 1e3, (e - d) * 1e3, kms, (e - a) * 1e3))
            m = np.median(np.array(rec[2:]), axis=0)
            out.append({"rows": n, "op": op, "set_ref_ms": m[0], "enqueue_ms": m[1], "wait_ms": m[2],
                        "retrieve_ms": m[3], "kernel_ms": m[4], "total_ms": m[5]})
            log(f"{op:9s} rows {n:4d}: set_ref {m[0]:.3f}  enqueue {m[1]:.3f}  wait {m[2]:.3f}  "
                f"retrieve {m[3]:.3f}  kernels {m[4]:.3f}  total {m[5]:.3f} ms")
    top1 = next(r for r in out if r["rows"] == 1 and r["op"] == "topk")
    rec = base_record(args, "serving request time split (1 GPU)", top1["total_ms"], "ms", top1["total_ms"] / 1e3,
                      {"workload": f"single requests (top-50, threshold 0.8) on the {ndb}-molecule resident "
                                   f"database, 1 to 512 query molecules"}, steps=1, warmup=2)
    rec["higher_is_better"] = False
    rec["split"] = out
    db.close()
    return rec


def workload_serve(args):
    rates = [float(r) for r in args.rates.split(",")]
    part = serving_part(args, rates, lambda r: args.requests)
    best = max(part["rates"], key=lambda r: r["pairs_per_s"] if r["p99_ms"] < 50 else -1)
    rec = base_record(args, "Tanimoto pairs/s (1 GPU, bit-exact), serving loop", best["pairs_per_s"], "pairs/s",
                      best["wall_s"], {"workload": "serving: " + part["workload"]}, steps=1, warmup=1)
    rec.update(part)
    return rec


def workload_default(args):
    """The all-pairs line, with the batched and serving parts unless
    --quick."""
    rec = workload_allpairs(args)
    if args.quick:
        return rec
    peaks = device_peaks()
    t0 = time.perf_counter()
    rec["batched"] = batched_part(args, peaks=peaks)
    log(f"batched part: {time.perf_counter() - t0:.0f} s")
    t0 = time.perf_counter()
    rec["serving"] = serving_part(args, [30.0, 1000.0, 2500.0, 5000.0, 10000.0],
                                  lambda r: int(min(600, max(120, 3 * r))))
    log(f"serving part: {time.perf_counter() - t0:.0f} s")
    return rec


# counter passes of a profile (rocprofv3 --pmc, one directory each) and the
# name each counter takes in the `counters:` line
COUNTER_PASSES = {
    "fetch": {"FETCH_SIZE": "fetch_kib"},
    "write": {"WRITE_SIZE": "write_kib"},
    "sq1": {"SQ_INSTS_VALU": "valu_insts", "SQ_INSTS_LDS": "lds_insts", "SQ_LDS_BANK_CONFLICT": "lds_bank_conflict",
            "SQ_WAVE_CYCLES": "wave_cycles", "SQ_WAVES": "waves", "SQ_INSTS_SALU": "salu_insts"},
    "sq2": {"SQ_LDS_IDX_ACTIVE": "lds_idx_active", "SQ_LDS_ADDR_CONFLICT": "lds_addr_conflict",
            "SQ_WAIT_ANY": "wait_any", "SQ_INSTS_VMEM_RD": "vmem_rd_insts", "SQ_INSTS_VMEM_WR": "vmem_wr_insts"},
    "sq3": {"SQ_ACTIVE_INST_VALU": "valu_active", "SQ_ACTIVE_INST_VALU2": "valu_active2", "SQ_CYCLES": "sq_cycles"},
}


def counters_line(directory, kernel_regex):
    """The `counters:` line of a profile from its rocprofv3 --pmc passes
    (subdirectories named as COUNTER_PASSES): the last dispatch whose kernel
    name matches kernel_regex, per pass, with the source hash of this tree."""
    import csv
    import re
    pat = re.compile(kernel_regex)
    out = {"sources_sha": sources_sha(), "fetch_factor": 2.0, "write_factor": 1.0, "kernel": None}
    for sub, names in COUNTER_PASSES.items():
        files = glob.glob(os.path.join(directory, sub, "*counter_collection.csv"))
        if not files:
            raise SystemExit(f"no counter_collection.csv under {os.path.join(directory, sub)}")
        per, kname = {}, {}
        with open(files[0]) as f:
            for r in csv.DictReader(f):
                if not pat.search(r["Kernel_Name"]):
                    continue
                d = int(r["Dispatch_Id"])
                per.setdefault(d, {})
                per[d][r["Counter_Name"]] = per[d].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
                kname[d] = r["Kernel_Name"]
        if not per:
            raise SystemExit(f"no dispatch of {kernel_regex!r} in {files[0]}")
        last = max(per)
        out["kernel"] = kname[last]
        for counter, key in names.items():
            out[key] = per[last][counter]
    return "counters: " + json.dumps(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--counters-dir", default="",
                    help="only print the `counters:` line of the rocprofv3 passes under this directory")
    ap.add_argument("--counters-kernel", default="k_bits", help="with --counters-dir: kernel name regex")
    ap.add_argument("--workload", choices=["default", "allpairs", "zinc", "cpu4096", "batched", "threshold", "serve",
                                           "request-split"],
                    default="default")
    ap.add_argument("--quick", action="store_true", help="default workload: the all-pairs line only")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_ALLPAIRS, help="all-pairs: molecules (rows = columns)")
    ap.add_argument("--db", type=int, default=0, help="database molecules (batched/threshold)")
    ap.add_argument("--db-serve", type=int, default=0, help="database molecules of the serving loop")
    ap.add_argument("--sets", type=int, default=2000, help="batched: query sets")
    ap.add_argument("--requests", type=int, default=600, help="serve: requests per arrival rate")
    ap.add_argument("--rates", default="30,300,1000,2000,2500", help="serve: Poisson arrival rates, batches/s")
    ap.add_argument("--slots", type=int, default=2, help="serve: concurrent comparators")
    ap.add_argument("--max-rows", type=int, default=1024, help="serve: query molecules per launch")
    ap.add_argument("--mem-budget", type=int, default=16, help="serve: admission budget, GiB")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle checks")
    ap.add_argument("--graphs", type=int, default=1,
                    help="batched/serve: also run the calls in graph mode (siml_set_graphs) and report both")
    ap.add_argument("--autotune", type=int, default=1,
                    help="batched/serve: launch shapes chosen by the autotuner (1) or the defaults (0)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-plain", action="store_true", help="batched: only the batched call (profiling runs)")
    ap.add_argument("--group-tables", type=int, default=2, choices=[1, 2, 4],
                    help="most row groups one workgroup of the bit-sliced kernels serves")
    ap.add_argument("--launch-shape", default="",
                    help="batched: fixed launch shape 'workgroups,topk_wave_cap,topk_seed' instead of the autotuner")
    ap.add_argument("--policy", choices=["auto", "merge", "bits", "sparse"], default="auto",
                    help="auto: bit-sliced 32-row kernels, and for the float32 matrix the sparse (inverted-index) "
                         "path when the query set's postings are short; bits: never the sparse path; sparse: "
                         "always it; merge: one-row-per-workgroup merge kernel")
    ap.add_argument("--layout", choices=["block", "natural", "sorted"], default="block",
                    help="query storage order: size-sorted 256-column blocks, the caller's order, or fully "
                         "size-sorted")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_calibration(args.copy_bytes)
    if args.counters_dir:
        print(counters_line(args.counters_dir, args.counters_kernel))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the box has one); --gpus must be 1. "
                                   "Sharding across GPUs is described in DESIGN.md."}))
        return 2
    rec = {"default": workload_default, "allpairs": workload_allpairs, "zinc": workload_zinc,
           "cpu4096": workload_cpu4096, "batched": workload_batched, "threshold": workload_threshold,
           "serve": workload_serve, "request-split": workload_request_split}[args.workload](args)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
