"""Benchmark: FP64 SpGEMM throughput on one MI355X.

Default workload: C = A * A^T with A the R-MAT matrix of n = 2^20 rows,
16,000,000 edge draws, (a,b,c,d) = (.57,.19,.19,.05), duplicates merged,
values U[-1,1], seed 1 (nnz(A) = 15,361,802; 19,250,644,998 intermediate
products; nnz(C) ~ 9.1e9, so C's row pointers are int64).

A step is one full SpGEMM through the device API -- symbolic phase (row
pointers of C, nnz(C)) and numeric phase (columns and values) -- with A and
B = A^T resident in HBM; C's column/value buffers are allocated once and
rewritten every step.  GFLOP/s = 2 * intermediate products / step time
(main_spgemm-col.cu:4787).

Prints one JSON line (rank 0 only; this build runs on one GPU):
  value         whole-job GFLOP/s
  roofline      the dominant stage's algorithmic bytes / its duration per step
                (HIP events on the library's stream; on R-MAT the big-row
                numeric stage, three launches by item size class) against
                8 TB/s
  cpu_baseline  the reference's own CPU SpGEMM (oracle/_ref, OpenMP) on a
                bounded, evenly strided sample of A's rows, same B
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

# The served section runs the request server's estimator and three workers,
# each on a stream of its own.  With HIP's default of 4 hardware queues per
# process those streams can land on a shared queue, and a worker's host
# round trips then wait behind another worker's kernels: saturation fell
# from ~4,800 to 580-1,220 req/s in 4 of 5 runs that followed the R-MAT
# product in one process.  The runtime reads the setting when it starts, so
# it is set before torch is imported (a value given by the caller stands).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from tilespgemm_amd import gen
    if name == "rmat20_aat":
        A = gen.rmat(20, 16_000_000, seed=1)
        return dict(name="R-MAT n=2^20, 16M draws, (.57,.19,.19,.05), U[-1,1], seed 1; C=A*A^T",
                    m=1 << 20, n=1 << 20, A=A, aat=True)
    if name.startswith("rmat") and name.endswith("_aat"):
        # rmat<scale>_aat (16 draws per row) or rmat<scale>e<factor>_aat
        spec = name[4:-4]
        s, ef = (int(x) for x in spec.split("e")) if "e" in spec else (int(spec), 16)
        A = gen.rmat(s, ef << s, seed=1)
        return dict(name=f"R-MAT n=2^{s}, {ef << s} draws, seed 1; C=A*A^T",
                    m=1 << s, n=1 << s, A=A, aat=True)
    if name == "laplacian256_aa":
        A = gen.laplacian2d(256)
        return dict(name="2D 5-point Laplacian 256x256; C=A*A", m=65536, n=65536, A=A, aat=False)
    if name == "stencil27_64_aa":
        A = gen.stencil27(64, seed=0)
        return dict(name="27-point stencil 64^3, N(0,1) values, seed 0; C=A*A", m=262144,
                    n=262144, A=A, aat=False)
    raise SystemExit(f"unknown workload {name}")


def csr_bytes(rows, nnz, ptr_bytes):
    return (rows + 1) * ptr_bytes + nnz * 12


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, B, m, n, k, target_s=15.0, gpu_rows=None):
    """The reference thmkl_dcsrmultcsr (compiled from its sources into
    oracle/_ref) on every stride-th row of A, timed as its main() does
    (both phases plus the malloc between, thmkl_dcsrmultcsr.c:190-210).
    With gpu_rows (rows -> the GPU's (ic, jc, c) for those rows of C), every
    8th row of the sample is also checked against the reference's output:
    structure exactly, values bit for bit, and the max relative error
    |gpu - cpu| / sum |a_ik b_kj| is reported."""
    from oracle import py_oracle as O
    kind = "reference" if O.ref_available("fma") else "port"
    ia, ja, a = (np.asarray(x) for x in A)
    blen = np.diff(np.asarray(B[0], dtype=np.int64))

    def sample(stride):
        rows = np.arange(0, m, stride)
        lens = (ia[rows + 1] - ia[rows]).astype(np.int64)
        sp = np.zeros(rows.size + 1, dtype=np.int64)
        np.cumsum(lens, out=sp[1:])
        idx = np.concatenate([np.arange(ia[r], ia[r + 1]) for r in rows]) if rows.size else \
            np.zeros(0, np.int64)
        return rows.size, (sp, ja[idx], a[idx]), int(blen[ja[idx]].sum())

    last = {}

    def run(ms, As):
        t0 = time.perf_counter()
        if kind == "reference":
            out = O.ref_thmkl(ms, n, k, As, B, "fma")
        else:
            out = O.csr_spgemm(ms, k, As, B)
        t = time.perf_counter() - t0
        last["out"] = out
        return t

    # the reference keeps C's row pointers in int32 (common.h:15-17, no overflow
    # check): keep the sample's intermediate products (>= its nnz(C)) < 1.5e9
    total_prods = int(blen[ja].sum())
    min_stride = max(1, -(-total_prods // 1_500_000_000))
    stride = max(min_stride, m // 256)
    while True:
        ms, As, prods = sample(stride)
        t = run(ms, As)
        if t >= target_s / 3 or stride == min_stride:
            break
        # grow the sample towards target_s (at most 64x per round)
        stride = max(min_stride, int(stride * max(t, target_s / 64) / target_s))
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    res = {"value": 2.0 * prods / t / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": kind,
           "cpu_model": cpu_model(),
           "build": "gcc -O3 -march=x86-64-v3 -fopenmp (FMA contraction, as the reference's "
                    "-march=native build)" if kind == "reference" else "oracle/csr_oracle.c",
           "sample": f"every {stride}th row of A ({ms} rows, {prods} intermediate products), "
                     f"full B; {t:.2f} s"}
    if gpu_rows is not None:
        ric, rjc, rc = (np.asarray(x) for x in last["out"])
        sel = np.arange(0, ms, 8)
        rows = np.arange(0, m, stride)[sel]
        gic, gjc, gc = gpu_rows(rows)
        lens = (ric[sel + 1] - ric[sel]).astype(np.int64)
        wic = np.zeros(sel.size + 1, dtype=np.int64)
        np.cumsum(lens, out=wic[1:])
        pos = np.concatenate([np.arange(ric[i], ric[i + 1]) for i in sel])
        wjc, wc = rjc[pos], rc[pos]
        same_struct = np.array_equal(np.asarray(gic, np.int64), wic) and \
            np.array_equal(np.asarray(gjc, np.int64), wjc.astype(np.int64))
        bitwise = same_struct and np.array_equal(np.asarray(gc).view(np.int64),
                                                 np.ascontiguousarray(wc).view(np.int64))
        err = None
        if same_struct:
            ia_ = np.asarray(A[0])
            sub_lens = (ia_[rows + 1] - ia_[rows]).astype(np.int64)
            sp_ = np.zeros(rows.size + 1, dtype=np.int64)
            np.cumsum(sub_lens, out=sp_[1:])
            idx = np.concatenate([np.arange(ia_[r], ia_[r + 1]) for r in rows])
            As_ = (sp_, np.asarray(A[1])[idx], np.asarray(A[2])[idx])
            scale = O.csr_absdot(rows.size, k, As_, B, wic, wjc)
            d = np.abs(np.asarray(gc) - wc)
            with np.errstate(divide="ignore", invalid="ignore"):
                r = np.where(scale > 0, d / np.where(scale > 0, scale, 1.0),
                             np.where(d > 0, np.inf, 0.0))
            err = float(r.max()) if r.size else 0.0
        res["parity"] = {"rows": int(rows.size), "outputs": int(pos.size),
                         "structure_equal": bool(same_struct), "bitwise": bool(bitwise),
                         "max_rel_err": err,
                         "what": "every 8th row of the CPU sample: GPU C rows against the "
                                 "reference's output for them; error scale sum |a_ik b_kj|"}
    return res


def _pcts(x_ms):
    x = np.asarray(x_ms, dtype=np.float64)
    return round(float(np.percentile(x, 50)), 3), round(float(np.percentile(x, 99)), 3)


def batched_section(args, graph_extra=True, graph_tune=True):
    """Batched form: 512 independent ER products C_i = A_i * A_i (n
    log-uniform in [1000, 65536], mean degree U[4,32]) as one block-diagonal
    product -- symbolic and numeric phases of the whole batch every step --
    against plain per-call launches of the same products on the same handle,
    each call's result waited for as a client would.  All requests are
    present at t=0: a request's latency is its batch's completion (batched)
    or its own call's completion (per-call).  With graph_extra, the numeric
    phase alone captured once in a HIP graph and replayed is reported as a
    labelled extra (it skips the symbolic phase, so it is not the batched
    figure)."""
    import torch
    import tilespgemm_amd as tsg
    from tilespgemm_amd import gen
    t0 = time.time()
    mats = gen.er_batch(args.batch, seed=7)
    log(f"[bench] batch of {len(mats)} ER matrices, gen {time.time() - t0:.1f}s, "
        f"rows {sum(n for n, _ in mats)}, nnz {sum(X[1].size for _, X in mats)}")
    dev = "cuda:0"
    pairs = []
    products = 0
    for n, X in mats:
        d = tsg.DeviceCsr.from_host(n, n, *X, device=dev)
        pairs.append((d, d))
        blen = np.diff(X[0].astype(np.int64))
        products += int(blen[X[1]].sum())
    method = {"auto": tsg.METHOD_AUTO, "tile": tsg.METHOD_TILE, "row": tsg.METHOD_ROW}[args.method]
    h = tsg.Handle(0, method=method)
    s = torch.cuda.Stream()
    h.set_stream(s)
    torch.cuda.synchronize()

    tuned = None
    if args.row_bins is not None:
        h.set_tuning({"row_bins": tuple(int(x) for x in args.row_bins.split(","))})
    if not args.no_autotune:
        from tilespgemm_amd import autotune as AT
        t0 = time.time()
        with torch.cuda.stream(s):
            tuned = AT.tune(h, AT.batch_run(h, pairs), reps=2, log=log)
        log(f"[bench] autotune {time.time() - t0:.1f}s: tile_density {tuned['tile_density']}, "
            f"row_bins {tuned['row_bins']}, {tuned['default_ms']:.2f} -> {tuned['ms']:.2f} ms")

    def batched_step():
        arr, ic, per, roff, tot = h.batch_symbolic(pairs)
        jc = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(tot, 1), dtype=torch.float64, device=dev)
        h.batch_numeric(arr, ic, jc, c)
        return arr, ic, jc, c, tot

    steps = max(1, args.batch_steps)
    with torch.cuda.stream(s):
        for _ in range(max(1, args.warmup)):
            arr, ic, jc, c, tot = batched_step()
        s.synchronize()
        times = []
        for _ in range(steps):
            t1 = time.perf_counter()
            arr, ic, jc, c, tot = batched_step()
            s.synchronize()
            times.append(time.perf_counter() - t1)
    t_batch = float(np.mean(times))
    log(f"[bench] batched: {t_batch * 1e3:.2f} ms per batch, nnz(C) total {tot}")

    extra = None
    if graph_extra:
        # numeric phase captured once, replayed; the graph replays the
        # numeric phase alone, so its tuning is searched for that phase (the
        # symbolic plan is then rebuilt with it before capture)
        tuned_g = None
        if graph_tune and not args.no_autotune:
            with torch.cuda.stream(s):
                tuned_g = AT.tune(h, None, log=log,
                                  measure=AT.batch_numeric_measure(h, pairs, reps=2))
                log(f"[bench] numeric-phase autotune: row_bins {tuned_g['row_bins']}, "
                    f"{tuned_g['default_ms']:.2f} -> {tuned_g['ms']:.2f} ms")
                del arr, ic, jc, c
                arr, ic, per, roff, tot = h.batch_symbolic(pairs)
                jc = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
                c = torch.empty(max(tot, 1), dtype=torch.float64, device=dev)
                h.batch_numeric(arr, ic, jc, c)
            s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            h.batch_numeric(arr, ic, jc, c)
        g.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize()
        t_graph = (time.perf_counter() - t1) / steps
        log(f"[bench] graph-replayed numeric phase: {t_graph * 1e3:.2f} ms per batch")
        del g
        # like for like: per-call numeric phases of the same products, each
        # product's symbolic plan kept in a handle of its own (same tuning,
        # same stream), each call waited for as a client would
        hs, outs = [], []
        with torch.cuda.stream(s):
            for d, _ in pairs:
                hh = tsg.Handle(0, method=method)
                if tuned_g is not None:
                    hh.set_tuning({"tile_density": tuned_g["tile_density"],
                                   "row_bins": tuple(tuned_g["row_bins"]),
                                   "block_pointer_budget": tuned_g["block_pointer_budget"]})
                ic1, nnz1 = hh.symbolic(d, d)
                jc1 = torch.empty(max(nnz1, 1), dtype=torch.int32, device=dev)
                c1 = torch.empty(max(nnz1, 1), dtype=torch.float64, device=dev)
                hh.numeric(d, d, ic1, jc1, c1)
                hs.append(hh)
                outs.append((ic1, jc1, c1))
            s.synchronize()
            t1 = time.perf_counter()
            for _ in range(steps):
                for (d, _), hh, (ic1, jc1, c1) in zip(pairs, hs, outs):
                    hh.numeric(d, d, ic1, jc1, c1)
                    s.synchronize()
            t_pcn = (time.perf_counter() - t1) / steps
        for hh in hs:
            hh.close()
        del hs, outs
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log(f"[bench] per-call numeric phases (plans kept): {t_pcn * 1e3:.2f} ms for the batch")
        extra = {"what": "numeric phase only (symbolic plans reused) on both sides: the batch's "
                         "numeric phase captured once in a HIP graph and replayed, against "
                         "per-call numeric phases of the same products with each product's plan "
                         "kept; not comparable with per-call full products",
                 "req_s": round(len(pairs) / t_graph, 2),
                 "per_call_numeric_req_s": round(len(pairs) / t_pcn, 2),
                 "speedup_vs_per_call_numeric": round(t_pcn / t_graph, 2),
                 "gflops": round(2.0 * products / t_graph / 1e9, 2),
                 "autotune": None if tuned_g is None else {
                     "row_bins": tuned_g["row_bins"], "tile_density": tuned_g["tile_density"],
                     "block_pointer_budget": tuned_g["block_pointer_budget"],
                     "default_ms": round(tuned_g["default_ms"], 3),
                     "tuned_ms": round(tuned_g["ms"], 3)}}
        if tuned is not None:
            h.set_tuning({"tile_density": tuned["tile_density"],
                          "row_bins": tuple(tuned["row_bins"]),
                          "block_pointer_budget": tuned["block_pointer_budget"]})

    # plain per-call launches of the same products, each waited for
    with torch.cuda.stream(s):
        for d, _ in pairs[:4]:
            h.spgemm(d, d)
        s.synchronize()
        done = []
        t1 = time.perf_counter()
        service = []
        for d, _ in pairs:
            ta = time.perf_counter()
            h.spgemm(d, d)
            s.synchronize()
            tb = time.perf_counter()
            service.append((tb - ta) * 1e3)
            done.append((tb - t1) * 1e3)
        t_single = time.perf_counter() - t1
    log(f"[bench] per-call: {t_single * 1e3:.2f} ms for {len(pairs)} products")
    h.close()
    nreq = len(pairs)
    b50, b99 = _pcts(np.array(times) * 1e3)
    p50, p99 = _pcts(done)
    s50, s99 = _pcts(service)
    return {
        "workload": f"{nreq} independent ER C=A*A, n log-uniform [1000, 65536], mean degree "
                    "U[4,32], U[-1,1] values, seed 7; all requests present at t=0",
        "requests": nreq, "intermediate_products": products, "nnz_c": int(tot),
        "req_s": round(nreq / t_batch, 2),
        "gflops": round(2.0 * products / t_batch / 1e9, 2),
        "ms_per_batch": round(t_batch * 1e3, 3),
        "latency_p50_ms": b50, "latency_p99_ms": b99,
        "per_call_req_s": round(nreq / t_single, 2),
        "per_call_latency_p50_ms": p50, "per_call_latency_p99_ms": p99,
        "per_call_service_p50_ms": s50, "per_call_service_p99_ms": s99,
        "speedup_vs_per_call": round(t_single / t_batch, 2),
        "autotune": None if tuned is None else {
            "tile_density": tuned["tile_density"], "row_bins": tuned["row_bins"],
            "block_pointer_budget": tuned["block_pointer_budget"],
            "default_ms": round(tuned["default_ms"], 3),
            "tuned_ms": round(tuned["ms"], 3), "trials": len(tuned["trials"])},
        "graph_replayed_numeric_extra": extra,
    }


def bench_batch(args):
    """--workload er_batch: the batched section alone as its own line."""
    b = batched_section(args)
    out = {
        "metric": "batched req/s", "value": b["req_s"], "unit": "req/s",
        "n_gpus": 1, "steps": args.batch_steps, "warmup": args.warmup,
        "ms_per_step": b["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": b,
    }
    print(json.dumps(out), flush=True)


def served_section(args, served_sweep=(0.5, 0.8)):
    """Served form: an open-loop Poisson request stream (gen.served_mix:
    70% ER, 25% Laplacian, 5% R-MAT 2^18 A*A^T) against the request server
    (device-resident operands, admission by the intermediate-product bound,
    continuous batching on several workers) and against the same server with
    one worker and batches of one (plain per-call launches).  Latency of a
    request = completion - scheduled arrival.  Saturation (every request at
    t=0), then Poisson rates at 0.25/0.5/0.8 of the per-call saturation rate
    for both, and at served_sweep fractions of the served saturation rate for
    the server alone."""
    import threading
    import queue as pyqueue
    import torch
    import tilespgemm_amd as tsg
    from tilespgemm_amd import gen
    t0 = time.time()
    free_b, total_b = torch.cuda.mem_get_info()
    log(f"[bench] served section: device memory free {free_b / 2**30:.1f} of "
        f"{total_b / 2**30:.1f} GiB")
    nreq = args.requests
    kinds, pick, er, lap, rm = gen.served_mix(nreq, seed=11, edge_factor=args.edge_factor)
    dev = "cuda:0"
    h = tsg.Handle(0)
    d_er = [tsg.DeviceCsr.from_host(X[0].size - 1, X[0].size - 1, *X, device=dev) for X in er]
    d_lap = [tsg.DeviceCsr.from_host(s * s, s * s, *X, device=dev) for s, X in lap]
    d_rm = []
    for X in rm:
        dA = tsg.DeviceCsr.from_host(1 << 18, 1 << 18, *X, device=dev)
        d_rm.append((dA, h.transpose(dA)))
    torch.cuda.synchronize()
    ops = []
    products = 0
    for kd, p in zip(kinds, pick):
        if kd == 0:
            ops.append((d_er[p], d_er[p]))
        elif kd == 1:
            ops.append((d_lap[p], d_lap[p]))
        else:
            ops.append(d_rm[p])
    prod_cache = {}
    for A, B in ops:
        key = (id(A), id(B))
        if key not in prod_cache:
            prod_cache[key] = h.intermediate_products(A, B)
        products += prod_cache[key]
    h.close()
    log(f"[bench] served mix: {nreq} requests (ER {int((kinds == 0).sum())}, Laplacian "
        f"{int((kinds == 1).sum())}, R-MAT {int((kinds == 2).sum())}), "
        f"{products} intermediate products, setup {time.time() - t0:.1f}s")

    # one server per mode, created and warmed once and kept for every run,
    # as a serving process keeps its workers' workspaces and output pool (a
    # fresh server per run allocated them inside the timed region)
    servers = {}

    def run(workers, max_batch, rate):
        """rate None: every request submitted at t=0 (saturation)."""
        key = (workers, max_batch)
        if key not in servers:
            servers[key] = tsg.Server(0, workers=workers, max_batch=max_batch,
                                      mem_budget=args.mem_budget << 30)
        srv = servers[key]
        rng = np.random.default_rng(123)
        if rate is None:
            sched = np.zeros(nreq)
        else:
            sched = np.cumsum(rng.exponential(1.0 / rate, nreq))
        lat = np.zeros(nreq)
        bsz = np.zeros(nreq, dtype=np.int64)
        q = pyqueue.Queue()
        sub_lag = np.zeros(nreq)
        errors = []

        def waiter():
            for _ in range(nreq):
                i, t = q.get()
                try:
                    res = srv.wait(t)
                    lat[i] = sub_lag[i] + res.total_ms * 1e-3
                    bsz[i] = res.batch_size
                    srv.release(t)
                except Exception as e:  # noqa: BLE001 - reported below
                    errors.append(e)
                    return
        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for i, (A, B) in enumerate(ops):
            target = start + sched[i]
            while True:
                d = target - time.perf_counter()
                if d <= 0:
                    break
                time.sleep(d if d > 2e-3 else 0)
            ts = time.perf_counter()
            sub_lag[i] = ts - target
            q.put((i, srv.submit(A, B)))
        th.join()
        span = time.perf_counter() - start
        if errors:
            raise errors[0]
        return dict(req_s=nreq / span, p50_ms=float(np.percentile(lat, 50) * 1e3),
                    p99_ms=float(np.percentile(lat, 99) * 1e3),
                    mean_batch=float(bsz.mean()), span_s=span)

    served = (args.workers, args.max_batch)
    percall = (1, 1)
    run(*served, None)  # warm the device pool and kernels
    run(*percall, None)
    sat_s = run(*served, None)
    sat_p = run(*percall, None)

    # Where the time goes: the same requests split into the R-MAT ones and
    # the rest, each stream alone at saturation in both modes, and the whole
    # mix as plain device-API calls in arrival order (no server).  Served
    # time is at least the R-MAT stream's, which batching cannot shorten
    # (one product each, already filling the GPU), so the served speed-up is
    # at most (per-call R-MAT + per-call rest) / served R-MAT.
    all_ops = ops
    split = {}
    for label, pick_kind in (("rmat", lambda kd: kd == 2), ("rest", lambda kd: kd != 2)):
        ops = [op for op, kd in zip(all_ops, kinds) if pick_kind(kd)]
        nreq = len(ops)
        split[label] = {"requests": nreq, "served_s": run(*served, None)["span_s"],
                        "per_call_s": run(*percall, None)["span_s"]}
    ops = all_ops
    nreq = len(ops)
    hp = tsg.Handle(0)
    for A, B in ops[:8]:
        hp.spgemm(A, B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for A, B in ops:
        ic_, jc_, c_ = hp.spgemm(A, B)
        torch.cuda.synchronize()
        del ic_, jc_, c_
    plain_s = time.perf_counter() - t1
    hp.close()
    bound = (split["rmat"]["per_call_s"] + split["rest"]["per_call_s"]) / split["rmat"]["served_s"]
    log(f"[bench] split: {split}, plain per-call {plain_s:.3f} s, speed-up bound {bound:.2f}")
    log(f"[bench] saturation: served {sat_s['req_s']:.0f} req/s (mean batch "
        f"{sat_s['mean_batch']:.1f}), per-call {sat_p['req_s']:.0f} req/s")
    sweep = []
    for frac in (0.25, 0.5, 0.8):
        rate = frac * sat_p["req_s"]
        rs = run(*served, rate)
        rp = run(*percall, rate)
        sweep.append({"rate": round(rate, 1), "served_p50_ms": round(rs["p50_ms"], 3),
                      "served_p99_ms": round(rs["p99_ms"], 3),
                      "per_call_p50_ms": round(rp["p50_ms"], 3),
                      "per_call_p99_ms": round(rp["p99_ms"], 3),
                      "served_mean_batch": round(rs["mean_batch"], 2)})
        log(f"[bench] rate {rate:.0f}/s: served p50 {rs['p50_ms']:.2f} p99 {rs['p99_ms']:.2f} ms"
            f" | per-call p50 {rp['p50_ms']:.2f} p99 {rp['p99_ms']:.2f} ms")
    for frac in served_sweep:
        rate = frac * sat_s["req_s"]
        rs = run(*served, rate)
        sweep.append({"rate": round(rate, 1), "served_p50_ms": round(rs["p50_ms"], 3),
                      "served_p99_ms": round(rs["p99_ms"], 3),
                      "served_mean_batch": round(rs["mean_batch"], 2)})
        log(f"[bench] rate {rate:.0f}/s: served p50 {rs['p50_ms']:.2f} p99 {rs['p99_ms']:.2f} ms")
    for srv in servers.values():
        srv.close()
    return {
        "workload": f"served mix of {nreq} requests: 70% ER C=A*A (n U[1000,16000], deg 8), "
                    "25% 2D Laplacian C=A*A (side 64-512), 5% R-MAT 2^18 C=A*A^T "
                    f"({args.edge_factor}*2^18 draws); Poisson arrivals",
        "workers": args.workers, "max_batch": args.max_batch, "mem_budget_gib": args.mem_budget,
        "intermediate_products": products,
        "req_s": round(sat_s["req_s"], 1),
        "gflops": round(2.0 * products * sat_s["req_s"] / nreq / 1e9, 2),
        "per_call_req_s": round(sat_p["req_s"], 1),
        "speedup_vs_per_call": round(sat_s["req_s"] / sat_p["req_s"], 2),
        "saturation_p50_ms": round(sat_s["p50_ms"], 2),
        "saturation_p99_ms": round(sat_s["p99_ms"], 2),
        "saturation_per_call_p50_ms": round(sat_p["p50_ms"], 2),
        "saturation_per_call_p99_ms": round(sat_p["p99_ms"], 2),
        "mean_batch": round(sat_s["mean_batch"], 2),
        "plain_per_call_req_s": round(nreq / plain_s, 1),
        "breakdown": {"what": "each sub-stream alone at saturation (s); served time >= the "
                              "R-MAT stream's, so speed-up <= (per-call R-MAT + per-call rest) "
                              "/ served R-MAT",
                      "rmat": {k: round(v, 4) if isinstance(v, float) else v
                               for k, v in split["rmat"].items()},
                      "rest": {k: round(v, 4) if isinstance(v, float) else v
                               for k, v in split["rest"].items()},
                      "speedup_bound": round(bound, 2)},
        "sweep": sweep,
    }


def bench_served(args):
    """--workload served: the served section alone as its own line."""
    v = served_section(args)
    out = {
        "metric": "served req/s", "value": v["req_s"], "unit": "req/s",
        "n_gpus": 1, "steps": args.requests, "warmup": args.requests,
        "ms_per_step": round(1e3 / v["req_s"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": v,
    }
    print(json.dumps(out), flush=True)


def bench_copy(args):
    """Calibration: a plain device copy of 4 GiB (float64) -- a kernel whose
    HBM traffic is known (4 GiB read + 4 GiB written per launch), used to
    scale the FETCH_SIZE / WRITE_SIZE counters of the SpGEMM kernels; also
    reports the copy's bandwidth."""
    import torch
    nbytes = 4 << 30
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    out = {"metric": "copy GB/s", "value": round(2 * nbytes / (ms * 1e-3) / 1e9, 1),
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "device copy of 4 GiB (read 4 GiB, write 4 GiB)",
                      "bytes_read": nbytes, "bytes_written": nbytes}}
    print(json.dumps(out), flush=True)


def bench_host(args):
    """The reference-compatible host entry end to end (tsg_dcsrmultcsr_i64:
    host CSR in, request=1, the caller's allocation of jc and c, request=0,
    host CSR out -- the sequence thmkl_dcsrmultcsr.c:190-210 times), so the
    PCIe copies of A, B and C and the first touch of C's host pages are
    inside the time.  A rate beside the device-resident one, never instead
    of it.  C runs whole on the device, then streamed in row slabs of each
    --slab-entries cap."""
    import tilespgemm_amd as tsg
    from tilespgemm_amd import gen
    t0 = time.time()
    W = make_workload(args.host_of)
    m, n = W["m"], W["n"]
    A = tuple(np.ascontiguousarray(x) for x in W["A"])
    B = tuple(np.ascontiguousarray(x) for x in gen.transpose(m, n, *A)) if W["aat"] else A
    k = m if W["aat"] else n
    blen = np.diff(np.asarray(B[0], dtype=np.int64))
    products = int(blen[A[1]].sum())
    log(f"[bench] host entry on {args.host_of}: nnz(A)={A[1].size}, {products} products, "
        f"setup {time.time() - t0:.1f}s")
    caps = [0] + [int(x) for x in args.slab_entries.split(",") if x]
    runs = []
    try:
        for cap in caps:
            tsg.set_host_slab_entries(cap)
            best = None
            for rep in range(max(1, args.steps)):
                ta = time.perf_counter()
                ic = tsg.dcsrmultcsr_i64(1, m, n, k, A, B)
                tb = time.perf_counter()
                jc, c = tsg.dcsrmultcsr_i64(0, m, n, k, A, B, ic=ic)
                tc = time.perf_counter()
                nnz_c = int(ic[-1])
                slabs = tsg.host_last_slabs()
                del jc, c
                r = dict(slab_cap=cap, slabs=slabs, request1_s=round(tb - ta, 3),
                         request0_s=round(tc - tb, 3), total_s=round(tc - ta, 3),
                         gflops=round(2 * products / (tc - ta) / 1e9, 2), nnz_c=nnz_c)
                log(f"[bench] host entry cap {cap} rep {rep}: {r}")
                if best is None or r["total_s"] < best["total_s"]:
                    best = r
            runs.append(best)
    finally:
        tsg.set_host_slab_entries(0)
    whole = runs[0]
    out = {"metric": "FP64 GFLOP/s, host entry incl. PCIe copies", "value": whole["gflops"],
           "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
           "ms_per_step": round(whole["total_s"] * 1e3, 1), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": W["name"] + " through tsg_dcsrmultcsr_i64 (host buffers)",
                      "intermediate_products": products, "runs": runs}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat20_aat")
    ap.add_argument("--method", default="auto", choices=["auto", "tile", "row"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--row-bins", default=None,
                    help="row-path bins b1,b2,b3 (tsg_set_row_bins; default: the library's)")
    ap.add_argument("--item-order", type=int, default=None,
                    help="tsg_set_item_order threshold (big-row items run in first-column "
                         "order from this many items; default: the library's)")
    ap.add_argument("--batch", type=int, default=512, help="products in the er_batch workload")
    ap.add_argument("--batch-steps", type=int, default=5, help="timed batches (batched section)")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the batched and served sections")
    ap.add_argument("--no-autotune", action="store_true",
                    help="er_batch: keep the default tuning parameters")
    ap.add_argument("--requests", type=int, default=2000, help="requests in the served workload")
    ap.add_argument("--edge-factor", type=int, default=4, help="R-MAT draws / n in the served mix")
    ap.add_argument("--workers", type=int, default=3, help="server workers (served workload)")
    ap.add_argument("--max-batch", type=int, default=64, help="server batch cap (served)")
    ap.add_argument("--mem-budget", type=int, default=32, help="server budget, GiB (served)")
    ap.add_argument("--host-of", default="rmat20_aat",
                    help="host workload: the product run through the host entry")
    ap.add_argument("--slab-entries", default="2147483648",
                    help="host workload: comma-separated slab caps (C entries) run after C whole")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a build target.", file=sys.stderr)
        raise SystemExit(2)

    if args.workload == "er_batch":
        return bench_batch(args)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "host":
        return bench_host(args)

    import torch
    import tilespgemm_amd as tsg

    t0 = time.time()
    W = make_workload(args.workload)
    m, n = W["m"], W["n"]
    A = W["A"]
    log(f"[bench] workload {args.workload}: nnz(A)={A[1].size} gen {time.time() - t0:.1f}s")
    dev = "cuda:0"
    dA = tsg.DeviceCsr.from_host(m, n, *A, device=dev)
    method = {"auto": tsg.METHOD_AUTO, "tile": tsg.METHOD_TILE, "row": tsg.METHOD_ROW}[args.method]
    h = tsg.Handle(0, method=method, timing=True)
    if args.item_order is not None:
        h.set_item_order(args.item_order)
    if args.row_bins is not None:
        h.set_tuning({"row_bins": tuple(int(x) for x in args.row_bins.split(","))})
    if W["aat"]:
        dB = h.transpose(dA)
        k = m
    else:
        dB = dA
        k = n
    torch.cuda.synchronize()
    products = h.intermediate_products(dA, dB)
    log(f"[bench] intermediate products {products}")

    # first symbolic: size C
    ic, nnz_c = h.symbolic(dA, dB)
    log(f"[bench] nnz(C)={nnz_c} method={h.method_used} "
        f"workspace={h.workspace_bytes / 2**30:.2f} GiB")
    jc = torch.empty(max(nnz_c, 1), dtype=torch.int32, device=dev)
    c = torch.empty(max(nnz_c, 1), dtype=torch.float64, device=dev)

    def step():
        _, nz = h.symbolic(dA, dB, ic=ic)
        assert nz == nnz_c
        h.numeric(dA, dB, ic, jc, c)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i} done")
    stage_sum = {}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
        for kname, v in h.stage_times().items():   # waits for the step's events
            stage_sum[kname] = stage_sum.get(kname, 0.0) + v
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed * 1e3 / args.steps
    gflops = 2.0 * products / (ms_step * 1e-3) / 1e9
    stages = {kname: v / args.steps for kname, v in stage_sum.items()}
    method_used = h.method_used
    log(f"[bench] {ms_step:.2f} ms/step, {gflops:.1f} GFLOP/s, stages {stages}")

    # roofline of the dominant kernel
    dom = max(stages, key=stages.get)
    ptr_c = 8
    bytes_a = csr_bytes(m, A[1].size, 4)
    bytes_b = csr_bytes(n, dB.nnz, 4)
    bytes_c = csr_bytes(m, nnz_c, ptr_c)
    whole = bytes_a + bytes_b + bytes_c
    # the numeric kernels write C's columns and values and read A and B once
    dom_bytes = {"row.big_numeric": bytes_a + bytes_b + 12 * nnz_c,
                 "tile.step3": bytes_a + bytes_b + 12 * nnz_c}.get(dom, whole)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel come from a committed
    # rocprofv3 counter profile of this workload (they cannot be counted
    # inside the run); the line says so
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get(dom)
        if traffic is not None:
            traffic_src = (f"profiles/traffic_{args.workload}.md: rocprofv3 --pmc FETCH_SIZE, "
                           "then WRITE_SIZE, per dispatch, summed over the stage's dispatches in "
                           "one step, scaled by the same counters on a plain 4 GiB copy; "
                           "profiled, not measured in this run")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": dom, "kernel_ms": round(stages[dom], 3),
            "algorithmic_bytes": dom_bytes,
            "whole_step_GBs": round(whole / (ms_step * 1e-3) / 1e9, 1)}

    def gpu_rows(rows):
        """The GPU's C restricted to the given rows (host arrays)."""
        ic_h = ic.cpu().numpy()
        lens = ic_h[rows + 1] - ic_h[rows]
        gic = np.zeros(rows.size + 1, dtype=np.int64)
        np.cumsum(lens, out=gic[1:])
        pos = np.concatenate([np.arange(ic_h[r], ic_h[r + 1]) for r in rows])
        pos_d = torch.from_numpy(pos).to(jc.device)
        return gic, jc[pos_d].cpu().numpy(), c[pos_d].cpu().numpy()

    cpu = None
    if not args.no_cpu_baseline:
        Bh = (dB.indptr.cpu().numpy(), dB.indices.cpu().numpy(), dB.data.cpu().numpy())
        try:
            cpu = cpu_baseline(A, Bh, m, n, k, gpu_rows=gpu_rows)
            log(f"[bench] cpu baseline {cpu}")
        except Exception as e:  # noqa: BLE001
            log(f"[bench] cpu baseline failed: {e}")

    batched = served = None
    if not args.no_extras and args.workload == "rmat20_aat":
        # free the product's device memory before the other two parts of the metric
        h.close()
        del ic, jc, c, dA, dB
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        served = served_section(args, served_sweep=())
        torch.cuda.empty_cache()
        batched = batched_section(args, graph_extra=True, graph_tune=False)
        torch.cuda.empty_cache()

    out = {
        "metric": "FP64 GFLOP/s", "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": W["name"], "nnz_a": int(A[1].size), "nnz_c": int(nnz_c),
                   "intermediate_products": int(products),
                   "method": {1: "tile", 2: "row"}.get(method_used, "?"),
                   "stages_ms": {kk: round(v, 3) for kk, v in stages.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if batched is not None:
        out["batched"] = batched
    if served is not None:
        out["served"] = served
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
