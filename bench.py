#!/usr/bin/env python3
"""Benchmark: effective FP64 GFLOPS (2mnk/t) of 2-level <2,2,2;7> Strassen at
16384^3 on one MI355X, beside the project's classical FP64 MFMA GEMM.

One step = one C := A*B + C product through the C ABI (fmmhip_dgemm_fmm) on
device-resident, seeded uniform[-1,1] operands (column-major).  W untimed
warm-up steps, then K timed steps between device synchronisations.  Every
step is bracketed by HIP events on the library's stream (latency
percentiles), and the library's launch timing records events around the
summation stage and the MFMA kernel inside each step (roofline).

Prints ONE JSON line.  --profile runs one warm-up and one step of the same
workload plus a 1 GiB device copy (the calibration kernel for the HBM byte
counters) and nothing else, for rocprofv3.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fmm-blis_amd"))

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 matrix peak
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
INT8_PEAK_TOPS = 5000.0       # dense int8 MFMA, tera-operations per second
SWEEP_TOL = 5e-15            # FP64 FMM normalised error bound of the sweep (two levels of the deepest table)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def leaf_flops(info, m, n, k):
    """Executed leaf flops 2 R^L ceil(m/m~^L) ceil(k/k~^L) ceil(n/n~^L)."""
    lm = -(-m // info["mt"])
    lk = -(-k // info["kt"])
    ln = -(-n // info["nt"])
    return 2.0 * info["R"] * lm * lk * ln


def read_traffic(tag):
    """HBM bytes per launch of the MFMA kernel for this workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by
    the same counters on a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(seconds_target=15.0, runs=5):
    """The oracle's 1-level Strassen (222) and its classical product at
    2048^3 (the plain configuration's size and seed) on the host cores: the
    CPU port of the reference's path.  The reference itself (BLIS + the
    fmm_blis plugin) is not built: it compiles only inside BLIS's
    configure-generated tree."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 2048
    A, B, C = O.uniform(n, n, 0, 0), O.uniform(n, n, 0, 1), O.uniform(n, n, 0, 2)
    alg = O.parse(os.path.join(ROOT, "tests", "golden", "algorithms", "222.txt"))

    def best(call):
        times = []
        t_end = time.time() + seconds_target / 2
        while True:
            Cw = C.copy(order="F")
            t0 = time.perf_counter()
            call(Cw)
            times.append(time.perf_counter() - t0)
            if time.time() > t_end or len(times) >= runs:
                return min(times), len(times)
    t_fmm, r_fmm = best(lambda Cw: O.fmm_dgemm(alg, 1.0, A, B, 1.0, Cw))
    t_cls, r_cls = best(lambda Cw: O.dgemm(1.0, A, B, 1.0, Cw))
    return {"value": round(2.0 * n ** 3 / t_fmm / 1e9, 2), "unit": "GFLOPS", "cores": O.threads(), "kind": "port",
            "sample": f"oracle 1-level <2,2,2;7> Strassen FP64, 2048^3, C += A*B, min of {r_fmm} runs",
            "classical": {"value": round(2.0 * n ** 3 / t_cls / 1e9, 2), "unit": "GFLOPS",
                          "sample": f"oracle classical FP64 product, 2048^3, C += A*B, min of {r_cls} runs"}}


def batch_shapes(count, seed):
    """(dtype, m, n, k) of the batching workload: m, n, k log-uniform in
    [256, 8192], multiples of 64; the first half FP64, the rest FP32."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def lu():
        x = float(np.exp(rng.uniform(np.log(256.0), np.log(8192.0))))
        return int(min(8192, max(256, round(x / 64.0) * 64)))
    return [("f8" if i < count // 2 else "f4", lu(), lu(), lu()) for i in range(count)]


def run_batch(args, F):
    """The batching form: one captured graph over all problems with the
    tuner's choices, beside the same graph with every problem classical and
    beside one classical call per problem (no graph)."""
    uniform = args.batch_uniform > 0
    shapes = [("f8", args.batch_uniform, args.batch_uniform, args.batch_uniform)] * args.batch_count if uniform \
        else batch_shapes(args.batch_count, args.seed)
    mats = []
    for i, (dt, m, n, k) in enumerate(shapes):
        A = F.DeviceMatrix(m, k, dtype=dt).fill_uniform(args.seed, 3 * i)
        B = F.DeviceMatrix(k, n, dtype=dt).fill_uniform(args.seed, 3 * i + 1)
        C = F.DeviceMatrix(m, n, dtype=dt).fill_uniform(args.seed, 3 * i + 2)
        mats.append((A, B, C))
    F.sync()
    flops = sum(2.0 * m * n * k for (_, m, n, k) in shapes)
    flops64 = sum(2.0 * m * n * k for (dt, m, n, k) in shapes if dt == "f8")

    def make(cand, tile=0):
        b = F.Batch()
        for i, ((dt, m, n, k), (A, B, C)) in enumerate(zip(shapes, mats)):
            b.add(m, n, k, 1.0, A, B, 1.0, C, cand=cand)
            if tile:
                b.set_tile(i, tile)
        t0 = time.perf_counter()
        b.plan()
        return b, time.perf_counter() - t0

    def timed(run, steps, warmup):
        for _ in range(warmup):
            run()
        F.sync()
        # and at least 250 ms of warm-up: a batch of small problems replays in
        # about a millisecond, and two replays leave the clocks ramping
        t_end = time.perf_counter() + 0.25
        while time.perf_counter() < t_end:
            run()
            F.sync()
        evs = [(F.Event(), F.Event()) for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record()
            run()
            e1.record()
        F.sync()
        wall = time.perf_counter() - t0
        return wall * 1e3 / steps, [e0.elapsed_ms(e1) for e0, e1 in evs]

    tuned, plan_s = make(-1)
    info = tuned.info()
    fast_items = tuned.fast_items()
    choices = {"f64": [0, 0, 0], "f32": [0, 0, 0]}
    tiles = {"f64": {"128": 0, "64": 0}, "f32": {"128": 0, "64": 0}}
    leaf64 = leaf32 = 0.0
    for i, (dt, m, n, k) in enumerate(shapes):
        c = tuned.choice(i)
        choices["f64" if dt == "f8" else "f32"][c] += 1
        tiles["f64" if dt == "f8" else "f32"][str(tuned.tile(i))] += 1
        d = 2 ** c
        fl = 2.0 * (7 ** c) * -(-m // d) * -(-n // d) * -(-k // d)
        if dt == "f8":
            leaf64 += fl
        else:
            leaf32 += fl
    ms, lat = timed(tuned.run, args.steps, args.warmup)

    cls, _ = make(0)
    cls_ms, _ = timed(cls.run, max(1, args.classical_steps), 1) if args.classical_steps > 0 else (None, None)
    cls.close()
    # the same choices with every problem on 128 x 128 tiles: what the
    # per-problem tile choice gains
    t128_ms = None
    if args.classical_steps > 0:
        t128, _ = make(-1, 128)
        t128_ms, _ = timed(t128.run, max(1, args.classical_steps), 1)
        t128.close()

    seq_ms = None
    if args.classical_steps > 0:
        def seq():
            for (dt, m, n, k), (A, B, C) in zip(shapes, mats):
                (F.dgemm if dt == "f8" else F.sgemm)(m, n, k, 1.0, A, B, 1.0, C)
        seq_ms, _ = timed(seq, 1, 1)

    # roofline of the persistent kernels: executed leaf flops over the step
    # time against the flop-weighted MFMA peak of the two datatypes
    t_peak = leaf64 / (FP64_PEAK_TFLOPS * 1e12) + leaf32 / (FP32_PEAK_TFLOPS * 1e12)
    achieved = (leaf64 + leaf32) / (ms * 1e-3) / 1e12
    peak = (leaf64 + leaf32) / t_peak / 1e12
    eff = flops / (ms * 1e-3) / 1e9
    out = {
        "metric": ("effective GFLOPS (2mnk/t) of a uniform FP64 batch, tuned FMM vs classical" if uniform else
                   "effective GFLOPS (2mnk/t) of a 512-GEMM uneven FP64/FP32 batch, tuned FMM vs classical"),
        "value": round(eff, 1),
        "unit": "GFLOPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if uniform else "f64+f32",
        "data": f"synthetic (SplitMix64 uniform[-1,1], seed {args.seed}, device-resident)",
        "config": {
            "workload": (f"batch of {len(shapes)} FP64 GEMMs C += A*B, m=n=k={args.batch_uniform}, "
                         f"one captured HIP graph" if uniform else
                         f"batch of {len(shapes)} GEMMs C += A*B, m,n,k log-uniform [256,8192] multiples of 64, "
                         f"first half FP64, rest FP32, one captured HIP graph"),
            "problems": len(shapes), "total_flops": flops, "fp64_flops": flops64,
            "choices_classical_1level_2level": choices,
            "tile_edges": tiles,
            "items_f64": info["items_f64"], "items_f32": info["items_f32"],
            "fast_kernel_items_f64_f32": list(fast_items),
            "workspace_bytes": info["workspace_bytes"], "tuner_est_ms": round(info["est_ms"], 3),
            "plan_s": round(plan_s, 3),
            "classical_batch_ms": round(cls_ms, 3) if cls_ms else None,
            "classical_batch_gflops": round(flops / (cls_ms * 1e-3) / 1e9, 1) if cls_ms else None,
            "tuned_over_classical": round(cls_ms / ms, 4) if cls_ms else None,
            "all_128_tiles_batch_ms": round(t128_ms, 3) if t128_ms else None,
            "sequential_classical_calls_ms": round(seq_ms, 3) if seq_ms else None,
        },
        "latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": round(peak, 2), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None, "kernel": "fmm_batch_kernel (f64, f32)",
                     "leaf_flops_f64": leaf64, "leaf_flops_f32": leaf32},
    }
    print(json.dumps(out), flush=True)
    tuned.close()


def run_sweep(args, F):
    """FP64 C += A*B with every algorithm at 1 and 2 levels (automatic pack
    mode) and the classical kernel on one shape: effective GFLOPS and the
    normalised error against the classical product of the same operands;
    where the automatic mode materialised the products, the rate with the
    W epilogue fused (pack 1) as well."""
    m = n = args.size
    k = args.k or args.size
    A = F.DeviceMatrix(m, k).fill_uniform(0, 0)
    B = F.DeviceMatrix(k, n).fill_uniform(0, 1)
    C = F.DeviceMatrix(m, n)
    Cr = F.DeviceMatrix(m, n).fill_uniform(0, 2)
    F.dgemm(m, n, k, 1.0, A, B, 1.0, Cr)
    F.sync()
    nA, nB = F.fro(A), F.fro(B)
    flops = 2.0 * m * n * k

    def timed(call, steps):
        call()
        F.sync()
        evs = [(F.Event(), F.Event()) for _ in range(steps)]
        for e0, e1 in evs:
            e0.record()
            call()
            e1.record()
        F.sync()
        return sum(e0.elapsed_ms(e1) for e0, e1 in evs) / steps

    cls_ms = timed(lambda: F.dgemm(m, n, k, 1.0, A, B, 1.0, C), args.steps)
    rows = [{"alg": "classical", "levels": 0, "ms": round(cls_ms, 3), "gflops": round(flops / cls_ms / 1e6, 1),
             "rel_err": 0.0, "vs_classical": 1.0}]
    for name in ["222", "232", "322", "333", "444"]:
        f = F.strassen() if name == "222" else \
            F.new_fmm(os.path.join(ROOT, "tests", "golden", "algorithms", f"{name}.txt"))
        for levels in (1, 2):
            alg = F.Alg([f] * levels)
            mode = alg.pack_for(m, n, k)
            ms = timed(lambda: F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, C), args.steps)
            fused_w = None
            if mode == 2:
                # the same call with the W epilogue fused (no product in HBM)
                alg.set_pack(1)
                fused_w = round(flops / timed(lambda: F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, C), args.steps) / 1e6, 1)
                alg.set_pack(-1)
            C.fill_uniform(0, 2)
            F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, C)
            F.sync()
            err = F.fro_diff(C, Cr) / (nA * nB)
            rows.append({"alg": name, "levels": levels, "ms": round(ms, 3), "gflops": round(flops / ms / 1e6, 1),
                         "rel_err": err, "vs_classical": round(cls_ms / ms, 4), "pack": mode,
                         "gflops_fused_w_epilogue": fused_w if mode == 2 else round(flops / ms / 1e6, 1)})
            alg.close()
            F.check(F.lib().fmmhip_workspace_release(), "workspace_release")
    best = max(rows, key=lambda r: r["gflops"])
    bad = [f"{r['alg']} L{r['levels']}: {r['rel_err']:.2e}" for r in rows if not r["rel_err"] <= SWEEP_TOL]
    out = {
        "metric": "effective FP64 GFLOPS (2mnk/t), FMM algorithm sweep",
        "value": best["gflops"], "unit": "GFLOPS", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (SplitMix64 uniform[-1,1], seed 0, device-resident)",
        "config": {"workload": f"FP64 C += A*B, m={m} n={n} k={k}, column-major, every algorithm at 1 and 2 levels",
                   "m": m, "n": n, "k": k, "best": f"{best['alg']} L{best['levels']}", "results": rows,
                   "error_bound": SWEEP_TOL, "over_bound": bad},
    }
    print(json.dumps(out), flush=True)
    if bad:
        raise SystemExit(f"sweep: normalised error above {SWEEP_TOL:g}: {', '.join(bad)}")


def run_precision(args, F):
    """The precision form: classical, 1- and 2-level Strassen with FP64,
    FP32, split BF16 and int8-slice leaves at 8192^3 and 16384^3 -- bf16x3
    (four planes, ten plane products), bf16x3_8 (three planes, eight),
    bf16x3_6 (three planes, six) and int8x4 (four int8 slices, ten slice
    products) -- on uniform[-1,1] inputs and
    on the stress variant (rows of A and columns of B scaled by
    10^U(-4,4)); effective GFLOPS and the normalised error against the
    classical FP64 product of the same operands."""
    import numpy as np
    leaves = [("f64", F.LEAF_NATIVE), ("f32", F.LEAF_F32), ("bf16x3", F.LEAF_BF16X3), ("bf16x3_8", F.LEAF_BF16X3_8),
              ("bf16x3_6", F.LEAF_BF16X3_6), ("int8x4", F.LEAF_INT8X4)]
    rows = []
    for n in (8192, 16384):
        for variant in ("uniform", "stress"):
            A = F.DeviceMatrix(n, n).fill_uniform(0, 0)
            B = F.DeviceMatrix(n, n).fill_uniform(0, 1)
            if variant == "stress":
                rng = np.random.default_rng(5)
                rs, cs = F.DeviceBuffer(8 * n), F.DeviceBuffer(8 * n)
                rs.upload(10.0 ** rng.uniform(-4, 4, n))
                cs.upload(10.0 ** rng.uniform(-4, 4, n))
                F.check(F.lib().fmmhip_scale_rows_cols(A.ptr, n, n, A.ld, rs.ptr, None, None), "scale")
                F.check(F.lib().fmmhip_scale_rows_cols(B.ptr, n, n, B.ld, None, cs.ptr, None), "scale")
            Cr = F.DeviceMatrix(n, n)
            F.dgemm(n, n, n, 1.0, A, B, 0.0, Cr)
            C = F.DeviceMatrix(n, n)
            F.sync()
            nA, nB = F.fro(A), F.fro(B)
            for name, levels in (("classical", 0), ("222", 1), ("222", 2)):
                alg = F.Alg(classical=True) if levels == 0 else F.Alg([F.strassen()] * levels)
                for lname, leaf in leaves:
                    F.dgemm_fmm(alg, n, n, n, 1.0, A, B, 0.0, C, leaf=leaf)
                    F.sync()
                    err = F.fro_diff(C, Cr) / (nA * nB)
                    evs = [(F.Event(), F.Event()) for _ in range(2)]
                    for e0, e1 in evs:
                        e0.record()
                        F.dgemm_fmm(alg, n, n, n, 1.0, A, B, 0.0, C, leaf=leaf)
                        e1.record()
                    F.sync()
                    ms = sum(e0.elapsed_ms(e1) for e0, e1 in evs) / len(evs)
                    rows.append({"n": n, "inputs": variant, "alg": name, "levels": levels, "leaf": lname,
                                 "ms": round(ms, 3), "effective_gflops": round(2.0 * n ** 3 / ms / 1e6, 1),
                                 "rel_err": err})
                alg.close()
                F.check(F.lib().fmmhip_workspace_release(), "workspace_release")
            for X in (A, B, C, Cr):
                X.free()
    # a combination qualifies when every size and input variant stays within 1e-10
    def key(r):
        return (r["alg"], r["levels"], r["leaf"])
    worst = {}
    for r in rows:
        worst[key(r)] = max(worst.get(key(r), 0.0), r["rel_err"])
    ok = [r for r in rows if r["n"] == 16384 and r["inputs"] == "uniform" and worst[key(r)] <= 1e-10]
    best = max(ok, key=lambda r: r["effective_gflops"])
    out = {
        "metric": "precision form: effective GFLOPS (2mnk/t) at 16384^3 of the fastest algorithm and leaf whose "
                  "normalised error stays <= 1e-10 at 8192^3 and 16384^3, uniform and stress inputs",
        "value": best["effective_gflops"], "unit": "GFLOPS", "n_gpus": 1, "steps": 2, "warmup": 1,
        "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 operands; f64/f32/bf16 leaves",
        "data": "synthetic (SplitMix64 uniform[-1,1], seed 0; stress: rows of A, columns of B x 10^U(-4,4), seed 5)",
        "config": {"workload": "C = A*B, m=n=k in {8192, 16384}, classical / 222 L1 / 222 L2 x leaves f64, f32, "
                               "bf16x3 (4 planes, 10 plane products), bf16x3_8 (3, 8), bf16x3_6 (3, 6), "
                               "int8x4 (4 int8 slices, 10 slice products)",
                   "best": f"{best['alg']} L{best['levels']} {best['leaf']}",
                   "worst_error": {f"{a} L{l} {f}": e for (a, l, f), e in worst.items()}, "results": rows},
    }
    print(json.dumps(out), flush=True)


def serve_stream(count, rate, seed, max_dim):
    """(dtype, m, n, k, arrival offset s): m, n, k log-uniform in
    [128, max_dim] rounded to multiples of 32, 70% FP64, Poisson arrivals."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        dims = [int(min(max_dim, max(128, round(float(np.exp(rng.uniform(np.log(128.0), np.log(max_dim)))) / 32.0) * 32)))
                for _ in range(3)]
        out.append(("f8" if rng.uniform() < 0.7 else "f4", dims[0], dims[1], dims[2], t))
    return out


def run_serve(args, F):
    """The serving form: a Poisson stream of mixed FP64/FP32 GEMMs over
    device-resident operand pools, admitted by HBM budget onto `slots`
    streams by the library's request loop; latency from arrival to observed
    completion."""
    import ctypes
    D = args.max_dim
    pools = {"f8": (F.DeviceMatrix(D, D).fill_uniform(args.seed, 0), F.DeviceMatrix(D, D).fill_uniform(args.seed, 1)),
             "f4": (F.DeviceMatrix(D, D, dtype="f4").fill_uniform(args.seed, 2),
                    F.DeviceMatrix(D, D, dtype="f4").fill_uniform(args.seed, 3))}
    F.sync()
    free_b, tot_b = ctypes.c_size_t(), ctypes.c_size_t()
    F.check(F.lib().fmmhip_mem_info(ctypes.byref(free_b), ctypes.byref(tot_b)), "mem")
    budget = int(args.budget_gib * (1 << 30)) if args.budget_gib > 0 else int(min(0.85 * free_b.value, 256 << 30))

    def run(srv, stream, t0):
        for dt, m, n, k, t in stream:
            A, B = pools[dt]
            req = F.Request()
            req.is_double = 1 if dt == "f8" else 0
            req.m, req.n, req.k, req.alpha, req.beta = m, n, k, 1.0, 0.0
            req.A, req.lda = A.ptr.value, A.ld
            req.B, req.ldb = B.ptr.value, B.ld
            req.C, req.ldc = None, 0
            req.cand, req.keep, req.arrival_s = -1, 0, t0 + t
            srv.submit_request(req)
        srv.drain()
        return srv.stats()

    # warm-up stream (kernels loaded, tuner and arena paths exercised)
    warm = F.Server(budget, slots=args.slots)
    run(warm, serve_stream(max(args.warmup, 1) * 8, args.rate, args.seed + 1, min(D, 2048)), warm.now() + 0.01)
    warm.close()

    stream = serve_stream(args.requests, args.rate, args.seed, D)
    srv = F.Server(budget, slots=args.slots)
    t_wall = time.perf_counter()
    st = run(srv, stream, srv.now() + 0.05)
    wall = time.perf_counter() - t_wall
    srv.close()
    flops = sum(2.0 * m * n * k for (_, m, n, k, _) in stream)
    flops64 = sum(2.0 * m * n * k for (dt, m, n, k, _) in stream if dt == "f8")
    span = st["span_s"]
    out = {
        "metric": "serving form: p99 GEMM latency (arrival to completion) at a Poisson request rate",
        "value": round(st["p99_ms"], 3),
        "unit": "ms",
        "n_gpus": 1, "steps": args.requests, "warmup": max(args.warmup, 1) * 8,
        "ms_per_step": round(span * 1e3 / max(st["completed"], 1), 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "data": f"synthetic (SplitMix64 uniform[-1,1] operand pools {D}x{D}, seed {args.seed}, device-resident)",
        "config": {
            "workload": f"{args.requests} GEMM requests C = A*B, Poisson arrivals at {args.rate:g}/s, m,n,k "
                        f"log-uniform [128,{D}] multiples of 32, 70% FP64 / 30% FP32, {args.slots} streams, "
                        f"HBM budget {budget / (1 << 30):.1f} GiB, algorithm per request by the tuner",
            "rate_per_s": args.rate, "requests": args.requests, "slots": args.slots, "budget_bytes": budget,
            "completed": st["completed"], "failed": st["failed"],
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "mean_ms": round(st["mean_ms"], 3),
            "max_ms": round(st["max_ms"], 3),
            "throughput_req_per_s": round(st["completed"] / span, 2) if span > 0 else None,
            "effective_gflops": round(flops / span / 1e9, 1) if span > 0 else None,
            "fp64_flop_share": round(flops64 / flops, 4),
            "span_s": round(span, 4), "wall_s": round(wall, 3),
            "arena_peak_bytes": st["arena_peak"], "max_inflight": st["max_inflight"],
        },
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--alg", default="222", help="222, 232, 322, 333, 444 or classical")
    ap.add_argument("--levels", type=int, default=2)
    ap.add_argument("--leaf", default="f64", choices=["f64", "f32", "bf16x3", "bf16x3_8", "bf16x3_6", "int8x4"])
    ap.add_argument("--pack", type=int, default=-1,
                    help="0 fused, 1 pre-summed, 2 pre-summed with the products materialised, -1 automatic")
    ap.add_argument("--classical-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy check against the classical kernel")
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--form", default="single", choices=["single", "batch", "serve", "sweep", "precision"],
                    help="single: one m=n=k product; batch: 512 uneven FP64/FP32 GEMMs in one graph; "
                         "serve: Poisson request stream through the continuous-batching server; "
                         "sweep: every algorithm at 1 and 2 levels on one FP64 shape; "
                         "precision: FP64/FP32/BF16x3 leaves at 8192^3 and 16384^3, uniform and scaled inputs")
    ap.add_argument("--k", type=int, default=0, help="sweep/single: k if not --size (rank-k shapes)")
    ap.add_argument("--no-bf16x3", action="store_true", help="single: skip the BF16x3 line beside the FP64 one")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-uniform", type=int, default=0,
                    help="batch form: every problem FP64 of this size cubed (default: the uneven 512-GEMM mix)")
    ap.add_argument("--rate", type=float, default=500.0, help="serve: arrivals per second")
    ap.add_argument("--requests", type=int, default=1000, help="serve: requests in the timed stream")
    ap.add_argument("--slots", type=int, default=8, help="serve: streams (requests in flight)")
    ap.add_argument("--budget-gib", type=float, default=0.0, help="serve: HBM budget (0: 85%% of free memory)")
    ap.add_argument("--max-dim", type=int, default=16384, help="serve: largest m, n, k")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tuner-table", default="", help="batch/serve: load this tuner table (fmmhip_tuner_save format)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the multi-GPU sharding is a "
              f"design note, not built)", file=sys.stderr)
        sys.exit(2)

    import fmm_blis_amd as F
    F.check(F.lib().fmmhip_init(), "init")
    if args.tuner_table:
        F.tuner_load(args.tuner_table)
    if args.form == "batch":
        return run_batch(args, F)
    if args.form == "serve":
        return run_serve(args, F)
    if args.form == "sweep":
        return run_sweep(args, F)
    if args.form == "precision":
        return run_precision(args, F)
    n = args.size
    m = k = n
    if args.alg == "classical":
        args.levels = 1
        f = F.FmmT(1, 1, 1, 1)
        alg = F.Alg(classical=True)
    else:
        f = F.strassen() if args.alg == "222" else \
            F.new_fmm(os.path.join(ROOT, "tests", "golden", "algorithms", f"{args.alg}.txt"))
        alg = F.Alg([f] * args.levels)
    alg.set_pack(args.pack)
    info = alg.info()
    leaf = {"f64": F.LEAF_NATIVE, "f32": F.LEAF_F32, "bf16x3": F.LEAF_BF16X3, "bf16x3_8": F.LEAF_BF16X3_8,
            "bf16x3_6": F.LEAF_BF16X3_6, "int8x4": F.LEAF_INT8X4}[args.leaf]
    tag = f"{args.alg}L{args.levels}_{args.leaf}_{n}"

    A = F.DeviceMatrix(m, k).fill_uniform(0, 0)
    B = F.DeviceMatrix(k, n).fill_uniform(0, 1)
    C = F.DeviceMatrix(m, n).fill_uniform(0, 2)
    F.sync()
    # the mode the calls below take (automatic: decided from the free memory now)
    pack_mode = alg.pack_for(m, n, k, leaf)

    def step():
        F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, C, leaf=leaf)

    if args.profile:
        step()
        step()
        gib = 1 << 30
        src, dst = F.DeviceBuffer(gib), F.DeviceBuffer(gib)
        F.check(F.lib().fmmhip_copy_kernel(dst.ptr, src.ptr, gib, None), "copy")
        F.sync()
        print(json.dumps({"profile": tag, "steps": 2, "calibration_copy_bytes": gib}), flush=True)
        return

    for _ in range(args.warmup):
        step()
    F.sync()
    evs = [(F.Event(), F.Event()) for _ in range(args.steps)]
    F.timing_read()                       # drop any earlier record
    F.timing_enable(True)
    F.sync()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        step()
        e1.record()
    F.sync()
    wall = time.perf_counter() - t0
    F.timing_enable(False)
    launches, kern_total, presum_total, kern_max = F.timing_read()
    lat = [e0.elapsed_ms(e1) for e0, e1 in evs]
    ms_step = wall * 1e3 / args.steps
    kern_ms = kern_total / max(launches, 1)
    presum_ms = presum_total / max(launches, 1)
    eff = 2.0 * m * n * k / (ms_step * 1e-3) / 1e9

    # classical FP64 MFMA GEMM on the same operands (its own C)
    cls_gflops = None
    if args.classical_steps > 0:
        Cc = F.DeviceMatrix(m, n).fill_uniform(0, 2)
        F.dgemm(m, n, k, 1.0, A, B, 1.0, Cc)
        F.sync()
        ce = [(F.Event(), F.Event()) for _ in range(args.classical_steps)]
        for e0, e1 in ce:
            e0.record()
            F.dgemm(m, n, k, 1.0, A, B, 1.0, Cc)
            e1.record()
        F.sync()
        cls_ms = sum(e0.elapsed_ms(e1) for e0, e1 in ce) / len(ce)
        cls_gflops = 2.0 * m * n * k / (cls_ms * 1e-3) / 1e9
        Cc.free()

    # the other pack modes of an FMM call, for the record beside the chosen
    # one: operand sums formed while staging (fused, no workspace) and the W
    # scatter fused into the MFMA kernel (pre-summed, no product in HBM)
    modes = None
    if args.alg != "classical" and args.pack == -1 and args.leaf == "f64" and args.classical_steps > 0:
        modes = {}
        Cm = F.DeviceMatrix(m, n).fill_uniform(0, 2)
        for pk, name in ((0, "fused"), (1, "pre-summed"), (2, "pre-summed, products materialised")):
            a_m = F.Alg([f] * args.levels)
            a_m.set_pack(pk)
            F.dgemm_fmm(a_m, m, n, k, 1.0, A, B, 1.0, Cm)
            F.sync()
            me = [(F.Event(), F.Event()) for _ in range(2)]
            for e0, e1 in me:
                e0.record()
                F.dgemm_fmm(a_m, m, n, k, 1.0, A, B, 1.0, Cm)
                e1.record()
            F.sync()
            mms = sum(e0.elapsed_ms(e1) for e0, e1 in me) / len(me)
            modes[name] = {"ms": round(mms, 3), "effective_gflops": round(2.0 * m * n * k / (mms * 1e-3) / 1e9, 1)}
            a_m.close()
            F.check(F.lib().fmmhip_workspace_release(), "workspace_release")
        Cm.free()

    # accuracy: one fresh FMM and one fresh classical product from the same C
    rel = None
    if not args.no_check:
        Cf = F.DeviceMatrix(m, n).fill_uniform(0, 2)
        Cr = F.DeviceMatrix(m, n).fill_uniform(0, 2)
        F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, Cf, leaf=leaf)
        F.dgemm(m, n, k, 1.0, A, B, 1.0, Cr)
        F.sync()
        rel = F.fro_diff(Cf, Cr) / (F.fro(A) * F.fro(B))
        Cf.free()

    # the precision form beside the FP64 line, on the same operands: the
    # BF16x3 leaf (four planes, ten plane products) and the INT8x4 leaf
    # (four int8 slices, ten slice products) with this run's algorithm and
    # with the classical one, each with its error against the classical
    # FP64 product
    bf3 = None
    if args.leaf == "f64" and not args.no_bf16x3 and not args.no_check:
        bf3 = []
        nAB = F.fro(A) * F.fro(B)
        for lname, lid, desc, label, a_run in (
                ("bf16x3", F.LEAF_BF16X3, "4 planes, 10 plane products", f"{args.alg} L{args.levels}", alg),
                ("bf16x3", F.LEAF_BF16X3, "4 planes, 10 plane products", "classical", F.Alg(classical=True)),
                ("int8x4", F.LEAF_INT8X4, "4 int8 slices, 10 slice products", f"{args.alg} L{args.levels}", alg),
                ("int8x4", F.LEAF_INT8X4, "4 int8 slices, 10 slice products", "classical",
                 F.Alg(classical=True))):
            Cb = F.DeviceMatrix(m, n).fill_uniform(0, 2)
            F.dgemm_fmm(a_run, m, n, k, 1.0, A, B, 1.0, Cb, leaf=lid)
            F.sync()
            b_err = F.fro_diff(Cb, Cr) / nAB
            be = [(F.Event(), F.Event()) for _ in range(3)]
            for e0, e1 in be:
                e0.record()
                F.dgemm_fmm(a_run, m, n, k, 1.0, A, B, 1.0, Cb, leaf=lid)
                e1.record()
            F.sync()
            b_ms = sum(e0.elapsed_ms(e1) for e0, e1 in be) / len(be)
            bf3.append({"leaf": f"{lname} leaf, {label} ({desc})",
                        "ms": round(b_ms, 3), "effective_gflops": round(2.0 * m * n * k / (b_ms * 1e-3) / 1e9, 1),
                        "rel_err_vs_classical": b_err, "meets_1e-10": b_err <= 1e-10})
            Cb.free()
            F.check(F.lib().fmmhip_workspace_release(), "workspace_release")

    fl = leaf_flops(info, m, n, k)
    peak = {"f64": FP64_PEAK_TFLOPS, "f32": FP32_PEAK_TFLOPS, "int8x4": INT8_PEAK_TOPS}.get(args.leaf,
                                                                                         BF16_PEAK_TFLOPS)
    # the split BF16 leaves run 10, 8 or 6 bf16 plane products per leaf
    # product, the INT8x4 leaf 10 int8 slice products (operations, not flops)
    mfma_fl = fl * {"bf16x3": 10, "bf16x3_8": 8, "bf16x3_6": 6, "int8x4": 10}.get(args.leaf, 1)
    achieved = mfma_fl / (kern_ms * 1e-3) / 1e12
    pack = {0: "fused", 1: "pre-summed", 2: "pre-summed, products materialised"}[pack_mode]
    # the counter summary of this workload and mode (profiles/traffic_<tag>.txt)
    tag = f"{args.alg}L{args.levels}{ {0: 'fused', 1: '', 2: 'mat'}[pack_mode] if args.alg != 'classical' else ''}" \
          f"_{args.leaf}_{n}"
    if args.leaf.startswith("bf16"):
        pack = "bf16 planes (split stage)" + (", products materialised" if pack_mode == 2 else "")
    if args.leaf == "int8x4":
        pack = "int8 slices (row-maximum and slice stages)" + (", products materialised" if pack_mode == 2 else "")
    out = {
        "metric": "effective FP64 GFLOPS (2mnk/t) FMM vs classical at 16384^3; p99 GEMM latency",
        "value": round(eff, 1),
        "unit": "GFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.leaf == "f64" else f"f64 ({args.leaf} leaf)",
        "data": "synthetic (SplitMix64 uniform[-1,1], seed 0, device-resident)",
        "config": {
            "workload": f"{args.levels}-level <{f.m_tilde},{f.k_tilde},{f.n_tilde};{f.R}> FMM, FP64 C += A*B, "
                        f"m=n=k={n}, column-major",
            "m": m, "n": n, "k": k, "alg": args.alg, "levels": args.levels, "leaf": args.leaf, "pack": pack,
            "composed": [info["mt"], info["kt"], info["nt"], info["R"]],
            "classical_fp64_gflops": round(cls_gflops, 1) if cls_gflops else None,
            "fmm_over_classical": round(eff / cls_gflops, 4) if cls_gflops else None,
            "rel_err_vs_classical": rel,
            "summation_ms_per_step": round(presum_ms, 3),     # materialised: the W stage included
            "precision_form": bf3,
            "pack_modes": modes,
        },
        "latency_ms": {"p50": round(percentile(lat, 50), 3), "p99": round(percentile(lat, 99), 3),
                       "mean": round(sum(lat) / len(lat), 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
                     "unit": "TOP/s" if args.leaf == "int8x4" else "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": read_traffic(tag),
                     "kernel": ("bf16x3_gemm_kernel" if args.leaf.startswith("bf16") else
                                "int8x4_gemm_kernel" if args.leaf == "int8x4" else
                                "fmm_fast_kernel" if alg.last_shape() == 3 else "fmm_gemm_kernel"),
                     "kernel_ms": round(kern_ms, 3),
                     "mfma_flops_per_launch": mfma_fl, "effective_over_peak": round(eff / 1e3 / peak, 4)},
    }
    if args.leaf == "f64" and not args.profile:
        # the device's own FP64 MFMA rate and copy bandwidth, beside the published peaks
        t_meas, gbs = F.measure_peaks()
        c_gbs, r_gbs, w_gbs = F.measure_hbm()
        out["roofline"]["measured_peak"] = round(t_meas, 2)
        out["roofline"]["frac_of_measured"] = round(achieved / t_meas, 4)
        out["roofline"]["measured_copy_gbs"] = round(gbs, 1)
        out["roofline"]["measured_hbm_gbs"] = {"copy": round(c_gbs, 1), "read": round(r_gbs, 1),
                                               "write": round(w_gbs, 1)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
