#!/usr/bin/env python3
"""Bench: 40 independent plonk-like proofs of N = 2^14 (Poseidon2, 200 FRI
queries), plain per-instance GPU proving vs all 40 proved together in
batched launches, bit-exact.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one pass of the prove phase over the 40-instance batch (synthetic
seeded instances, inputs already in host memory in Montgomery form; the
trace columns are copied to device memory before the timed region and the
prove calls read them there, lg_prove_device).  The same batch proved from
host buffers (PCIe upload inside the call, lg_prove) is reported beside it.  `value` is batched proofs/s; the plain path, p50/p99 prove
latencies, the kernel roofline and the CPU baseline (the numpy oracle on a
bounded sample, one 2^14 proof) are reported beside it.  One JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=40)
    ap.add_argument("--logn", type=int, default=14)
    ap.add_argument("--queries", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-oracle-check", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="prove with the default launch choices (no autotuning)")
    ap.add_argument("--requests", type=int, default=100, help="serve: number of requests")
    ap.add_argument("--qps", type=float, default=40.0, help="serve: Poisson arrival rate")
    ap.add_argument("--slots", type=int, default=3, help="serve: concurrent proving slots")
    ap.add_argument("--workload", default="plonk40", choices=["plonk40", "uneven64", "wide", "serve"],
                    help="plonk40: the 40 x 2^14 batch (default); uneven64: one proof of 64 plonk-like "
                         "modules of 2^10..2^20 rows with range lookups; wide: one proof of a 2^22-row, "
                         "96-column Fibonacci/cubic trace")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the proofs are independent, so N GPUs would each run this "
              "same batch in a process of their own (one context per GPU, no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        n = args.copy_probe << 18
        x = torch.ones(n, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "copies": 5}))
        return
    from loom_amd import workloads, prover
    if args.workload == "serve":
        return serve_bench(args, torch, workloads, prover)
    if args.workload != "plonk40":
        return single_proof_bench(args, torch, workloads, prover)

    B, logn, nq = args.instances, args.logn, args.queries
    print("[bench] building %d instances of 2^%d" % (B, logn), file=sys.stderr, flush=True)
    inst = [workloads.plonk_instance(logn, seed) for seed in range(B)]
    items = [prover.Prepared(p, t, nq) for p, t in inst]
    ctx = prover.Context(0)
    # inputs resident in HBM before the timed region (lg_prove_device)
    dev_all = prover.DeviceInputs(items, 0)
    dev_one = [prover.DeviceInputs([it], 0) for it in items]

    # autotuning (untimed): launch choices for the batched and the plain path,
    # picked by timing whole proves of these inputs (lg_plan_autotune)
    tuning = None
    if not args.no_tune:
        print("[bench] autotuning", file=sys.stderr, flush=True)
        tb = prover.autotune(ctx, items, dev_inputs=dev_all, reps=7)
        tp = prover.autotune(ctx, items[:1], dev_inputs=dev_one[0], reps=7)
        tuning = {"batched": tb, "plain": tp}
    # warmup (plan compile, buffer pools, graph recording) for both paths
    for _ in range(max(1, args.warmup)):
        prover.prove_prepared(ctx, items[:1], dev_inputs=dev_one[0])
        prover.prove_prepared(ctx, items, dev_inputs=dev_all)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # plain: one proof per call, sequentially
    plain_lat = []
    plain_words = None
    t0 = time.perf_counter()
    for step in range(args.steps):
        out = []
        for it, dv in zip(items, dev_one):
            a = time.perf_counter()
            out.append(prover.prove_prepared(ctx, [it], dev_inputs=dv)[0])
            plain_lat.append((time.perf_counter() - a) * 1e3)
        plain_words = out
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    print("[bench] plain done %.3fs" % plain_s, file=sys.stderr, flush=True)

    # batched: all instances in one call
    batch_lat = []
    batch_words = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(args.steps):
        a = time.perf_counter()
        batch_words = prover.prove_prepared(ctx, items, dev_inputs=dev_all)
        batch_lat.append((time.perf_counter() - a) * 1e3)
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t0
    print("[bench] batched done %.3fs" % batch_s, file=sys.stderr, flush=True)

    # batched from host buffers (upload over PCIe inside the call), reported beside
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(args.steps):
        host_words = prover.prove_prepared(ctx, items)
    torch.cuda.synchronize()
    host_s = time.perf_counter() - t0
    host_same = all(np.array_equal(a, b) for a, b in zip(host_words, batch_words))

    identical = all(np.array_equal(a, b) for a, b in zip(plain_words, batch_words))
    graphs = prover.graph_stats(ctx, items)  # recordings, replays, eager fallbacks of the plan

    # kernel timing pass (untimed): HIP events around every launch of one batched step
    ctx.set_timing(True)
    prover.prove_prepared(ctx, items, dev_inputs=dev_all)
    ctx.set_timing(False)
    stats = ctx.timing_report()
    tot_ms = sum(v[1] for v in stats.values())
    dom = max(stats, key=lambda k: stats[k][1])
    cnt, ms, byts, perms = stats[dom]
    kernels = {k: {"launches": v[0], "ms": round(v[1], 4), "share": round(v[1] / tot_ms, 4)}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    # roofline line: the rate-4 LDE of the committed columns (reedsolomon.Encode),
    # the streaming kernel of the commit phase; algorithmic bytes 20 N per limb-column
    hk = "ntt_lde"
    hc, hms, hbytes, _ = stats[hk]
    achieved = (hbytes / hc) / ((hms / hc) * 1e-3) / 1e9
    roofline = {"kernel": hk, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes_per_launch": hbytes / hc, "avg_launch_us": round(hms / hc * 1e3, 3)}
    traffic_file = os.path.join(ROOT, "profiles", "traffic_%s.json" % hk)
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    p2_rate = sum(v[3] for v in stats.values()) / (sum(v[1] for k, v in stats.items() if v[3] > 0) * 1e-3)
    # VALU roofline of the dominant kernel (Poseidon2 leaf hashing): 834
    # modular multiplications per permutation (6*24*2 full-round S-box + 21*2
    # partial-round S-box + 21*24 internal diagonal) against the Montgomery
    # product rate this device sustains (issue-rate probe, measured here)
    probe = ctx.issue_probe()
    lc, lms, _, lperms = stats["leaf_hash"]
    valu_achieved = lperms * 834 / (lms * 1e-3) / 1e12
    valu_peak = probe["montgomery"] / 1e3
    valu_roofline = {"kernel": "leaf_hash", "bound": "valu", "achieved": round(valu_achieved, 3),
                     "peak": round(valu_peak, 3), "unit": "T modmul/s", "frac": round(valu_achieved / valu_peak, 4),
                     "permutations_per_launch": lperms / lc, "avg_launch_us": round(lms / lc * 1e3, 3)}
    probe = {k: round(v, 2) for k, v in probe.items()}

    # batched verification of the batch's proofs (lg_verify), untimed for the
    # headline: proofs/s and every verdict
    for _ in range(2):
        vres = prover.verify_prepared(ctx, items, batch_words)
    vt = []
    for _ in range(5):
        a = time.perf_counter()
        vres = prover.verify_prepared(ctx, items, batch_words)
        vt.append(time.perf_counter() - a)
    verify = {"proofs_per_s": round(B / float(np.median(vt)), 1), "ms_per_batch": round(float(np.median(vt)) * 1e3, 3),
              "all_valid": all(r == (0, "") for r in vres)}

    # CPU baseline: the oracle on a bounded sample (one proof of instance 0)
    cpu = None
    oracle_ok = None
    if not args.no_cpu:
        from oracle import prover as OP
        print("[bench] cpu baseline (oracle, 1 proof of 2^%d, %d queries)" % (logn, nq), file=sys.stderr,
              flush=True)
        a = time.perf_counter()
        _, enc, _ = OP.prove(inst[0][0], inst[0][1], nq)
        dt = time.perf_counter() - a
        cpu = {"value": round(1.0 / dt, 4), "unit": "proofs/s", "cores": 1, "kind": "port",
               "sample": "1 proof of instance 0 (2^%d plonk-like, %d queries) by the numpy oracle, %.1f s"
                         % (logn, nq, dt)}
        if not args.no_oracle_check:
            oracle_ok = bool(np.array_equal(enc, batch_words[0]))
        from oracle import verifier as OV
        a = time.perf_counter()
        OV.verify(inst[0][0], nq, batch_words[0])
        verify["cpu_oracle_s_per_proof"] = round(time.perf_counter() - a, 3)

    n_proofs = B * args.steps
    value = n_proofs / batch_s
    plain_rate = n_proofs / plain_s
    line = {
        "metric": "proofs/s + p50/p99 prove ms, 40x2^14 plonk, batched vs plain GPU, bit-exact",
        "value": round(value, 3),
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (KoalaBear Montgomery, E6)",
        "data": "synthetic seeded plonk-like instances (seeds 0..%d)" % (B - 1),
        "config": {"workload": "%d independent plonk-like proofs, N=2^%d, Poseidon2, %d FRI queries, rate 4"
                               % (B, logn, nq),
                   "instances": B, "log_n": logn, "queries": nq, "parallelism": "batched launches, 1 GPU"},
        "plain_proofs_per_s": round(plain_rate, 3),
        "batched_over_plain": round(value / plain_rate, 3),
        "plain_p50_ms": round(pct(plain_lat, 50), 3),
        "plain_p99_ms": round(pct(plain_lat, 99), 3),
        "batched_p50_ms": round(pct(batch_lat, 50), 3),
        "batched_p99_ms": round(pct(batch_lat, 99), 3),
        "bit_exact_batched_vs_plain": identical and host_same,
        "batched_host_inputs_proofs_per_s": round(n_proofs / host_s, 3),
        "bit_exact_vs_oracle": oracle_ok,
        "graphs": graphs,
        "verify": verify,
        "poseidon2_perms_per_s": round(p2_rate, 1),
        "dominant_kernel": dom,
        "kernels": kernels,
        "tuning": tuning,
        "roofline": roofline,
        "valu_roofline": valu_roofline,
        "issue_probe_gops": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


def serve_bench(args, torch, workloads, prover):
    """Serving: a Poisson stream of requests, each 1-40 plonk-like instances of
    one size, log2 N geometric on 12..20 (P(12) = 1/2, halving per step),
    against serial plain proving of the same requests (one instance per call,
    in arrival order; its latencies are those of a FIFO queue in front of it)."""
    from loom_amd import serve
    rng = np.random.Generator(np.random.PCG64(12345))
    sizes = list(range(12, 21))
    probs = np.array([2.0 ** -(k + 1) for k in range(len(sizes))])
    probs /= probs.sum()
    pool = {}
    for ln in sizes:  # a few distinct instances per size, drawn by the requests
        cnt = 6 if ln <= 16 else 2
        pool[ln] = [prover.Prepared(*workloads.plonk_instance(ln, 1000 * ln + s), args.queries) for s in range(cnt)]
    print("[bench] serve: instance pool ready", file=sys.stderr, flush=True)
    reqs = []
    t = 0.0
    for _ in range(args.requests):
        t += rng.exponential(1.0 / args.qps)
        ln = int(rng.choice(sizes, p=probs))
        k = int(rng.integers(1, 41))
        reqs.append((t, ln, [pool[ln][int(rng.integers(0, len(pool[ln])))] for _ in range(k)]))
    n_inst = sum(len(r[2]) for r in reqs)
    ctx = prover.Context(0)
    # warm every plan on the plain context and on the server slots
    for ln in sizes:
        prover.prove_prepared(ctx, pool[ln][:1])
    srv = serve.Server(0, slots=args.slots, max_batch=64)
    for ln in sizes:
        srv.wait(srv.submit(pool[ln][:1]))
    print("[bench] serve: warm, %d requests, %d instances" % (len(reqs), n_inst), file=sys.stderr, flush=True)
    # serving run
    pending = {}
    lat = []
    results = {}
    t0 = time.perf_counter()
    i = 0
    while i < len(reqs) or pending:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            pending[srv.submit(reqs[i][2])] = (i, reqs[i][0])
            i += 1
        for tk in list(pending):
            got = srv.wait(tk, 0)
            if got is not None:
                ri, ta = pending.pop(tk)
                lat.append((time.perf_counter() - t0 - ta) * 1e3)
                results[ri] = got
        time.sleep(0.0005)
    serve_s = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    print("[bench] serve: served in %.2fs" % serve_s, file=sys.stderr, flush=True)
    # serial plain proving of the same requests
    service = []
    same = True
    for ri, (ta, ln, items) in enumerate(reqs):
        a = time.perf_counter()
        for j, it in enumerate(items):
            w = prover.prove_prepared(ctx, [it])[0]
            same = same and np.array_equal(w, results[ri][j])
        service.append(time.perf_counter() - a)
    serial_s = sum(service)
    fin = 0.0
    slat = []
    for (ta, _, _), sv in zip(reqs, service):
        fin = max(fin, ta) + sv
        slat.append((fin - ta) * 1e3)
    ctx.close()
    line = {
        "metric": "serving: proofs/s and p50/p99 request latency vs serial plain GPU proving",
        "value": round(n_inst / serve_s, 2), "unit": "proofs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "u32 (KoalaBear Montgomery, E6)", "data": "synthetic seeded",
        "config": {"workload": "%d requests, Poisson arrivals at %.1f req/s, 1-40 plonk-like instances each, "
                               "log2 N geometric on 12..20, %d FRI queries, %d slots"
                               % (len(reqs), args.qps, args.queries, args.slots)},
        "instances": n_inst, "served_s": round(serve_s, 3), "p50_ms": round(pct(lat, 50), 2),
        "p99_ms": round(pct(lat, 99), 2), "serial_plain_proofs_per_s": round(n_inst / serial_s, 2),
        "serial_plain_p50_ms": round(pct(slat, 50), 2), "serial_plain_p99_ms": round(pct(slat, 99), 2),
        "bit_exact_served_vs_plain": same, "server_stats": st,
    }
    print(json.dumps(line))


def single_proof_bench(args, torch, workloads, prover):
    """One large proof per step (the uneven-module and wide-trace shapes)."""
    if args.workload == "uneven64":
        sizes = workloads.uneven_sizes(64, 0, 10, 20)
        print("[bench] uneven64: module log sizes %s" % sizes, file=sys.stderr, flush=True)
        prog, tr = workloads.uneven_range_instance(sizes, 0, 256)
        desc = "one proof: 64 plonk-like modules, log2 N uniform in 10..20 (total %d rows), each with a " \
               "logup range lookup into range_256, Poseidon2, %d FRI queries" % (sum(1 << x for x in sizes), args.queries)
    else:
        print("[bench] wide: generating 2^22 x 96 trace", file=sys.stderr, flush=True)
        prog, tr = workloads.wide_instance(22, 0)
        desc = "one proof: N=2^22, 96 base columns (32 Fibonacci pairs, 32 cubic chains), Poseidon2, " \
               "%d FRI queries" % args.queries
    item = prover.Prepared(prog, tr, args.queries)
    del tr
    ctx = prover.Context(0)
    dev = prover.DeviceInputs([item], 0)
    print("[bench] inputs on device", file=sys.stderr, flush=True)
    tuning = None if args.no_tune else prover.autotune(ctx, [item], dev_inputs=dev, reps=1)
    # the first prove after tuning runs eagerly and the second records the
    # HIP graph of the device phase: both stay out of the timed steps
    for _ in range(max(2, args.warmup)):
        first = prover.prove_prepared(ctx, [item], dev_inputs=dev)[0]
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    lat = []
    same = True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        w = prover.prove_prepared(ctx, [item], dev_inputs=dev)[0]
        lat.append((time.perf_counter() - a) * 1e3)
        same = same and np.array_equal(w, first)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print("[bench] timed steps done", file=sys.stderr, flush=True)
    ctx.set_timing(True)
    prover.prove_prepared(ctx, [item], dev_inputs=dev)
    ctx.set_timing(False)
    stats = ctx.timing_report()
    tot_ms = sum(v[1] for v in stats.values())
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3), "share": round(v[1] / tot_ms, 4),
                   "GB_per_s": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[2] else None}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    roofline = None
    if "ntt_lde" in stats:
        hc, hms, hbytes, _ = stats["ntt_lde"]
        ach = hbytes / (hms * 1e-3) / 1e9
        roofline = {"kernel": "ntt_lde", "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "algorithmic_bytes": hbytes, "launches": hc, "ms": round(hms, 3)}
    line = {
        "metric": "proofs/s (single large proof)", "value": round(args.steps / dt, 4), "unit": "proofs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "u32 (KoalaBear Montgomery, E6)", "data": "synthetic seeded",
        "config": {"workload": desc}, "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
        "deterministic": same, "kernel_ms_total": round(tot_ms, 2), "kernels": kernels, "roofline": roofline,
        "tuning": tuning,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
