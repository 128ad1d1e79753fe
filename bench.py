#!/usr/bin/env python3
"""Benchmark: decoded info Gbit/s of the layered offset min-sum decode.

Workload (one "step" = one batched decode launch over resident inputs):
5G NR BG1, Z=384, rate 1/3 -- K = 8448 info bits, 68Z mother code with
the first 2Z columns punctured, 25,344 transmitted bits -- uniform random
info bits, BPSK over AWGN at Eb/N0 = 1.5 dB, f32 channel LLRs, at most 8
layered iterations with syndrome early exit, 4,096 codewords per launch.
Each codeword's result is exactly QcLdpcDecoder::decode_5g on it alone
(posterior, hard decisions, iterations).

Reported beside the headline number:
* the plain per-config GPU decoder (runtime-structured kernel, messages
  streamed through HBM) on the same batch, and the ratio;
* per-launch latency p50/p99 and single-codeword decode latency p50/p99;
* `roofline`: algorithmic bytes of the dominant kernel per launch over its
  average launch time (HIP events on the launch stream), against MI355X's
  8 TB/s HBM peak, and the HBM traffic measured by rocprofv3 counters
  when a profile summary is committed;
* `cpu_baseline`: the CPU restatement of the reference decode path timed
  on this host on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "info Gbit/s (BG1 Z=384, 8 it) batched vs plain; p50/p99 decode latency"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fast_f32_bg1_z384_b4096_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(dec, batch, ebno, seed, threads=16):
    """Uniform random info bits, encoded, BPSK/AWGN on the 66Z transmitted
    positions (first 2Z punctured -> LLR 0)."""
    from syndrome_amd import awgn_transmit
    n, k, z = dec.variable_node_count(), dec.info_bit_count_5g(), 384
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, size=(batch, k), dtype=np.uint8)
    cws = dec.encode_batch(info, threads=threads)
    llr = np.zeros((batch, n), np.float32)
    rate = k / (n - 2 * z)

    def work(t):
        for i in range(t, batch, threads):
            llr[i, 2 * z:] = awgn_transmit(ebno, rate, seed * 1000003 + i, cws[i, 2 * z:])

    ts = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return info, cws, llr


def time_launches(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([a.elapsed_time(b) for a, b in ev])  # ms, kernel on the launch stream
    return wall, per


def cpu_baseline(llr, iters, beta, budget_s=10.0):
    """The reference's CPU path restated (AVX2 f32 kernels + the multi-worker
    SPSC pipeline, oracle/cpu_baseline.cpp) on this host's cores, on a
    bounded sample of the same codewords: one worker to calibrate, then all
    workers on a sample sized for about `budget_s` seconds."""
    from oracle import orc
    code = orc.code_3gpp(0, 384)
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    t1, _, _ = orc.pipeline_run(code, beta, llr[:96], iters, 1)
    one = 96 / t1
    count = int(min(max(one * workers * budget_s * 0.7, 256), 20000))
    sample = llr[np.arange(count) % llr.shape[0]]
    el, its, _ = orc.pipeline_run(code, beta, sample, iters, workers)
    gbps = count * code.k / el / 1e9
    return {"value": round(gbps, 6), "unit": "Gbit/s", "cores": workers, "kind": "port",
            "sample": f"{count} codewords of the same workload through an AVX2 restatement of the reference decode "
                      f"and its SPSC worker pipeline, {workers} pinned workers, {el:.1f} s; 1 worker: "
                      f"{96 * code.k / t1 / 1e9:.6f} Gbit/s; mean iterations {float(its.mean()):.3f}"}


def summarize_profiles(trace_dir, fetch_dir, write_dir, out_prefix, kernel_substr, cmd_trace, cmd_pmc):
    """Turn rocprofv3 CSV output (kernel trace + two --pmc runs) into the
    committed text summary and the traffic JSON the roofline reads.

    HBM bytes per launch = FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of
    the decode kernel, each scaled by what the same counter reads for the
    device copy of the input batch (known bytes) in the same run."""
    import csv
    from collections import defaultdict

    def rows(d, name):
        return list(csv.DictReader(open(os.path.join(d, name))))

    tr = rows(trace_dir, "run_kernel_trace.csv")
    by_grid = defaultdict(list)
    for r in tr:
        if kernel_substr in r["Kernel_Name"]:
            by_grid[int(r["Grid_Size_X"]) if "Grid_Size_X" in r else int(r.get("Grid_Size", 0))].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    stats = rows(trace_dir, "run_kernel_stats.csv")
    counters = {}
    for nm, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        ker, cpy = [], []
        for r in rows(d, "run_counter_collection.csv"):
            v = float(r["Counter_Value"])
            if kernel_substr in r["Kernel_Name"] and int(r["Grid_Size"]) == max(by_grid):
                ker.append(v)
            elif "copyBuffer" in r["Kernel_Name"] and v > 1000:
                cpy.append(v)
        counters[nm] = (float(np.median(ker)), float(np.median(cpy)))
    calib = json.loads(open(os.path.join(fetch_dir, "..", "bench_copycal.json")).read())["copy_calibration_bytes"] \
        if os.path.exists(os.path.join(fetch_dir, "..", "bench_copycal.json")) else None
    lines = [f"# rocprofv3 summary: {kernel_substr}", "", "Kernel trace command:", "", f"    {cmd_trace}", "",
             "Counter commands (one counter per run):", "", f"    {cmd_pmc}", "", "## kernel trace (stats CSV, all dispatches)", "",
             "| kernel | calls | average ns | min ns | max ns |", "|---|---|---|---|---|"]
    for r in stats[:6]:
        lines.append(f"| {r['Name'][:70]} | {r['Calls']} | {float(r['AverageNs']):.0f} | {r['MinNs']} | {r['MaxNs']} |")
    lines += ["", "## decode kernel by grid size (trace CSV)", "", "| grid (work-items) | launches | average ms |",
              "|---|---|---|"]
    for g, v in sorted(by_grid.items()):
        lines.append(f"| {g} | {len(v)} | {np.mean(v):.4f} |")
    out = {"kernel": kernel_substr, "grid_work_items": max(by_grid),
           "average_ms_full_batch": float(np.mean(by_grid[max(by_grid)]))}
    if calib:
        fk, fc = counters["FETCH_SIZE"]
        wk, wc = counters["WRITE_SIZE"]
        f_scale = calib / (fc * 1024.0)
        w_scale = calib / (wc * 1024.0)
        fetch_b = fk * 1024.0 * f_scale
        write_b = wk * 1024.0 * w_scale
        out.update({"fetch_kib_raw": fk, "write_kib_raw": wk, "copy_fetch_kib_raw": fc, "copy_write_kib_raw": wc,
                    "copy_bytes": calib, "fetch_scale": f_scale, "write_scale": w_scale,
                    "hbm_fetch_bytes_per_launch": fetch_b, "hbm_write_bytes_per_launch": write_b,
                    "hbm_bytes_per_launch": fetch_b + write_b})
        lines += ["", "## HBM traffic per full-batch launch (--pmc)", "",
                  f"FETCH_SIZE {fk:.0f} KiB raw; the input copy ({calib} B) reads {fc:.0f} KiB raw -> scale {f_scale:.4f}",
                  f"WRITE_SIZE {wk:.0f} KiB raw; the input copy writes {wc:.0f} KiB raw -> scale {w_scale:.4f}",
                  f"=> {fetch_b / 1e6:.1f} MB fetched + {write_b / 1e6:.1f} MB written = "
                  f"{(fetch_b + write_b) / 1e6:.1f} MB per launch"]
    os.makedirs(os.path.dirname(out_prefix), exist_ok=True)
    open(out_prefix + ".md", "w").write("\n".join(lines) + "\n")
    open(out_prefix + "_traffic.json", "w").write(json.dumps(out, indent=1) + "\n")
    print("\n".join(lines))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-profiles":
        # bench.py --summarize-profiles TRACE_DIR FETCH_DIR WRITE_DIR OUT_PREFIX KERNEL CMD_TRACE CMD_PMC
        summarize_profiles(*sys.argv[2:9])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--ebno", type=float, default=1.5)
    ap.add_argument("--beta", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=200,
                    help="single-codeword launches timed for latency (0 = skip)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run device copies of the input batch (known bytes) to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has exactly one MI355X, and the decode "
              "shards by codeword with no exchange, so N>1 is not measured here.", file=sys.stderr)
        sys.exit(2)

    import torch
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, _lib as L

    syndrome_amd.lib()
    dev = QcLdpcDecoder.with_lifting_size(syndrome_amd.BG1, 384, args.beta)
    n, k = dev.variable_node_count(), dev.info_bit_count_5g()
    ez = dev.required_edge_buffer()
    B = args.batch
    log(f"[bench] generating {B} codewords (BG1 Z=384 R=1/3, {args.ebno} dB)")
    t = time.perf_counter()
    info, cws, llr = make_workload(dev, B, args.ebno, args.seed)
    log(f"[bench] workload ready in {time.perf_counter() - t:.1f} s")

    d_in = torch.from_numpy(llr).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD

    def fast():
        dev.decode_batch_f32(d_in, d_post, d_hard, d_it, args.iters, flags=flags)

    log("[bench] timing the batched decoder")
    wall, per = time_launches(fast, args.steps, args.warmup, torch)
    ms = wall * 1e3 / args.steps
    hard = d_hard.cpu().numpy()
    its = d_it.cpu().numpy()
    bit_err = int((hard[:, :k] != info).sum())
    bler = float(((hard[:, :k] != info).any(axis=1)).mean())
    mean_it = float(its.mean())
    value = B * k / (ms * 1e-3) / 1e9

    # roofline of the dominant kernel (fast_f32<BG1>), algorithmic bytes
    # per codeword: N*4 (input) + N*4 (posterior) + N*1 (hard) + t*2*E*Z*4 (R read+write per iteration)
    alg_bytes = float(B * (n * 4 + n * 4 + n) + int(its.sum()) * 2 * ez * 4)
    kern_ms = float(per.mean())
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_PROFILE):
        try:
            traffic = json.load(open(TRAFFIC_PROFILE)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "syn::fast_f32<BG1>", "kernel_ms": round(kern_ms, 4),
            "algorithmic_bytes_per_launch": alg_bytes}

    out = {}
    if args.copy_calibrate:
        d_copy = torch.empty_like(d_in)
        for _ in range(3):
            d_copy.copy_(d_in)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = int(d_in.numel() * 4)
        del d_copy
        cal = os.environ.get("SYNDROME_COPYCAL_OUT")
        if cal:
            open(cal, "w").write(json.dumps({"copy_calibration_bytes": out["copy_calibration_bytes"]}))
    if not args.no_plain:
        log("[bench] timing the plain per-config decoder")
        d_edge = torch.empty((B, ez), dtype=torch.float32, device="cuda")
        d_post2 = torch.empty_like(d_in)
        d_hard2 = torch.empty_like(d_hard)
        d_it2 = torch.empty_like(d_it)

        def plain():
            dev.decode_batch_f32(d_in, d_post2, d_hard2, d_it2, args.iters, flags=flags | L.PLAIN, d_edge=d_edge)

        pwall, pper = time_launches(plain, args.plain_steps, 1, torch)
        pms = pwall * 1e3 / args.plain_steps
        same = bool(torch.equal(d_hard2, d_hard) and torch.equal(d_it2, d_it))
        out["plain"] = {"value": round(B * k / (pms * 1e-3) / 1e9, 4), "unit": "Gbit/s", "ms_per_step": round(pms, 3),
                        "identical_decoded_bits": same}
        out["batched_vs_plain"] = round(value / out["plain"]["value"], 3)
        del d_edge

    # single code block latency (one codeword per launch, inputs resident)
    one = lambda: dev.decode_batch_f32(d_in[:1], d_post[:1], d_hard[:1], d_it[:1], args.iters, flags=flags)
    lat1 = time_launches(one, args.latency_samples, 10, torch)[1] if args.latency_samples > 0 else np.zeros(1)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(llr, args.iters, args.beta)

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gbit/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded uniform info bits, BG1 encoder, BPSK/AWGN 1.5 dB, first 2Z punctured",
        "config": {"workload": "5G NR BG1 Z=384 rate 1/3 (K=8448, 25,344 transmitted), f32 LLRs, max 8 iterations "
                               "with syndrome early exit, batch 4,096 codewords per launch",
                   "batch": B, "max_iterations": args.iters, "ebno_db": args.ebno, "offset_beta": args.beta},
        "latency_ms": {"batch_launch_p50": round(float(np.percentile(per, 50)), 4),
                       "batch_launch_p99": round(float(np.percentile(per, 99)), 4),
                       "single_codeword_p50": round(float(np.percentile(lat1, 50)), 4),
                       "single_codeword_p99": round(float(np.percentile(lat1, 99)), 4)},
        "mean_iterations": round(mean_it, 4), "bler": bler, "bit_errors": bit_err,
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
