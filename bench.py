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


def _cpu_model():
    """The host CPU as lscpu names it (model, family/model/stepping, sockets,
    cores, threads), or /proc/cpuinfo's model name without lscpu."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        keys = ("Model name", "Vendor ID", "CPU family", "Model", "Stepping", "Socket(s)", "Core(s) per socket",
                "Thread(s) per core", "CPU(s)", "CPU max MHz")
        got = {k: kv[k] for k in keys if k in kv}
        if got:
            # the kernel's own name for the part, and the AMD family's core
            # generation (family 0x19 = Zen 3 / Zen 4, 0x1A = Zen 5)
            try:
                for line in open("/proc/cpuinfo"):
                    if line.startswith("model name"):
                        got["/proc/cpuinfo model name"] = line.split(":", 1)[1].strip()
                        break
            except OSError:
                pass
            fam = {"25": "Zen 3 / Zen 4", "26": "Zen 5"}.get(got.get("CPU family", ""))
            if got.get("Vendor ID") == "AuthenticAMD" and fam:
                got["core generation (from CPU family)"] = fam
            return got
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return {"Model name": line.split(":", 1)[1].strip()}
    except OSError:
        pass
    return {"Model name": "unknown"}


def _host_cores():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n))


CPU_DESC = ("the reference's CPU decode restated: its AVX2 {kind} kernels (simd_avx2.rs) driven through its "
            "SPSC worker pipeline (ldpc_pipeline.rs), workers pinned, tested bit-exact against the oracle "
            "(tests/test_cpu_baseline.py)")


def cpu_pipeline_baseline(codes, frame_code, offs, arena, info_bits, iters, beta, int8=False, n_filler=None,
                          budget_s=8.0, seed=7, points=None):
    """The CPU pipeline (oracle/cpu_baseline.cpp) on a bounded sample of a
    workload: frames drawn at random from the workload's own decoder inputs
    (arena, per-frame code and offset), a worker sweep over `points` (default
    1 and every core this process may use, at most 16), each point sized for
    about budget_s / len(points) seconds.  `value` = info Gbit/s of the best
    point (info_bits per frame, summed over the stream)."""
    from oracle import orc
    cores = _host_cores()
    points = sorted(set(points or (1, cores)))
    rng = np.random.default_rng(seed)
    nfr = len(frame_code)
    probe = rng.integers(0, nfr, 32).astype(np.int32)
    t1, _, _ = orc.pipeline_mixed(codes, frame_code, offs, arena, iters, 1, order=probe, int8=int8, beta=beta,
                                  n_filler=n_filler)
    per_frame = t1 / len(probe)
    sweep, best = {}, None
    for w in points:
        count = int(min(max(budget_s / len(points) / per_frame * w * 0.8, 64), 200000))
        order = rng.integers(0, nfr, count).astype(np.int32)
        el, its, _ = orc.pipeline_mixed(codes, frame_code, offs, arena, iters, w, order=order, int8=int8,
                                        beta=beta, n_filler=n_filler)
        bits = float(np.asarray(info_bits, np.float64)[order].sum())
        gbps = bits / el / 1e9
        sweep[str(w)] = round(gbps, 6)
        if best is None or gbps > best[0]:
            best = (gbps, w, count, el, float(its[np.unique(order)].mean()))
    gbps, w, count, el, mit = best
    return {"value": round(gbps, 6), "unit": "Gbit/s", "cores": w, "kind": "port",
            "sample": f"{count} frames drawn at random from this workload's own decoder inputs, through "
                      + CPU_DESC.format(kind="int8" if int8 else "f32") + f"; {w} workers, {el:.1f} s; "
                      f"mean iterations {mit:.3f}",
            "worker_sweep_gbps": sweep, "cpu_model": _cpu_model(), "host_cores_available": cores}


def cpu_baseline(llr, iters, beta, budget_s=12.0):
    """The reference's CPU path restated (AVX2 f32 kernels + the multi-worker
    SPSC pipeline, oracle/cpu_baseline.cpp) on this host's cores, on bounded
    samples of the same codewords: a worker sweep 1, 2, 4, ... up to the
    cores this process may use (at most 16), each sample sized for about
    budget_s / (sweep points) seconds; `value` is the best point."""
    from oracle import orc
    code = orc.code_3gpp(0, 384)
    cores = _host_cores()
    t1, _, _ = orc.pipeline_run(code, beta, llr[:96], iters, 1)
    one = 96 / t1  # codewords per second on one worker
    points = sorted({w for w in (1, 2, 4, 8, 16) if w <= cores} | {cores})
    sweep, best, best_w, best_n, best_el, mean_it = {}, 0.0, 1, 0, 0.0, 0.0
    for w in points:
        count = int(min(max(one * w * budget_s / len(points) * 0.7, 128), 40000))
        sample = llr[np.arange(count) % llr.shape[0]]
        el, its, _ = orc.pipeline_run(code, beta, sample, iters, w)
        gbps = count * code.k / el / 1e9
        sweep[str(w)] = round(gbps, 6)
        mean_it = float(its.mean())
        if gbps > best:
            best, best_w, best_n, best_el = gbps, w, count, el
    return {"value": round(best, 6), "unit": "Gbit/s", "cores": best_w, "kind": "port",
            "sample": f"{best_n} codewords of the same workload through " + CPU_DESC.format(kind="f32")
                      + f"; {best_w} workers, {best_el:.1f} s; mean iterations {mean_it:.3f}",
            "worker_sweep_gbps": sweep, "cpu_model": _cpu_model(), "host_cores_available": cores}


# ---------------------------------------------------------------------------
# roofline: the decode kernels are bound by VALU issue, not by HBM
# ---------------------------------------------------------------------------

# MI355X VALU peak in lane-operations per second: 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz = 39.3 T (the published 157.3 TFLOP/s fp32 counts an FMA
# as two operations and packed-FP32 dual issue as two more).
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# Algorithmic operations per edge-slot-iteration ((edge, z) pair per layered
# pass), counted from the reference's scalar kernel with nothing done twice
# (its pass 2 recomputes |q| and q's sign, counted once):
#  f32 (qc_ldpc.rs:1092, 1155-1178, 1181-1204): Q-build sub 1; pass 1: |q|
#   mask, sign mask, sign xor, a <= m1, a < m2 (5); pass 2: exclusive sign
#   xor, a == m1, select, - beta, mag < 0, select, or sign, R' - R, L + d (9):
#   15 in all;
#  int8 (qc_ldpc.rs:1639-1640, 1679-1690, 1708-1742): Q-build sub, clamp
#   max, min (3); pass 1: abs, sign mask, xor, min, max, min (6); pass 2:
#   exclusive sign xor, a == m1, select, - beta_q, max 0, sign as neg =
#   s >> 7, xor, sub (3), delta, saturating add, clamp max, min (13): 22.
ALG_OPS_PER_EDGE_SLOT = {"f32": 15, "i8": 22}
# The int8 kernel for BG2 runs two codewords per lane in the 16-bit halves
# of packed instructions (v_pk_*_i16/u16): one lane-instruction performs two
# of the operations above, so its roof is twice the lane-instruction peak.
PACKED_OPS_PER_LANE_INSTRUCTION = 2


def _profile_json(name, suffix):
    path = os.path.join(ROOT, "profiles", f"{name}_{suffix}.json")
    if os.path.exists(path):
        try:
            return json.load(open(path))
        except (OSError, ValueError):
            return None
    return None


def valu_roofline(kernel, dtype, edge_slot_iters, kernel_ms, compulsory_bytes, profile, packed=False):
    """Roofline of one decode launch: achieved algorithmic lane-operations
    (ALG_OPS_PER_EDGE_SLOT x executed edge-slot-iterations) over the
    launch's average duration against the VALU peak, which is the binding
    roof; beside it the VALU issue share and VALU instructions per
    edge-slot-iteration that the SQ counters of profiles/<profile>.md read,
    and HBM: the compulsory bytes (input, posterior, hard decisions) over the
    same time, with the counter traffic of that profile.  packed: the kernel
    does its arithmetic in packed 16-bit instructions, whose peak is
    PACKED_OPS_PER_LANE_INSTRUCTION operations per lane-instruction."""
    ops = ALG_OPS_PER_EDGE_SLOT[dtype] * float(edge_slot_iters)
    t = kernel_ms * 1e-3
    tops = ops / t / 1e12
    peak = VALU_PEAK_TOPS * (PACKED_OPS_PER_LANE_INSTRUCTION if packed else 1)
    traffic = (_profile_json(profile, "traffic") or {}).get("hbm_bytes_per_launch")
    sq = _profile_json(profile, "sq") or {}
    hbm = compulsory_bytes / t / 1e9
    return {"bound": "valu", "achieved": round(tops, 3), "peak": round(peak, 2),
            "unit": "Top/s (16-bit lanes of packed instructions)" if packed else "Tlane-op/s",
            "frac": round(tops / peak, 4), "traffic": traffic, "kernel": kernel,
            "kernel_ms": round(kernel_ms, 4), "alg_ops_per_edge_slot_iteration": ALG_OPS_PER_EDGE_SLOT[dtype],
            "edge_slot_iterations_per_launch": float(edge_slot_iters),
            "valu_issue_frac_profiled": sq.get("valu_issue_frac"),
            "valu_per_edge_slot_iteration_profiled": sq.get("valu_per_edge_slot_iteration"),
            "hbm": {"compulsory_bytes_per_launch": float(compulsory_bytes), "achieved": round(hbm, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(hbm / HBM_PEAK_GBS, 4),
                    "traffic_achieved": round(traffic / t / 1e9, 1) if traffic else None},
            "profile": f"profiles/{profile}.md"}


def _launch_key(r):
    """(grid work-items, workgroup size, kernel) of a rocprofv3 CSV row: two
    launches of one kernel can share a grid size with different workgroups,
    and two kernels a launch shape."""
    g = r["Grid_Size_X"] if "Grid_Size_X" in r else r.get("Grid_Size", 0)
    w = r["Workgroup_Size_X"] if "Workgroup_Size_X" in r else r.get("Workgroup_Size", 0)
    return int(g), int(w), r["Kernel_Name"].split("(")[0].replace("void ", "")


def _kernel_match(kernel_substr, name):
    """A kernel name matches any of the '|'-separated substrings."""
    return any(k in name for k in kernel_substr.split("|"))


def summarize_profiles(trace_dir, fetch_dir, write_dir, out_prefix, kernel_substr, cmd_trace, cmd_pmc):
    """Turn rocprofv3 CSV output (kernel trace + two --pmc runs) into the
    committed text summary and the traffic JSON the roofline reads.

    HBM bytes per launch = FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of
    the profiled kernel, each scaled by what the same counter reads for the
    calibration copy of known bytes in the same run (the run's largest
    device copy).  A workload whose step launches the kernel at several grid
    sizes (a plan's two launches) also gets per-step figures: the sums over
    its grids of the per-grid averages."""
    import csv
    from collections import defaultdict

    def rows(d, name):
        return list(csv.DictReader(open(os.path.join(d, name))))

    tr = rows(trace_dir, "run_kernel_trace.csv")
    by_grid = defaultdict(list)
    for r in tr:
        if _kernel_match(kernel_substr, r["Kernel_Name"]):
            by_grid[_launch_key(r)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    stats = rows(trace_dir, "run_kernel_stats.csv")
    counters = {}
    for nm, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        ker, cpy = defaultdict(list), []
        for r in rows(d, "run_counter_collection.csv"):
            v = float(r["Counter_Value"])
            if _kernel_match(kernel_substr, r["Kernel_Name"]):
                ker[_launch_key(r)].append(v)
            elif "copyBuffer" in r["Kernel_Name"]:
                cpy.append((int(r["Grid_Size"]), v))
        big = max(g for g, _ in cpy)
        counters[nm] = ({g: float(np.median(v)) for g, v in ker.items()},
                        float(np.median([v for g, v in cpy if g == big])))
    calib_path = os.path.join(fetch_dir, "..", "bench_copycal.json")
    calib = json.loads(open(calib_path).read())["copy_calibration_bytes"] if os.path.exists(calib_path) else None
    lines = [f"# rocprofv3 summary: {kernel_substr}", "", "Kernel trace command:", "", f"    {cmd_trace}", "",
             "Counter commands (one counter per run):", "", f"    {cmd_pmc}", "", "## kernel trace (stats CSV, all dispatches)", "",
             "| kernel | calls | average ns | min ns | max ns |", "|---|---|---|---|---|"]
    for r in stats[:6]:
        lines.append(f"| {r['Name'][:70]} | {r['Calls']} | {float(r['AverageNs']):.0f} | {r['MinNs']} | {r['MaxNs']} |")
    lines += ["", "## profiled kernel by launch shape (trace CSV)", "",
              "| kernel | grid (work-items) | workgroup | launches | average ms |", "|---|---|---|---|---|"]
    for g, v in sorted(by_grid.items()):
        lines.append(f"| {g[2]} | {g[0]} | {g[1]} | {len(v)} | {np.mean(v):.4f} |")
    top = max(by_grid)
    out = {"kernel": kernel_substr, "grid_work_items": top[0], "workgroup": top[1],
           "average_ms_full_batch": float(np.mean(by_grid[top])),
           "average_ms_per_step_all_grids": float(sum(np.mean(v) for v in by_grid.values()))}
    if len(by_grid) > 1:
        lines += ["", f"One step launches every shape above once: {out['average_ms_per_step_all_grids']:.4f} ms of "
                      f"kernel time per step in all (launches that run concurrently overlap within it)."]
    if calib:
        fk, fc = counters["FETCH_SIZE"]
        wk, wc = counters["WRITE_SIZE"]
        f_scale = calib / (fc * 1024.0)
        w_scale = calib / (wc * 1024.0)
        fetch_b = fk[top] * 1024.0 * f_scale
        write_b = wk[top] * 1024.0 * w_scale
        step_b = sum(fk.values()) * 1024.0 * f_scale + sum(wk.values()) * 1024.0 * w_scale
        out.update({"fetch_kib_raw": fk[top], "write_kib_raw": wk[top], "copy_fetch_kib_raw": fc,
                    "copy_write_kib_raw": wc, "copy_bytes": calib, "fetch_scale": f_scale, "write_scale": w_scale,
                    "hbm_fetch_bytes_per_launch": fetch_b, "hbm_write_bytes_per_launch": write_b,
                    "hbm_bytes_per_launch": fetch_b + write_b if len(by_grid) == 1 else step_b,
                    "hbm_bytes_full_batch_launch": fetch_b + write_b, "hbm_bytes_per_step_all_grids": step_b})
        lines += ["", "## HBM traffic (--pmc)", "",
                  f"The calibration copy ({calib} B) reads FETCH_SIZE {fc:.0f} KiB raw -> scale {f_scale:.4f}, "
                  f"and WRITE_SIZE {wc:.0f} KiB raw -> scale {w_scale:.4f}.", "",
                  "| kernel | grid (work-items) | workgroup | FETCH_SIZE KiB raw | WRITE_SIZE KiB raw | HBM MB (scaled) |",
                  "|---|---|---|---|---|---|"]
        for g in sorted(fk):
            mb = (fk[g] * f_scale + wk.get(g, 0.0) * w_scale) * 1024.0 / 1e6
            lines.append(f"| {g[2]} | {g[0]} | {g[1]} | {fk[g]:.0f} | {wk.get(g, 0.0):.0f} | {mb:.1f} |")
        lines += ["", f"=> {out['hbm_bytes_per_launch'] / 1e6:.1f} MB of HBM traffic per step."]
    os.makedirs(os.path.dirname(out_prefix), exist_ok=True)
    open(out_prefix + ".md", "w").write("\n".join(lines) + "\n")
    open(out_prefix + "_traffic.json", "w").write(json.dumps(out, indent=1) + "\n")
    print("\n".join(lines))


# ---------------------------------------------------------------------------
# the other configurations, reported beside the headline number
# ---------------------------------------------------------------------------

RATES = (1 / 3, 1 / 2, 2 / 3, 3 / 4, 5 / 6, 8 / 9)
# Eb/N0 (dB) of 10 % code-block failures by (base graph, rate) for BPSK/AWGN,
# 8 iterations, Z uniform over the 51 sizes, as measured on MI355X by
# `python bench.py --calibrate-ebno` (calibrate_ebno below); the uneven
# workload draws each block's Eb/N0 uniformly within +-1 dB of its point.
EBNO_10PCT = {(0, RATES[0]): 1.30, (0, RATES[1]): 1.77, (0, RATES[2]): 2.55, (0, RATES[3]): 2.98,
              (0, RATES[4]): 3.57, (0, RATES[5]): 4.15,
              (1, RATES[0]): 1.42, (1, RATES[1]): 1.96, (1, RATES[2]): 2.63, (1, RATES[3]): 3.02,
              (1, RATES[4]): 3.68, (1, RATES[5]): 4.70}
VALID_Z = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 44, 48, 52,
           56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384)


def _timed(fn, reps, torch):
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def bench_int8(torch, args):
    """5G NR BG2 Z=208 rate 1/5 (K=2080, 50Z sent), Eb/N0 ~ U[0,3] dB per
    frame, quantize_llr_i16 at scale 8 fused into the load, int8 min-sum,
    max 8 iterations, 16,384 frames per launch."""
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, QuantParams, traffic_blocks, _lib as L
    B, z = 16384, 208
    dec = QcLdpcDecoder.with_lifting_size(syndrome_amd.BG2, z, args.beta)
    n, k = dec.variable_node_count(), dec.info_bit_count_5g()
    rng = np.random.default_rng(args.seed + 3)
    e_llr, e_off, info, _ = traffic_blocks(np.ones(B), np.full(B, z), np.zeros(B), np.full(B, 50 * z),
                                           rng.uniform(0, 3, B), np.full(B, k / (50 * z)), args.seed + 3)
    llr = np.zeros((B, n), np.float32)
    llr[:, 2 * z:] = e_llr.reshape(B, 50 * z)  # E = Ncb at rv 0, no filler: identity placement
    d_in = torch.from_numpy(llr).cuda()
    d_post = torch.empty((B, n), dtype=torch.int16, device="cuda")
    d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    q = QuantParams(8.0, 32767)
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD | L.QUANTIZE
    fast = lambda: dec.decode_batch_i8(d_in, d_post, d_hard, d_it, args.iters, flags=flags, quant=q)
    wall, per = time_launches(fast, 10, 2, torch)
    t = wall / 10
    hard = d_hard.cpu().numpy()
    bler = float(((hard[:, :k] != info.reshape(B, k)).any(axis=1)).mean())
    its = d_it.cpu().numpy()
    d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
    plain = lambda: dec.decode_batch_i8(d_in, d_post, d_hard, d_it, args.iters, flags=flags | L.PLAIN,
                                        d_edge=d_edge, quant=q)
    tp = _timed(plain, 2, torch) if not args.no_plain else float("nan")
    out = {"workload": "BG2 Z=208 rate 1/5, Eb/N0 U[0,3] dB, int8 (scale 8, fused quantizer), max 8 it, "
                       "batch 16,384", "value": round(B * k / t / 1e9, 4), "unit": "Gbit/s",
           "ms_per_launch": round(t * 1e3, 4), "plain_value": round(B * k / tp / 1e9, 4),
           "mean_iterations": round(float(its.mean()), 3), "bler": bler,
           # compulsory bytes per codeword: f32 input, i16 posterior, u8 hard decisions
           "roofline": valu_roofline("syn::pack_i8<1, 208> (BG2, Z=208, two codewords per lane)", "i8",
                                     float(its.sum()) * 197 * z, float(np.mean(per)), float(B) * n * 7,
                                     "pack_i8_bg2_z208_b16384", packed=True)}
    if not args.no_cpu:
        from oracle import orc
        code = orc.code_3gpp(1, z)
        out["cpu_baseline"] = cpu_pipeline_baseline([code], np.zeros(B, np.int32), np.arange(B, dtype=np.int64) * n,
                                                    llr.ravel(), np.full(B, k), args.iters, args.beta, int8=True)
    return out


def calibrate_ebno(torch, args, blocks=2048, steps=8):
    """The Eb/N0 (dB) at which 10 % of code blocks fail, per base graph and
    rate of the uneven workload: Z uniform over the 51 lifting sizes, K' = K,
    E = round(K / R), rv 0, Qm 1, f32 decode with 8 iterations; bisection on
    [-2, 8] dB with `blocks` fresh blocks per probe.  Prints the table."""
    from syndrome_amd import DecodePlan, traffic_blocks, _lib as L
    table = {}
    for bg in (0, 1):
        for ri, r in enumerate(RATES):
            lo, hi = -2.0, 8.0
            for step in range(steps):
                mid = 0.5 * (lo + hi)
                rng = np.random.default_rng(1000 * bg + 100 * ri + step)
                z = np.array(VALID_Z)[rng.integers(0, 51, blocks)]
                k = (22 if bg == 0 else 10) * z
                e_len = np.maximum(np.round(k / r).astype(np.int64), 1)
                bgs = np.full(blocks, bg)
                e_llr, e_off, info, i_off = traffic_blocks(bgs, z, np.zeros(blocks), e_len, np.full(blocks, mid),
                                                           k / e_len, 77 + step + 10 * ri + 100 * bg)
                n = (68 if bg == 0 else 52) * z
                offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
                plan = DecodePlan(bgs, z, np.zeros(blocks), offs, args.beta)
                plan.set_dematch(e_off, e_len)
                d_llr = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
                plan.dematch(torch.from_numpy(e_llr).cuda(), d_llr)
                d_hard = torch.empty(int(n.sum()), dtype=torch.uint8, device="cuda")
                d_it = torch.empty(blocks, dtype=torch.int32, device="cuda")
                plan.run(d_llr, None, d_hard, d_it, args.iters, L.INIT_5G | L.WRITE_HARD)
                hard = d_hard.cpu().numpy()
                fails = sum(bool((hard[offs[i]:offs[i] + k[i]] != info[i_off[i]:i_off[i] + k[i]]).any())
                            for i in range(blocks))
                if fails / blocks > 0.1:
                    lo = mid
                else:
                    hi = mid
            table[f"{bg},{r:.4f}"] = round(0.5 * (lo + hi), 2)
            log(f"[calibrate] BG{bg + 1} R={r:.3f}: {table[f'{bg},{r:.4f}']} dB")
    print(json.dumps(table), flush=True)
    return table


def bench_uneven(torch, args):
    """32,768 code blocks: BG1/BG2 at p=1/2, Z uniform over the 51 lifting
    sizes, rate uniform over {1/3,...,8/9} via rate matching (rv 0, Qm 1),
    Eb/N0 at the (base graph, rate)'s measured 10% BLER point +-1 dB;
    de-matched on device and decoded in one persistent launch (f32 and
    int8)."""
    from syndrome_amd import DecodePlan, QcLdpcDecoder, QuantParams, traffic_blocks, _lib as L
    B = 32768
    rng = np.random.default_rng(args.seed + 4)
    bg = rng.integers(0, 2, B)
    z = np.array(VALID_Z)[rng.integers(0, 51, B)]
    r = np.array(RATES)[rng.integers(0, 6, B)]
    k = np.where(bg == 0, 22, 10) * z
    e_len = np.maximum(np.round(k / r).astype(np.int64), 1)
    ebno = np.array([EBNO_10PCT[(int(b), float(x))] for b, x in zip(bg, r)]) + rng.uniform(-1, 1, B)
    e_llr, e_off, info, i_off = traffic_blocks(bg, z, np.zeros(B), e_len, ebno, k / e_len, args.seed + 4)
    n = np.where(bg == 0, 68, 52) * z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, np.zeros(B), offs, args.beta)
    plan.set_dematch(e_off, e_len)
    d_e = torch.from_numpy(e_llr).cuda()
    d_llr = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    d_post = torch.empty_like(d_llr)
    d_hard = torch.empty(int(n.sum()), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    plan.dematch(d_e, d_llr)
    block = plan.tune(d_llr, d_post, d_hard, d_it, args.iters, flags) if not args.no_tune else plan.shape()["block"]
    full = lambda: (plan.dematch(d_e, d_llr), plan.run(d_llr, d_post, d_hard, d_it, args.iters, flags))
    t = _timed(full, 5, torch)
    _, per_dec = time_launches(lambda: plan.run(d_llr, d_post, d_hard, d_it, args.iters, flags), 5, 1, torch)
    t_dec = float(np.mean(per_dec)) / 1e3
    # the de-match kernel alone (HIP events on the launch stream); its
    # algorithmic bytes: E f32 read + N f32 written per block
    _, per_dm = time_launches(lambda: plan.dematch(d_e, d_llr), 20, 2, torch)
    t_dm = float(np.mean(per_dm)) / 1e3
    dm_bytes = 4.0 * float(e_len.sum() + n.sum())
    hard = d_hard.cpu().numpy()
    its = d_it.cpu().numpy()
    errs = sum(bool((hard[offs[i]:offs[i] + k[i]] != info[i_off[i]:i_off[i] + k[i]]).any()) for i in range(B))
    # per-config launches: the same blocks grouped by (BG, Z), one launch each
    groups = {}
    for i in range(B):
        groups.setdefault((int(bg[i]), int(z[i])), []).append(i)
    per = []
    for (b, zz), ids in groups.items():
        dec = QcLdpcDecoder.with_lifting_size(b, zz, args.beta)
        nn = (68 if b == 0 else 52) * zz
        idx = torch.from_numpy(np.concatenate([np.arange(offs[i], offs[i] + nn) for i in ids])).cuda()
        x = d_llr[idx].view(len(ids), nn).contiguous()
        per.append((dec, x, torch.empty_like(x), torch.empty(x.shape, dtype=torch.uint8, device="cuda"),
                    torch.empty(len(ids), dtype=torch.int32, device="cuda")))
    def per_config():
        for dec, x, p, h, it in per:
            dec.decode_batch_f32(x, p, h, it, args.iters, flags=flags)
    t_per = _timed(per_config, 3, torch) if not args.no_plain else float("nan")
    # int8 variant on the same blocks (quantizer fused)
    plan8 = DecodePlan(bg, z, np.zeros(B), offs, args.beta, int8=True)
    d_post8 = torch.empty(int(n.sum()), dtype=torch.int16, device="cuda")
    q = QuantParams(8.0, 32767)
    f8 = flags | L.QUANTIZE
    if not args.no_tune:
        plan8.tune(d_llr, d_post8, d_hard, d_it, args.iters, f8, quant=q)
    t8 = _timed(lambda: plan8.run(d_llr, d_post8, d_hard, d_it, args.iters, f8, quant=q), 5, torch)
    bits = float(k.sum())
    edges = np.where(bg == 0, 316, 197) * z
    roof = valu_roofline("syn::persist_decode<false, *> and syn::persistx_f32<0, ...> (every launch of the plan)", "f32",
                         float((edges * its).sum()), t_dec * 1e3, float(n.sum()) * 9, "persist_f32_uneven_b32768")
    cpu = cpu8 = None
    if not args.no_cpu:
        from oracle import orc
        keys = sorted({(int(b), int(zz)) for b, zz in zip(bg, z)})
        kidx = {kk: i for i, kk in enumerate(keys)}
        codes = [orc.code_3gpp(b, zz) for b, zz in keys]
        fcode = np.array([kidx[(int(b), int(zz))] for b, zz in zip(bg, z)], np.int32)
        host = d_llr.cpu().numpy()
        cpu = cpu_pipeline_baseline(codes, fcode, offs, host, k, args.iters, args.beta)
        cpu8 = cpu_pipeline_baseline(codes, fcode, offs, host, k, args.iters, args.beta, int8=True)
    return {"workload": "32,768 code blocks, BG1/BG2, Z over the 51 sizes, rates 1/3..8/9 via rate matching, "
                        "Eb/N0 at the measured 10% BLER point +-1 dB, max 8 it",
            "roofline": roof, "cpu_baseline": cpu, "cpu_baseline_int8": cpu8,
            "value": round(bits / t / 1e9, 4), "unit": "Gbit/s", "ms_dematch_plus_decode": round(t * 1e3, 3),
            "decode_only_value": round(bits / t_dec / 1e9, 4), "per_config_launches_value": round(bits / t_per / 1e9, 4),
            "per_config_launches": len(groups), "int8_value": round(bits / t8 / 1e9, 4), "tuned_block": block,
            "shape": plan.shape(), "mean_iterations": round(float(its.mean()), 3), "bler": round(errs / B, 4),
            "dematch": {"ms_per_launch": round(t_dm * 1e3, 4), "algorithmic_bytes": dm_bytes,
                        "roofline": {"bound": "hbm", "achieved": round(dm_bytes / t_dm / 1e9, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(dm_bytes / t_dm / 1e9 / HBM_PEAK_GBS, 4)}}}


def serving_traffic(torch, args, n_tb=96):
    """The serving workload's pool: n_tb transport blocks (TBS log-uniform
    in [1e3, 5e5] bits, MCS rate in {1/3,1/2,2/3,3/4,5/6}, Qm 1): random
    payload, CRC-24A attached, segmented into code blocks with their
    CRC-24B (compute_segmentation / segment), encoded, rate matched, sent
    over BPSK/AWGN at 2.0 dB and de-matched on the device into decoder
    inputs.  Returns a dict of the per-block and per-TB arrays and the
    inputs in HBM (d_llr) and in pinned host memory (host_pool)."""
    from syndrome_amd import DecodePlan, HostPool, traffic_blocks_info
    from syndrome_amd.transport import attach_crc24, compute_segmentation, segment
    rng = np.random.default_rng(args.seed + 5)
    cb_bg, cb_z, cb_nf, cb_e, cb_kp, tb_first, tb_num, tbs, infos = [], [], [], [], [], [], [], [], []
    for _ in range(n_tb):
        a = int(np.exp(rng.uniform(np.log(1000), np.log(500000))))
        r = float((1 / 3, 1 / 2, 2 / 3, 3 / 4, 5 / 6)[rng.integers(0, 5)])
        sg = compute_segmentation(a, r)
        blocks = segment(attach_crc24(rng.integers(0, 2, a).astype(np.uint8)), sg)
        tb_first.append(len(cb_bg))
        tb_num.append(sg.c)
        tbs.append(a)
        e = int(round((a + 24) / r / sg.c))
        for i in range(sg.c):
            cb_bg.append(sg.bg)
            cb_z.append(sg.z)
            cb_nf.append(sg.n_filler)
            cb_e.append(max(e, 1))
            cb_kp.append(sg.k_prime)
            infos.append(blocks[i])
    cb_bg, cb_z, cb_nf, cb_e, cb_kp = map(np.array, (cb_bg, cb_z, cb_nf, cb_e, cb_kp))
    nb = len(cb_bg)
    i_off = np.concatenate([[0], np.cumsum(cb_kp)[:-1]]).astype(np.int64)
    e_llr, e_off = traffic_blocks_info(cb_bg, cb_z, cb_nf, cb_e, np.full(nb, 2.0), cb_kp / cb_e,
                                       np.concatenate(infos), i_off, args.seed + 5)
    n = np.where(cb_bg == 0, 68, 52) * cb_z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    plan = DecodePlan(cb_bg, cb_z, cb_nf, offs, args.beta)
    plan.set_dematch(e_off, cb_e)
    d_llr = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    plan.dematch(torch.from_numpy(e_llr).cuda(), d_llr)
    host_pool = HostPool(int(n.sum()))
    host_pool.array[:] = d_llr.cpu().numpy()
    # info bits per code block: the TB payload spread over its blocks
    info_bits = np.zeros(nb, np.int64)
    for t in range(n_tb):
        info_bits[tb_first[t]:tb_first[t] + tb_num[t]] = tbs[t] // tb_num[t]
    tb_out_off = np.concatenate([[0], np.cumsum(tbs)[:-1]]).astype(np.int64)
    return {"n_tb": n_tb, "nb": nb, "offs": offs, "bg": cb_bg, "z": cb_z, "nf": cb_nf, "info_bits": info_bits,
            "tb_first": tb_first, "tb_num": tb_num, "tb_size": np.array(tbs), "tb_out_off": tb_out_off,
            "d_tb_out": torch.zeros(int(sum(tbs)), dtype=torch.uint8, device="cuda"),
            "mean_tb_bits": float(np.mean(tbs)), "d_llr": d_llr, "host_pool": host_pool}


def bench_serving(torch, args, peak_ref_gbps):
    """The serving workload (serving_traffic) offered as Poisson arrivals of
    transport blocks to the continuous-batching server; offered load swept
    against the saturated peak, inputs resident in HBM and in pinned host
    memory."""
    from syndrome_amd import Server, serve_loadgen
    tr = serving_traffic(torch, args)
    n_tb, nb, offs, cb_bg, cb_z, cb_nf = tr["n_tb"], tr["nb"], tr["offs"], tr["bg"], tr["z"], tr["nf"]
    info_bits, tb_first, tb_num, mean_tb_bits = tr["info_bits"], tr["tb_first"], tr["tb_num"], tr["mean_tb_bits"]
    d_llr, host_pool = tr["d_llr"], tr["host_pool"]
    out = {"workload": "Poisson TB arrivals, TBS log-U[1e3,5e5], rate in {1/3..5/6}, BPSK/AWGN 2.0 dB, "
                       f"{n_tb} TBs / {nb} code blocks cycled, HBM budget 1 GiB",
           "modes": {"resident": "one grid stays on the device and pulls code blocks from a queue",
                     "lanes": "batches of what is pending, launched on 3 streams",
                     "resident_transport_blocks": "whole transport blocks submitted; CRC-24B, CRC-24A and payload "
                                                  "decided on the device; latency per transport block"}}

    def pct(x, q):
        return round(float(np.percentile(x, q)) / 1e6, 4)

    if not args.no_cpu:
        # the same code blocks through the CPU pipeline: its saturated
        # throughput (a latency sweep at the GPU's offered loads would only
        # measure an unbounded queue: they are far above this peak)
        from oracle import orc
        keys = sorted({(int(b), int(zz)) for b, zz in zip(cb_bg, cb_z)})
        kidx = {kk: i for i, kk in enumerate(keys)}
        codes = [orc.code_3gpp(b, zz) for b, zz in keys]
        fcode = np.array([kidx[(int(b), int(zz))] for b, zz in zip(cb_bg, cb_z)], np.int32)
        out["cpu_baseline"] = cpu_pipeline_baseline(codes, fcode, offs, host_pool.array, info_bits, args.iters,
                                                    args.beta, n_filler=cb_nf, budget_s=6.0)

    for mode, streams, pools in (("resident", 0, (("inputs_in_hbm", d_llr), ("inputs_in_pinned_host", host_pool))),
                                 ("lanes", 3, (("inputs_in_hbm", d_llr),))):
        srv = Server(args.beta, args.iters, hbm_budget_bytes=1 << 30, max_batch=1024, streams=streams)
        for where, pool in pools:
            run = lambda rate, dur: serve_loadgen(srv, pool, offs, cb_bg, cb_z, cb_nf, info_bits, tb_first, tb_num,
                                                  rate, dur, seed=args.seed)
            run(0, 0.5)  # warm
            cbl, tbl, peak_bps, _ = run(0, args.serve_seconds)
            peak_tb_rate = peak_bps / mean_tb_bits
            res = {"peak_info_gbps": round(peak_bps / 1e9, 4), "peak_cb_per_s": round(len(cbl) / args.serve_seconds, 1),
                   "loads": []}
            for load in (0.1, 0.3, 0.5, 0.7, 0.9, 1.0):
                cbl, tbl, bps, waits = run(load * peak_tb_rate, args.serve_seconds)
                res["loads"].append({"offered_fraction_of_peak": load, "achieved_info_gbps": round(bps / 1e9, 4),
                                     "cb_latency_ms_p50": pct(cbl, 50), "cb_latency_ms_p99": pct(cbl, 99),
                                     "tb_latency_ms_p50": pct(tbl, 50), "tb_latency_ms_p99": pct(tbl, 99),
                                     "code_blocks": int(len(cbl)), "admission_waits": int(waits)})
            out[f"{mode}_{where}"] = res
        if mode == "resident":
            # transport blocks end to end: CRC-24B per block, CRC-24A and the
            # payload decided on the device, latency per transport block
            from syndrome_amd import serve_loadgen_tb
            runt = lambda rate, dur: serve_loadgen_tb(srv, d_llr, offs, cb_bg, cb_z, cb_nf, tb_first, tb_num,
                                                      tr["tb_size"], tr["d_tb_out"], tr["tb_out_off"], rate, dur,
                                                      seed=args.seed)
            runt(0, 0.5)
            tbl, ncb, peak_bps, fails, _, peak_good, _ = runt(0, args.serve_seconds)
            peak_tb_rate = peak_bps / mean_tb_bits
            res = {"peak_info_gbps": round(peak_bps / 1e9, 4), "peak_goodput_gbps": round(peak_good / 1e9, 4),
                   "peak_tb_per_s": round(len(tbl) / args.serve_seconds, 1),
                   "crc_failures_at_peak": int(fails), "loads": []}
            split = lambda x, m, q: pct(x[m], q) if m.any() else None
            for load in (0.1, 0.3, 0.5, 0.7, 0.9):
                tbl, ncb, bps, fails, waits, good, okf = runt(load * peak_tb_rate, args.serve_seconds)
                res["loads"].append({"offered_fraction_of_peak": load, "achieved_info_gbps": round(bps / 1e9, 4),
                                     "goodput_gbps": round(good / 1e9, 4),
                                     "tb_latency_ms_p50": pct(tbl, 50), "tb_latency_ms_p99": pct(tbl, 99),
                                     "tb_latency_ms_p99_crc_passed": split(tbl, okf, 99),
                                     "tb_latency_ms_p99_crc_failed": split(tbl, ~okf, 99),
                                     "transport_blocks": int(len(tbl)), "code_blocks": int(ncb),
                                     "crc_failures": int(fails)})
            out["resident_transport_blocks_inputs_in_hbm"] = res
        if mode == "resident":
            # the batched (work-queue) decoder's peak on this same traffic:
            # its code blocks, copied 16 times into one arena, decoded by one
            # plan (inputs in HBM), and the server offered 70 % of that
            R = 16
            tot = int(offs[-1] + (68 if cb_bg[-1] == 0 else 52) * cb_z[-1])
            from syndrome_amd import DecodePlan, _lib as L
            rb = np.tile(cb_bg, R)
            rplan = DecodePlan(rb, np.tile(cb_z, R), np.tile(cb_nf, R),
                               (np.tile(offs, R) + np.repeat(np.arange(R, dtype=np.int64) * tot, nb)), args.beta)
            r_in = d_llr[:tot].repeat(R)
            r_post = torch.empty_like(r_in)
            r_hard = torch.empty(r_in.numel(), dtype=torch.uint8, device="cuda")
            r_it = torch.empty(R * nb, dtype=torch.int32, device="cuda")
            fl = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
            _, rper = time_launches(lambda: rplan.run(r_in, r_post, r_hard, r_it, args.iters, fl), 5, 1, torch)
            bpk = float(info_bits.sum()) * R / (float(rper.mean()) * 1e-3) / 1e9
            del r_in, r_post, r_hard, rplan
            cbl, tbl, bps, waits = serve_loadgen(srv, d_llr, offs, cb_bg, cb_z, cb_nf, info_bits, tb_first, tb_num,
                                                 0.7 * bpk * 1e9 / mean_tb_bits, args.serve_seconds, seed=args.seed)
            out["batched_peak_same_traffic"] = {"info_gbps": round(bpk, 4), "ms_per_launch": round(float(rper.mean()), 4),
                                                "code_blocks_per_launch": int(R * nb)}
            out["resident_inputs_in_hbm_at_70pct_of_batched_peak_same_traffic"] = {
                "offered_info_gbps": round(0.7 * bpk, 4), "achieved_info_gbps": round(bps / 1e9, 4),
                "cb_latency_ms_p50": pct(cbl, 50), "cb_latency_ms_p99": pct(cbl, 99),
                "tb_latency_ms_p99": pct(tbl, 99)}
        if mode == "resident" and peak_ref_gbps is not None:
            # the same traffic offered at 70% of the batched BG1 Z=384
            # decoder's peak (measured by the headline run); above this
            # traffic's own batched peak, its mix of codes being slower to
            # decode than BG1 Z=384, this load cannot be carried
            cbl, tbl, bps, waits = serve_loadgen(srv, d_llr, offs, cb_bg, cb_z, cb_nf, info_bits, tb_first, tb_num,
                                                 0.7 * peak_ref_gbps * 1e9 / mean_tb_bits, args.serve_seconds, seed=args.seed)
            out["resident_inputs_in_hbm_at_70pct_of_batched_bg1_peak"] = {
                "offered_info_gbps": round(0.7 * peak_ref_gbps, 4), "achieved_info_gbps": round(bps / 1e9, 4),
                "cb_latency_ms_p50": pct(cbl, 50), "cb_latency_ms_p99": pct(cbl, 99),
                "tb_latency_ms_p99": pct(tbl, 99)}
        srv.close()
    return out


def bench_wifi(torch, args):
    """Wi-Fi Z=27 rate 1/2 (n=648, k=324), seeded random info bits, BPSK/AWGN
    2.0 dB, f32, the crate's default offset (0.5), max 10 iterations with
    early exit, 10,000 frames in one launch; beside the CPU path."""
    from syndrome_amd import QcLdpcDecoder, awgn_transmit, _lib as L
    F = 10000
    dec = QcLdpcDecoder.wifi(27, 1, 2, 0.5)
    n, k = 648, 324
    info = np.random.default_rng(args.seed + 1).integers(0, 2, (F, k), dtype=np.uint8)
    cws = dec.encode_batch(info)
    llr = np.stack([awgn_transmit(2.0, 0.5, f, cws[f]) for f in range(F)])
    d_in = torch.from_numpy(llr).cuda()
    d_post, d_hard = torch.empty_like(d_in), torch.empty((F, n), dtype=torch.uint8, device="cuda")
    d_it, d_edge = torch.empty(F, dtype=torch.int32, device="cuda"), torch.empty((F, dec.required_edge_buffer()),
                                                                                     device="cuda")
    run = lambda: dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 10, flags=L.WRITE_POST | L.WRITE_HARD,
                                       d_edge=d_edge)
    t = _timed(run, 10, torch)
    its = round(float(d_it.float().mean()), 3)
    plain = lambda: dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 10, flags=L.WRITE_POST | L.WRITE_HARD | L.PLAIN,
                                         d_edge=d_edge)
    tp = _timed(plain, 5, torch)
    out = {"workload": "Wi-Fi Z=27 R=1/2, 2.0 dB, beta 0.5, max 10 it, 10,000 frames", "value": round(F * k / t / 1e9, 4),
           "unit": "Gbit/s", "ms_per_launch": round(t * 1e3, 3), "mean_iterations": its,
           "plain_value": round(F * k / tp / 1e9, 4)}
    if not args.no_cpu:
        from oracle import orc
        code = orc.code_wifi(27, 1, 2)
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        el, its, _ = orc.pipeline_run(code, 0.5, llr, 10, workers, init5g=False)
        out["cpu_pipeline"] = {"value": round(F * k / el / 1e9, 5), "unit": "Gbit/s", "cores": workers,
                               "kind": "port"}
    return out


def bench_encode(torch, args):
    """Device encoder: 5G NR BG1 Z=384 (K=8448, N=26,112), 4,096 seeded
    random info blocks in HBM -> codewords in HBM, one launch.  Algorithmic
    HBM bytes per block: K read + N written (one byte per bit)."""
    from syndrome_amd import QcLdpcDecoder
    B = 4096
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.5)
    k, n = dec.info_bit_count_5g(), dec.variable_node_count()
    info = np.random.default_rng(args.seed + 2).integers(0, 2, (B, k), dtype=np.uint8)
    d_info = torch.from_numpy(info).cuda()
    d_cw = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    run = lambda: dec.encode_batch_device(d_info, d_cw)
    if getattr(args, "copy_calibrate", False):  # known bytes to scale HBM counters
        copy_calibrate(torch, int(d_cw.numel()))
    _, per = time_launches(run, 50, 5, torch)
    t = float(np.mean(per)) / 1e3
    sample = 256
    if not np.array_equal(d_cw[:sample].cpu().numpy(), dec.encode_batch(info[:sample])):
        raise RuntimeError("device codewords differ from the host encoder")
    gbs = B * (k + n) / t / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "encode_stream_bg1_z384_b4096_traffic.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
    out = {"workload": "BG1 Z=384 systematic encode, 4,096 blocks per launch, u8 bits in HBM",
           "value": round(B * k / t / 1e9, 3), "unit": "info Gbit/s", "ms_per_launch": round(t * 1e3, 4),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "kernel": "syn::encode_stream", "algorithmic_bytes_per_launch": B * (k + n),
                        "profile": "profiles/encode_stream_bg1_z384_b4096.md"}}
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    dec.encode_batch(info[:1024], threads=threads)
    th = time.perf_counter() - t0
    out["host_encoder"] = {"value": round(1024 * k / th / 1e9, 4), "unit": "info Gbit/s", "cores": threads,
                           "sample": "1,024 blocks"}
    if not args.no_cpu:
        from oracle import orc
        code = orc.code_3gpp(0, 384)
        t0 = time.perf_counter()
        for i in range(64):
            orc.encode(code, info[i])
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(64 * k / tc / 1e9, 5), "unit": "info Gbit/s", "cores": 1,
                               "kind": "port", "sample": "64 blocks through the oracle encoder"}
    return out


def extra_benches(torch, args, peak_gbps):
    """(name, thunk) of every configuration reported beside the headline."""
    return (("int8_bg2_z208", lambda: bench_int8(torch, args)),
            ("uneven_persistent", lambda: bench_uneven(torch, args)),
            ("serving", lambda: bench_serving(torch, args, peak_gbps)),
            ("wifi_z27", lambda: bench_wifi(torch, args)),
            ("encode_bg1_z384", lambda: bench_encode(torch, args)))


def summarize_sq(sq_dirs, out_md, kernel_substr, cmds, edge_slot_iters, clock_ghz=2.4, cus=256, kernel_ms=None,
                 lds_bytes=4, grid=None):
    """Append the SQ counters of the profiled kernel's full-batch dispatches
    (the largest grid, or `grid`; mean over them) to a profile summary, with
    what they imply, and write them to <summary>_sq.json for bench.py:
      * VALU issue share = VALU instructions x 4 cycles over cus x 4 SIMDs x
        the kernel time (the traffic JSON's average, or kernel_ms);
      * VALU lane-instructions per edge-slot-iteration = VALU instructions x
        64 lanes over the launch's executed edge-slot-iterations
        (edge_slot_iters: (edge, z) pairs x layered passes, summed over its
        codewords), comparable with the algorithmic count in bench.py;
      * LDS throughput, every LDS instruction moving lds_bytes per lane."""
    import csv
    from collections import defaultdict
    vals = defaultdict(list)
    per_grid = defaultdict(lambda: defaultdict(list))
    durations = defaultdict(dict)  # launch shape -> dispatch -> ms, from the counter runs' own timestamps
    for d in sq_dirs:
        rows = list(csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        ker = [r for r in rows if _kernel_match(kernel_substr, r["Kernel_Name"])]
        big = max(_launch_key(r) for r in ker)
        for r in ker:
            key = _launch_key(r)
            per_grid[r["Counter_Name"]][key].append(float(r["Counter_Value"]))
            if "Start_Timestamp" in r:
                durations[key][(d, r["Dispatch_Id"])] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
            if key == big:
                vals[r["Counter_Name"]].append(float(r["Counter_Value"]))
    if grid == "all":  # one step = one launch of every grid: sum the per-grid means
        mean = {k: float(sum(np.mean(v) for v in g.values())) for k, g in per_grid.items()}
    else:
        mean = {k: float(np.mean(v)) for k, v in vals.items()}
    if kernel_ms is None:
        tj = json.load(open(out_md[:-3] + "_traffic.json"))
        kernel_ms = tj["average_ms_per_step_all_grids"] if grid == "all" else tj["average_ms_full_batch"]
    elif kernel_ms < 0:  # the counter runs' own dispatch durations (mean per dispatch)
        ds = [v for dd in durations.values() for v in dd.values()]
        kernel_ms = float(np.mean(ds)) if ds else 0.0
    kern_s = kernel_ms / 1e3
    valu = mean.get("SQ_INSTS_VALU", 0.0)
    issue = valu * 4 / (cus * 4 * kern_s * clock_ghz * 1e9) if kern_s > 0 else None
    per_slot = valu * 64 / float(edge_slot_iters) if edge_slot_iters else None
    lines = ["", f"## SQ counters of the same kernel (full-batch dispatches, mean of "
                 f"{max((len(v) for v in vals.values()), default=0)})",
             "", "Commands (counters within the per-block limits, one pass each):", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "| counter | value |", "|---|---|"]
    for k in sorted(mean):
        lines.append(f"| {k} | {mean[k]:,.0f} |")
    if issue is not None:
        lines += ["", f"VALU issue share: {valu:.4e} VALU instructions x 4 cycles over {cus} CUs x 4 SIMDs and the "
                      f"{kern_s * 1e3:.4f} ms kernel ({clock_ghz} GHz) = {issue * 100:.1f} % of the device's VALU "
                      f"issue slots."]
    else:
        lines += [""]
    if per_slot:
        lines += [f"VALU lane-instructions per edge-slot-iteration: {valu:.4e} x 64 lanes over "
                  f"{float(edge_slot_iters):.4e} executed edge-slot-iterations = {per_slot:.2f}."]
    waits, cyc = mean.get("SQ_WAIT_ANY"), mean.get("SQ_WAVE_CYCLES")
    if waits and cyc:
        lines += [f"Waves waiting: SQ_WAIT_ANY / SQ_WAVE_CYCLES = {waits / cyc * 100:.1f} % of wave cycles."]
    lds = mean.get("SQ_INSTS_LDS", 0.0)
    if lds and kern_s > 0:
        lds_tbs = lds * 64 * lds_bytes / kern_s / 1e12
        lds_peak = cus * 128 * clock_ghz * 1e9 / 1e12
        lines += [f"LDS throughput: {lds:.3e} LDS instructions x 64 lanes x {lds_bytes} B over the kernel = "
                  f"{lds_tbs:.1f} TB/s, {lds_tbs / lds_peak * 100:.0f} % of the {lds_peak:.1f} TB/s the CUs' LDS "
                  f"can serve (128 B per CU per cycle)."]
    open(out_md, "a").write("\n".join(lines) + "\n")
    json.dump({"kernel": kernel_substr, "kernel_ms": kernel_ms, "counters": mean,
               "valu_issue_frac": round(issue, 4) if issue is not None else None,
               "valu_per_edge_slot_iteration": round(per_slot, 3) if per_slot else None,
               "edge_slot_iterations": float(edge_slot_iters)},
              open(out_md[:-3] + "_sq.json", "w"), indent=1)
    print("\n".join(lines))


def summarize_trace(trace_dir, out_md, title, cmd, extra_json=None):
    """A kernel-trace-only profile summary: the stats table and, per kernel,
    launches by grid size with their average duration; with extra_json (the
    `--only` output of the same run) the measured figures beside it."""
    import csv
    from collections import defaultdict
    stats = list(csv.DictReader(open(os.path.join(trace_dir, "run_kernel_stats.csv"))))
    by = defaultdict(list)
    for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_trace.csv"))):
        g = int(r["Grid_Size_X"]) if "Grid_Size_X" in r else int(r.get("Grid_Size", 0))
        by[(r["Kernel_Name"][:70], g)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    lines = [f"# rocprofv3 summary: {title}", "", "Command:", "", f"    {cmd}", "",
             "## kernel trace (stats CSV, all dispatches)", "",
             "| kernel | calls | average ns | min ns | max ns |", "|---|---|---|---|---|"]
    for r in stats[:8]:
        lines.append(f"| {r['Name'][:70]} | {r['Calls']} | {float(r['AverageNs']):.0f} | {r['MinNs']} | {r['MaxNs']} |")
    lines += ["", "## by kernel and grid size (trace CSV, the 8 longest in total)", "",
              "| kernel | grid (work-items) | launches | average ms |", "|---|---|---|---|"]
    for (k, g), v in sorted(by.items(), key=lambda kv: -sum(kv[1]))[:8]:
        lines.append(f"| {k} | {g} | {len(v)} | {np.mean(v):.4f} |")
    if extra_json and os.path.exists(extra_json):
        lines += ["", "## the same run's figures (`--only` JSON line)", "", "```",
                  open(extra_json).read().strip()[:1500], "```"]
    os.makedirs(os.path.dirname(out_md), exist_ok=True)
    open(out_md, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def copy_calibrate(torch, nbytes=256 << 20):
    """Device copies of a known size (three of nbytes), so that the HBM
    counters of a profiled run can be scaled by what they read for known
    bytes; the size goes to $SYNDROME_COPYCAL_OUT when set."""
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    cal = os.environ.get("SYNDROME_COPYCAL_OUT")
    if cal:
        open(cal, "w").write(json.dumps({"copy_calibration_bytes": int(nbytes)}))
    return int(nbytes)


def _number_arg(x):
    """A number, or @FILE:KEY.PATH naming one inside a JSON line this script
    printed (the last line of FILE), e.g.
    @out.json:int8_bg2_z208.roofline.edge_slot_iterations_per_launch."""
    if not x.startswith("@"):
        return float(x)
    path, key = x[1:].split(":", 1)
    obj = json.loads(open(path).read().strip().splitlines()[-1])
    for part in key.split("."):
        obj = obj[part]
    return float(obj)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-trace":
        # bench.py --summarize-trace TRACE_DIR OUT_MD TITLE CMD [ONLY_JSON]
        summarize_trace(*sys.argv[2:7])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-sq":
        # bench.py --summarize-sq [--kernel-ms MS] [--lds-bytes B] [--grid G] OUT_MD KERNEL EDGE_SLOT_ITERS
        #          SQ_DIR1 CMD1 [SQ_DIR2 CMD2 ...]
        a = sys.argv[2:]
        opt = {}
        while a and a[0] in ("--kernel-ms", "--lds-bytes", "--grid"):
            opt[a[0]] = a[1]
            a = a[2:]
        rest = a[3:]
        summarize_sq(rest[0::2], a[0], a[1], rest[1::2], _number_arg(a[2]),
                     kernel_ms=_number_arg(opt["--kernel-ms"]) if "--kernel-ms" in opt else None,
                     lds_bytes=int(opt.get("--lds-bytes", 4)), grid=opt.get("--grid"))
        return
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
    ap.add_argument("--no-plain", action="store_true", help="skip the plain and per-config comparison decoders")
    ap.add_argument("--no-tune", action="store_true", help="plans keep their default launch shape (profiling)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="duration of each serving load point")
    ap.add_argument("--quick", action="store_true", help="headline only (no int8 / uneven / serving / Wi-Fi extras)")
    ap.add_argument("--calibrate-ebno", action="store_true",
                    help="measure the 10%% BLER Eb/N0 of each (base graph, rate) of the uneven workload and exit")
    ap.add_argument("--latency-samples", type=int, default=200,
                    help="single-codeword launches timed for latency (0 = skip)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="also run device copies of the input batch (known bytes) to scale HBM counters")
    ap.add_argument("--only", default=None, help="run one extra by name and print its JSON (no headline line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU. The decode shards by codeword with no exchange between GPUs, "
              "so a multi-GPU run is one independent process per GPU; N>1 is not measured by this script.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, _lib as L

    syndrome_amd.lib()
    if args.calibrate_ebno:
        calibrate_ebno(torch, args)
        return
    if args.only:
        fn = dict(extra_benches(torch, args, None)).get(args.only)
        if fn is None:
            sys.exit(f"unknown extra {args.only!r}")
        if args.copy_calibrate and args.only != "encode_bg1_z384":
            copy_calibrate(torch)
        print(json.dumps({args.only: fn()}), flush=True)
        return
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

    # roofline of the dominant kernel, fastx_f32<BG1, 384>: VALU-bound (the
    # check-to-variable messages stay in registers, so HBM sees the
    # compulsory N*4 input + N*4 posterior + N*1 hard-decision bytes, and
    # the 42 extension columns' posteriors, which this kernel keeps in the
    # posterior buffer: one 4-byte read and one write per extension
    # variable and executed iteration, mostly served by L2)
    kern_ms = float(per.mean())
    roof = valu_roofline("syn::fastx_f32<0, 384> (BG1, Z=384)", "f32", float(its.sum()) * ez, kern_ms,
                         float(B) * n * 9, "fastx_f32_bg1_z384_b4096")
    roof["hbm"]["extension_bytes_per_launch"] = float(its.sum()) * 42 * 384 * 8

    out = {}
    if args.copy_calibrate:
        out["copy_calibration_bytes"] = copy_calibrate(torch, int(d_in.numel() * 4))
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

    # the same batch through the work-queue (persistent) decoder: a plan of
    # 4,096 BG1 Z=384 blocks, one persistent launch
    from syndrome_amd import DecodePlan
    plan = DecodePlan(np.zeros(B, np.int32), np.full(B, 384), np.zeros(B), np.arange(B, dtype=np.int64) * n,
                      args.beta)
    d_post3, d_hard3, d_it3 = torch.empty_like(d_in), torch.empty_like(d_hard), torch.empty_like(d_it)
    if not args.no_tune:
        plan.tune(d_in, d_post3, d_hard3, d_it3, args.iters, flags)
    qwall, qper = time_launches(lambda: plan.run(d_in, d_post3, d_hard3, d_it3, args.iters, flags), args.steps,
                                args.warmup, torch)
    qms = qwall * 1e3 / args.steps
    out["persistent"] = {"value": round(B * k / (qms * 1e-3) / 1e9, 4), "unit": "Gbit/s", "ms_per_step": round(qms, 4),
                         "kernel": "syn::persistx_f32<0, 384> (the two-codeword kernel's work-queue form)",
                         "shape": plan.shape(),
                         "vs_per_config_kernel": round(B * k / (qms * 1e-3) / 1e9 / value, 4),
                         "identical_decoded_bits": bool(torch.equal(d_hard3, d_hard) and torch.equal(d_it3, d_it) and
                                                        torch.equal(d_post3.view(torch.int32), d_post.view(torch.int32)))}
    if "plain" in out:
        out["persistent_vs_plain"] = round(out["persistent"]["value"] / out["plain"]["value"], 3)
    del plan, d_post3, d_hard3

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
    if not args.quick:
        extras = {}
        for name, fn in extra_benches(torch, args, value):
            log(f"[bench] extra: {name}")
            t0 = time.perf_counter()
            try:
                extras[name] = fn()
            except Exception as e:  # an extra never hides the headline line
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] extra {name} done in {time.perf_counter() - t0:.1f} s")
        line["extras"] = extras
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
