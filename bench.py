"""Benchmark: the list-vs-list subset check on one MI355X.

Workload (the subset-check microbench): a forward list of 2^20 state sets and
a backward list of 2^20 state sets over n = 256 states (4 x uint64 each);
forward states present with probability 0.08, backward with 0.75; one forward
set in 10^5 is replaced by its intersection with a random backward set (a
planted subset).  One step = one pass of the subset kernel over all 2^40
pairs, producing the existence flag and the exact count of contained pairs.
Inputs are resident in HBM before the timed region.

Also reported (not the headline value): time to the exact shortest reset
length of a random binary DFA with n = 300 through the configs/exact.json
plan, GPU against the reference binary on the host CPU when it is present
(oracle/_ref/), and against the oracle otherwise.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "time-to-exact reset length (s) + subset checks/s; learned vs plain bound, 1 GPU"
HBM_PEAK_GBS = 8000.0


LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12  # 256 CUs x 128 B/clk x 2.4 GHz
IN_FLIGHT = 8  # plans in flight in the time-to-exact workload's second measurement (one context each)


def lds_roofline(ctx, small, tile_count: int, kernel_ms: float, indexed: bool) -> dict:
    """LDS figure of the subset kernel.  Algorithmic LDS bytes: every state of
    a stream set costs one bit per tile set, sum(|S|) * |T| / 8 bytes.  The
    peak is the measured read rate of the access the kernel uses: 16-byte
    lanes (indexed kernel) or 8-byte lanes (bit-walking kernel)."""
    bits = int(np.unpackbits(small.view(np.uint8)).sum())
    lds_bytes = bits * tile_count / 8
    achieved = lds_bytes / (kernel_ms / 1000.0) / 1e12
    measured = {b: ctx.calibrate_lds(b) for b in (8, 16)}
    peak = measured[16] if indexed else measured[8]
    return {"achieved": achieved, "peak": peak, "unit": "TB/s", "frac": achieved / peak,
            "lds_bytes_per_launch": lds_bytes,
            "peak_measured_b64": measured[8], "peak_measured_b128": measured[16], "peak_planning": LDS_PEAK_TBS,
            "note": "peak = measured LDS read rate (sw_calibrate_lds) with the kernel's own lane width: "
                    "16 B (indexed kernel) or 8 B (bit-walking kernel); planning figure 256 CUs x 128 B/clk x 2.4 GHz"}


def measured_traffic(path: str):
    """HBM bytes per launch of the subset kernel from a committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size),
    or None when there is none."""
    try:
        for line in Path(path).read_text().splitlines():
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def make_lists(count: int, n: int, p_small: float, p_large: float, seed: int):
    """Bernoulli bitsets; bit q of set i is bit q%64 of word q/64."""
    rng = np.random.default_rng(seed)
    W = (n + 63) // 64
    def gen(p):
        out = np.empty((count, W), dtype=np.uint64)
        thr = np.uint16(int(round(p * 65536)))
        weights = (np.uint64(1) << np.arange(64, dtype=np.uint64))
        step = 1 << 15
        for lo in range(0, count, step):
            hi = min(count, lo + step)
            bits = rng.integers(0, 65536, size=(hi - lo, W * 64), dtype=np.uint16) < thr
            bits[:, n:] = False
            b = bits.reshape(hi - lo, W, 64).astype(np.uint64)
            out[lo:hi] = (b * weights).sum(axis=2, dtype=np.uint64)
        return out
    small = gen(p_small)
    large = gen(p_large)
    planted = []
    for i in range(0, count, 100000):
        j = int(rng.integers(0, count))
        small[i] &= large[j]
        planted.append((i, j))
    return small, large, planted


GOLDEN_COUNT = ROOT / "tests" / "golden" / "c2_count.json"


def host_cpu() -> dict:
    """CPU model and the hardware threads this process may use."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "threads_available": len(os.sched_getaffinity(0))}


def cpu_subset_baseline(ctx, small, large, rows: int, golden: dict | None) -> dict:
    """The oracle's all-pairs count (a port of the checker's semantics,
    subsets_checker_kernel.cu:8-33) over every host thread, on the first
    `rows` forward sets x all backward sets; the GPU counts the same rows and
    both are compared, and with the committed golden count when it covers
    the sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    cpu = host_cpu()
    threads = cpu["threads_available"]
    t0 = time.perf_counter()
    total = orc.subset_count_mt(small[:rows], large, threads=threads)
    dt = time.perf_counter() - t0
    pairs = rows * large.shape[0]
    gpu_count = ctx.subset_check(small.shape[1] * 64, small[:rows], large, count=True)[1]
    out = {"value": pairs / dt, "unit": "subset checks/s", "cores": threads, "kind": "port",
           "cpu_model": cpu["model"],
           "sample": f"first {rows} forward sets x all {large.shape[0]} backward sets ({pairs} pairs, "
                     f"{dt:.1f} s, oracle count {total}, GPU count of the same rows {gpu_count})",
           "sample_counts_equal": total == gpu_count}
    if golden and large.shape[0] == golden["count"] and rows % golden["block_rows"] == 0:
        want = sum(golden["block_counts"][:rows // golden["block_rows"]])
        out["sample_count_equals_golden"] = total == want
    assert total == gpu_count, (total, gpu_count)
    return out


def time_to_exact(ctx, n: int, seed: int) -> dict:
    from synchrowords_amd import gen
    d = gen.random_dfa(n, 2, seed)
    ctx.pipeline(n, 2, d)  # warm: tables, allocations, kernels loaded
    t0 = time.perf_counter()
    r = ctx.pipeline(n, 2, d)
    gpu_s = time.perf_counter() - t0
    out = {"workload": f"random binary DFA n={n} seed={seed}, configs/exact.json plan",
           "length": r.upper, "gpu_s": round(gpu_s, 4),
           "gpu_stage_ms": {k: round(v, 2) for k, v in r.times_ms},
           "exact_search": {k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.exact.items()}}
    ref = ROOT / "oracle" / "_ref" / f"ref_{n}_2_1"
    if ref.exists():
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(gen.to_text(n, 2, d))
            path = f.name
        try:
            t0 = time.perf_counter()
            res = subprocess.run([str(ref), path], capture_output=True, text=True, timeout=300)
            cpu_s = time.perf_counter() - t0
        finally:
            os.unlink(path)
        line = res.stdout.strip()
        out.update({"cpu_reference_s": round(cpu_s, 4), "cpu_reference_line": line, "cpu_threads": 1,
                    "speedup": round(cpu_s / gpu_s, 2)})
    ref_mt = ROOT / "oracle" / "_ref" / f"ref_{n}_2_16"
    if ref_mt.exists():
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(gen.to_text(n, 2, d))
            path = f.name
        try:
            t0 = time.perf_counter()
            subprocess.run([str(ref_mt), path], capture_output=True, text=True, timeout=300)
            cpu16 = time.perf_counter() - t0
        finally:
            os.unlink(path)
        out.update({"cpu_reference_16_threads_s": round(cpu16, 4), "speedup_vs_16_threads": round(cpu16 / gpu_s, 2)})
    return out


def run_reference(n: int, k: int, threads: int, automata) -> dict | None:
    """The reference binary (oracle/_ref/ref_N_K_T, the reference's own
    sources built for N, K and T threads) on a file of automata: per
    automaton its [lower, upper] and the plan's time (JIT compile excluded,
    as in the reference's output), or None when it was not built."""
    from synchrowords_amd import gen
    exe = ROOT / "oracle" / "_ref" / f"ref_{n}_{k}_{threads}"
    if not exe.exists():
        return None
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        for d in automata:
            f.write(gen.to_text(n, k, d))
        path = f.name
    try:
        t0 = time.perf_counter()
        res = subprocess.run([str(exe), path], capture_output=True, text=True, timeout=1800)
        wall = time.perf_counter() - t0
    finally:
        os.unlink(path)
    rows = []
    for line in res.stdout.splitlines():
        m = {kv.split("=")[0]: kv.split("=")[1] for kv in line.split("|")[-1].split() if "=" in kv}
        if "total_us" in m:
            rows.append((int(m["lo"]), int(m["hi"]), float(m["total_us"]) / 1e6))
    return {"rows": rows, "s_total": round(sum(r[2] for r in rows), 4), "wall_s": round(wall, 3)}


def bench_times(ctx) -> dict:
    """Time to the exact length through the configs/exact.json plan, per
    workload: random binary DFAs with n = 48 (seeds 0..19), n = 200 (seeds
    0..9), n = 300 (seeds 1000..1004), and the Cerny automaton C_64 (length
    63^2 = 3969).  GPU: one plan call per automaton after a warm call, wall
    clock.  Reference: its own binary on the host CPU at one thread and at
    16 threads (min(64, host threads) on the box), its plan time per
    automaton.  Every length is compared."""
    import concurrent.futures as cf
    import synchrowords_amd as sw
    from synchrowords_amd import gen
    cpu = host_cpu()
    work = [("random n=48 k=2, seeds 0..19", 48, [gen.random_dfa(48, 2, s) for s in range(20)]),
            ("random n=200 k=2, seeds 0..9", 200, [gen.random_dfa(200, 2, s) for s in range(10)]),
            ("Cerny C_64", 64, [gen.cerny(64)]),
            ("random n=300 k=2, seeds 1000..1004", 300, [gen.random_dfa(300, 2, s) for s in range(1000, 1005)])]
    out = {"workload": "time to exact reset length, configs/exact.json plan", "cpu_model": cpu["model"],
           "host_threads": cpu["threads_available"], "rows": []}
    for name, n, autos in work:
        ctx.pipeline(n, 2, autos[0])  # warm
        gpu, lengths, stages = [], [], {}
        for d in autos:
            t0 = time.perf_counter()
            r = ctx.pipeline(n, 2, d)
            gpu.append(time.perf_counter() - t0)
            lengths.append((r.lower, r.upper))
            for st, ms in r.times_ms:
                stages[st] = stages.get(st, 0.0) + ms
            assert r.word is not None and len(r.word) == r.upper and replay_ok(n, d, r.word)
        row = {"workload": name, "automata": len(autos), "gpu_s_total": round(sum(gpu), 4),
               "gpu_s_max": round(max(gpu), 4), "gpu_stage_ms": {k: round(v, 2) for k, v in stages.items()}}
        # The same automata with several plans in flight: one context (own
        # stream) per host thread, each taking every j-th automaton in turn;
        # the reference's app runs a file one automaton at a time.
        if len(autos) > 1:
            j = min(IN_FLIGHT, len(autos))
            ctxs = [sw.Context(0) for _ in range(j)]
            for c in ctxs:
                c.pipeline(n, 2, autos[0])  # warm
            got = [None] * len(autos)

            def work(w):
                for i in range(w, len(autos), j):
                    r = ctxs[w].pipeline(n, 2, autos[i])
                    got[i] = (r.lower, r.upper)

            t0 = time.perf_counter()
            with cf.ThreadPoolExecutor(j) as ex:
                list(ex.map(work, range(j)))
            wall = time.perf_counter() - t0
            for c in ctxs:
                c.close()
            assert got == lengths
            row[f"gpu_s_total_{j}_in_flight"] = round(wall, 4)
        for t in (1, 16):
            ref = run_reference(n, 2, t, autos)
            if ref is None:
                continue
            same = [tuple(x[:2]) for x in ref["rows"]] == [tuple(x) for x in lengths]
            row[f"reference_{t}t_s_total"] = ref["s_total"]
            row[f"speedup_vs_{t}t"] = round(ref["s_total"] / sum(gpu), 2)
            if f"gpu_s_total_{IN_FLIGHT}_in_flight" in row:
                row[f"speedup_in_flight_vs_{t}t"] = round(ref["s_total"] / row[f"gpu_s_total_{IN_FLIGHT}_in_flight"], 2)
            row[f"lengths_equal_{t}t"] = same
        print(f"[bench] times {row}", file=sys.stderr, flush=True)
        out["rows"].append(row)
    return out


def bench_reduce(ctx) -> dict:
    """The configs/exact_reduce.json plan (Reduce on the device, then Exact on
    the reduced automaton from the Reduce's list) against configs/exact.json
    on the same automata: random binary DFAs n = 300 (seeds 1000..1004) and
    n = 200 (seeds 0..9).  Best of two warm calls each; the lengths must
    agree with each other and the reduced N with the reference's
    (tests/golden/reduce_plan.json)."""
    from synchrowords_amd import gen
    sys.path.insert(0, str(ROOT / "tests"))
    from cli_helpers import EXACT_PLAN, EXACT_REDUCE_PLAN
    plans = {"exact": json.dumps(EXACT_PLAN), "exact_reduce": json.dumps(EXACT_REDUCE_PLAN)}
    golden = {c["name"]: c for c in json.loads((ROOT / "tests" / "golden" / "reduce_plan.json").read_text())["cases"]}
    work = [(f"random300_s{s}", 300, gen.random_dfa(300, 2, s)) for s in range(1000, 1005)] + \
           [(f"random200_s{s}", 200, gen.random_dfa(200, 2, s)) for s in range(10)]
    ctx.pipeline(300, 2, work[0][2], plan=plans["exact_reduce"])  # warm
    rows, tot = [], {"exact": 0.0, "exact_reduce": 0.0}
    for name, n, d in work:
        row = {"automaton": name}
        for key, plan in plans.items():
            best = None
            for _ in range(2):
                t0 = time.perf_counter()
                r = ctx.pipeline(n, 2, d, plan=plan)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            tot[key] += best
            row[f"{key}_ms"] = round(best * 1e3, 2)
            row[f"{key}_length"] = r.upper
            if key == "exact_reduce":
                row["reduced_n"], row["reduce_steps"] = r.reduced_n, r.reduce_steps
                row["stage_ms"] = {a: round(b, 2) for a, b in r.times_ms}
                g = golden.get(name)
                if g is not None:
                    row["reduced_n_equals_reference"] = r.reduced_n == g["reduced_n"]
            assert r.word is not None and len(r.word) == r.upper and replay_ok(n, d, r.word)
        assert row["exact_length"] == row["exact_reduce_length"]
        rows.append(row)
        print(f"[bench] reduce {row}", file=sys.stderr, flush=True)
    return {"workload": "configs/exact_reduce.json against configs/exact.json, random binary DFAs n=300 "
                        "(seeds 1000..1004) and n=200 (seeds 0..9)",
            "exact_s_total": round(tot["exact"], 4), "exact_reduce_s_total": round(tot["exact_reduce"], 4),
            "rows": rows}


def replay_ok(n: int, d, word) -> bool:
    import synchrowords_amd as sw
    return sw.replay(n, 2, d, word) == 1


def bench_frontier(ctx, steps: int, warmup: int, log2_frontier: int = 23) -> dict:
    """One search step at scale (sw_frontier_step): a random binary DFA with
    n = 300 (W = 5, the size of the n=300 exact searches), a frontier of
    2^23 random sets (one in eight repeated), a visited list of the images
    of the frontier's first half, so that about half of the candidates are
    visited and an eighth are duplicates.  Per kernel: HIP-event time on the
    library's stream, algorithmic bytes and the fraction of 8 TB/s; for the
    hash insert, probes per second."""
    import synchrowords_amd as sw
    n, k = 300, 2
    W = (n + 63) // 64
    m = 1 << log2_frontier
    d = np.array(sw.random_dfa(n, k, 7), dtype=np.uint32)
    rng = np.random.default_rng(20251204)
    f = rng.integers(0, 2**64 - 1, size=(m, W), dtype=np.uint64, endpoint=True)
    f &= rng.integers(0, 2**64 - 1, size=(m, W), dtype=np.uint64, endpoint=True)  # p = 1/4 per state
    f[:, W - 1] &= np.uint64((1 << (n - 64 * (W - 1))) - 1)
    rep = m // 8
    f[m - rep:] = f[rng.integers(0, m - rep, size=rep)]
    df = sw.DeviceArray.from_host(ctx, f)
    del f
    half = m // 2
    dv = sw.DeviceArray(ctx, half * k * W * 8)
    nv, _ = ctx.frontier_step_device(n, k, d, df, half, None, 0, dv)
    dout = sw.DeviceArray(ctx, m * k * W * 8)
    dpar = sw.DeviceArray(ctx, m * k * 4)
    dlet = sw.DeviceArray(ctx, m * k)
    for _ in range(warmup):
        ctx.frontier_step_device(n, k, d, df, m, dv, nv, dout, dpar, dlet)
    runs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        nnew, st = ctx.frontier_step_device(n, k, d, df, m, dv, nv, dout, dpar, dlet)
        runs.append(st)
    wall = (time.perf_counter() - t0) / max(1, steps)
    for a in (df, dv, dout, dpar, dlet):
        a.free()
    med = lambda key: float(np.median([r[key] for r in runs]))
    cand = k * m
    probes = runs[-1]["probes"]
    rb = 8 * W
    kernels = {
        # parents read once per letter's pass are L2 hits; children written
        "k_expand": (med("ms_expand"), m * rb + cand * rb, "parent 8W read + child 8W written"),
        # candidate read, slot index written, one 4-byte slot read per probe
        "k_hash_insert": (med("ms_insert"), cand * (rb + 4) + probes * 4,
                          "candidate 8W read + slot_of 4 B written + 4 B per probe"),
        # slot_of read, owner slot read, keep and position written
        "keep_scan": (med("ms_scan"), cand * 16, "slot_of 4 B + slot 4 B read, keep and pos 4 B each written"),
        # keep flags read; per new set: position read, set read and written, parent and letter written
        "k_commit": (med("ms_commit"), cand * 4 + nnew * (4 + 2 * rb + 4 + 1),
                     "keep 4 B per candidate; per new set pos 4 B, set 8W read + 8W written, parent 4 B, letter 1 B"),
    }
    out = {"workload": f"one search step, random binary DFA n={n} (W={W}), frontier 2^{log2_frontier} sets "
                       f"(1 in 8 repeated), visited list {nv} sets (images of the first half)",
           "candidates": cand, "new_sets": nnew, "visited": nv, "probes": probes,
           "mean_probes_per_candidate": round(probes / cand, 4), "table_slots": runs[-1]["table_slots"],
           "steps": steps, "wall_ms_per_step": round(wall * 1e3, 3), "kernels": {}}
    for name, (ms, nbytes, unit) in kernels.items():
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out["kernels"][name] = {"ms": round(ms, 4), "algorithmic_bytes": int(nbytes), "unit_bytes": unit,
                                "GB_s": round(gbs, 1), "frac_of_8TBs": round(gbs / 8000.0, 4)}
    ins = kernels["k_hash_insert"][0]
    out["probes_per_s"] = probes / (ins * 1e-3) if ins > 0 else 0.0
    out["candidates_per_s_step"] = cand / (sum(v[0] for v in kernels.values()) * 1e-3)
    return out


def bench_batch(ctx, count: int, cap_rows: int = 0, max_pairs: int = 0) -> dict:
    """The 4096-instance batch (n in [16,160], k in {2,3}) in one persistent launch."""
    from synchrowords_amd import gen
    import synchrowords_amd as sw
    specs = gen.batch_specs(4096, 2025)[:count]
    inst = [(n, k, sw.random_dfa(n, k, seed)) for n, k, seed in specs]
    ctx.batch(inst[:64], cap_rows=cap_rows, max_pairs=max_pairs)  # warm
    t0 = time.perf_counter()
    res, st = ctx.batch(inst, cap_rows=cap_rows, max_pairs=max_pairs)
    wall = time.perf_counter() - t0
    out = {"workload": f"batch of {count} random DFAs, n in [16,160], k in {{2,3}} (batch_specs(4096, 2025))",
           "wall_s": round(wall, 4), "instances_per_s": count / wall, "stats": st}
    fx = ROOT / "tests" / "golden" / "c4_batch.json"
    if fx.exists():
        cases = json.loads(fx.read_text())["cases"][:count]
        bad = [c["i"] for c, r in zip(cases, res)
               if r["non_synchro"] != bool(c["non_synchro"]) or (not c["non_synchro"] and r["length"] != c["upper"])]
        out["mismatches_vs_reference"] = len(bad)
        out["reference_cpu_s_build_container"] = round(sum(c["ref_us"] for c in cases) / 1e6, 2)
    bad_words = sum(1 for (n, k, d), r in zip(inst, res)
                    if not r["non_synchro"] and n > 1 and (r["word"] is None or sw.replay(n, k, d, r["word"]) != 1))
    out["words_not_synchronizing"] = bad_words
    return out


def bench_overflow(ctx, count: int, in_flight: int) -> dict:
    """The batch's overflow path on its own: every instance of the batch as a
    whole-device search under its Eppstein bound, one at a time; the ones
    with more than 2^16 sets on a side (those that leave the persistent
    launch) again, one at a time and `in_flight` at a time (one context and
    host thread each; 0 skips that leg)."""
    import threading
    import synchrowords_amd as sw
    from synchrowords_amd import gen
    specs = gen.batch_specs(4096, 2025)[:count]
    inst = [(n, k, sw.random_dfa(n, k, seed)) for n, k, seed in specs]
    nsy, bounds, _ = ctx.eppstein_batch(inst)
    todo = [i for i, (n, _, _) in enumerate(inst) if not nsy[i] and n > 1]
    ctx.exact(*inst[todo[0]], upper=int(bounds[todo[0]]))  # warm
    t0 = time.perf_counter()
    stats = {}
    for i in todo:
        stats[i] = ctx.exact(*inst[i], upper=int(bounds[i])).stats
    t_all = time.perf_counter() - t0
    big = [i for i in todo if max(stats[i]["sets_forward"], stats[i]["sets_backward"]) > (1 << 16)]
    big.sort(key=lambda i: (-inst[i][1], -inst[i][0]))
    t0 = time.perf_counter()
    for i in big:
        ctx.exact(*inst[i], upper=int(bounds[i]))
    t_big = time.perf_counter() - t0
    print(f"[bench] overflow: {len(todo)} searches in {t_all:.2f} s, {len(big)} large in {t_big:.2f} s",
          file=sys.stderr, flush=True)
    if in_flight <= 0:
        ctxs = []
    else:
        ctxs = [sw.Context(0) for _ in range(in_flight)]
    for c in ctxs:
        c.exact(*inst[big[0]], upper=int(bounds[big[0]]))  # warm each context
    nxt, lock = [0], threading.Lock()

    def work(c):
        while True:
            with lock:
                j = nxt[0]
                nxt[0] += 1
            if j >= len(big):
                return
            c.exact(*inst[big[j]], upper=int(bounds[big[j]]))
    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(c,)) for c in ctxs]
    for x in th:
        x.start()
    for x in th:
        x.join()
    t_in_flight = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    agg = lambda key: round(sum(stats[i][key] for i in big) / 1e3, 3)
    return {"workload": f"whole-device searches of the batch of {count} (batch_specs(4096, 2025)), Eppstein bounds",
            "searches": len(todo), "serial_s_all": round(t_all, 3), "large": len(big),
            "large_serial_s": round(t_big, 3), "in_flight": in_flight,
            "large_in_flight_s": round(t_in_flight, 3) if ctxs else None,
            "large_expand_s": agg("ms_expand"), "large_goal_s": agg("ms_goal"),
            "large_median_sets_per_side": int(np.median([max(stats[i]["sets_forward"], stats[i]["sets_backward"])
                                                          for i in big]))}


def bench_brute(ctx, count: int, no_cpu: bool) -> dict:
    """Power-set BFS: a batch of small automata in one launch (one workgroup
    each, visited set in LDS) and single automata of 24 states (per-level
    launches), beside the oracle's BFS on one host core (a sample)."""
    import synchrowords_amd as sw
    rng = np.random.default_rng(2026)
    inst = []
    for j in range(count):
        n = int(rng.integers(8, 21))
        k = int(rng.integers(2, 4))
        inst.append((n, k, sw.random_dfa(n, k, 60000 + j)))
    ctx.brute_batch(inst[:64])  # warm
    t0 = time.perf_counter()
    got = ctx.brute_batch(inst)
    wall = time.perf_counter() - t0
    out = {"workload": f"batch of {count} random DFAs, n in [8,20], k in {{2,3}} (seeded), power-set BFS",
           "wall_s": round(wall, 4), "instances_per_s": count / wall}
    # Slowly synchronizing ones visit far more subsets: the hard family of
    # 14..20 states (lengths 49..82), states relabelled at random.
    from synchrowords_amd import gen

    def relabel(n, k, d, perm):
        inv = np.argsort(perm)
        return [int(perm[d[int(inv[q]) * k + a]]) for q in range(n) for a in range(k)]

    hard = []
    for j in range(min(count, 1024)):
        n = int(rng.integers(14, 21))
        hard.append((n, 2, relabel(n, 2, gen.hard(n), rng.permutation(n))))
    ctx.brute_batch(hard[:16])
    t0 = time.perf_counter()
    got_hard = ctx.brute_batch(hard)
    wall_h = time.perf_counter() - t0
    out["hard_batch"] = {"workload": f"{len(hard)} relabelled hard DFAs (gen.hard), n in [14,20], k=2",
                         "wall_s": round(wall_h, 4), "instances_per_s": len(hard) / wall_h}
    big = [(24, 2, sw.random_dfa(24, 2, 70000 + s)) for s in range(3)]
    ctx.brute(*big[0])  # warm
    t0 = time.perf_counter()
    big_len = [ctx.brute(*b) for b in big]
    out["n24_k2_s_each"] = round((time.perf_counter() - t0) / len(big), 4)
    out["n24_k2_lengths"] = big_len
    if not no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        t0 = time.perf_counter()
        done = 0
        mism = 0
        while done < count and time.perf_counter() - t0 < 15.0:
            n, k, d = inst[done]
            mism += int(orc.brute(n, k, d) != got[done])
            done += 1
        cpu = time.perf_counter() - t0
        t1 = time.perf_counter()
        big_cpu = [orc.brute(*b) for b in big]
        big_cpu_s = (time.perf_counter() - t1) / len(big)
        out["cpu_baseline"] = {"value": done / cpu, "unit": "instances/s", "cores": 1, "kind": "port",
                               "sample": f"first {done} automata of the batch, oracle power-set BFS"}
        out["mismatches_vs_oracle_sample"] = mism
        out["speedup_batch"] = round(out["instances_per_s"] / (done / cpu), 1)
        out["n24_k2_cpu_s_each"] = round(big_cpu_s, 4)
        t2 = time.perf_counter()
        done_h = 0
        mism_h = 0
        while done_h < len(hard) and time.perf_counter() - t2 < 15.0:
            n, k, d = hard[done_h]
            mism_h += int(orc.brute(n, k, d) != got_hard[done_h])
            done_h += 1
        cpu_h = time.perf_counter() - t2
        out["hard_batch"]["cpu_baseline"] = {"value": done_h / cpu_h, "unit": "instances/s", "cores": 1,
                                             "kind": "port", "sample": f"first {done_h} of the hard batch"}
        out["hard_batch"]["mismatches_vs_oracle_sample"] = mism_h
        out["hard_batch"]["speedup"] = round(out["hard_batch"]["instances_per_s"] / (done_h / cpu_h), 1)
        out["n24_mismatches"] = sum(int(a != b) for a, b in zip(big_len, big_cpu))
    return out


def bench_eppstein(ctx, count: int, no_cpu: bool) -> dict:
    """Eppstein's bound (pair BFS + greedy word) for the 4096-instance batch in
    one launch, and for single n=300 automata, beside the oracle's restatement
    on one host core (a sample)."""
    import synchrowords_amd as sw
    from synchrowords_amd import gen
    specs = gen.batch_specs(4096, 2025)[:count]
    inst = [(n, k, sw.random_dfa(n, k, seed)) for n, k, seed in specs]
    ctx.eppstein_batch(inst[:64])  # warm
    t0 = time.perf_counter()
    ns, bounds, _ = ctx.eppstein_batch(inst)
    wall = time.perf_counter() - t0
    out = {"workload": f"Eppstein bound of {count} random DFAs, n in [16,160], k in {{2,3}} (batch_specs(4096, 2025))",
           "wall_s": round(wall, 4), "instances_per_s": count / wall, "non_synchro": int(ns.sum())}
    fx = ROOT / "tests" / "golden" / "c4_batch.json"
    if fx.exists():
        cases = json.loads(fx.read_text())["cases"][:count]
        out["non_synchro_mismatches_vs_reference"] = sum(int(bool(c["non_synchro"]) != bool(a))
                                                         for c, a in zip(cases, ns))
    singles = [sw.random_dfa(300, 2, 1000 + s) for s in range(5)]
    ctx.eppstein(300, 2, singles[0])
    t0 = time.perf_counter()
    sb = [ctx.eppstein(300, 2, d)[1] for d in singles]
    out["n300_s_each"] = round((time.perf_counter() - t0) / len(singles), 5)
    out["n300_bounds"] = sb
    if not no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        t0 = time.perf_counter()
        done = mism = 0
        while done < count and time.perf_counter() - t0 < 15.0:
            n, k, d = inst[done]
            w = orc.eppstein(n, k, d, (n - 1) ** 2 + 1)
            mism += int(w[0] != bool(ns[done]) or w[1] != int(bounds[done]))
            done += 1
        cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / cpu, "unit": "instances/s", "cores": 1, "kind": "port",
                               "sample": f"first {done} automata of the batch, oracle pair BFS + greedy"}
        out["mismatches_vs_oracle_sample"] = mism
        out["speedup_batch"] = round(out["instances_per_s"] / (done / cpu), 1)
        t1 = time.perf_counter()
        ob = [orc.eppstein(300, 2, d, 299 * 299 + 1)[1] for d in singles]
        out["n300_cpu_s_each"] = round((time.perf_counter() - t1) / len(singles), 5)
        out["n300_mismatches"] = sum(int(a != b) for a, b in zip(sb, ob))
    return out


def bench_learned(ctx, scorer: str, seeds, width: int) -> dict:
    """Learned-bound against plain-bound exact search on the GPU (n=300, k=2).

    Both pipelines run Eppstein, a device beam of the same width and the
    device exact search capped by the smaller bound, one after the other;
    they differ only in the beam's order: cardinality (the reference's) or
    the trained network.  The configs/exact.json plan (the device Beam of the
    reference's dynamic width, its first pass beside Eppstein) is timed
    beside them."""
    import synchrowords_amd as sw
    from synchrowords_amd.train_scorer import load_scorer
    m = load_scorer(ctx, scorer)
    n, k = m.n, m.k
    rows = []
    plan_no_beam = {"upper_bound": "1ULL * (AUT_N - 1) * (AUT_N - 1) + 1", "algorithms": [
        {"name": "Brute", "config": {"max_n": "15"}},
        {"name": "Eppstein", "config": {"find_word": True}}]}

    def run(d, mlp):
        t0 = time.perf_counter()
        re = ctx.pipeline(n, k, d, plan=plan_no_beam)
        ub = re.upper
        tb = time.perf_counter()
        b, w = ctx.beam(n, k, d, width, ub, mlp=mlp)
        t_beam = time.perf_counter() - tb
        ex = ctx.exact(n, k, d, upper=min(b, ub), upper_word=w if b <= ub else None)
        return ex.length, b, time.perf_counter() - t0, t_beam

    run(sw.random_dfa(n, k, seeds[0]), m)  # warm
    for s in seeds:
        d = sw.random_dfa(n, k, s)
        t0 = time.perf_counter()
        rp = ctx.pipeline(n, k, d)
        t_plan = time.perf_counter() - t0
        len_c, b_c, t_c, tb_c = run(d, None)
        len_l, b_l, t_l, tb_l = run(d, m)
        rows.append({"seed": s, "length": rp.upper, "length_plain": len_c, "length_learned": len_l,
                     "bound_cardinality": b_c, "bound_learned": b_l, "bound_plan_beam": rp.beam_bound,
                     "plain_s": round(t_c, 4), "learned_s": round(t_l, 4), "plan_s": round(t_plan, 4),
                     "beam_s_cardinality": round(tb_c, 4), "beam_s_learned": round(tb_l, 4)})
        print(f"[bench] learned seed {s}: {rows[-1]}", file=sys.stderr, flush=True)
    m.close()
    same = all(r["length_plain"] == r["length_learned"] == r["length"] for r in rows)
    tot = lambda key: round(sum(r[key] for r in rows), 4)
    gap = lambda key: float(np.mean([r[key] - r["length"] for r in rows]))
    return {"workload": f"held-out random DFAs n={n} k={k}, seeds {seeds[0]}..{seeds[-1]}, beam width {width}",
            "identical_lengths": same,
            "plain_s_total": tot("plain_s"), "learned_s_total": tot("learned_s"), "plan_s_total": tot("plan_s"),
            "mean_gap_cardinality": gap("bound_cardinality"), "mean_gap_learned": gap("bound_learned"),
            "mean_gap_plan_beam": gap("bound_plan_beam"),
            "exact_bound_cardinality": sum(r["bound_cardinality"] == r["length"] for r in rows),
            "exact_bound_learned": sum(r["bound_learned"] == r["length"] for r in rows),
            "rows": rows}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["subset", "batch", "learned", "exact", "brute", "eppstein", "times",
                                           "frontier", "reduce", "overflow"],
                    default="subset",
                    help="subset: the headline line; batch / learned: the batched and learned-bound reports; "
                         "exact: one verbose exact search (--exact-n, --exact-seed); brute: power-set BFS; "
                         "eppstein: Eppstein's bound; frontier: one search step at 2^23 frontier sets; "
                         "reduce: the configs/exact_reduce.json plan against configs/exact.json")
    ap.add_argument("--exact-n", type=int, default=300)
    ap.add_argument("--exact-seed", type=int, default=1)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--in-flight", type=int, default=8, help="overflow workload: searches at a time (0: skip)")
    ap.add_argument("--batch-cap-rows", type=int, default=0, help="sets per side inside a workgroup (0: default)")
    ap.add_argument("--batch-max-pairs", type=int, default=0, help="largest goal check inside a workgroup (0: default)")
    ap.add_argument("--scorer", type=str, default=str(ROOT / "synchrowords_amd" / "models" / "scorer_n300_k2.npz"))
    ap.add_argument("--learned-width", type=str, default="64,2048",
                    help="learned workload: beam widths, comma-separated (one report each)")
    ap.add_argument("--learned-seeds", type=int, default=20)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (hardware-counter calibration)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="smaller lists (2^16) for a fast check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-exact", action="store_true", help="skip the time-to-exact leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shape)")
    ap.add_argument("--subset-kernel", choices=["auto", "indexed", "bitwalk"], default="auto",
                    help="which subset kernel the microbench launches (auto: indexed for large launches)")
    ap.add_argument("--traffic-profile", type=str,
                    default=str(ROOT / "profiles" / "k_vsub_ix_subset_microbench_pmc_traffic.md"),
                    help="rocprofv3 --pmc summary holding 'traffic_bytes_per_launch: N' for the roofline")
    ap.add_argument("--shape", type=str, default="", metavar="T,G,E,M",
                    help="use this launch shape (threads, target workgroups, early exit, mask bits) instead of tuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the search does not shard "
              "across GPUs in this build)", file=sys.stderr)
        return 2

    import synchrowords_amd as sw

    if args.copy_calibration:
        ctx = sw.Context(0)
        nbytes = args.copy_calibration << 20
        src = sw.DeviceArray(ctx, nbytes)
        dst = sw.DeviceArray(ctx, nbytes)
        import ctypes
        for _ in range(max(1, args.steps)):
            sw._check(sw.load().sw_dev_copy(ctx.handle, dst.ptr, src.ptr, nbytes))
        print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": max(1, args.steps)}))
        src.free()
        dst.free()
        ctx.close()
        return 0

    if args.workload != "subset":
        ctx = sw.Context(0)
        if args.workload == "batch":
            line = bench_batch(ctx, args.batch_count, args.batch_cap_rows, args.batch_max_pairs)
        elif args.workload == "overflow":
            line = bench_overflow(ctx, args.batch_count, args.in_flight)
        elif args.workload == "brute":
            line = bench_brute(ctx, args.batch_count, args.no_cpu)
        elif args.workload == "eppstein":
            line = bench_eppstein(ctx, args.batch_count, args.no_cpu)
        elif args.workload == "times":
            line = bench_times(ctx)
        elif args.workload == "frontier":
            line = bench_frontier(ctx, args.steps, args.warmup)
        elif args.workload == "reduce":
            line = bench_reduce(ctx)
        elif args.workload == "exact":
            d = sw.random_dfa(args.exact_n, 2, args.exact_seed)
            ub = ctx.pipeline(args.exact_n, 2, d).beam_bound  # warm; the plan's Beam bound caps the search
            r = ctx.exact(args.exact_n, 2, d, upper=ub, cfg=sw.ExactConfig(verbose=True))
            line = {"workload": f"exact search, random binary DFA n={args.exact_n} seed={args.exact_seed}, "
                                f"upper bound {ub} (Beam)", "length": r.length, "stats": r.stats}
        else:
            seeds = list(range(1000, 1000 + args.learned_seeds))
            widths = [int(w) for w in args.learned_width.split(",")]
            reports = [bench_learned(ctx, args.scorer, seeds, w) for w in widths]
            line = reports[0] if len(reports) == 1 else {
                "workload": "learned against plain bound at several beam widths", "widths": widths,
                "reports": reports}
        ctx.close()
        print(json.dumps(line))
        return 0

    n = 256
    count = (1 << 16) if args.quick else (1 << 20)
    t_gen = time.perf_counter()
    small, large, planted = make_lists(count, n, 0.08, 0.75, seed=20251203)
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] generated {count} x {count} sets over n={n} in {t_gen:.1f}s", file=sys.stderr, flush=True)

    ctx = sw.Context(0)
    if args.subset_kernel != "auto":
        ctx.set_subset_kernel(args.subset_kernel == "indexed", 1)
    info = ctx.info()
    ds = sw.DeviceArray.from_host(ctx, small)
    dl = sw.DeviceArray.from_host(ctx, large)
    ctx.sync()

    tuned = None
    if args.shape:
        th, wg, ee, mb = (int(v) for v in args.shape.split(","))
        ctx.set_subset_shape(n, count, count, th, wg, ee, mb)
        tuned = {"threads": th, "target_wgs": wg, "early_exit": ee, "mask_bits": mb, "source": "--shape"}
    elif not args.no_tune:
        # The autotuner times every launch shape on seeded lists of the same
        # sizes and densities and keeps the fastest for this size class in
        # the context.
        t_tune = time.perf_counter()
        tuned = ctx.autotune_subset(n, count, count, 0.08, 0.75, reps=1)
        print(f"[bench] autotuned {tuned} in {time.perf_counter() - t_tune:.1f}s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        cnt, anyf = ctx.subset_check_device(n, ds, count, dl, count, count=True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    counts = []
    for _ in range(args.steps):
        cnt, anyf = ctx.subset_check_device(n, ds, count, dl, count, count=True)
        counts.append(cnt)
    kernel_ms = ctx.timer_stop() / args.steps
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    assert anyf, "planted subsets not found"
    assert len(set(counts)) == 1, "count differs between steps"
    golden = json.loads(GOLDEN_COUNT.read_text()) if GOLDEN_COUNT.exists() else None
    count_ok = None
    if golden and count == golden["count"]:
        # The oracle's count of all 2^40 pairs, computed once over every host
        # thread (tests/golden/make_c2_count.py).
        count_ok = counts[0] == golden["pair_count"]
        assert count_ok, (counts[0], golden["pair_count"])
    pairs = count * count
    value = pairs / (ms_per_step / 1000.0)
    algo_bytes = 2 * count * 32  # both lists read once: (|X| + |Y|) * 8W
    achieved = algo_bytes / (kernel_ms / 1000.0) / 1e9
    print(f"[bench] count {counts[0]} any {anyf} kernel {kernel_ms:.3f} ms step {ms_per_step:.3f} ms",
          file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "subset checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded Bernoulli bitsets, planted subsets)",
        "config": {"workload": f"subset-check microbench: {count} x {count} sets, n={n} (W=4), p=0.08/0.75, "
                               f"1 in 1e5 planted",
                   "pair_count": counts[0], "pair_count_equals_oracle": count_ok,
                   "exists": bool(anyf), "device": info["name"], "cus": info["cus"],
                   "launch_shape": tuned or "default", "subset_kernel": args.subset_kernel},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.traffic_profile),
                     "note": "count of 2^40 pairs: integer VALU/LDS bound, not HBM bound; "
                             "achieved = algorithmic bytes (|X|+|Y|)*32 per launch / kernel time"},
        "pair_tests_per_s_kernel": pairs / (kernel_ms / 1000.0),
        # The bound that applies: each stream state costs one 64-bit LDS row
        # read per lane, so one launch reads sum(|S|) * |T| / 8 bytes of LDS.
        "lds_roofline": lds_roofline(ctx, small, count, kernel_ms, args.subset_kernel != "bitwalk"),
    }
    if not args.no_cpu:
        # The whole 2^40-pair count over every host thread: about 16 s on
        # the box's 16 threads.
        rows = 256 if args.quick else count
        line["cpu_baseline"] = cpu_subset_baseline(ctx, small, large, min(rows, count), golden)
    if not args.no_exact:
        try:
            line["time_to_exact"] = time_to_exact(ctx, 300, 1)
        except Exception as e:  # reported, never hidden
            line["time_to_exact"] = {"error": repr(e)}
    ds.free()
    dl.free()
    ctx.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
