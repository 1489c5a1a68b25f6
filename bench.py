#!/usr/bin/env python3
"""Residue keys joined per second on one MI355X: batched vs plain, bit-exact match sets.

Workload (default): the lambda-family batch -- every ksmall instance of the 118 divisor moduli of
lambda(N) = 155 016 423 578 016 000 (bound U = N) at radii 4..7 with r <= rmax(D): 280 instances,
1.60e9 keys (nD + nI), 79 pairs. One step = one pass of the join over the whole batch:
  batched -- all 280 instances in one solve, captured once into a HIP graph (cf_batch_capture)
             and replayed as one graph launch per step, the timed value;
  plain   -- one solve per instance, each an eager sequence of launches, timed beside it.
Both legs prepare their instances first (descriptors and exact bound tables uploaded by
cf_batch_create), so the timed region starts with every input resident in HBM.
Every instance's (nD, nI, pairs) is checked against the reference's logged counts, and the pair sets of
the CPU-baseline sample are checked against the CPU join. Inputs (instance descriptors) are host
values of a few KB; everything the join touches is generated on the device.

JSON line fields beyond the required ones: plain (value, ms_per_step), batched_vs_plain, roofline
(dominant kernel, live HIP-event timing on the kernel's stream over the timed steps, plus a per-kernel
breakdown from one extra untimed step with every kernel timed), cpu_baseline, verify.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "carmichael-frontier_amd"))

PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes per I record of the two kernels that dominate a step:
#   k_part_scatter writes each I record (8 bytes: bucket offset, fingerprint, leaf index) once;
#   k_part_probe reads each 8-byte record once. Its 16-byte probe-table reads hit L2 by design (an
#   XCD probes one partition at a time and the partition's table slice fits its L2), so they are
#   not HBM bytes; the measured traffic (FETCH_SIZE) shows what they cost when they do miss.
KERNEL_BYTES_PER_I = {"k_part_scatter": 8, "k_part_probe": 8}
NODE_BYTES = 32 + 8   # a last-level tree node (LeafNode) and its entry of the leaf-offset scan


def run_batched(batch):
    """One batched step: the whole batch in one solve."""
    res, st = batch.solve()
    return res, st


def run_plain(singles):
    """One plain step: one solve per instance (each instance prepared on its own)."""
    return [b.solve()[0][0] for b in singles]


def verify_counts(insts, res):
    bad = []
    for inst, r in zip(insts, res):
        lg = inst.meta.get("logged")
        if lg is None or (lg["nD"], lg["nI"], lg["pairs"]) != (r.nD, r.nI, len(r.pairs_d)):
            bad.append((inst.tag, lg, (r.nD, r.nI, len(r.pairs_d))))
    return bad


def cpu_baseline(insts, res_batched, threads, budget_keys):
    """The CPU restatement of the reference join (hash the D residues, T threads striped over the first
    insertion index; src/mitm64.cpp) over a bounded sample of the same instances, pair sets compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    order = sorted(range(len(insts)), key=lambda j: insts[j].meta["logged"]["nD"] + insts[j].meta["logged"]["nI"])
    sample, keys = [], 0
    for j in order[::2]:
        k = insts[j].meta["logged"]["nD"] + insts[j].meta["logged"]["nI"]
        if keys + k > budget_keys:
            continue
        sample.append(j)
        keys += k
    t_sum = 0.0
    mismatches = 0
    for j in sample:
        x = insts[j]
        r = oracle.cpu_join(x.M, x.r, x.P0 % x.M, x.Icap, x.Dmin, x.B0, x.INS, x.split_idx, x.hmax, x.rank_ids,
                            threads=threads)
        t_sum += r["td"] + r["ti"]
        if not np.array_equal(oracle.pair_set(r["pairs_d"], r["pairs_i"]), res_batched[j].pair_array()):
            mismatches += 1
    return dict(value=keys / t_sum if t_sum else 0.0, unit="residue keys joined/s", cores=threads, kind="port",
                sample=f"{len(sample)} of the {len(insts)} batch instances (every other one in size order, "
                       f"{keys:.3g} keys), D build + I scan time (oracle/oracle.c or_cpu_join)",
                seconds=round(t_sum, 2), pair_set_mismatches=mismatches)


def host_cpu_model() -> dict:
    """The host CPU: model name, logical CPUs this process may use, physical cores, SMT state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from calibrate_cpu import host_cpu
    return host_cpu()


def cpu_calibration():
    """The port's speed against the reference's own CPU join (src/mitm64.cpp compiled from its
    sources) on identical inputs, as measured by oracle/calibrate_cpu.py on this kind of host."""
    path = os.path.join(ROOT, "oracle", "calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    return {"port_over_reference_speed": round(c["port_over_reference_speed"], 3), "threads": c["threads"],
            "host": c["host_cpu"].get("model"), "anchors": {a["instance"]: {
                "reference_keys_per_s": round(a["reference_mitm64"]["keys_per_s"]),
                "port_keys_per_s": round(a["port_or_cpu_join"]["keys_per_s"]),
                "same_records": a["same_records"],
                "reference_matches_within_port_pairs": a["reference_matches_within_port_pairs"]}
                for a in c["anchors"]}, "source": "oracle/calibration.json (oracle/calibrate_cpu.py)"}


def copy_probe(ctx, nbytes):
    """A plain device-to-device copy of nbytes: the calibration point for FETCH_SIZE / WRITE_SIZE."""
    from cfjoin import xp
    n = nbytes // 8
    src = xp.empty(n)
    dst = xp.empty(n)
    ctx.d2d(dst.ptr, src.ptr, n * 8)
    ctx.synchronize()


def kernel_name(s):
    s = s.replace("(anonymous namespace)::", "").replace("void ", "")
    return s.split("(")[0].split("<")[0]


def counter_pass(counters, family, setting):
    """One rocprofv3 counter pass (`counters`, at most 8 SQ and 4 TCC counters) over a child run of
    this script (one eager solve of the batch with the same tuning, then a 1 GiB device copy).
    Returns ({kernel: {counter: mean per dispatch}, '_dispatches': {kernel: n}}, None) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    tune_arg = ",".join(str(setting.get(k, 0)) for k in ("slice_kib", "d_per_bucket", "seg_fill_pct", "probe_blocks"))
    with tempfile.TemporaryDirectory(prefix="cf_pmc_", dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", tmp, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--pmc-child", "--family", family, "--tuning", tune_arg,
               "--copy-probe", str(1 << 30)]
        try:
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=180, check=True)
        except (subprocess.SubprocessError, OSError) as e:
            return None, f"{'+'.join(counters)} pass failed: {type(e).__name__}"
        files = [os.path.join(r, f) for r, _, fs in os.walk(tmp) for f in fs if f.endswith("counter_collection.csv")]
        if not files:
            return None, f"{'+'.join(counters)} pass wrote no counters"
        sums, seen = {}, {}
        for r in csv.DictReader(open(files[0])):
            k = "copy" if "copyBuffer" in r["Kernel_Name"] else kernel_name(r["Kernel_Name"])
            c = r["Counter_Name"]
            v = float(r["Counter_Value"])
            if k == "copy":   # the calibration copy: its largest dispatch (the 1 GiB one)
                sums.setdefault(k, {})[c] = max(sums.get(k, {}).get(c, 0.0), v)
                seen[k] = 1
                continue
            sums.setdefault(k, {})[c] = sums.get(k, {}).get(c, 0.0) + v
            seen.setdefault(k, set()).add(r["Dispatch_Id"])
    out = {k: {c: v / (1 if k == "copy" else len(seen[k])) for c, v in d.items()} for k, d in sums.items()}
    out["_dispatches"] = {k: (1 if k == "copy" else len(v)) for k, v in seen.items()}
    return out, None


def measure_traffic(kernel, family, setting):
    """HBM bytes per launch of `kernel`: two counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch); each counter is scaled by what it reads for the 1 GiB calibration copy (the copy reads
    and writes 1 GiB). Returns (bytes per launch, details) or (None, reason)."""
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        res, why = counter_pass([ctr], family, setting)
        if res is None:
            return None, why
        if kernel not in res or "copy" not in res or res["copy"].get(ctr, 0) <= 0:
            return None, f"{ctr}: no dispatches of {kernel} or of the copy"
        scale = (1 << 20) / res["copy"][ctr]      # KiB the copy moved / KiB the counter read for it
        out[ctr] = (res[kernel][ctr] * 1024 * scale, scale, res["_dispatches"][kernel])
    total = out["FETCH_SIZE"][0] + out["WRITE_SIZE"][0]
    return total, {"fetch_bytes": round(out["FETCH_SIZE"][0]), "write_bytes": round(out["WRITE_SIZE"][0]),
                   "fetch_scale": round(out["FETCH_SIZE"][1], 4), "write_scale": round(out["WRITE_SIZE"][1], 4),
                   "launches": out["FETCH_SIZE"][2]}


ALU_COUNTERS = ["SQ_INSTS_VALU", "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_INT64", "SQ_INSTS_SALU", "SQ_WAVES"]


def valu_peaks(ctx):
    """This GPU's vector-ALU issue rate (wave instructions per second) for v_add_u32 (full rate) and
    v_mul_lo_u32 / v_mad_u64_u32 (the multiply forms), from the library's microbenchmark."""
    import ctypes
    from cfjoin import lib
    peaks = {}
    for op, name, iters in ((0, "v_add_u32", 65536), (1, "v_mul_lo_u32", 16384), (3, "v_mad_u64_u32", 8192)):
        n, ms = lib.c_u64(), ctypes.c_double()
        best = 0.0
        for _ in range(2):
            lib.check(lib.load().cf_peak_valu(ctx.handle, op, iters, ctypes.byref(n), ctypes.byref(ms)), "cf_peak_valu")
            best = max(best, n.value / (ms.value / 1e3))
        peaks[name] = best
    return peaks


def measure_alu(kernels, breakdown, family, setting, peaks):
    """Integer-ALU roofline of `kernels`: VALU wave instructions per dispatch (SQ_INSTS_VALU, of which
    SQ_INSTS_VALU_INT32 / _INT64 are integer arithmetic) over the kernel's mean launch time (HIP events,
    eager step), against the measured v_add_u32 issue rate. A wave instruction is 64 lane operations."""
    res, why = counter_pass(ALU_COUNTERS, family, setting)
    if res is None:
        return {"error": why}
    peak = peaks["v_add_u32"]
    out = {"counters": ALU_COUNTERS, "peak_wave_instr_per_s": {k: round(v, -6) for k, v in peaks.items()},
           "peak_basis": "v_add_u32 issue rate measured by cf_peak_valu", "kernels": {}}
    for k in kernels:
        if k not in res or k not in breakdown or not breakdown[k][1]:
            continue
        ms = breakdown[k][0] / breakdown[k][1]
        c = res[k]
        rate = c["SQ_INSTS_VALU"] / (ms / 1e3)
        out["kernels"][k] = {"valu_per_launch": round(c["SQ_INSTS_VALU"]), "int32": round(c["SQ_INSTS_VALU_INT32"]),
                             "int64": round(c["SQ_INSTS_VALU_INT64"]), "salu": round(c["SQ_INSTS_SALU"]),
                             "waves": round(c["SQ_WAVES"]), "avg_launch_ms": round(ms, 4),
                             "achieved_wave_instr_per_s": round(rate, -6),
                             "achieved_lane_ops_per_s": round(rate * 64, -9),
                             "frac": round(rate / peak, 4)}
    return out


def c4_main(args):
    """The large-radius streamed join: lambda(N) = 155 016 423 578 016 000, ksmall base, r = 8, the
    affordable pool of 256 primes (h <= 1). Full D side: all C(64,8) = 4 426 165 368 D records
    (Dmin = 1), bucketed in HBM; I side (471 754 092 records) generated and probed in chunks.
    Beside it the reference's own run of the same instance with the D side filtered by Dmin (its
    logged counts are checked). The headline solve takes the library's default chunking (chunks of
    up to min(2^31, free device memory / 48 B) I leaves: the whole I side here); a second solve streams
    the I side in chunks of --c4-chunk leaves and must return the same pair set. Every pair is
    re-checked exactly on the host."""
    import copy
    from cfjoin import lib, pipeline
    from cfjoin import workloads as W
    logged = W.logged_table(os.path.join(ROOT, "tests", "golden", "logged_instances.json"))
    filt = W.anchor("lambdaN", 8, logged)
    full = copy.deepcopy(filt)
    full.Dmin = 1
    full.tag = "lambdaN:anchor:r8:full-D"
    ctx = lib.default_context()
    lg = filt.meta["logged"]
    t0 = time.perf_counter()
    rf = pipeline.solve(filt, ctx, max_chunk=args.c4_chunk)
    t_f = time.perf_counter() - t0
    print(f"[c4] D-filtered: nD={rf.nD} nI={rf.nI} pairs={len(rf.pairs_d)} in {t_f:.2f} s (logged {lg})",
          file=sys.stderr, flush=True)
    runs = []
    for chunk in (0, args.c4_chunk):
        # one untimed solve first: the device pool then holds the solve's blocks, so the timed one
        # measures the join and not the first allocation of ~150 GB
        ctx.release_pool()
        pipeline.solve(full, ctx, max_chunk=chunk)
        ctx.synchronize()
        t0 = time.perf_counter()
        r = pipeline.solve(full, ctx, max_chunk=chunk)
        t = time.perf_counter() - t0
        runs.append((chunk, r, t))
        print(f"[c4] full D, chunk {chunk or 'default'}: nD={r.nD} nI={r.nI} pairs={len(r.pairs_d)} in {t:.2f} s "
              f"(device {r.stats['ms_total']:.1f} ms, peak {r.stats['peak_bytes'] / 2**30:.1f} GiB)", file=sys.stderr,
              flush=True)
    (_, r1, t1), (_, r2, t2) = runs
    checks = pipeline.recheck(full, r1)
    keys = r1.nD + r1.nI
    line = {
        "metric": "large-radius streamed join: residue keys joined/s, full C(64,8) D side in HBM",
        "value": keys / (r1.stats["ms_total"] / 1e3), "unit": "residue keys joined/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "u64", "data": "synthetic: the anchor instance rebuilt from the census",
        "config": {"workload": "lambda(N) anchor, ksmall base, r=8, |INS|=256, h<=1, Dmin=1 (full D side)",
                   "nD": r1.nD, "nI": r1.nI, "i_chunk": "library default (the whole I side in one chunk here)"},
        "device_ms": r1.stats["ms_total"], "wall_s": round(t1, 3), "peak_device_bytes": r1.stats["peak_bytes"],
        "pairs": len(r1.pairs_d),
        "streamed": {"i_chunk": args.c4_chunk, "device_ms": r2.stats["ms_total"], "wall_s": round(t2, 3),
                     "value": keys / (r2.stats["ms_total"] / 1e3),
                     "peak_device_bytes": r2.stats["peak_bytes"]},
        "d_filtered": {"nD": rf.nD, "nI": rf.nI, "pairs": len(rf.pairs_d), "wall_s": round(t_f, 3),
                       "logged": {"nD": lg["nD"], "nI": lg["nI"], "pairs": lg["pairs"]}},
        "verify": {"d_filtered_matches_logged": (rf.nD, rf.nI, len(rf.pairs_d)) == (lg["nD"], lg["nI"], lg["pairs"]),
                   "full_d_records_equal_C64_8": r1.nD == 4426165368,
                   "same_pairs_both_chunkings": bool(np.array_equal(r1.pair_array(), r2.pair_array())),
                   "pairs_rechecked_exact": len(checks), "pairs_failing_recheck": sum(not c["integral"] for c in checks)},
    }
    print(json.dumps(line), flush=True)


def rows_main(args):
    """The reference's own composition, stage by stage, on the M0 anchor: M0 = 1 768 248 177 696 000,
    r = 6, ksmall base, the affordable pool of 375 primes, h <= 2 (logged: nD 8 340 300, nI 228 944 388).
    The driver's GPU path (gpu/mitm_gpu_mstar_driver.py:101-118,209-238) through the drop-in modules:
    D side filtered by Dmin (the GPU stand-in of mitm64_ddump, a plan), sorted once (sort_side); I
    side from generate_chunks in 2e8-record chunks, each sorted and joined against D (join_sorted).
    Beside it: the same instance as one fused solve (cf_solve), the full C(64,6) D side by
    rank-range generation (generate, the DFILTER=0 path), and the CPU join on 16 host threads.
    Stage times are wall times between device synchronizations; counts are checked against the log."""
    sys.path.insert(0, os.path.join(ROOT, "carmichael-frontier_amd", "gpu"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mitm_gpu2 as G
    import igen_gpu as IG
    import oracle
    from cfjoin import lib, pipeline, xp
    from cfjoin import workloads as W
    logged = W.logged_table(os.path.join(ROOT, "tests", "golden", "logged_instances.json"))
    inst = W.anchor("M0", 6, logged)
    lg = inst.meta["logged"]
    ctx = lib.default_context()

    def timed(f):
        ctx.synchronize()
        t = time.perf_counter()
        out = f()
        ctx.synchronize()
        return out, time.perf_counter() - t

    # warm-up: code objects and the device pool
    pipeline.solve(inst, ctx)

    def composition():
        st = {}
        plan, st["d_plan_s"] = timed(lambda: lib.Plan.deletion(inst.B0, inst.r, inst.M, inst.Dmin, lib.CF_ID_PACK_DESC))
        nD = plan.total()
        Dk, Did = xp.empty(nD), xp.empty(nD)
        _, st["d_write_s"] = timed(lambda: plan.write(0, len(inst.B0), Dk.ptr, Did.ptr))
        plan.close()
        (Dks, Dids), st["d_sort_s"] = timed(lambda: G.sort_side(Dk, Did))
        del Dk, Did
        gen = IG.generate_chunks(inst.INS, inst.r, inst.M, inst.P0, inst.Icap, inst.split, inst.hmax,
                                 max_records=200_000_000)
        t_gen = t_sort = t_join = 0.0
        nI = pairs = chunks = 0
        total = None
        while True:
            ctx.synchronize()
            t = time.perf_counter()
            try:
                Ik, Iid = next(gen)
            except StopIteration as e:
                total = e.value
                break
            ctx.synchronize()
            t_gen += time.perf_counter() - t
            nI += int(Ik.size)
            chunks += 1
            (Iks, Iids), dt = timed(lambda: G.sort_side(Ik, Iid))
            t_sort += dt
            del Ik, Iid
            (pd, pi), dt = timed(lambda: G.join_sorted(Dks, Dids, Iks, Iids))
            t_join += dt
            pairs += int(pd.size)
            del Iks, Iids, pd, pi
            print(f"[rows] I chunk {chunks}: {nI} records so far", file=sys.stderr, flush=True)
        del Dks, Dids
        return st, nD, t_gen, t_sort, t_join, nI, pairs, chunks, total

    # the composition twice: the first run fills the device pool (what the first instance of a
    # family run pays), the second is the steady state every later instance sees
    cold = composition()
    st, nD, t_gen, t_sort, t_join, nI, pairs, chunks, total = composition()
    ctx.release_pool()
    cold_total = sum(cold[0].values()) + cold[2] + cold[3] + cold[4]
    t_comp = st["d_plan_s"] + st["d_write_s"] + st["d_sort_s"] + t_gen + t_sort + t_join
    res, t_fused = timed(lambda: pipeline.solve(inst, ctx))
    (fk, fi), t_full = timed(lambda: G.generate(inst.B0, inst.r, inst.M, 1, bits=6))
    n_full = int(fk.size)
    del fk, fi
    ctx.release_pool()
    cpu = None
    if not args.no_cpu:
        r = oracle.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                            inst.hmax, inst.rank_ids, threads=16)
        cpu = dict(value=(r["nD"] + r["nI"]) / (r["td"] + r["ti"]), unit="residue keys joined/s", cores=16,
                   kind="port", sample="the whole instance (oracle/oracle.c or_cpu_join: D hash build + I scan)",
                   seconds=round(r["td"] + r["ti"], 2),
                   pairs_equal=bool(np.array_equal(oracle.pair_set(r["pairs_d"], r["pairs_i"]), res.pair_array())))
    keys = nD + nI
    line = {
        "metric": "reference call surface, stage by stage: residue keys joined/s",
        "value": keys / t_comp, "unit": "residue keys joined/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "u64", "data": "synthetic: the anchor instance rebuilt from the census",
        "config": {"workload": "M0 anchor, ksmall base, r=6, |INS|=375, h<=2, D filtered by Dmin, I chunks of 2e8",
                   "nD": nD, "nI": nI, "i_chunks": chunks},
        "stages_s": {"d_plan": round(st["d_plan_s"], 4), "d_write": round(st["d_write_s"], 4),
                     "d_sort": round(st["d_sort_s"], 4), "i_generate_chunks": round(t_gen, 4),
                     "i_sort_side": round(t_sort, 4), "join_sorted": round(t_join, 4), "total": round(t_comp, 4)},
        "first_run": {"total_s": round(cold_total, 4), "value": keys / cold_total,
                      "note": "the same composition on an empty device pool (its allocations included)"},
        "rates": {"i_records_per_s": nI / t_gen, "sort_keys_per_s": nI / t_sort, "join_probes_per_s": nI / t_join},
        "fused_solve": {"value": keys / t_fused, "wall_s": round(t_fused, 4), "device_ms": res.stats["ms_total"],
                        "pairs": len(res.pairs_d)},
        "generate_full_C64_6": {"records": n_full, "wall_s": round(t_full, 4), "records_per_s": n_full / t_full},
        "cpu_baseline": cpu,
        "verify": {"nD_nI_pairs_equal_logged": (nD, nI, pairs) == (lg["nD"], lg["nI"], lg["pairs"]),
                   "first_run_equal": (cold[1], cold[5], cold[6]) == (nD, nI, pairs),
                   "generate_chunks_total": total == nI, "fused_equals_composition": len(res.pairs_d) == pairs,
                   "full_generate_equals_C64_6": n_full == 74974368},
    }
    print(json.dumps(line), flush=True)


def audit_main(args):
    """Every logged instance of both divisor families (1 412 lines of the reference's run logs),
    rebuilt from the census and solved on the device in batches; (nD, nI, pairs) compared with the
    log line by line, every pair re-checked exactly on the host (cfjoin/audit.py)."""
    from cfjoin import audit
    rep = audit.audit(os.path.join(ROOT, "tests", "golden", "logged_instances.json"),
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    line = {
        "metric": "coverage audit: logged instances reproduced on the device",
        "value": rep["keys"] / (rep["device_ms"] / 1e3) if rep["device_ms"] else 0.0,
        "unit": "residue keys joined/s", "n_gpus": 1, "higher_is_better": True, "dtype": "u64",
        "data": "synthetic: every logged instance rebuilt from the census",
        "config": {"workload": "both divisor families, every logged (row, radius)", "batches": rep["batches"]},
        "instances": rep["instances"], "solved_on_device": rep["solved"], "keys": rep["keys"],
        "pairs": rep["pairs"], "device_ms": round(rep["device_ms"], 2), "wall_s": round(rep["wall_s"], 2),
        "mismatches": rep["mismatches"], "recheck_failures": rep["recheck_failures"],
        "beyond_device": rep["beyond_device"],
    }
    print(json.dumps(line), flush=True)


def serve_main(args):
    """The serving measurement: one set of requests (both families, radii 4..8, up to 5e8 records
    each), served (a) by the continuous-batching server and (b) one request at a time (one worker,
    batches of one, FIFO): first all arriving at once (each mode's capacity), then as a Poisson stream
    at 0.9 of the one-at-a-time capacity (latency at a fixed rate). Every result is checked against the
    reference's logged counts and the runs' pair sets against each other."""
    from cfjoin import serve
    from cfjoin import workloads as W
    logged = W.logged_table(os.path.join(ROOT, "tests", "golden", "logged_instances.json"))
    insts = W.serving_requests(args.requests, seed=args.seed, logged=logged)
    keys = sum(x.meta["logged"]["nD"] + x.meta["logged"]["nI"] for x in insts)
    print(f"[serve] {len(insts)} requests, {keys:.4g} keys", file=sys.stderr, flush=True)
    # warm-up: a few requests one at a time (fills the device pools)
    serve.run_one_at_a_time(insts[:min(16, len(insts))])

    def stream(workers, max_batch, arrivals):
        srv = serve.Server(workers=workers, max_batch=max_batch)
        t0 = time.perf_counter()
        try:
            out = serve.run_stream(srv, insts, arrivals)
        finally:
            srv.close()
        return sorted(out, key=lambda r: r.tag), time.perf_counter() - t0

    # capacity of each mode: every request arrives at once
    burst = np.zeros(len(insts))
    single_burst, wall_sb = stream(1, 1, burst)
    batched_burst, wall_bb = stream(args.workers, args.max_batch, burst)
    print(f"[serve] all at once: one at a time {wall_sb:.3f} s, continuous batching {wall_bb:.3f} s", file=sys.stderr,
          flush=True)
    # fixed-rate streams: Poisson arrivals at 0.9, 1.5 and 3 times what one-at-a-time service
    # sustains; above 1 the one-at-a-time queue grows through the stream, while continuous batching
    # folds the requests that arrive during a solve into the next batch
    cap_rate = len(insts) / wall_sb
    factors = [args.rate / cap_rate] if args.rate else [0.9, 1.5, 3.0]
    rates, same, bad = [], True, []
    for j, f in enumerate(factors):
        rate = f * cap_rate
        arrivals = serve.poisson_arrivals(len(insts), rate, seed=args.seed + 1 + j)
        batched, wall_b = stream(args.workers, args.max_batch, arrivals)
        single, wall_s = stream(1, 1, arrivals)
        sb, ss = serve.summary(batched, wall_b), serve.summary(single, wall_s)
        print(f"[serve] Poisson {f:.1f}x ({rate:.0f}/s): continuous batching {wall_b:.3f} s p50 {sb['p50_ms']:.2f} ms "
              f"p99 {sb['p99_ms']:.2f} ms batch {sb['mean_batch']:.1f}; one at a time {wall_s:.3f} s p50 "
              f"{ss['p50_ms']:.2f} ms p99 {ss['p99_ms']:.2f} ms", file=sys.stderr, flush=True)
        rates.append({"rate_vs_one_at_a_time_capacity": f, "arrivals_per_s": round(rate, 2),
                      "continuous_batching": sb, "one_at_a_time": ss,
                      "p99_ratio": ss["p99_ms"] / sb["p99_ms"] if sb["p99_ms"] else None})
        bad += [x.tag for x, r in zip(insts, batched)
                if (r.nD, r.nI, len(r.pairs_d)) != (x.meta["logged"]["nD"], x.meta["logged"]["nI"],
                                                    x.meta["logged"]["pairs"])]
        same &= all(np.array_equal(a.pair_array(), b.pair_array()) and (a.nD, a.nI) == (b.nD, b.nI)
                    for runs in ((batched, single), (batched, batched_burst), (batched, single_burst))
                    for a, b in zip(*runs))
    sbb, ssb = serve.summary(batched_burst, wall_bb), serve.summary(single_burst, wall_sb)
    top = rates[-1]
    line = {
        "metric": "serving: residue keys joined/s and p50/p99 request latency at a fixed Poisson arrival rate",
        "value": top["continuous_batching"]["keys_per_s"], "unit": "residue keys joined/s", "n_gpus": 1,
        "higher_is_better": True,
        "p50_ms": round(top["continuous_batching"]["p50_ms"], 3), "p99_ms": round(top["continuous_batching"]["p99_ms"], 3),
        "dtype": "u64", "data": "synthetic: instances rebuilt from the reference's census rows",
        "config": {"workload": f"{len(insts)} requests drawn (seed {args.seed}) from the logged instances of both "
                               f"families, radii 4..8, <= 5e8 keys each; {keys} keys",
                   "arrival_rate_per_s": top["arrivals_per_s"],
                   "rate_vs_one_at_a_time_capacity": top["rate_vs_one_at_a_time_capacity"],
                   "workers": args.workers, "max_batch": args.max_batch},
        "rates": rates,
        "all_at_once": {"continuous_batching": sbb, "one_at_a_time": ssb,
                        "throughput_ratio": sbb["keys_per_s"] / ssb["keys_per_s"]},
        "verify": {"logged_count_mismatches": len(bad), "same_results_in_mix_and_alone": bool(same)},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--family", default="lambdaN", choices=["lambdaN", "M0"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain leg (profiling the batched path alone)")
    ap.add_argument("--no-graph", action="store_true", help="time the batched leg as eager launches, not a graph")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tuning (skip the autotuner)")
    ap.add_argument("--tuning", default="", metavar="SLICE,DPB,FILL,GRID",
                    help="pin the tuning (slice KiB, D per bucket, segment fill %%, probe blocks) instead of tuning")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes for traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve", action="store_true",
                    help="serving mode: a Poisson stream of modulus solves through the continuous-batching server, "
                         "beside the same stream served one request at a time")
    ap.add_argument("--requests", type=int, default=240, help="serving mode: requests in the stream")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving mode: one arrival rate per second (0: 0.9, 1.5 and 3 times what one-at-a-time "
                         "service sustains)")
    ap.add_argument("--workers", type=int, default=4, help="serving mode: server workers (streams)")
    ap.add_argument("--max-batch", type=int, default=64, help="serving mode: requests per batch at most")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--audit", action="store_true",
                    help="every logged instance of both families solved and compared with the log")
    ap.add_argument("--rows", action="store_true",
                    help="the reference's own composition (generate_chunks, sort_side, join_sorted) stage by stage")
    ap.add_argument("--c4", action="store_true",
                    help="large-radius mode: the lambda(N) anchor at r = 8 with the full D side (C(64,8) records), "
                         "I side streamed in chunks, beside the reference's D-filtered run")
    ap.add_argument("--c4-chunk", type=int, default=1 << 28,
                    help="large-radius mode: I records per chunk of the streamed comparison solve")
    ap.add_argument("--cpu-keys", type=float, default=8.0e8, help="CPU-baseline sample size in keys")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="after the run, one device-to-device copy of BYTES (calibrates byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cfjoin import lib, pipeline
    from cfjoin import workloads as W

    if args.serve:
        serve_main(args)
        return
    if args.c4:
        c4_main(args)
        return
    if args.rows:
        rows_main(args)
        return
    if args.audit:
        audit_main(args)
        return

    logged = W.logged_table(os.path.join(ROOT, "tests", "golden", "logged_instances.json"))
    insts = W.batch_instances(args.family, radii=(4, 5, 6, 7), logged=logged)
    ctx = lib.default_context()
    print(f"[bench] {len(insts)} instances of the {args.family} family, radii 4..7", file=sys.stderr, flush=True)

    # instance preparation (descriptors and bound tables onto the device) happens before timing
    batch = pipeline.Batch(insts)
    pinned = [int(x) for x in args.tuning.split(",")] if args.tuning else None
    if pinned:
        batch.set_tuning(*pinned)
    if args.pmc_child:
        # one eager solve of the batch, then the calibration copy: what the counter passes read
        run_batched(batch)
        copy_probe(ctx, args.copy_probe or (1 << 30))
        return
    singles = [] if args.no_plain else [pipeline.Batch([x]) for x in insts]
    # warm-up doubles as the bit-exactness check of both paths
    for _ in range(max(args.warmup, 1)):
        res_b, _ = run_batched(batch)
    res_p = res_b if args.no_plain else run_plain(singles)
    bad_b = verify_counts(insts, res_b)
    bad_p = verify_counts(insts, res_p)
    same = all((a.nD, a.nI) == (b.nD, b.nI) and (a.pair_array() == b.pair_array()).all() for a, b in zip(res_b, res_p))
    keys = sum(r.nD + r.nI for r in res_b)
    nI_total = sum(r.nI for r in res_b)
    print(f"[bench] warm-up done: {keys:.4g} keys, logged-count mismatches batched={len(bad_b)} plain={len(bad_p)}, "
          f"batched == plain pair sets: {same}", file=sys.stderr, flush=True)

    # ---- autotuner (untimed): slice per partition, D records per bucket, segment fill, probe grid;
    # each candidate timed the way the step runs (graph replays unless --no-graph)
    tuning = {"setting": dict(zip(("slice_kib", "d_per_bucket", "seg_fill_pct", "probe_blocks"), pinned)),
              "pinned": True} if pinned else None
    if not args.no_tune and not pinned:
        from cfjoin import autotune
        tr = autotune.tune(batch, reps=5, graph=not args.no_graph)
        tuning = {"setting": tr["best"], "ms": round(tr["ms"], 3), "default_ms": round(tr["default_ms"], 3),
                  "trials": len(tr["trials"]), "timed_as": "eager solve" if args.no_graph else "graph replay"}
        print(f"[bench] tuned: {tr['best']} {tr['ms']:.2f} ms (default {tr['default_ms']:.2f} ms)", file=sys.stderr,
              flush=True)

    # ---- kernel breakdown: one untimed eager step with every kernel bracketed by HIP events on the
    # stream it runs on; it names the dominant kernel (the one of the two I-record kernels that
    # takes longer) and gives its launch times (a graph replay carries no per-kernel events; the
    # replayed launches are the same kernels on the same data, checked with rocprofv3)
    ctx.prof_enable(None)
    ctx.prof_reset()
    _, st_prof = run_batched(batch)
    breakdown = ctx.prof_dump()
    ctx.prof_disable()
    dominant = max(KERNEL_BYTES_PER_I, key=lambda k: breakdown.get(k, (0.0, 0))[0])
    k_ms, k_launches = breakdown.get(dominant, (0.0, 0))

    # ---- the eager batched solve (the same launches issued one by one from the host), timed over the
    # same steps, beside the captured graph's replays
    ctx.synchronize()
    te0 = time.perf_counter()
    for _ in range(args.steps):
        run_batched(batch)
    ctx.synchronize()
    eager_ms = (time.perf_counter() - te0) * 1e3 / args.steps

    # ---- the persistent launch: the batch's solve captured into one HIP graph; one replay is
    # checked against the eager result
    graph_same = None
    if not args.no_graph:
        batch.capture()
        res_g, _ = run_batched(batch)
        graph_same = all((a.nD, a.nI) == (b.nD, b.nI) and (a.pair_array() == b.pair_array()).all()
                         for a, b in zip(res_b, res_g))
    print(f"[bench] graph replay == eager: {graph_same}", file=sys.stderr, flush=True)

    # ---- timed: batched
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_batched(batch)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = keys / ((t1 - t0) / args.steps)

    # ---- plain, one step
    plain_ms = plain_value = None
    if not args.no_plain:
        ctx.synchronize()
        t2 = time.perf_counter()
        run_plain(singles)
        ctx.synchronize()
        t3 = time.perf_counter()
        plain_ms = (t3 - t2) * 1e3
        plain_value = keys / (t3 - t2)

    per_rec = KERNEL_BYTES_PER_I[dominant]
    bytes_per_step = per_rec * nI_total
    achieved = bytes_per_step / (k_ms / 1e3) / 1e9 if k_ms else 0.0
    roofline = dict(kernel=dominant, bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=None,
                    algorithmic=f"{per_rec} B per I record x {nI_total} I records per step",
                    avg_launch_ms=round(k_ms / max(k_launches, 1), 4), launches_per_step=k_launches,
                    kernel_ms_per_step=round(k_ms, 3),
                    kernel_timing="HIP events around each launch on its stream, eager profiled step",
                    breakdown_ms_per_step={k: round(v[0], 3) for k, v in
                                           sorted(breakdown.items(), key=lambda kv: -kv[1][0])})

    # the device work is done: give the memory back before the counter passes run their own solve
    batch.close()
    for x in singles:
        x.close()
    ctx.release_pool()
    if not args.no_traffic:
        setting = (tuning or {}).get("setting") or {}
        traffic, detail = measure_traffic(dominant, args.family, setting)
        roofline["traffic"] = round(traffic) if traffic else None
        roofline["traffic_detail"] = detail
        if traffic:
            roofline["traffic_vs_algorithmic"] = round(traffic / (per_rec * nI_total / max(k_launches, 1)), 3)
            if dominant == "k_part_scatter" and detail:
                # the scatter also reads each last-level I node (32 bytes) and its leaf offset (8 bytes):
                # the reads its record writes do not account for
                pb = NODE_BYTES * st_prof["nI_parents"] / max(k_launches, 1)
                roofline["parents"] = {"nodes_per_step": st_prof["nI_parents"], "bytes_per_launch": round(pb),
                                       "share_of_fetch": round(pb / detail["fetch_bytes"], 3),
                                       "traffic_vs_algorithmic_with_parents":
                                           round(traffic / (per_rec * nI_total / max(k_launches, 1) + pb), 3)}
        roofline["alu"] = measure_alu(["k_part_scatter", "k_part_probe", "k_expand_write_count", "k_dwin_sort",
                                       "k_part_scatter_d", "k_dwin_scatter"], breakdown, args.family, setting,
                                      valu_peaks(ctx))

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(insts, res_b, threads, args.cpu_keys)
        cpu["host_cpu"] = host_cpu_model()
        cpu["calibration"] = cpu_calibration()

    line = {
        "metric": "residue keys joined/s (batched vs plain, 1 GPU), bit-exact match sets",
        "value": value,
        "unit": "residue keys joined/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: instances rebuilt from the reference's census rows; no record data is loaded",
        "config": {"workload": f"{args.family}-family batch: {len(insts)} instances (moduli x radii 4..7 <= rmax), "
                               f"{keys} keys (nD+nI)", "instances": len(insts), "keys_per_step": keys,
                   "parallelism": "single GPU", "tuning": tuning},
        "plain": {"value": plain_value, "ms_per_step": plain_ms},
        "eager_batched": {"value": keys / (eager_ms / 1e3), "ms_per_step": eager_ms},
        "batched_vs_plain": value / plain_value if plain_value else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "graph": not args.no_graph,
        "verify": {"logged_count_mismatches_batched": len(bad_b),
                   "graph_replay_equals_eager": graph_same,
                   "logged_count_mismatches_plain": None if args.no_plain else len(bad_p),
                   "batched_equals_plain": None if args.no_plain else bool(same),
                   "pairs": sum(len(r.pairs_d) for r in res_b)},
    }
    if cpu and cpu["value"]:
        line["vs_cpu"] = value / cpu["value"]
    # the whole path against the materialized join's cost: a 16-byte (key, id) record written once and
    # read once per key, 32 bytes; this path moves fewer (8-byte I records, D records in LDS-sorted
    # windows), so the figure is what the reference's join shape would need to reach this rate
    line["path_roofline"] = {"bytes_per_key": 32, "achieved": round(value * 32 / 1e9, 1), "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": round(value * 32 / 1e9 / PEAK_HBM_GBS, 4)}
    print(json.dumps(line), flush=True)

    if args.copy_probe:
        copy_probe(ctx, args.copy_probe)


if __name__ == "__main__":
    main()
