"""Benchmark of the Davidson sigma step (H.C) on one MI355X.

Workload (default --config syn36d2h): the 36-orbital / 36-electron synthetic
Hamiltonian with D2h irreps round-robin (integrals breaking the symmetry set
to zero), S = 0, at bond dimension M = 1000.  The run first grows the blocks
by a noisy two-site sweep ramp (M 250 -> 500 -> 1000, density-matrix noise
1e-4 -> 1e-5, FP64 sigma, composite operators built in HBM), checkpointing the
sweep state after every sweep.  Then:

* step: one H.C (multiply_h, davidson.C:17-22) of the FP64 GPU path on the
  mid-chain two-site superblock rebuilt from the ramp's final blocks; operator
  blocks and wavefunction resident in HBM; the timed region is exactly --steps
  applications bracketed by stream synchronisation, HIP events on the sigma
  stream.  `value` is its FP64-equivalent algorithmic TFLOP/s.
* mixed_split_bf16: the same superblock in split precision (three BF16
  slices, FP32 MFMA steps, FP64 sums), its speed-up and errors.
* batched_roots: four vectors in one launch set against four serial sigmas.
* sweep: one full two-site sweep (forward and backward) at M from the ramp's
  checkpoint, once with FP64 sigmas and once with the precision ladder
  (split-precision sigmas until the Davidson residual norm is below 1e-5,
  FP64 after), sweep times, phases and |dE|.
* cpu_baseline: the oracle's restatement of multiplyH with the reference's
  OpenMP scheme on all host cores, on a bounded sample of the same H.C.

Prints one JSON line; with --gpus > 1 it exits with a message (the path runs
on one GPU).  --summarize-trace / --summarize-traffic / --summarize-sweep turn
rocprofv3 CSVs of such a run into the tables kept with the profiles.
"""
import argparse
import ctypes
import json
import os
import re
import resource
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import block_amd  # noqa: E402
from block_amd import integrals as gen  # noqa: E402

METRIC = "H·C sigma TFLOP/s + sweep time, mixed vs FP64 GPU at |ΔE|≤1e-8 Eh"
PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
COLS = block_amd.SWEEP_COLS


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sb, seconds_budget):
    """The oracle's restatement of multiplyH with the reference's parallel
    scheme (OpenMP over operator pairs, per-thread sigma copies summed at the
    end, operatorloops.h:87-97, distribute.h:143-147) and the host's blocked
    DGEMM for the sector products, on every host core this process may use:
    full H.C of the same superblock, repeated within the time budget."""
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    lib = ctypes.CDLL(path)
    f = lib.orc_sigma_fast_timed
    dp = ctypes.POINTER(ctypes.c_double)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, dp]
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    c = np.random.default_rng(1).standard_normal(sb.wave_size)
    v = np.zeros(sb.wave_size)
    total, n = 0.0, 0
    while n == 0 or total < seconds_budget:
        t = ctypes.c_double()
        if f(sb._h, c.ctypes.data_as(dp), v.ctypes.data_as(dp), threads, ctypes.byref(t)) != 0:
            raise RuntimeError("oracle sigma failed")
        total += t.value
        n += 1
        if total > seconds_budget / 2:
            break
    return {"value": sb.flops * n / total / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{n} full H.C of the benchmark superblock ({sb.flops:.3e} flop each), {total:.1f} s"}


def _rows(sw, first, last):
    return sw[(sw[:, COLS.index("sweep")] >= first) & (sw[:, COLS.index("sweep")] <= last)]


def ramp_leg(ints, cfg, M, ck, verbose=1):
    """FP64 sweeps growing the blocks to M with density-matrix noise, the
    sweep state saved after every sweep (the timed legs start from it)."""
    sched = [tuple(r) for r in cfg["ramp"]]
    n = cfg["ramp_sweeps"]
    t0 = time.time()
    e, sw = block_amd.dmrg(ints, nelec=cfg["nelec"], schedule=sched, max_sweeps=n, sweep_tol=0.0, checkpoint=ck,
                           verbose=verbose)
    fwd, bwd, fsec, bsec = block_amd.state_block_dims(ints, ck)
    k = ints.norb
    return {"schedule": [list(r) for r in sched], "sweeps": n, "seconds": round(time.time() - t0, 1),
            "energy": float(e[0]),
            "sweep_seconds": [round(float(x), 2) for x in sw[:, COLS.index("seconds")]],
            "kept_states_fwd": fwd, "kept_states_bwd": bwd, "sectors_fwd": fsec, "sectors_bwd": bsec,
            "kept_states_mid": fwd[k // 2 - 2], "max_kept": max(fwd + bwd)}


def sweep_leg(ints, cfg, M, ck, tmp, verbose=1):
    """One full two-site sweep at M from the ramp's saved state, in FP64 and
    with the precision ladder (its split phase on the superblocks with at
    least half of their flops on BF16 slices); every step starts from the
    previous step's transformed wavefunction.  The FP64 leg saves its final state, whose
    blocks give the kept states of the timed sweep."""
    n = cfg["ramp_sweeps"]
    sched = [tuple(r) for r in cfg["ramp"]] + [(n, M, cfg["timed_tol"], 0.0)]
    out = {"from_sweep": n, "davidson_tol": cfg["timed_tol"], "timed_sweeps": [n]}
    after = os.path.join(tmp, "timed.state")
    for name, ladder in (("fp64", 0.0), ("mixed", 1e-5)):
        t0 = time.time()
        e, sw = block_amd.dmrg(ints, nelec=cfg["nelec"], schedule=sched, max_sweeps=n + 1, sweep_tol=0.0,
                               ladder=ladder, ladder_from=n, resume=ck, verbose=verbose,
                               checkpoint=after if name == "fp64" else None)
        wall = time.time() - t0
        rows = _rows(sw, n, n)
        col = lambda c: float(rows[:, COLS.index(c)].sum())
        out[name] = {"sweep_seconds": round(col("seconds"), 3), "sigma_seconds": round(col("sigma_seconds"), 3),
                     "sigma_calls": int(col("sigma_calls")), "split_calls": int(col("split_calls")),
                     "build_seconds": round(col("build_seconds"), 3),
                     "davidson_other_seconds": round(col("davidson_seconds"), 3),
                     "renorm_seconds": round(col("renorm_seconds"), 3),
                     "sigma_tflops": round(col("sigma_flops") / max(col("sigma_seconds"), 1e-9) / 1e12, 3),
                     "max_discarded": float(rows[:, COLS.index("discarded")].max()),
                     "energy": float(e[0]), "run_seconds": round(wall, 2)}
        log(f"sweep {name}: {out[name]['sweep_seconds']} s, E = {out[name]['energy']:.12f}")
    fwd, bwd, fsec, bsec = block_amd.state_block_dims(ints, after)
    out["kept_states_after"] = {"fwd": fwd, "bwd": bwd, "fwd_sectors": fsec, "bwd_sectors": bsec}
    out["sweep_speedup_mixed_vs_fp64"] = round(out["fp64"]["sweep_seconds"] / out["mixed"]["sweep_seconds"], 3)
    out["sigma_speedup_mixed_vs_fp64"] = round(out["fp64"]["sigma_seconds"] / max(out["mixed"]["sigma_seconds"], 1e-9), 3)
    out["abs_dE_Eh"] = abs(out["fp64"]["energy"] - out["mixed"]["energy"])
    return out


def batched_leg(sb, steps, warmup, nvec=4):
    """nvec vectors in one launch set (every tile once per root, a tile's
    roots on one XCD) against nvec serial sigmas, FP64 and split."""
    h = block_amd.multiply_h(sb, max_vec=nvec)
    c = np.stack([np.random.default_rng(10 + i).standard_normal(sb.wave_size) for i in range(nvec)])
    out = {"nvec": nvec}
    for name, prec in (("fp64", block_amd.FP64), ("split", block_amd.SPLIT_BF16)):
        h.precision = prec
        res = {}
        for label, nb in (("serial", 1), ("batched", nvec)):
            h.set_batch(nb)
            wall, _, disp = h.timed(c, warmup, steps)
            res[label] = {"ms_per_step": round(wall / steps, 4), "dispatches_per_step": disp / steps,
                          "tflops": round(h.flops * nvec * steps / (wall * 1e-3) / 1e12, 3)}
        res["speedup_vs_serial"] = round(res["serial"]["ms_per_step"] / res["batched"]["ms_per_step"], 3)
        out[name] = res
    h.precision = block_amd.FP64
    return out


def traffic_from_profile(config, M, wave_size, pairs):
    """HBM bytes per H.C of the fused kernels, from the committed counter
    summary of this workload (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    passes, each scaled by the same counter on a plain copy of known size).
    None when no summary for this exact superblock is committed."""
    path = os.path.join(ROOT, "profiles", f"k_sigma_fp64_{config}_M{M}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic:"):
            kv = dict(t.split("=") for t in line.split()[1:])
            if int(kv["wave_size"]) == wave_size and int(kv["gemm_pairs"]) == pairs:
                return int(kv["bytes_per_step"])
    return None


def copy_calibration(n, reps):
    """A plain device copy of n doubles (bsg_copy), `reps` times: the known
    traffic (8n bytes read, 8n written per launch) that scales the HBM
    counters of a profile run."""
    lib = block_amd.sigma_lib
    vp = ctypes.c_void_p
    lib.bsg_init.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    lib.bsg_malloc.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.bsg_copy.argtypes = [vp, vp, vp, ctypes.c_int64]
    lib.bsg_synchronize.argtypes = [vp]
    lib.bsg_mfree.argtypes = [vp, vp]
    lib.bsg_free.argtypes = [vp]
    ctx, x, y = vp(), vp(), vp()
    assert lib.bsg_init(ctypes.byref(ctx), 0) == 0
    assert lib.bsg_malloc(ctx, 8 * n, ctypes.byref(x)) == 0
    assert lib.bsg_malloc(ctx, 8 * n, ctypes.byref(y)) == 0
    for _ in range(reps):
        assert lib.bsg_copy(ctx, x, y, n) == 0
    assert lib.bsg_synchronize(ctx) == 0
    lib.bsg_mfree(ctx, x)
    lib.bsg_mfree(ctx, y)
    lib.bsg_free(ctx)
    print(json.dumps({"copy_calibration": {"doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n,
                                           "launches": reps}}))
    return 0


def _short(name):
    """k_sigma<1, 4, 2> from the demangled signature (or from a mangled name
    the profiler left as it was, e.g. a kernel with __bf16 arguments)."""
    m = re.match(r"_ZN\w*?\d+(k_[a-z_]+)I((?:Li-?\d+E)+)E", name)
    if m:
        return f"{m.group(1)}<{', '.join(re.findall(r'Li(-?[0-9]+)E', m.group(2)))}>"
    name = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return name.split("(")[0]


def _csv_rows(d, suffix):
    import csv
    import glob
    files = glob.glob(os.path.join(d, f"*{suffix}"))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    return list(csv.DictReader(open(files[0])))


def _last_by_kernel(rows, last, pattern):
    by = {}
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        k = _short(r["Kernel_Name"])
        if re.search(pattern, k):
            by.setdefault(k, []).append(r)
    return {k: v[-last:] for k, v in by.items()}


def summarize_trace(d, last, pattern):
    """Markdown table of the last `last` dispatches of every kernel matching
    `pattern` in a rocprofv3 --kernel-trace CSV (the bench's own calls are the
    last ones of a run: the superblock build before them also runs these
    kernels on smaller superblocks)."""
    print("| kernel | dispatches | average ms | min ms | max ms | grid | workgroup | VGPR | AGPR | SGPR | LDS B | scratch B |")
    print("|---|---|---|---|---|---|---|---|---|---|---|---|")
    by = _last_by_kernel(_csv_rows(d, "kernel_trace.csv"), last, pattern)
    for k, rs in by.items():
        t = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in rs]
        r = rs[-1]
        print(f"| {k} | {len(t)} | {sum(t) / len(t):.4f} | {min(t):.4f} | {max(t):.4f} | {r['Grid_Size_X']} | "
              f"{r['Workgroup_Size_X']} | {r['VGPR_Count']} | {r['Accum_VGPR_Count']} | {r['SGPR_Count']} | "
              f"{r['LDS_Block_Size']} | {r['Scratch_Size']} |")
    # the fused kernels of one H.C run concurrently on two streams: the span
    # from the first start to the last end of the i-th last dispatch of each
    # (the small-block kernel runs in both precision modes: counted only in a
    # trace without split-precision calls)
    any_split = any("k_sigma_bf" in k for k in by)
    fp64_names = ("k_sigma<", "k_wave<") + (() if any_split else ("k_small<",))
    for label, names in (("FP64", fp64_names), ("split", ("k_sigma_bf",))):
        groups = [rs for k, rs in by.items() if any(n in k for n in names)]
        if not groups or min(len(g) for g in groups) == 0:
            continue
        n = min(len(g) for g in groups)
        spans = [(max(int(g[-1 - i]["End_Timestamp"]) for g in groups) -
                  min(int(g[-1 - i]["Start_Timestamp"]) for g in groups)) / 1e6 for i in range(n)]
        print(f"\n{label} fused kernels, first start to last end per call: average {sum(spans) / n:.4f} ms over {n} calls")


def summarize_counters(d, last, pattern):
    """Mean of every counter over the last `last` dispatches per kernel
    (rocprofv3 --pmc CSV)."""
    rows = _csv_rows(d, "counter_collection.csv")
    out = {}
    by = {}
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        k = _short(r["Kernel_Name"])
        if re.search(pattern, k):
            by.setdefault((k, r["Counter_Name"]), []).append(float(r["Counter_Value"]))
    for (k, cn), v in by.items():
        out.setdefault(k, {})[cn] = sum(v[-last:]) / len(v[-last:])
    return out


def summarize_traffic(dirs, last, pattern, copy_doubles):
    """HBM bytes per dispatch from a FETCH_SIZE and a WRITE_SIZE pass (KiB per
    dispatch), each scaled by the same counter on the copy calibration (a
    copy of copy_doubles doubles: 8 copy_doubles bytes read and written)."""
    fetch, write, cfetch, cwrite = (summarize_counters(d, last if i < 2 else 5, pattern if i < 2 else "k_copy")
                                    for i, d in enumerate(dirs))
    kib = 8 * copy_doubles / 1024
    ff = kib / fetch_val(cfetch, "FETCH_SIZE")
    fw = kib / fetch_val(cwrite, "WRITE_SIZE")
    print(f"copy calibration: {copy_doubles} doubles = {kib:.0f} KiB each way; FETCH_SIZE reads "
          f"{fetch_val(cfetch, 'FETCH_SIZE'):.0f} KiB (factor {ff:.3f}), WRITE_SIZE reads "
          f"{fetch_val(cwrite, 'WRITE_SIZE'):.0f} KiB (factor {fw:.3f})")
    print()
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read bytes (scaled) | written bytes (scaled) |")
    print("|---|---|---|---|---|")
    tot = 0.0
    for k in sorted(fetch):
        fr, wr = fetch[k]["FETCH_SIZE"], write[k]["WRITE_SIZE"]
        tot += (fr * ff + wr * fw) * 1024
        print(f"| {k} | {fr:.0f} | {wr:.0f} | {fr * ff * 1024:.4g} | {wr * fw * 1024:.4g} |")
    print()
    print(f"total per step: {tot:.4g} bytes")


def summarize_sweep(d):
    """GPU time of a whole profiled run (rocprofv3 --kernel-trace
    --memory-copy-trace --stats CSVs) by kernel and by copy direction,
    against the wall span of its kernel trace."""
    fam = {}
    for r in _csv_rows(d, "kernel_stats.csv"):
        k = _short(r["Name"])
        t, n = fam.get(k, (0, 0))
        fam[k] = (t + int(r["TotalDurationNs"]), n + int(r["Calls"]))
    trace = _csv_rows(d, "kernel_trace.csv")
    span = (max(int(r["End_Timestamp"]) for r in trace) - min(int(r["Start_Timestamp"]) for r in trace)) / 1e9
    ktot = sum(t for t, _ in fam.values()) / 1e9
    print(f"wall span of the kernel trace: {span:.2f} s; kernels busy {ktot:.2f} s (summed over streams)")
    print()
    print("| kernel | calls | total s | share of kernel time |")
    print("|---|---|---|---|")
    for k, (t, n) in sorted(fam.items(), key=lambda x: -x[1][0]):
        if t / 1e9 >= 0.01:
            print(f"| {k} | {n} | {t / 1e9:.3f} | {t / 1e9 / ktot:.1%} |")
    print()
    print("| copies | calls | total s |")
    print("|---|---|---|")
    for r in _csv_rows(d, "memory_copy_stats.csv"):
        print(f"| {r['Name']} | {r['Calls']} | {int(r['TotalDurationNs']) / 1e9:.3f} |")


def fetch_val(d, name):
    return next(iter(d.values()))[name]



# V block dimension ranges of the tile tuner's buckets (blocks up to 32 x 32
# run in the small-block kernel, which has no tile choice)
TUNE_EDGES = [(1, 32), (33, 64), (65, 128), (129, 256), (257, 512), (513, 1 << 30)]


def tuner_start(buckets, start):
    """The tuner's first choices: (fp64, split) per bucket (row, column edge
    indices into TUNE_EDGES) from the start rules whose ranges are exactly a
    bucket's, -1 elsewhere; the start rules of other ranges are kept as
    they are."""
    choice = {b: [-1, -1] for b in buckets}
    edges = {(TUNE_EDGES[bm][0], TUNE_EDGES[bm][1], TUNE_EDGES[bn][0], TUNE_EDGES[bn][1]): (bm, bn) for bm, bn in buckets}
    keep = []
    for r in start:
        if tuple(r[:4]) in edges:
            choice[edges[tuple(r[:4])]] = [r[4], r[5]]
        else:
            keep.append(tuple(r))
    return choice, keep


def tuner_rules(choice, keep):
    """Tile table rules of the tuner's current choices (buckets on the model
    in both precisions give no rule), then the kept rules."""
    return [(TUNE_EDGES[bm][0], TUNE_EDGES[bm][1], TUNE_EDGES[bn][0], TUNE_EDGES[bn][1], f, sp)
            for (bm, bn), (f, sp) in choice.items() if f >= 0 or sp >= 0] + list(keep)


def tune_tiles(sbs, out_path, steps, warmup, label, top=8, min_gain=0.005, start=(), passes=1):
    """Measured tile table over one or more superblocks: the V blocks are
    bucketed by (rows, cols) ranges; for the buckets with most of the flops
    (about m n (m + n) per block, each superblock's share normalised), every
    candidate geometry of the FP64 and of the split-precision kernels is
    timed on the whole sigma of every superblock (the other buckets on their
    current choice), scored by the sum over superblocks of the time relative
    to the MFMA-count model's, and a bucket keeps the best candidate when it
    beats the model by min_gain.  Writes the rules to out_path and checks
    that the tuned plans give the same sigma.  start: rules to begin from
    (a bucket's rule is its first choice, rules of other ranges are kept);
    passes > 1 repeats the sweep over the buckets (coordinate descent), each
    bucket decided again with the others on their latest choice."""
    weight = {}
    for sb in sbs:
        w = {}
        for m, n in sb.block_shapes():
            if max(m, n) <= 32:
                continue
            bm = next(i for i, (lo, hi) in enumerate(TUNE_EDGES) if lo <= m <= hi)
            bn = next(i for i, (lo, hi) in enumerate(TUNE_EDGES) if lo <= n <= hi)
            w[(bm, bn)] = w.get((bm, bn), 0.0) + float(m) * n * (m + n)
        tot = sum(w.values()) or 1.0
        for k, x in w.items():
            weight[k] = weight.get(k, 0.0) + x / tot / len(sbs)
    buckets = sorted(weight, key=lambda b: -weight[b])[:top]
    hs = [block_amd.multiply_h(sb) for sb in sbs]
    cs = [np.random.default_rng(1).standard_normal(sb.wave_size) for sb in sbs]
    block_amd.set_tile_table([])
    for h in hs:
        h.replan()
    v0 = [h(c) for h, c in zip(hs, cs)]
    choice, keep = tuner_start(buckets, start)

    def rules():
        return tuner_rules(choice, keep)

    def times():
        block_amd.set_tile_table(rules())
        out = []
        for h, c in zip(hs, cs):
            h.replan()
            out.append(min(h.timed(c, warmup, steps)[0] for _ in range(3)) / steps)
        return out

    log_rows = []
    for slot, prec in [x for _ in range(passes) for x in ((0, block_amd.FP64), (1, block_amd.SPLIT_BF16))]:
        for h in hs:
            h.precision = prec
        for b in buckets:
            choice[b][slot] = -1
            base = times()
            score = lambda t: sum(x / y for x, y in zip(t, base)) / len(base)
            best, best_s, best_t = -1, 1.0, base
            for cand in range(block_amd.tile_candidates(prec)):
                choice[b][slot] = cand
                t = times()
                if score(t) < best_s:
                    best, best_s, best_t = cand, score(t), t
            choice[b][slot] = best if best_s < 1.0 - min_gain else -1
            log_rows.append((slot, b, round(weight[b], 3), [round(x, 4) for x in base], best,
                             [round(x, 4) for x in best_t]))
            log(f"tune {'fp64' if slot == 0 else 'split'} bucket m {TUNE_EDGES[b[0]]} n {TUNE_EDGES[b[1]]} "
                f"({weight[b]:.1%} of flops): model {[round(x, 4) for x in base]} ms, best candidate {best} "
                f"{[round(x, 4) for x in best_t]} ms")
    block_amd.set_tile_table(rules())
    err = 0.0
    for h, c, v in zip(hs, cs, v0):
        h.precision = block_amd.FP64
        h.replan()
        err = max(err, float(np.linalg.norm(h(c) - v) / np.linalg.norm(v)))
    if err > 1e-12:
        raise RuntimeError(f"tuned plan changed sigma: relative difference {err:.2e}")
    with open(out_path, "w") as f:
        f.write("# m_lo m_hi n_lo n_hi fp64 split: tile geometry per V block shape range (-1: MFMA-count model)\n")
        f.write(f"# measured by bench.py --tune on {label} ({', '.join(str(sb.wave_size) for sb in sbs)}-element\n"
                f"# superblocks): every superblock's whole sigma timed per candidate (min of 3 x {steps}\n"
                f"# applications), kept if the mean time relative to the model's is >{min_gain:.1%} lower;\n"
                f"# fp64 candidate {block_amd.tile_candidates(block_amd.FP64) - 1} is the per-wave tile kernel\n")
        for r in rules():
            f.write(" ".join(str(x) for x in r) + "\n")
    return {"rules": rules(), "trials": log_rows, "sigma_rel_diff_tuned_vs_model": err}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="syn36d2h")
    ap.add_argument("--M", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision timing")
    ap.add_argument("--no-sweep", action="store_true", help="skip the timed sweeps")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-roots leg")
    ap.add_argument("--state", metavar="FILE", help="start from this saved sweep state instead of the ramp")
    ap.add_argument("--from-warmup", action="store_true",
                    help="superblock of the infinite-system warm-up at M (no ramp, no saved state, no timed "
                         "sweep): the configurations whose sweep states do not fit on disk")
    ap.add_argument("--keep-state", metavar="FILE", help="copy the ramp's saved state here")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run a device copy of N doubles (HBM counter calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR", help="table of a rocprofv3 kernel trace (no GPU run)")
    ap.add_argument("--summarize-sweep", metavar="DIR",
                    help="GPU time by kernel and copy direction of a profiled run (no GPU run)")
    ap.add_argument("--summarize-traffic", nargs=4, metavar=("FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="HBM traffic from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (no GPU run)")
    ap.add_argument("--settle", type=int, default=None,
                    help="noise-free sweeps at M after the ramp, before the timed sweep (default: the config's)")
    ap.add_argument("--tune", metavar="OUT", help="measure the tile table on this superblock and write it to OUT")
    ap.add_argument("--tune-passes", type=int, default=1, help="passes of the tuner over the buckets")
    ap.add_argument("--tune-from", default="", metavar="TABLE",
                    help="tile table the tuner starts from (its bucket rules are first choices, the rest kept)")
    ap.add_argument("--tune-with", default="", metavar="CONFIGS",
                    help="comma-separated further configurations whose superblocks the tuner also times")
    ap.add_argument("--verbose", type=int, default=1, help="sweep log level (2: per step)")
    ap.add_argument("--last", type=int, default=6, help="dispatches per kernel to summarise (the last ones)")
    ap.add_argument("--kernels", default="k_sigma|k_wave|k_small|k_reduce", help="regex of the kernel names to summarise")
    args = ap.parse_args()
    if args.summarize_trace:
        return summarize_trace(args.summarize_trace, args.last, args.kernels) or 0
    if args.summarize_sweep:
        return summarize_sweep(args.summarize_sweep) or 0
    if args.summarize_traffic:
        return summarize_traffic(args.summarize_traffic, args.last, args.kernels, args.copy_calibration) or 0
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 5)
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    cfg = gen.CONFIGS[args.config]
    M = args.M or cfg["M"]
    if args.M:
        cfg = dict(cfg, ramp=[(s, min(m, M), t, n) for (s, m, t, n) in cfg["ramp"]])
    settle = cfg.get("settle", 0) if args.settle is None else args.settle
    if settle > 0:
        # noise-free sweeps at M that converge the state before the timed sweep
        n0 = cfg["ramp_sweeps"]
        cfg = dict(cfg, ramp=[tuple(r) for r in cfg["ramp"]] + [(n0, M, cfg.get("settle_tol", 1e-10), 0.0)],
                   ramp_sweeps=n0 + settle)
    tmp = tempfile.mkdtemp(prefix="blkbench_")
    try:
        return run_bench(args, cfg, M, settle, tmp)
    finally:
        # sweep states are GBs at M >= 800: never leave them behind
        shutil.rmtree(tmp, ignore_errors=True)


def run_bench(args, cfg, M, settle, tmp):
    norb = cfg["norb"]
    fcidump = os.path.join(tmp, f"{args.config}.FCIDUMP")
    gen.make_config_fcidump(args.config, fcidump)
    ints = block_amd.Integrals(fcidump)
    ck = args.state or os.path.join(tmp, "ramp.state")
    ramp = None
    if args.from_warmup:
        args.no_sweep = True
        log(f"infinite-system warm-up to M={M} on {args.config}")
        t0 = time.time()
        # per-step lines (verbose 2) keep a long warm-up visibly alive
        sb = block_amd.SuperBlock.from_dmrg_warmup(ints, cfg["nelec"], 0, ints.isym, M, verbose=max(2, args.verbose))
        log(f"warm-up done in {time.time() - t0:.1f} s")
    elif not args.state:
        log(f"ramp to M={M} on {args.config}")
        ramp = ramp_leg(ints, cfg, M, ck, args.verbose)
        log(f"ramp done in {ramp['seconds']} s, kept states at mid-chain {ramp['kept_states_mid']}")
        if args.keep_state:
            shutil.copyfile(ck, args.keep_state)
    t0 = time.time()
    if not args.from_warmup:
        sb = block_amd.SuperBlock.from_sweep_state(ints, ck, cfg["nelec"], 0, ints.isym)
    t_build = time.time() - t0
    if args.tune:
        sbs = [sb]
        for extra in [x for x in args.tune_with.split(",") if x]:
            ecfg = gen.CONFIGS[extra]
            epath = os.path.join(tmp, f"{extra}.FCIDUMP")
            gen.make_config_fcidump(extra, epath)
            eints = block_amd.Integrals(epath)
            eck = os.path.join(tmp, f"{extra}.state")
            log(f"ramp on {extra} for the tuner")
            ramp_leg(eints, ecfg, ecfg["M"], eck, args.verbose)
            sbs.append(block_amd.SuperBlock.from_sweep_state(eints, eck, ecfg["nelec"], 0, eints.isym))
        label = f"--config {args.config}" + (f" --M {args.M}" if args.M else "") + \
            (f" --tune-with {args.tune_with}" if args.tune_with else "") + \
            (f" --tune-from {args.tune_from}" if args.tune_from else "")
        start = block_amd.read_tile_table(args.tune_from) if args.tune_from else ()
        if args.tune_passes > 1:
            label += f" --tune-passes {args.tune_passes}"
        print(json.dumps({"tune": tune_tiles(sbs, args.tune, args.steps, args.warmup, label, start=start,
                                             passes=args.tune_passes)}))
        return 0
    h = block_amd.multiply_h(sb, device=0)
    c = np.random.default_rng(0).standard_normal(sb.wave_size)
    c /= np.linalg.norm(c)
    log(f"superblock: wave {sb.wave_size}, {h.pairs} GEMM pairs, {h.flops:.3e} flop per H.C")
    wall_ms, kern_ms, disp = h.timed(c, args.warmup, args.steps)
    flops = h.flops
    f1, f2 = h.phase_flops()
    ms_per_step = wall_ms / args.steps
    value = flops * args.steps / (wall_ms * 1e-3) / 1e12
    # dominant kernels: the fused GEMM-pair kernels of one H.C (grouped and
    # per-wave tiles on the sigma stream, small blocks and basic tiles beside
    # them), timed
    # fork to join with HIP events on the sigma stream
    achieved = (f1 + f2) * args.steps / (kern_ms[0] * 1e-3) / 1e12
    traffic = traffic_from_profile(args.config, M, sb.wave_size, h.pairs)
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_MFMA, 5), "traffic": traffic, "kernel": "k_sigma+k_wave+k_small",
                "kernel_ms_per_step": round(kern_ms[0] / args.steps, 4),
                "dispatches_per_step": disp / args.steps,
                "kernel_ms_per_step_all": {"fused": round(kern_ms[0] / args.steps, 4),
                                           "k_reduce": round(kern_ms[1] / args.steps, 4),
                                           "side_stream": round(kern_ms[2] / args.steps, 4)}}
    # the same applications as one captured HIP graph launch each
    h.set_graph(True)
    g_wall, _, _ = h.timed(c, args.warmup, args.steps)
    h.set_graph(False)
    graph = {"ms_per_step": round(g_wall / args.steps, 4), "speedup_vs_launches": round(wall_ms / g_wall, 3)}
    mixed = None
    if not args.no_split:
        v64 = h(c)
        # errors of the split sigma at the FP64 ground state of this
        # superblock (device Davidson, FP64): Ritz energy and residual norm
        d = sb.diagonal_cpu()
        g0 = np.zeros((1, sb.wave_size))
        g0[0, int(np.argmin(d))] = 1.0
        e_ref, x_ref, _, _ = block_amd.davidson(sb, g0, tol=1e-14)
        x = x_ref[0] / np.linalg.norm(x_ref[0])
        hx64 = h(x)
        h.precision = block_amd.SPLIT_BF16
        vs = h(c)
        hxs = h(x)
        s_wall, s_kern, s_disp = h.timed(c, args.warmup, args.steps)
        h.precision = block_amd.FP64
        mixed = {"split_share": round(h.split_share, 4),
                 "tflops": round(flops * args.steps / (s_wall * 1e-3) / 1e12, 4),
                 "ms_per_step": round(s_wall / args.steps, 4),
                 "kernel_ms_per_step": round(s_kern[0] / args.steps, 4),
                 "dispatches_per_step": s_disp / args.steps,
                 "speedup_vs_fp64": round(wall_ms / s_wall, 3),
                 "sigma_rel_err_vs_fp64": float(np.linalg.norm(vs - v64) / np.linalg.norm(v64)),
                 "rayleigh_err_vs_fp64_Eh": float(abs(c @ vs - c @ v64)),
                 "at_fp64_ground_state": {
                     "energy_Eh": float(e_ref[0]),
                     "ritz_energy_err_Eh": float(abs(x @ hxs - x @ hx64)),
                     "residual_norm_fp64": float(np.linalg.norm(hx64 - (x @ hx64) * x)),
                     "residual_norm_split": float(np.linalg.norm(hxs - (x @ hxs) * x))}}
    pairs, arena_gb = int(h.pairs), round(h.arena_bytes / 1e9, 3)
    del h
    batched = None if args.no_batch else batched_leg(sb, max(2, args.steps // 2), args.warmup)
    nl, nr, nb = sb.sectors()
    shapes = sb.block_shapes()
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: one H.C on the mid-chain two-site superblock, {norb} orbitals / "
                               f"{cfg['nelec']} electrons, S=0, {'D2h' if cfg.get('d2h') else 'C1'}, M={M}, blocks "
                               + ("from the infinite-system warm-up" if args.from_warmup else
                                  f"from a noisy two-site sweep ramp{f' and {settle} noise-free sweeps' if settle else ''}")
                               + ", FP64 sigma",
                   "wave_size": sb.wave_size, "flop_per_step": flops, "terms": sb.nterms,
                   "sectors_L_R_blocks": [nl, nr, nb], "largest_V_block": max(shapes, key=lambda s: s[0] * s[1]),
                   "gemm_pairs": pairs, "operator_arena_GB": arena_gb,
                   "superblock_build_s": round(t_build, 2),
                   "host_peak_rss_GB": round(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6, 2)},
        "roofline": roofline,
        "graph": graph,
    }
    if ramp:
        out["config"]["ramp"] = ramp
    if mixed:
        out["mixed_split_bf16"] = mixed
    if batched:
        out["batched_roots"] = batched
    if not args.no_sweep:
        out["sweep"] = sweep_leg(ints, cfg, M, ck, tmp, args.verbose)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(sb, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
