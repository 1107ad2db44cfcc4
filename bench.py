"""Benchmark of minibwa's seeding + extension hot path on one MI355X, as the end-to-end mapper.

Workload (2x150 bp pairs against a 3.1 Gbp reference): a 3.1 Gbp seeded synthetic reference
(uniform ACGT plus 8% repeats: copies of 300-6,000 bp segments at 2% divergence; 24 contigs of
129 Mbp) and 1M read pairs 2x150 bp (fragment ~ Normal(400, 50), 1% substitutions, 0.1% indels with
geometric length of mean 2) in two FASTQ files, splitmix64 seed 11.  The index is built on the device
before timing.  --ref-len 100000000 --contigs 4 --rep-frac 0.05 gives the 100 Mbp workload.

One step = one `minibwa map` pass over the 1M pairs, FASTQ in -> SAM out, by the batched GPU mapper
(minibwa_amd/bin/mbg-map-bench: the reference's own mb_map_file with the index loaded once and made
resident in HBM; seeding, anchors and chains in one device pass per 100 Mbp mini-batch, every DP and
mate-rescue call in device rounds, the rest the reference's host code on -t host threads).  The first
pass writes its SAM to a file; timed passes write to /dev/null.  value = read pairs per second of
the timed passes' wall time.

cpu_baseline: the compiled reference, `minibwa map -t<all host cores>` on the same FASTQ and index
files, timed as T_end - T_idx from its own stderr (main.c:86, map-main.c:561-562), output to
/dev/null.  A second reference run writes its SAM to a file, and the bench checks that it equals the
GPU mapper's byte for byte (minus the @PG line, which holds the command line).

--step runs the earlier diagnostic instead: one device pass of seeding, SA locate and a stand-in
extension job set over resident reads (mbg_map_step_dev), with its own serving loop.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned read pairs/s, 2x150bp vs 3.1Gbp synthetic ref (GPU vs CPU); p50/p99 lat"
METRIC_LONG = "mapped long reads/s, log-uniform 100-20000 bp vs 3.1Gbp synthetic ref, PAF output (GPU vs CPU)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--step", action="store_true",
                    help="run the device-pass diagnostic (seeding, SA locate, stand-in extension jobs) instead of the mapper")
    ap.add_argument("--threads", type=int, default=0, help="host threads of both mappers (default: all host cores)")
    ap.add_argument("--cpu-runs", type=int, default=1, help="timed runs of the reference mapper (median)")
    ap.add_argument("--ref-len", type=int, default=3_100_000_000)
    ap.add_argument("--contigs", type=int, default=24)
    ap.add_argument("--pre-len", type=int, default=15,
                    help="k-mer interval cache length on the device (the reference uses 10; results do not depend on it)")
    ap.add_argument("--rep-frac", type=float, default=0.08)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--cpu-pairs", type=int, default=1_000_000, help="read pairs in the CPU-baseline sample")
    ap.add_argument("--cpu-jobs", type=int, default=200_000, help="jobs in the --dp-batch CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--calib", type=int, default=0, metavar="MiB",
                    help="only run the plain copy kernel over MiB mebibytes (HBM byte-counter calibration)")
    ap.add_argument("--serve-qps", type=float, default=4000.0,
                    help="request rate of the serving phase (Poisson arrivals)")
    ap.add_argument("--serve-requests", type=int, default=1500, help="requests per offered rate in the serving phase (0: skip it)")
    ap.add_argument("--serve-sweep", default="250,500,1000,2000,3000",
                    help="offered request rates (requests/s) of the mapper's serving sweep")
    ap.add_argument("--gather", action="store_true",
                    help="after building the index, only measure the dependent random 64-B gather ceiling of its occ table")
    ap.add_argument("--dp-batch", type=int, default=0, metavar="N",
                    help="only run the extension-kernel batch: N independent (query, target) jobs, query 100-300 bp, "
                         "2%% substitutions, 0.5%% indels, 50 bp flanks, band 101, gaps 12/2 and 24/1; reports GCUPS")
    ap.add_argument("--ll-batch", type=int, default=0, metavar="N",
                    help="only run the mate-rescue local SW batch (ksw_ll, u8 core): N jobs of a 150 bp read against "
                         "a 550 bp window holding a 1%%-diverged copy of it in 80%% of the jobs; reports GCUPS")
    ap.add_argument("--rescue-batch", type=int, default=0, metavar="N",
                    help="only run the mate-rescue batch: N candidates of a 150 bp mate against a 550 bp window of a "
                         "100 Mbp reference, the mate a 1%%-diverged copy in 80%% of them; reports candidates/s")
    ap.add_argument("--anchor-batch", action="store_true",
                    help="only run mb_anchor and then mb_lchain_dp over the seeding output of the read pairs (default 1M "
                         "pairs of 2x150 bp on the 3.1 Gbp reference, max_occ 250, the mapper's chaining options for "
                         "150 bp); reports reads/s of each")
    ap.add_argument("--long", type=int, default=0, metavar="N",
                    help="with --anchor-batch: N single 10 kb reads (0.1%% substitutions) instead of the read pairs")
    ap.add_argument("--long-reads", type=int, default=0, metavar="N",
                    help="the mapper on N single-end long reads (log-uniform 100-20,000 bp, 0.1%% substitutions, "
                         "0.05%% indels) with PAF output (`map -f`) instead of the read pairs; reports reads/s")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner (launch shapes and batch sizes of the mapper) and write minibwa_amd/tuned.env")
    ap.add_argument("--fastq-batch", type=int, default=0, metavar="N",
                    help="only run the FASTQ reader (host): N 150 bp reads in one file, plain and gzip, in the CLI's "
                         "100M-base chunks; reports records/s beside the reference's mb_bseq_read")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Reads shard independently across GPUs with a replicated index; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    from minibwa_amd import mbg, synth

    if args.calib:
        ctx = mbg.Context(0)
        n = args.calib << 20
        a = ctx.to_device(np.zeros(n, np.uint8))
        b = ctx.to_device(np.zeros(n, np.uint8))
        for _ in range(3):
            mbg.check(ctx.L.mbg_stream_copy(ctx.h, None, mbg.C.c_void_p(b.ptr), mbg.C.c_void_p(a.ptr), n), "mbg_stream_copy")
        print(json.dumps({"calib_copy_bytes": n, "launches": 3}))
        return

    if args.fastq_batch:
        fastq_batch(args)
        return
    if args.autotune:
        autotune(args)
        return
    if args.dp_batch:
        dp_batch(args)
        return
    if args.ll_batch:
        ll_batch(args)
        return
    if args.rescue_batch:
        rescue_batch(args)
        return
    if not (args.step or args.anchor_batch or args.gather):
        mapper_bench(args)
        return

    t0 = time.time()
    ref = synth.reference(args.ref_len, seed=11, rep_frac=args.rep_frac, rep_min=300, rep_max=6000, rep_div=0.02)
    log(f"reference {len(ref)} bp in {time.time() - t0:.1f}s")
    work = args.workdir or tempfile.mkdtemp(prefix="mbg_bench_")
    os.makedirs(work, exist_ok=True)
    prefix = os.path.join(work, "ref")
    t0 = time.time()
    # the index is built by a child process (torch suffix sort on the GPU); this process only runs libmbg
    r = subprocess.run([sys.executable, "-m", "minibwa_amd.index", "--len", str(args.ref_len), "--seed", "11",
                        "--rep-frac", str(args.rep_frac), "--rep-min", "300", "--rep-max", "6000", "--rep-div", "0.02",
                        "--contigs", str(args.contigs), "--device", "cuda", "--out", prefix], cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        log("index build failed:", r.stderr[-2000:])
        sys.exit(1)
    log(r.stdout.strip(), f"(wall {time.time() - t0:.1f}s)")
    ctx = mbg.Context(0)
    t0 = time.time()
    ctx.idx_load(prefix, args.pre_len)
    info = ctx.info()
    log(f"index resident: {info['device_bytes'] / 1e9:.2f} GB (load, k-mer cache and SA grid in {time.time() - t0:.1f}s)")
    if args.gather:
        rows = []
        for chains in (262144, 1048576, 4194304):
            ctx.gather_bench(chains, 64, 256)
            ms, _ = ctx.gather_bench(chains, 64, 256)
            tx = chains * 64
            rows.append({"chains": chains, "steps": 64, "ms": round(ms, 3), "G_tx_per_s": round(tx / ms / 1e6, 2),
                         "GB_per_s_64B": round(tx * 64 / ms / 1e6, 1)})
        print(json.dumps({"gather_ceiling": rows, "occ_table_bytes": int((info["seq_len"] + 127) // 128 * 64),
                          "ref_bp": args.ref_len}), flush=True)
        ctx.close()
        return

    t0 = time.time()
    r1, r2 = synth.read_pairs(ref, args.pairs, seed=11, rlen=150, frag_mean=400, frag_sd=50, sub=0.01, indel=0.001)
    rl = 150
    reads = np.empty((2 * args.pairs, rl), np.uint8)
    reads[0::2] = r1
    reads[1::2] = r2
    n_reads = reads.shape[0]
    d_seq = ctx.to_device(reads.reshape(-1))
    d_off = ctx.to_device(np.arange(0, n_reads + 1, dtype=np.int64) * rl)
    d_len = ctx.to_device(np.full(n_reads, rl, np.int32))
    log(f"{args.pairs} pairs generated and resident in {time.time() - t0:.1f}s")
    opt = mbg.adaptive_opt(rl)
    if args.anchor_batch:
        if args.long:
            seq, _, _ = synth.single_reads(ref, args.long, seed=13, lmin=10000, lmax=10000, sub=0.001, indel=0.0005)
            reads = seq.reshape(args.long, 10000)
        anchor_batch(ctx, reads, prefix, args)
        ctx.close()
        return

    def step(stats=False):
        return ctx.map_step_dev(n_reads, d_off.ptr, d_len.ptr, d_seq.ptr, rl, opt, collect_stats=stats)

    # algorithmic work of one step (counters on; not timed)
    st_stats, _ = step(stats=True)
    work_stats = st_stats.as_dict()
    log("step work:", json.dumps({k: work_stats[k] for k in ("n_rec", "n_rows", "n_anchored", "n_fast", "n_jobs", "cells",
                                                            "occ_blocks", "kmer_lookups", "extend_steps", "lf_steps")}))
    for _ in range(args.warmup):
        step()
    per = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st, out = step()
        per.append(st.as_dict())
        log(f"step {i}: {st.ms_total:.2f} ms device")
    el = time.perf_counter() - t0  # every step synchronises its stream before returning
    value = args.steps * args.pairs / el
    # the same pass as one captured HIP graph against launch by launch, at serving batch sizes
    graph_cmp = graph_overhead(ctx, n_reads, d_off, d_len, d_seq, rl, opt)

    # roofline of the dominant kernel (seeding pass 1): algorithmic bytes = 64 B per occ block
    # fetched + 32 B per k-mer cache lookup, over the kernel's HIP-event duration
    ms_seed1 = float(np.mean([p["ms_seed1"] for p in per]))
    bytes_seed1 = 64 * work_stats["occ_blocks"][0] + 32 * work_stats["kmer_lookups"][0]
    achieved = bytes_seed1 / (ms_seed1 * 1e-3) / 1e9
    traffic = None
    # HBM bytes per launch of the same kernel from the committed FETCH_SIZE/WRITE_SIZE profile of this
    # workload (each counter scaled by what it reads for a plain copy of known size)
    tag = "cfg1" if args.ref_len == 100_000_000 else "3g" if args.ref_len == 3_100_000_000 else None
    pmc = os.path.join(ROOT, "profiles", f"step_{tag}_pmc.json") if tag else ""
    if os.path.exists(pmc):
        try:
            ks = json.load(open(pmc))["kernels"]
            traffic = [v for k, v in ks.items() if "smem_kernel<1>" in k][0]["hbm_bytes_per_launch"]
        except Exception:
            traffic = None
    phases = {k: round(float(np.mean([p[k] for p in per])), 3) for k in ("ms_seed1", "ms_seed2", "ms_rows", "ms_sa", "ms_jobs", "ms_dp", "ms_total")}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, out, per[-1], reads, prefix, args, work)
    latency = serve(ctx, ref, opt, args) if args.serve_requests > 0 else None

    res = {
        "metric": "device-pass read pairs/s (seeding, SA locate, stand-in extension jobs; diagnostic)",
        "value": round(value, 1),
        "unit": "read pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1000, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (FM-index ranks) + int8/int32 (DP)",
        "data": "synthetic (seeded splitmix64; index built on the device)",
        "config": {"workload": f"seed+locate+extend pass, {args.pairs} pairs 2x150 bp vs {args.ref_len / 1e9:g} Gbp synthetic "
                               f"reference ({args.rep_frac:.0%} repeats)",
                   "ref_bp": args.ref_len, "pairs": args.pairs, "read_len": rl, "contigs": args.contigs,
                   "min_len": 19, "max_occ": 250, "ksw_w": opt.ksw_w, "zdrop": opt.zdrop, "kmer_cache": args.pre_len},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "smem_kernel<1> (SMEM pass 1)", "algorithmic_bytes_per_launch": int(bytes_seed1),
                     "avg_launch_ms": round(ms_seed1, 4)},
        "cpu_baseline": cpu,
        "latency": latency,
        "phases_ms": phases,
        "graph_vs_direct": graph_cmp,
        "step_work": {k: work_stats[k] for k in ("n_rec", "n_rows", "n_anchored", "n_fast", "n_jobs", "cells", "lf_steps")},
    }
    print(json.dumps(res), flush=True)
    ctx.close()


def graph_overhead(ctx, n_reads, d_off, d_len, d_seq, rl, opt, sizes=(1000, 5000, 15000, 60000), reps=20):
    """Wall time per device pass (seed -> locate -> extend, synchronised) at serving batch sizes, as one
    captured graph and launch by launch: the launch overhead the graph removes."""
    rows = []
    for pairs in sizes:
        n = min(2 * pairs, n_reads)
        row = {"pairs": n // 2}
        for mode, on in (("graph", True), ("direct", False)):
            ctx.use_graphs(on)
            for _ in range(3):
                ctx.map_step_dev(n, d_off.ptr, d_len.ptr, d_seq.ptr, rl, opt)
            t0 = time.perf_counter()
            dev = 0.0
            for _ in range(reps):
                st, _ = ctx.map_step_dev(n, d_off.ptr, d_len.ptr, d_seq.ptr, rl, opt)
                dev += st.ms_total
            row[f"ms_{mode}"] = round((time.perf_counter() - t0) / reps * 1000, 4)
            row[f"dev_ms_{mode}"] = round(dev / reps, 4)
        ctx.use_graphs(False)
        row["saved_ms"] = round(row["ms_direct"] - row["ms_graph"], 4)
        rows.append(row)
        log("graph vs direct:", json.dumps(row))
    return rows


def _ref_timings(err):
    """(T_idx, T_end) from a reference-CLI stderr: the 'index loaded' and final 'Real time' stamps
    (map-main.c:561-562, main.c:86)."""
    t_idx = t_end = None
    for line in err.split("\n"):
        if line.endswith("index loaded") and line.startswith("[M::main_map::"):
            t_idx = float(line.split("::")[2].split("*")[0])
        elif line.startswith("[M::main] Real time:"):
            t_end = float(line.split()[3])
    if t_idx is None or t_end is None:
        raise RuntimeError("no timing lines in the mapper's stderr:\n" + err[-2000:])
    return t_idx, t_end


def _sam_body_md5(path):
    """md5 and line count of a SAM file without its @PG line (which holds the command line)."""
    import hashlib
    h, n = hashlib.md5(), 0
    with open(path, "rb") as f:
        for line in f:
            if line.startswith(b"@PG"):
                continue
            h.update(line)
            n += 1
    return h.hexdigest(), n


def _kv(line):
    w = line.replace("|", " ").split()
    out = {}
    for i in range(len(w) - 1):
        try:
            out[w[i]] = float(w[i + 1])
        except ValueError:
            pass
    return out


def mapper_index(args):
    """The workload's index (built on the device by a child process), reused from args.workdir when an
    earlier run built the same reference there; returns (workdir, index prefix)."""
    work = args.workdir or tempfile.mkdtemp(prefix="mbg_bench_")
    os.makedirs(work, exist_ok=True)
    prefix = os.path.join(work, "ref")
    t0 = time.time()
    params = json.dumps({"len": args.ref_len, "seed": 11, "rep_frac": args.rep_frac, "contigs": args.contigs})
    pfile = prefix + ".params.json"
    if os.path.exists(prefix + ".mbw") and os.path.exists(pfile) and open(pfile).read() == params:
        log(f"index of the same reference found in {work}")     # an earlier run with --workdir built it
    else:
        r = subprocess.run([sys.executable, "-m", "minibwa_amd.index", "--len", str(args.ref_len), "--seed", "11",
                            "--rep-frac", str(args.rep_frac), "--rep-min", "300", "--rep-max", "6000", "--rep-div", "0.02",
                            "--contigs", str(args.contigs), "--device", "cuda", "--out", prefix], cwd=ROOT, capture_output=True, text=True)
        if r.returncode != 0:
            log("index build failed:", r.stderr[-2000:])
            sys.exit(1)
        open(pfile, "w").write(params)
        log(r.stdout.strip(), f"(wall {time.time() - t0:.1f}s)")
    return work, prefix


# The autotuner's knobs: (environment setting, the built-in default, candidate values, workload,
# objective).  Objectives: "wall" the pass time, "seed" the device time of seeding + anchors (SA walks),
# "dp" that of the DP kernels.  Every knob changes launch shapes or batch sizes only, never results.
TUNE_KNOBS = [
    ("MBG_GATHER_BPC", "8", ["4", "6", "8", "12", "16"], "pairs", "seed"),   # 256-lane seeding/SA blocks per CU
    ("MBG_DP_WPC", "32", ["8", "16", "32", "64"], "pairs", "dp"),           # DP waves per CU
    ("MBG_FLUSH", "4096", ["2048", "4096", "8192"], "pairs", "wall"),      # pending DP requests that launch a batch
    ("MBG_PRE_LEN", "15", ["13", "14", "15"], "pairs", "seed"),             # k of the k-mer interval cache
    ("MBG_SEED_SEG", "512", ["0", "256", "512", "1024"], "long", "seed"),   # long-read segment length (0: one lane per read)
]
TUNE_MARGIN = 0.03        # a value replaces the default only when it beats it by this fraction


def autotune(args):
    """A coordinate search over TUNE_KNOBS on the mapper itself: for each knob in turn, every candidate
    value runs the mapper (3 passes of the workload, index loaded once; the objective is the mean of
    the last two passes) with the values chosen so far for the knobs before it, twice, keeping the
    better run; the value with the lowest objective is kept if it beats the default by TUNE_MARGIN
    (run-to-run noise is a few percent).  The choices go to minibwa_amd/tuned.env, which the mapper
    programs read at start-up (a setting in the environment wins), and every trial to the JSON line
    on stdout."""
    from minibwa_amd import synth
    exe = os.path.join(ROOT, "minibwa_amd", "bin", "mbg-map-bench")
    if not os.path.exists(exe):
        log(f"{exe} is not built (make -C minibwa_amd/integration)")
        sys.exit(1)
    cores = args.threads or host_cores()
    ref = synth.reference(args.ref_len, seed=11, rep_frac=args.rep_frac, rep_min=300, rep_max=6000, rep_div=0.02)
    work, prefix = mapper_index(args)
    r1, r2 = synth.read_pairs(ref, args.pairs, seed=11, rlen=150, frag_mean=400, frag_sd=50, sub=0.01, indel=0.001)
    f1, f2 = os.path.join(work, "r1.fq"), os.path.join(work, "r2.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="p_")
    del r1, r2
    n_long = 20000
    seq, off, ln = synth.single_reads(ref, n_long, seed=13, lmin=100, lmax=20000, sub=0.001, indel=0.0005)
    fl = os.path.join(work, "long_tune.fq")
    synth.write_fastq(fl, seq, off, ln, prefix="l_")
    del seq, off, ln, ref
    inputs = {"pairs": ([f1, f2], []), "long": ([fl], ["-f"])}

    def trial(settings, load):
        env = dict(os.environ, MBG_MAP_STATS="1", MBG_TUNE="none", **settings)
        fins, fopt = inputs[load]
        r = subprocess.run([exe] + fopt + ["-t", str(cores), "-P", "3", "-O", "/dev/null", prefix] + fins, capture_output=True,
                           text=True, env=env, timeout=900)
        if r.returncode != 0:
            log("trial failed:", settings, r.stderr[-1500:])
            return None
        pass_s = [float(l.split()[3]) for l in r.stderr.split("\n") if l.startswith("[mbg_bench] pass ")]
        summ = _kv([l for l in r.stderr.split("\n") if l.startswith("[mbg_map] summary:")][-1])
        seed_ms = (summ.get("dev_seed1_ms", 0) + summ.get("dev_seed2_ms", 0) + summ.get("dev_anchor_ms", 0)) / 3
        return {"wall": (pass_s[1] + pass_s[2]) / 2, "seed": seed_ms, "dp": summ.get("dev_dp_ms", 0) / 3}

    best, trials = {}, []
    for key, default, values, load, obj in TUNE_KNOBS:
        scores = {}
        for v in values:
            runs = [m for m in (trial(dict(best, **{key: v}), load) for _ in range(2)) if m is not None]
            if not runs:
                continue
            m = min(runs, key=lambda x: x[obj])
            scores[v] = m
            trials.append({"knob": key, "value": v, "workload": load, "objective": obj, **{k: round(x, 4) for k, x in m.items()}})
            log(f"tune {key}={v} ({load}): wall {m['wall']:.3f} s, seeding {m['seed']:.2f} ms, DP {m['dp']:.2f} ms")
        if scores:
            pick = min(scores, key=lambda v: scores[v][obj])
            if default in scores and scores[pick][obj] > (1 - TUNE_MARGIN) * scores[default][obj]:
                pick = default
            best[key] = pick
            log(f"tune {key}: {best[key]} (by {obj})")
    out = os.path.join(ROOT, "minibwa_amd", "tuned.env")
    with open(out, "w") as f:
        f.write("# Written by `python bench.py --autotune` (coordinate search over launch shapes and batch sizes,\n"
                f"# {args.pairs} pairs of 2x150 bp and {n_long} long reads on the {args.ref_len / 1e9:.1f} Gbp synthetic reference,\n"
                f"# {cores} host threads).  Read by the mapper programs at start-up; the environment wins.\n")
        for k, v in best.items():
            f.write(f"{k}={v}\n")
    log(f"wrote {out}")
    print(json.dumps({"autotune": best, "trials": trials, "workload": {"pairs": args.pairs, "long_reads": n_long,
                      "ref_bp": args.ref_len, "host_threads": cores}}))


def mapper_bench(args):
    """The end-to-end mapper: W + K passes of FASTQ -> SAM in one process (index loaded once), the
    reference CLI beside it, and a byte-for-byte SAM check between them."""
    from minibwa_amd import synth
    exe = os.path.join(ROOT, "minibwa_amd", "bin", "mbg-map-bench")
    cli = os.path.join(ROOT, "minibwa_amd", "bin", "minibwa-gpu")
    ref_cli = os.path.join(ROOT, "oracle", "_ref", "minibwa")
    for e in (exe, cli):
        if not os.path.exists(e):
            log(f"{e} is not built (make -C minibwa_amd/integration); the GPU mapper cannot run")
            sys.exit(1)
    cores = args.threads or host_cores()
    t0 = time.time()
    ref = synth.reference(args.ref_len, seed=11, rep_frac=args.rep_frac, rep_min=300, rep_max=6000, rep_div=0.02)
    log(f"reference {len(ref)} bp in {time.time() - t0:.1f}s")
    work, prefix = mapper_index(args)
    t0 = time.time()
    long_mode = args.long_reads > 0
    if long_mode:      # single-end long reads, PAF output
        seq, off, ln = synth.single_reads(ref, args.long_reads, seed=13, lmin=100, lmax=20000, sub=0.001, indel=0.0005)
        f1 = os.path.join(work, "long.fq")
        synth.write_fastq(f1, seq, off, ln, prefix="l_")
        n_units, tot_bases, inputs, fmt_opt = args.long_reads, int(ln.sum()), [f1], ["-f"]
        del seq, off, ln
        log(f"{n_units} long reads ({tot_bases / 1e6:.1f} Mbp) written to FASTQ in {time.time() - t0:.1f}s")
    else:
        r1, r2 = synth.read_pairs(ref, args.pairs, seed=11, rlen=150, frag_mean=400, frag_sd=50, sub=0.01, indel=0.001)
        f1, f2 = os.path.join(work, "r1.fq"), os.path.join(work, "r2.fq")
        synth.write_pairs_fastq(f1, f2, r1, r2, prefix="p_")
        n_units, tot_bases, inputs, fmt_opt = args.pairs, 300 * args.pairs, [f1, f2], []
        del r1, r2
        log(f"{args.pairs} pairs written to FASTQ in {time.time() - t0:.1f}s")
    unit = "long reads/s" if long_mode else "pairs/s"
    tuned = {}
    tpath = os.path.join(ROOT, "minibwa_amd", "tuned.env")
    if os.path.exists(tpath):          # the autotuner's settings, which the mapper programs read at start-up
        for l in open(tpath):
            if l.startswith("MBG_") and "=" in l:
                k, v = l.strip().split("=", 1)
                tuned[k] = v

    # the GPU mapper: warmup + timed passes in one process; pass 0 writes the SAM compared below
    gpu_sam = os.path.join(work, "gpu.sam")
    passes = args.warmup + args.steps
    n_front_per_pass = None
    env = dict(os.environ, MBG_MAP_STATS="1")
    # pass-1 algorithmic work is counted in the first (warmup) pass only: counting adds atomics
    fronts_pass = max(1, -(-tot_bases // 100_000_000))
    if args.warmup > 0:
        env["MBG_COUNT_WORK"] = str(fronts_pass)
    cmd = [exe] + fmt_opt + ["-t", str(cores), "-P", str(passes), "-o", gpu_sam, "-O", "/dev/null", prefix] + inputs
    log("gpu:", " ".join(cmd))
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=3600)
    if r.returncode != 0:
        log("GPU mapper failed:", r.stderr[-3000:])
        sys.exit(1)
    for l in r.stderr.split("\n"):
        if l.startswith("[mbg_map]") or l.startswith("[mbg_exec] front"):
            log(l)
    pass_s = [float(l.split()[3]) for l in r.stderr.split("\n") if l.startswith("[mbg_bench] pass ")]
    pass_cpu = [float(l.split()[5]) for l in r.stderr.split("\n") if l.startswith("[mbg_bench] pass ") and len(l.split()) > 5]
    load_s = [float(l.split()[2]) for l in r.stderr.split("\n") if l.startswith("[mbg_bench] load ")][0]
    fronts = [_kv(l) for l in r.stderr.split("\n") if l.startswith("[mbg_exec] front ")]
    summ = _kv([l for l in r.stderr.split("\n") if l.startswith("[mbg_map] summary:")][-1])
    assert len(pass_s) == passes, r.stderr[-2000:]
    for i, t in enumerate(pass_s):
        cpu_i = f", host CPU {pass_cpu[i]:.2f} s ({pass_cpu[i] / t:.1f} cores busy)" if i < len(pass_cpu) else ""
        log(f"pass {i}{' (warmup)' if i < args.warmup else ''}: {t:.3f} s = {n_units / t:,.0f} {unit}{cpu_i}")
    timed = pass_s[args.warmup:]
    el = sum(timed)
    value = args.steps * n_units / el
    n_front_per_pass = len(fronts) // passes
    tf = fronts[args.warmup * n_front_per_pass:]
    cf = fronts[:n_front_per_pass] if args.warmup > 0 else []

    # device time per timed pass, by kernel / phase (HIP events on the context's stream)
    per = lambda key: sum(f[key] for f in tf) / max(1, args.steps)
    seed1_ms, seed2_ms = per("seed1_ms"), per("seed2_ms")
    phases = {"smem_pass1_ms": seed1_ms, "smem_pass2_ms": seed2_ms, "group_ms": per("group_ms"), "anchor_sa_ms": per("anchor_ms"),
              "chain_ms": per("chain_ms"), "front_d2h_ms": per("copy_ms")}
    # DP and local SW: cumulative over all passes (every pass does the same work)
    dp_ms = summ.get("dev_dp_ms", 0.0) / passes
    phases["dp_ms"] = dp_ms
    phases["ll_ms"] = summ.get("dev_ll_ms", 0.0) / passes
    phases = {k: round(v, 3) for k, v in phases.items()}
    dev_total = sum(phases.values())
    occ1 = sum(f["seed1_occ_blocks"] for f in cf)
    kmer1 = sum(f["seed1_kmer_lookups"] for f in cf)
    bytes1 = 64 * occ1 + 32 * kmer1
    kernels = []
    if bytes1 > 0 and seed1_ms > 0:
        a = bytes1 / (seed1_ms * 1e-3) / 1e9
        # one pass 1 per front: smem_kernel<1>, and for reads of 1,024+ bases the segment lanes, fix-up
        # and fallback (smem_kernel<3>, <4>, <5>) after it on the same stream; the HIP events span them all
        ent = {"kernel": "SMEM pass 1 (smem_kernel<1>, with long reads' smem_kernel<3-5>)", "bound": "hbm",
               "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(a / HBM_PEAK_GBS, 4), "traffic": None,
               "algorithmic_bytes_per_launch": int(bytes1 / n_front_per_pass), "avg_launch_ms": round(seed1_ms / n_front_per_pass, 4),
               "share_of_device": round(seed1_ms / dev_total, 3) if dev_total else None}
        # dependent random transactions (occ blocks + k-mer lookups) per second, against the gather
        # ceiling measured on the same resident table (bench.py --gather, committed profile)
        ent["G_tx_per_s"] = round((occ1 + kmer1) / (seed1_ms * 1e-3) / 1e9, 2)
        gc = os.path.join(ROOT, "profiles", "gather_ceiling_3g.json")
        if args.ref_len == 3_100_000_000 and os.path.exists(gc):
            ceil = max(r["G_tx_per_s"] for r in json.load(open(gc))["gather_ceiling"])
            ent["gather_ceiling_G_tx_per_s"] = ceil
            ent["frac_of_gather_ceiling"] = round(ent["G_tx_per_s"] / ceil, 3)
        kernels.append(ent)
    cells = summ.get("dev_dp_cells", 0.0) / passes
    if cells > 0 and dp_ms > 0:
        ops = cells * 37 / (dp_ms * 1e-3) / 1e12      # about 37 integer operations per cell with traceback bits
        kernels.append({"kernel": "dp kernels (ksw_extd2 rounds)", "bound": "valu", "achieved": round(ops, 3), "peak": 39.3,
                        "unit": "Tops/s", "frac": round(ops / 39.3, 4), "traffic": None, "gcups": round(cells / (dp_ms * 1e-3) / 1e9, 1),
                        "cells_per_pass": int(cells), "launches_per_pass": round(summ.get("dev_dp_launches", 0) / passes, 1),
                        "avg_launch_ms": round(dp_ms / max(1.0, summ.get("dev_dp_launches", 0) / passes), 4),
                        "note": "dp_fast_kernel<true/false> (ksw_extd2 / ksw_extz2 batches, the ring path inline); ~37 integer "
                                "VALU ops per band cell with traceback bits; HIP events around each batch's launch",
                        "share_of_device": round(dp_ms / dev_total, 3) if dev_total else None})
    kernels.sort(key=lambda k: -(k["share_of_device"] or 0))
    roofline = kernels[0] if kernels else None
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload
    pmc = os.path.join(ROOT, "profiles", "mapper_3g_pmc.json")
    if args.ref_len == 3_100_000_000 and not long_mode and os.path.exists(pmc):
        ks = json.load(open(pmc))["kernels"]
        for k in kernels:
            key = "smem_kernel<1>" if k["kernel"].startswith("SMEM") else "dp_fast_kernel<true>"
            hit = [v for n, v in ks.items() if key in n]
            if hit:
                k["traffic"] = hit[0]["hbm_bytes_per_launch"]
                k["traffic_source"] = "profiles/mapper_3g_pmc.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)"

    # the same program, one `minibwa map` CLI run, timed the reference's way (T_end - T_idx)
    cli_run = None
    if not args.no_cpu:
        r = subprocess.run([cli, "map"] + fmt_opt + ["-t", str(cores), "-o", "/dev/null", prefix] + inputs, capture_output=True,
                           text=True, timeout=3600)
        if r.returncode == 0:
            ti, te = _ref_timings(r.stderr)
            cli_run = {"units_per_s": round(n_units / (te - ti), 1), "unit": unit, "map_s": round(te - ti, 3),
                       "index_load_s": round(ti, 3)}
            log("gpu CLI run:", json.dumps(cli_run))

    cpu, parity = None, None
    if not args.no_cpu and os.path.exists(ref_cli):
        runs = []
        for k in range(max(1, args.cpu_runs)):
            r = subprocess.run([ref_cli, "map"] + fmt_opt + ["-t", str(cores), "-o", "/dev/null", prefix] + inputs,
                               capture_output=True, text=True, timeout=3600)
            if r.returncode != 0:
                log("reference mapper failed:", r.stderr[-1000:])
                break
            ti, te = _ref_timings(r.stderr)
            runs.append(te - ti)
            log(f"reference run {k}: {te - ti:.3f} s mapping ({ti:.1f} s index load)")
        if runs:
            med = float(np.median(runs))
            cpu = {"value": round(n_units / med, 1), "unit": "long reads/s" if long_mode else "read pairs/s", "cores": cores,
                   "kind": "reference",
                   "sample": f"the whole workload: `minibwa map {' '.join(fmt_opt + ['-t' + str(cores)])}` (the reference built "
                             f"from source, -O3 -msse4.2, "
                             f"mimalloc) on the same FASTQ and index files, output to /dev/null; T_end - T_idx from its "
                             f"stderr, median of {len(runs)} run(s): {med:.3f} s; host: {cpu_model()}, {cores} cores, "
                             f"1 thread per core",
                   "runs_s": [round(x, 3) for x in runs]}
        cpu_sam = os.path.join(work, "cpu.sam")
        r = subprocess.run([ref_cli, "map"] + fmt_opt + ["-t", str(cores), "-o", cpu_sam, prefix] + inputs, capture_output=True,
                           text=True, timeout=3600)
        if r.returncode == 0:
            mg, ng = _sam_body_md5(gpu_sam)
            mc, nc = _sam_body_md5(cpu_sam)
            parity = {"sam_identical": mg == mc, "gpu_md5": mg, "cpu_md5": mc, "lines": ng, "cpu_lines": nc}
            log("parity:", json.dumps(parity))
    latency = None
    if args.serve_requests > 0 and not long_mode:
        latency = serve_mapper(args, ref, prefix, work, cores)

    if long_mode:
        workload = (f"end-to-end long-read mapping (FASTQ in, PAF out: map -f), {n_units} single-end reads log-uniform "
                    f"100-20,000 bp ({tot_bases / 1e6:.0f} Mbp) vs {args.ref_len / 1e9:g} Gbp synthetic reference "
                    f"({args.rep_frac:.0%} repeats)")
    else:
        workload = (f"end-to-end mapping (FASTQ in, SAM out), {args.pairs} pairs 2x150 bp vs "
                    f"{args.ref_len / 1e9:g} Gbp synthetic reference ({args.rep_frac:.0%} repeats)")
    res = {
        "metric": METRIC_LONG if long_mode else METRIC,
        "value": round(value, 1),
        "unit": "long reads/s" if long_mode else "read pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1000, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (FM-index ranks) + int8/int32 (DP)",
        "data": "synthetic (seeded splitmix64 reference and reads; index built on the device)",
        "config": {"workload": workload,
                   "ref_bp": args.ref_len, "pairs": 0 if long_mode else args.pairs, "long_reads": args.long_reads,
                   "read_bases": tot_bases, "read_len": "100-20000" if long_mode else 150, "contigs": args.contigs,
                   "host_threads": cores, "options": "minibwa map defaults (adaptive preset)", "mini_batches_per_pass": n_front_per_pass,
                   "tuned": tuned},
        "roofline": roofline,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "parity": parity,
        "latency": latency,
        "device_ms_per_pass": phases,
        "gpu_cli_single_run": cli_run,
        "index_load_s": round(load_s, 2),
        "device_batches_per_pass": round(summ.get("batches", 0) / passes, 1),
        "host_cpu_s_per_pass": round(sum(pass_cpu[args.warmup:]) / max(1, len(pass_cpu[args.warmup:])), 3) if pass_cpu else None,
    }
    print(json.dumps(res), flush=True)
    if not args.workdir:
        shutil.rmtree(work, ignore_errors=True)


def serve_mapper(args, ref, prefix, work, cores):
    """Served mapping on the batched mapper (minibwa_amd/bin/mbg-serve-bench, the serving loop of
    minibwa_amd/integration/mbg_serve.c): requests arrive as a Poisson process at each offered rate of
    the sweep; a request is 1-10,000 read pairs (log-uniform) of the 2x150 bp model, or (5% of the
    requests) one 10 kb read.  Requests are admitted continuously, in groups whose front passes run on
    the device while earlier groups' reads map in the same open run of the worker pool, under the
    device-memory budget (half the free HBM); pairs use predefined insert-size statistics (-I 400,50)
    so that every request's records are the same whatever else is in flight.  Latency = the request's
    records back on the host - its arrival."""
    from minibwa_amd import serve_workload as W
    exe = os.path.join(ROOT, "minibwa_amd", "bin", "mbg-serve-bench")
    if not os.path.exists(exe):
        log(f"{exe} is not built; no serving sweep")
        return None
    t0 = time.time()
    pools = W.write_pools(ref, work, 400_000, 200, seed=21)
    log(f"serving pools (400,000 pairs, 200 long reads of 10 kb) written in {time.time() - t0:.1f}s")
    points = []
    for k, qps in enumerate(float(x) for x in args.serve_sweep.split(",") if x):
        sched = W.schedule(args.serve_requests, qps, seed=100 + k)
        sp = os.path.join(work, "sched.txt")
        W.write_schedule(sp, sched)
        r = subprocess.run([exe, "-t", str(cores), prefix] + list(pools) + [sp], capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            log("serving run failed:", r.stderr[-2000:])
            break
        pt = json.loads(r.stdout.strip().split("\n")[-1])
        pt["offered_qps_target"] = qps
        log(f"serve @{qps:g} req/s: achieved {pt['achieved_qps']:.0f} req/s, {pt['pairs_per_s']:,.0f} pairs/s, "
            f"p50 {pt['p50_ms']:.1f} ms, p99 {pt['p99_ms']:.1f} ms, {pt['reqs_per_batch']:.1f} req/batch")
        points.append(pt)
    if not points:
        return None
    return {"workload": "Poisson arrivals; request = 1-10,000 pairs 2x150 bp (log-uniform) or, 5% of requests, one 10 kb "
                        "read; -I 400,50; latency = records back on the host - arrival",
            "requests_per_point": args.serve_requests, "points": points}



def serve(ctx0, ref, opt, args):
    """Served mapping: requests arrive as a Poisson process at args.serve_qps; a request is 1-10,000
    read pairs (log-uniform) of the workload's 2x150 bp model, or (5% of requests) one 10 kb read
    with 0.1% substitutions and 0.05% indels.  Three contexts share the resident index (two for pair
    requests, one for long reads), each on its own stream.  A free context takes every request of its
    class that has arrived, in arrival order, up to 65,536 pairs or 64 long reads, so requests that
    arrive while a batch is on the GPU are admitted by the next free context and overlap with it.
    Latency = completion of the request's step (results resident in HBM) - its arrival; the request's
    reads cross PCIe inside it."""
    import threading
    from minibwa_amd import mbg, synth
    rng = np.random.default_rng(29)
    n_req = args.serve_requests
    arrive = np.cumsum(rng.exponential(1.0 / args.serve_qps, n_req))
    size = np.exp(rng.uniform(0.0, np.log(10000.0), n_req)).astype(np.int64)
    long_req = rng.random(n_req) < 0.05
    size[long_req] = 1
    n_pairs = int(size[~long_req].sum())
    r1, r2 = synth.read_pairs(ref, n_pairs, seed=31)
    pairs = np.empty((2 * n_pairs, 150), np.uint8)
    pairs[0::2], pairs[1::2] = r1, r2
    del r1, r2
    n_long = int(long_req.sum())
    lseq, loff, llen = synth.single_reads(ref, n_long, seed=37, lmin=10000, lmax=10000, sub=0.001, indel=0.0005)
    assert (llen == 10000).all()
    first_read = np.zeros(n_req + 1, np.int64)                 # start of each pair request in `pairs` (reads)
    first_read[1:] = np.cumsum(np.where(long_req, 0, 2 * size))
    long_idx = np.cumsum(long_req) - 1
    cap_pairs, cap_long = 65536, 64
    opt_long = mbg.adaptive_opt(10000)
    queues = {False: np.nonzero(~long_req)[0], True: np.nonzero(long_req)[0]}

    class Worker:
        def __init__(self, ctx, is_long):
            self.ctx, self.is_long = ctx, is_long
            n = cap_long * 10000 if is_long else 2 * cap_pairs * 150
            self.seq = ctx.to_device(np.zeros(n, np.uint8))
            if is_long:
                self.off = ctx.to_device(np.arange(cap_long + 1, dtype=np.int64) * 10000)
                self.len = ctx.to_device(np.full(cap_long, 10000, np.int32))
            else:
                self.off = ctx.to_device(np.arange(2 * cap_pairs + 1, dtype=np.int64) * 150)
                self.len = ctx.to_device(np.full(2 * cap_pairs, 150, np.int32))

        def step(self, picked):
            if self.is_long:
                seq = np.concatenate([lseq[loff[int(long_idx[k])]:loff[int(long_idx[k]) + 1]] for k in picked])
                self.ctx.write_device(self.seq.ptr, seq)
                self.ctx.map_step_dev(len(picked), self.off.ptr, self.len.ptr, self.seq.ptr, 10000, opt_long)
            else:
                reads = pairs[np.concatenate([np.arange(first_read[k], first_read[k + 1]) for k in picked])]
                self.ctx.write_device(self.seq.ptr, reads)
                self.ctx.map_step_dev(len(reads), self.off.ptr, self.len.ptr, self.seq.ptr, 150, opt)

    extra = [mbg.Context(0), mbg.Context(0)]
    for c in extra:
        c.idx_attach(ctx0)
    workers = [Worker(ctx0, False), Worker(extra[0], False), Worker(extra[1], True)]
    for w in workers:                                          # grow each context's pools before the clock
        q = queues[w.is_long]
        if len(q):
            w.step(list(q[:cap_long]) if w.is_long else [int(k) for k in q[:1]])
    lock = threading.Lock()
    head = {False: 0, True: 0}
    done = np.zeros(n_req)
    steps = {False: [], True: []}
    err = []
    t0 = time.perf_counter() + 0.05

    def run(w):
        q = queues[w.is_long]
        try:
            while True:
                with lock:
                    h = head[w.is_long]
                    if h >= len(q):
                        return
                    now = time.perf_counter() - t0
                    wait = arrive[q[h]] - now
                    if wait <= 0:
                        picked, tot = [int(q[h])], int(size[q[h]])
                        e = h + 1
                        while e < len(q) and arrive[q[e]] <= now:
                            if w.is_long and len(picked) >= cap_long:
                                break
                            if not w.is_long and tot + size[q[e]] > cap_pairs:
                                break
                            picked.append(int(q[e]))
                            tot += int(size[q[e]])
                            e += 1
                        head[w.is_long] = e
                if wait > 0:
                    time.sleep(min(wait, 0.0005))
                    continue
                ts = time.perf_counter()
                w.step(picked)
                t = time.perf_counter()
                done[picked] = t - t0
                steps[w.is_long].append((len(picked), tot, (t - ts) * 1e3))
        except Exception as ex:  # surfaced after the join
            err.append(ex)

    th = [threading.Thread(target=run, args=(w,)) for w in workers]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for c in extra:
        c.close()
    if err:
        raise err[0]
    lat = (done - arrive) * 1e3
    span = done.max() - arrive[0]
    q = lambda x, p: round(float(np.percentile(x, p)), 3) if len(x) else None
    sp, sl = np.array(steps[False]).reshape(-1, 3), np.array(steps[True]).reshape(-1, 3)
    res = {"qps": args.serve_qps, "requests": n_req, "pairs": n_pairs, "long_reads": n_long,
           "p50_ms": q(lat, 50), "p99_ms": q(lat, 99),
           "p50_ms_pairs": q(lat[~long_req], 50), "p99_ms_pairs": q(lat[~long_req], 99),
           "p50_ms_long": q(lat[long_req], 50), "p99_ms_long": q(lat[long_req], 99),
           "served_pairs_per_s": round(n_pairs / span, 1), "offered_pairs_per_s": round(n_pairs / arrive[-1], 1),
           "pair_batches": len(sp), "mean_batch_pairs": round(float(sp[:, 1].mean()), 1) if len(sp) else None,
           "mean_pair_step_ms": round(float(sp[:, 2].mean()), 3) if len(sp) else None,
           "long_batches": len(sl), "mean_long_step_ms": round(float(sl[:, 2].mean()), 3) if len(sl) else None,
           "contexts": len(workers), "note": "latency = arrival -> step complete, H2D of the request's reads included"}
    log("serving:", json.dumps(res))
    return res


def dp_batch(args):
    """Extension-kernel batch: one launch of mbg_extd2_batch_dev over N resident jobs per mode, timed by
    the host around the launch and its completion; cells are the kernel's own band-cell counts."""
    from minibwa_amd import mbg, synth
    ctx = mbg.Context(0)
    mat = np.asarray(mbg.nt4_mat(2, 8, 0, 1), np.int8)
    out = {"metric": "extension GCUPS, ksw_extd2 band 101, 1M jobs q 100-300 bp", "n_jobs": args.dp_batch, "modes": []}
    for flag, zdrop, eb, two in ((0x40, 172, 10, False), (0x00, -1, 0, True)):
        qb, qoff, ql, tb, toff, tl = synth.dp_jobs(args.dp_batch, seed=11, two_sided=two)
        n = len(ql)
        jobs = np.zeros(n, mbg.DP_JOB_DTYPE)
        jobs["qoff"], jobs["toff"], jobs["qlen"], jobs["tlen"] = qoff[:-1], toff[:-1], ql, tl
        jobs["q"], jobs["e"], jobs["q2"], jobs["e2"] = 12, 2, 24, 1
        jobs["w"], jobs["zdrop"], jobs["end_bonus"], jobs["flag"] = 101, zdrop, eb, flag
        span = ql.astype(np.int64) + tl
        jobs["cigar_off"] = np.concatenate([[0], np.cumsum(span)[:-1]])
        d_jobs, d_q, d_t = ctx.to_device(jobs), ctx.to_device(qb), ctx.to_device(tb)
        d_res = ctx.to_device(np.zeros(n, mbg.DP_RES_DTYPE))
        d_cig = ctx.to_device(np.zeros(int(span.sum()), np.uint32))
        d_mat = mat
        times = []
        for it in range(3):
            t0 = time.perf_counter()
            mbg.check(ctx.L.mbg_extd2_batch_dev(ctx.h, None, n, mbg.C.c_void_p(d_jobs.ptr), mbg.C.c_void_p(d_q.ptr),
                                                mbg.C.c_void_p(d_t.ptr), 5, mbg._ptr(d_mat), mbg.C.c_void_p(d_res.ptr),
                                                mbg.C.c_void_p(d_cig.ptr), int(ql.max()), int(tl.max()), 101),
                      "mbg_extd2_batch_dev")
            res = ctx.to_host(d_res.ptr, mbg.DP_RES_DTYPE, n)      # waits for the kernel
            times.append(time.perf_counter() - t0)
        cells = int(res["cells"].astype(np.int64).sum())
        best = min(times)
        m = {"flag": flag, "zdrop": zdrop, "end_bonus": eb, "two_sided_flanks": two, "cells": cells,
             "s": round(best, 4), "GCUPS": round(cells / best / 1e9, 2)}
        cpu = dp_cpu_sample(jobs, qb, tb, res, mat, min(n, args.cpu_jobs))
        if cpu:
            m["cpu_baseline"] = cpu
        log("dp-batch", json.dumps(m))
        out["modes"].append(m)
    print(json.dumps(out), flush=True)
    ctx.close()


def ll_batch(args):
    """Mate-rescue local SW batch: one mbg_ll_batch_dev call over N resident jobs, timed by the host
    around the call (which waits for the kernel); cells = qlen x target rows processed."""
    from minibwa_amd import mbg, synth
    ctx = mbg.Context(0)
    mat = np.asarray(mbg.nt4_mat(1, 4, 0, 1), np.int8)    # pe.c:258-262 with -A2 -B8: match 1, mismatch -4
    gapo, gape, xtra = 6, 1, mbg.KSW_LL_SUBO | 19          # ceil(12/2), ceil(2/2), SUBO | min_len
    t0 = time.time()
    qb, qoff, ql, tb, toff, tl = synth.ll_jobs(args.ll_batch, seed=11)
    n = len(ql)
    jobs = np.zeros(n, mbg.LL_JOB_DTYPE)
    jobs["qoff"], jobs["toff"], jobs["qlen"], jobs["tlen"], jobs["size"] = qoff[:-1], toff[:-1], ql, tl, 1
    log(f"{n} ll jobs generated in {time.time() - t0:.1f}s")
    d_jobs, d_q, d_t = ctx.to_device(jobs), ctx.to_device(qb), ctx.to_device(tb)
    d_res = ctx.to_device(np.zeros(n, mbg.LL_RES_DTYPE))
    rows = int(tl.astype(np.int64).sum())
    times = []
    for it in range(3):
        t0 = time.perf_counter()
        mbg.check(ctx.L.mbg_ll_batch_dev(ctx.h, None, n, mbg.C.c_void_p(d_jobs.ptr), mbg.C.c_void_p(d_q.ptr),
                                         mbg.C.c_void_p(d_t.ptr), 5, mbg._ptr(mat), gapo, gape, xtra,
                                         mbg.C.c_void_p(d_res.ptr), int(ql.max()), rows), "mbg_ll_batch_dev")
        times.append(time.perf_counter() - t0)
    res = ctx.to_host(d_res.ptr, mbg.LL_RES_DTYPE, n)
    assert (res["score"] >= 0).all()
    cells = int((ql.astype(np.int64) * tl).sum())   # scores stay below 251: every row is processed
    best = min(times)
    out = {"metric": "mate-rescue local SW GCUPS (ksw_ll u8 core, 150 bp reads x 550 bp windows)",
           "value": round(cells / best / 1e9, 2), "unit": "GCUPS", "n_jobs": n, "cells": cells, "s": round(best, 4),
           "jobs_per_s": round(n / best), "rescued_frac": round(float((res["score"] >= 19).mean()), 4),
           "with_second_best": int((res["score2"] >= 0).sum())}
    cpu = ll_cpu_sample(jobs, qb, tb, mat, gapo, gape, xtra, min(n, args.cpu_jobs))
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    ctx.close()


def rescue_batch(args):
    """Mate-rescue batch: mbg_matesw_batch over N candidates on a resident 100 Mbp reference (l2b only);
    the value is candidates per second of device time (HIP events around the phases, window extraction
    to packed CIGARs); the host-inclusive rate (PCIe both ways) is reported beside it."""
    from minibwa_amd import mbg, synth
    ctx = mbg.Context(0)
    t0 = time.time()
    ref = synth.reference(100_000_000, seed=11)
    ctg_len, pac = synth.ref_to_l2b_arrays(ref, 1)
    ctx.l2b_upload(ctg_len, pac)
    reads, ts, te = synth.rescue_cands(ref, args.rescue_batch, seed=11)
    n = len(reads)
    log(f"{n} rescue candidates on a 100 Mbp reference in {time.time() - t0:.1f}s")
    # short-read preset (options.c:56-67) and defaults: -A2 -B8 -O12 -E2, band 100, z-drop 80, end bonus 10
    o = dict(a=2, b=8, b_ts=0, b_ambi=1, q=12, e=2, bw=100, zdrop=80, end_bonus=10, min_dp_max=30, min_len=19)
    opt = mbg.MateswOpt(**o, mt=0)
    min_sc = np.full(n, 20, np.int32)
    runs = []
    for it in range(3):
        st = mbg.MateswStats()
        t1 = time.perf_counter()
        res, _ = ctx.matesw(reads, np.zeros(n, np.int32), ts, te, min_sc, opt, st)
        runs.append((st.ms_total, time.perf_counter() - t1, st.as_dict()))
    best = min(runs, key=lambda r: r[0])
    st = best[2]
    out = {"metric": "mate-rescue candidates/s (mb_ungap + ksw_ll + ksw_extz2, 150 bp mates x 550 bp windows)",
           "value": round(n / (best[0] / 1e3)), "unit": "candidates/s", "n": n, "ms_device": round(best[0], 3),
           "host_inclusive_per_s": round(n / best[1]), "phases": st,
           "ll_GCUPS": round(st["ll_cells"] / (st["ms_ll"] / 1e3) / 1e9, 1),
           "rescued_frac": round(float(res["rescued"].mean()), 4)}
    cpu = rescue_cpu_sample(reads, ts, te, ref, o, min(n, args.cpu_jobs))
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    ctx.close()


def fastq_batch(args):
    """The FASTQ reader at the edge of the path (host code, no device): mbg_fq_read over one file of N
    150 bp reads (the synthetic pairs of the default workload, interleaved), in the mini-batches the CLI
    reads (100M bases, frag mode, min_cnt 40000: map-main.c:114-125), from the page cache; plain and
    gzip.  The value is records/s of the plain file; the reference's mb_bseq_read
    (oracle/_ref/cpu_hotpath bseq) on the same files is timed beside it."""
    from minibwa_amd import mbg, synth
    n = args.fastq_batch
    ref = synth.reference(10_000_000, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, (n + 1) // 2, seed=11, rlen=150)
    reads = np.empty((2 * r1.shape[0], 150), np.uint8)
    reads[0::2], reads[1::2] = r1, r2
    reads = reads[:n]
    work = tempfile.mkdtemp(prefix="mbg_fq_")
    fq = os.path.join(work, "reads.fq")
    synth.write_fastq(fq, reads.reshape(-1), np.arange(n + 1, dtype=np.int64) * 150, np.full(n, 150, np.int32), prefix="r")
    gz = fq + ".gz"
    with open(gz, "wb") as f:
        subprocess.run(["gzip", "-1", "-c", fq], stdout=f, check=True)
    cores = host_cores()
    chunk = 100_000_000
    out = {"metric": "FASTQ records/s (kseq parsing, U->T, nt4 codes), 150 bp reads, 100M-base chunks", "n_records": n,
           "file_bytes": os.path.getsize(fq), "gz_bytes": os.path.getsize(gz), "threads": cores}
    for tag, path in (("plain", fq), ("gz", gz)):
        for _ in range(2):                                  # the first pass warms the page cache
            r = mbg.FastxReader(path, cores)
            ch = mbg.FqChunk()
            t0 = time.perf_counter()
            n_rec = n_bases = 0
            while True:
                mbg.check(r.L.mbg_fq_read(r.h, chunk, 1, 0, 1, 40000, chunk, r.buf, mbg.C.byref(ch)), "mbg_fq_read")
                if ch.n == 0:
                    break
                n_rec += ch.n
                n_bases += int(ch.seq_off[ch.n])
            t = time.perf_counter() - t0
            r.close()
        assert n_rec == n and n_bases == 150 * n
        d = {"records_per_s": round(n / t), "MB_per_s": round(os.path.getsize(path) / t / 1e6, 1), "s": round(t, 3)}
        exe = os.path.join(ROOT, "oracle", "_ref", "cpu_hotpath")
        if not args.no_cpu and os.path.exists(exe):
            rr = subprocess.run([exe, "bseq", path, str(chunk)], capture_output=True, text=True, timeout=900)
            if rr.returncode == 0:
                c = json.loads(rr.stdout.strip().split("\n")[-1])
                d["cpu_baseline"] = {"value": round(c["records"] / c["bseq_s"]), "unit": "records/s", "cores": 1,
                                     "kind": "reference", "sample": f"mb_bseq_read over the same file, {c['bseq_s']:.2f}s"}
        out[tag] = d
    out["value"], out["unit"] = out["plain"]["records_per_s"], "records/s"
    if "cpu_baseline" in out["plain"]:
        out["cpu_baseline"] = out["plain"]["cpu_baseline"]
    shutil.rmtree(work, ignore_errors=True)
    print(json.dumps(out), flush=True)


def anchor_batch(ctx, reads, prefix, args):
    """mb_anchor then mb_lchain_dp batch: mbg_anchor_batch_dev over the SMEM intervals of every read (from
    mbg_seed_batch on the same resident index), then mbg_lchain_batch_dev over the anchors it left in HBM;
    the value is reads per second of host wall time around the anchor call and its completion, and the
    chaining rate is reported beside it.  The interval array is restored from a pristine copy before each
    anchor run (the call sorts it in place)."""
    from minibwa_amd import mbg
    n, rl = reads.shape
    t0 = time.time()
    flat, off = ctx.seed_intv_flat(reads, 19, 10)
    log(f"{n} reads seeded: {len(flat)} intervals in {time.time() - t0:.1f}s")
    max_occ = 250
    qlen = np.full(n, rl, np.int32)
    cs = np.zeros(len(flat) + 1, np.int64)
    np.cumsum(np.minimum(flat[:, 2], max_occ).astype(np.int64), out=cs[1:])
    aoff = np.ascontiguousarray(cs[off])
    n_slots = int(aoff[-1])
    d_u0, d_u = ctx.to_device(flat), ctx.to_device(flat)
    d_off, d_ql, d_aoff = ctx.to_device(off), ctx.to_device(qlen), ctx.to_device(aoff)
    d_rows = ctx.to_device(np.zeros(n_slots, np.uint64))
    d_anc = ctx.to_device(np.zeros(n_slots, mbg.ANCHOR_DTYPE))
    d_nanc, d_nu = ctx.to_device(np.zeros(n, np.int64)), ctx.to_device(np.zeros(n, np.int64))
    P = mbg.C.c_void_p
    times = []
    for it in range(4):
        mbg.check(ctx.L.mbg_stream_copy(ctx.h, None, P(d_u.ptr), P(d_u0.ptr), flat.nbytes), "mbg_stream_copy")
        t1 = time.perf_counter()
        mbg.check(ctx.L.mbg_anchor_batch_dev(ctx.h, None, n, n_slots, P(d_u.ptr), P(d_off.ptr), P(d_ql.ptr), max_occ,
                                             P(d_rows.ptr), P(d_anc.ptr), P(d_aoff.ptr), P(d_nanc.ptr), P(d_nu.ptr)),
                  "mbg_anchor_batch_dev")
        ctx.to_host(d_nanc.ptr, np.int64, 1)          # waits for the context stream
        times.append(time.perf_counter() - t1)
        log(f"anchor run {it}: {times[-1] * 1e3:.2f} ms")
    nanc = ctx.to_host(d_nanc.ptr, np.int64, n)
    nu = ctx.to_host(d_nu.ptr, np.int64, n)
    assert (nanc >= 0).all()
    best = min(times[1:])
    # chaining the resident anchors with the mapper's options for this read length
    copt = mbg.chain_opt(rl)
    wb = mbg.lchain_work_bytes(n_slots)
    d_work = ctx.to_device(np.zeros(wb, np.uint8))
    d_u = ctx.to_device(np.zeros(n_slots, np.uint64))
    d_nu_c = ctx.to_device(np.zeros(n, np.int32))
    d_cb = ctx.to_device(np.zeros(n_slots, mbg.ANCHOR_DTYPE))
    d_ncb = ctx.to_device(np.zeros(n, np.int64))
    ctimes = []
    for it in range(4):
        t1 = time.perf_counter()
        mbg.check(ctx.L.mbg_lchain_batch_dev(ctx.h, None, n, n_slots, P(d_anc.ptr), P(d_aoff.ptr), P(d_nanc.ptr),
                                             mbg.C.byref(copt), P(d_work.ptr), wb, P(d_u.ptr), P(d_nu_c.ptr), P(d_cb.ptr),
                                             P(d_ncb.ptr)), "mbg_lchain_batch_dev")
        ctx.to_host(d_ncb.ptr, np.int64, 1)
        ctimes.append(time.perf_counter() - t1)
        log(f"chain run {it}: {ctimes[-1] * 1e3:.2f} ms")
    n_chains = int(ctx.to_host(d_nu_c.ptr, np.int32, n).astype(np.int64).sum())
    cbest = min(ctimes[1:])
    out = {"metric": "mb_anchor reads/s (interval sort/dedup, SA locate, contig mapping, anchor sort/dedup)",
           "value": round(n / best), "unit": "reads/s", "n_reads": n, "ms": round(best * 1e3, 3),
           "n_intervals": len(flat), "n_intervals_dedup": int(nu.sum()), "n_slots": n_slots, "n_anchors": int(nanc.sum()),
           "ref_bp": args.ref_len, "max_occ": max_occ,
           "chain": {"value": round(n / cbest), "unit": "reads/s", "ms": round(cbest * 1e3, 3), "n_chains": n_chains,
                     "opt": {f: getattr(copt, f) for f, _ in mbg.ChainOpt._fields_ if f != "pad"}}}
    if not args.no_cpu:
        cpu = anchor_cpu_sample(prefix, flat, off, qlen, max_occ, copt, min(n, args.cpu_pairs))
        if cpu:
            out["cpu_baseline"], out["chain"]["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def anchor_cpu_sample(prefix, flat, off, qlen, max_occ, copt, n_s):
    """The reference's mb_anchor and mb_lchain_dp (oracle/_ref/cpu_hotpath anchor) on the first n_s reads'
    intervals -> (anchor baseline, chain baseline); each rate is n_s over that function's thread-seconds
    divided by the thread count."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_hotpath")
    if not os.path.exists(exe) or n_s <= 0:
        return None
    work = tempfile.mkdtemp(prefix="mbg_anchor_")
    fn = os.path.join(work, "intv.bin")
    with open(fn, "wb") as f:
        f.write(np.int64(n_s).tobytes())
        f.write(np.array([copt.max_dist_x, copt.max_dist_y, copt.bw, copt.max_skip, copt.max_iter, copt.min_sc],
                         np.int32).tobytes())
        f.write(np.float32(copt.pen_gap).tobytes())
        f.write(np.ascontiguousarray(qlen[:n_s]).tobytes())
        f.write(np.ascontiguousarray(off[:n_s + 1]).tobytes())
        f.write(np.ascontiguousarray(flat[:off[n_s]]).tobytes())
    cores = host_cores()
    r = subprocess.run([exe, "anchor", prefix, fn, str(max_occ), str(cores)], capture_output=True, text=True, timeout=900)
    shutil.rmtree(work, ignore_errors=True)
    if r.returncode != 0:
        log("anchor cpu sample failed:", r.stderr[-300:])
        return None
    d = json.loads(r.stdout.strip().split("\n")[-1])
    ta, tc = d["anchor_thread_s"] / cores, d["chain_thread_s"] / cores
    return ({"value": round(n_s / ta), "unit": "reads/s", "cores": cores, "kind": "reference",
             "sample": f"mb_anchor on the first {n_s} reads' intervals, {cores} threads, {d['anchor_thread_s']:.2f} "
                       f"thread-s ({d['wall_s']:.2f}s wall with chaining), {d['n_anchors']} anchors"},
            {"value": round(n_s / tc), "unit": "reads/s", "cores": cores, "kind": "reference",
             "sample": f"mb_lchain_dp on the same reads' anchors, {cores} threads, {d['chain_thread_s']:.2f} thread-s, "
                       f"{d['n_chains']} chains"})


def rescue_cpu_sample(reads, ts, te, ref, o, n_s):
    """The reference's rescue step (mb_ungap + mb_matesw_align, oracle/_ref/cpu_rescue) on the first n_s
    candidates, windows cut on the host."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_rescue")
    if not os.path.exists(exe) or n_s <= 0:
        return None
    rl = reads.shape[1]
    wl = int(te[0] - ts[0])
    win = np.asarray(ref, np.uint8)[ts[:n_s, None] + np.arange(wl)]
    c = np.zeros(n_s, [("qoff", "<i8"), ("woff", "<i8"), ("qlen", "<i4"), ("wlen", "<i4"), ("min_sc", "<i4"), ("pad", "<i4")])
    c["qoff"], c["woff"], c["qlen"], c["wlen"], c["min_sc"] = np.arange(n_s) * rl, np.arange(n_s) * wl, rl, wl, 20
    work = tempfile.mkdtemp(prefix="mbg_rescue_")
    fn = os.path.join(work, "cands.bin")
    p = [o[k] for k in ("a", "b", "b_ts", "b_ambi", "q", "e", "bw", "zdrop", "end_bonus", "min_dp_max", "min_len")]
    with open(fn, "wb") as f:
        f.write(np.array(p + [0], np.int32).tobytes())
        f.write(np.int64(n_s).tobytes())
        f.write(c.tobytes())
        f.write(np.int64(n_s * rl).tobytes()); f.write(np.ascontiguousarray(reads[:n_s]).tobytes())
        f.write(np.int64(n_s * wl).tobytes()); f.write(win.tobytes())
    cores = host_cores()
    r = subprocess.run([exe, fn, str(cores)], capture_output=True, text=True, timeout=600)
    shutil.rmtree(work, ignore_errors=True)
    if r.returncode != 0:
        log("rescue cpu sample failed:", r.stderr[-300:])
        return None
    d = json.loads(r.stdout.strip().split("\n")[-1])
    return {"value": round(n_s / d["rescue_s"]), "unit": "candidates/s", "cores": cores, "kind": "reference",
            "sample": f"mb_ungap + mb_matesw_align on the first {n_s} candidates, {cores} threads, {d['rescue_s']:.2f}s, "
                      f"{d['rescued']} rescued"}


def ll_cpu_sample(jobs, qb, tb, mat, gapo, gape, xtra, n_s):
    """ksw_ll_qinit + ksw_ll_u8_core of the compiled reference on the first n_s jobs (oracle/_ref/cpu_hotpath ll)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_hotpath")
    if not os.path.exists(exe) or n_s <= 0:
        return None
    work = tempfile.mkdtemp(prefix="mbg_ll_")
    jb = os.path.join(work, "ll_jobs.bin")
    j = jobs[:n_s]
    qend = int(j["qoff"][-1] + j["qlen"][-1])
    tend = int(j["toff"][-1] + j["tlen"][-1])
    with open(jb, "wb") as f:
        f.write(np.int64(n_s).tobytes())
        f.write(np.asarray(mat, np.int8).tobytes())
        f.write(np.array([gapo, gape, xtra], np.int32).tobytes())
        f.write(j.tobytes())
        f.write(np.int64(qend).tobytes()); f.write(qb[:qend].tobytes())
        f.write(np.int64(tend).tobytes()); f.write(tb[:tend].tobytes())
    cores = host_cores()
    r = subprocess.run([exe, "ll", jb, str(cores)], capture_output=True, text=True, timeout=600)
    shutil.rmtree(work, ignore_errors=True)
    if r.returncode != 0:
        log("ll cpu sample failed:", r.stderr[-300:])
        return None
    d = json.loads(r.stdout.strip().split("\n")[-1])
    cells = int((j["qlen"].astype(np.int64) * j["tlen"]).sum())
    return {"value": round(cells / d["ll_s"] / 1e9, 3), "unit": "GCUPS", "cores": cores, "kind": "reference",
            "sample": f"ksw_ll_qinit + ksw_ll_u8_core on the first {n_s} jobs, {cores} threads, {d['ll_s']:.2f}s"}


def dp_cpu_sample(jobs, qb, tb, res, mat, n_s):
    """ksw_extd2_sse of the compiled reference on the first n_s jobs (oracle/_ref/cpu_hotpath)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_hotpath")
    if not os.path.exists(exe) or n_s <= 0:
        return None
    work = tempfile.mkdtemp(prefix="mbg_dp_")
    rb, jb = os.path.join(work, "reads.bin"), os.path.join(work, "jobs.bin")
    with open(rb, "wb") as f:
        f.write(np.int64(0).tobytes())
    j = jobs[:n_s]
    qend = int(j["qoff"][-1] + j["qlen"][-1])
    tend = int(j["toff"][-1] + j["tlen"][-1])
    with open(jb, "wb") as f:
        f.write(np.int64(n_s).tobytes())
        f.write(np.asarray(mat, np.int8).tobytes())
        f.write(j.tobytes())
        f.write(np.int64(qend).tobytes()); f.write(qb[:qend].tobytes())
        f.write(np.int64(tend).tobytes()); f.write(tb[:tend].tobytes())
    cores = host_cores()
    r = subprocess.run([exe, "-", rb, jb, str(cores)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log("dp cpu sample failed:", r.stderr[-300:])
        return None
    d = json.loads(r.stdout.strip().split("\n")[-1])
    cells = int(res["cells"][:n_s].astype(np.int64).sum())
    return {"value": round(cells / d["dp_s"] / 1e9, 3), "unit": "GCUPS", "cores": cores, "kind": "reference",
            "sample": f"ksw_extd2_sse on the first {n_s} jobs, {cores} threads, {d['dp_s']:.2f}s"}


def cpu_baseline(ctx, out, st, reads, prefix, args, work):
    """Time the compiled reference's hot-path functions on a sample of the same reads."""
    from minibwa_amd import mbg
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_hotpath")
    if not os.path.exists(exe):
        log("cpu_baseline skipped: oracle/_ref/cpu_hotpath not built")
        return None
    n_s = min(args.cpu_pairs, args.pairs) * 2
    rl = reads.shape[1]
    rb = os.path.join(work, "cpu_reads.bin")
    with open(rb, "wb") as f:
        f.write(np.int64(n_s).tobytes())
        f.write(np.full(n_s, rl, np.int32).tobytes())
        f.write(np.ascontiguousarray(reads[:n_s]).tobytes())
    n_jobs = st["n_jobs"]
    jobs = ctx.to_host(out.jobs, mbg.DP_JOB_DTYPE, n_jobs)
    meta = ctx.to_host(out.meta, mbg.JOB_META_DTYPE, n_jobs)
    keep = np.nonzero(meta["rid"] < n_s)[0]
    jobs = jobs[keep]
    qb = ctx.to_host(out.qbuf, np.uint8, st["q_bytes"])
    tb = ctx.to_host(out.tbuf, np.uint8, st["t_bytes"])
    # repack the sample's jobs with their own sequence buffers
    qparts, tparts, qo, to = [], [], 0, 0
    j2 = jobs.copy()
    for i in range(len(jobs)):
        q = qb[jobs["qoff"][i]: jobs["qoff"][i] + jobs["qlen"][i]]
        t = tb[jobs["toff"][i]: jobs["toff"][i] + jobs["tlen"][i]]
        j2["qoff"][i], j2["toff"][i] = qo, to
        qparts.append(q), tparts.append(t)
        qo += len(q)
        to += len(t)
    qcat = np.concatenate(qparts) if qparts else np.zeros(0, np.uint8)
    tcat = np.concatenate(tparts) if tparts else np.zeros(0, np.uint8)
    jb = os.path.join(work, "cpu_jobs.bin")
    with open(jb, "wb") as f:
        f.write(np.int64(len(j2)).tobytes())
        f.write(np.asarray(mbg.nt4_mat(), np.int8).tobytes())
        f.write(j2.tobytes())
        f.write(np.int64(len(qcat)).tobytes()); f.write(qcat.tobytes())
        f.write(np.int64(len(tcat)).tobytes()); f.write(tcat.tobytes())
    cores = host_cores()
    log(f"cpu_baseline: {n_s // 2} pairs, {len(j2)} DP jobs, {cores} threads")
    r = subprocess.run([exe, prefix, rb, jb, str(cores)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log("cpu_baseline failed:", r.stderr[-500:])
        return None
    d = json.loads(r.stdout.strip().split("\n")[-1])
    tot = d["seed_anchor_s"] + d["dp_s"]
    return {"value": round((n_s // 2) / tot, 1), "unit": "read pairs/s", "cores": cores, "kind": "reference",
            "sample": f"first {n_s // 2} pairs of the workload: mb_seed_intv_batch + mb_anchor (incl. mb_bwt_sa_batch) "
                      f"in 48-read sub-batches, then ksw_extd2_sse on the sample's {len(j2)} extension jobs; "
                      f"{cores} threads, wall {tot:.2f}s",
            "seed_anchor_s": d["seed_anchor_s"], "dp_s": d["dp_s"]}


if __name__ == "__main__":
    main()
