#!/usr/bin/env python3
"""Benchmark of the MI355X read-matching path on the genome-scale workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload genome_se|uneven_se|paired]

Workload (the reference's GPU drop-in case): a synthetic 300 Mbp reference
of 24 chromosomes (i.i.d. bases, GC 41 %, CpG depleted to ~1 %), indexed
with the shipped constants (HTABSIZE 2^30, READLEN 100, WINLEN 2048,
SKIPMOD 2, KMERCUTOFF 1500), and 10 M single-end 100 bp bisulfite reads
(Beta(0.5,0.5) CpG methylation, 99.5 % non-CpG conversion, 0.5 %
substitutions, 0.05 % 1-bp indels) from both strands.

One step = the whole matching job over the 10 M reads as FAME's chunk loop
defines it (queryRoutine main.cpp:363-398): the first 300 000 reads matched in
both orientations, the strand decision, then the remaining reads in one
orientation.  Reads are resident in HBM before the timed region; index
construction and loading are outside it.  value = aligned reads/s
(succMatch / time).

Other workloads (--workload): uneven_se = the same genome with 3 000
segments copied 5-50 times (1 % divergence), 10 M single-end reads of
50-150 bp (30 % from the copied segments), READLEN 150; paired = 300 Mbp
with 20 % of the sequence in 300 bp repeat families (10 % divergence),
indexed without KMERCUTOFF (--no_loss), 5 M pairs of 2 x 150 bp with inner
gap uniform in [20, 400], indel rate 0.2 %, MINPDIST 20 / MAXPDIST 400,
READLEN 150 (value = matched reads / s, both mates counted, as FAME's
succMatch counts them).

serve = the genome_se reference and read model served as a request stream
(value = aligned reads/s at the full offered load; p50 / p99 chunk latency
per load, beside the plain form's at the same load).

Also reported: per-stage device time, the roofline of the dominant kernel,
and FAME's own OpenMP CPU path (built from its sources into oracle/_ref)
timed on this host on a bounded sample of the same workload, whose per-read
results and CpG report are compared with the HIP path's on that sample.
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

METRIC = ("aligned reads/s (batched vs plain GPU vs CPU), bit-exact CpG calls; p99 latency")
HBM_PEAK_GBS = 8000.0
CHUNKSIZE = 300000


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workdir_for(need_bytes, requested):
    cands = [requested] if requested else ["/dev/shm", os.environ.get("TMPDIR", "/tmp"), "/tmp"]
    for d in cands:
        if d and os.path.isdir(d):
            try:
                free = shutil.disk_usage(d).free
            except OSError:
                continue
            if free > need_bytes:
                return tempfile.mkdtemp(prefix="fame_bench_", dir=d)
    raise SystemExit(f"no folder with {need_bytes / 1e9:.1f} GB free for the index")


def ref_binary(htab, readlen, cores, nchrom, minpd=50, maxpd=550):
    name = f"fame_ref_h{htab}_r{readlen}_c{cores}_n{nchrom}_p{minpd}-{maxpd}_k{CHUNKSIZE}"
    p = os.path.join(ROOT, "oracle", "_ref", name)
    return p if os.path.exists(p) else None


# workloads: genome, index constants, reads
WORKLOADS = {
    "genome_se": dict(readlen=100, n=10_000_000, paired=False, noloss=False, minpd=50, maxpd=550,
                      desc="genome-scale single-end: 300 Mbp / 24 chromosomes, 10M x 100 bp bisulfite "
                           "reads, HTABSIZE 2^30"),
    "uneven_se": dict(readlen=150, n=10_000_000, paired=False, noloss=False, minpd=50, maxpd=550,
                      len_range=(50, 150), near_copies=3000, copied_frac=0.3,
                      desc="uneven single-end: 300 Mbp with 3000 segments copied 5-50 times, 10M reads of "
                           "50-150 bp (30% from the copies), READLEN 150"),
    "serve": dict(readlen=100, n=4_000_000, paired=False, noloss=False, minpd=50, maxpd=550, serve=True,
                  desc="serving loop: 300 Mbp index resident, chunks of single-end 100 bp reads arriving as a "
                       "Poisson stream, log-normal sizes (median 20k, sigma 1, clipped to [1k, 300k]), offered "
                       "load swept 20-100% of the batched peak, continuous batching vs the plain form"),
    "paired": dict(readlen=150, n=5_000_000, paired=True, noloss=True, minpd=20, maxpd=400,
                   repeat_frac=0.2, gap=(20, 400), indel=0.002,
                   desc="paired-end repeat-heavy: 300 Mbp with 20% in 300 bp repeat families (10% "
                        "divergence), --no_loss index, 5M pairs of 2 x 150 bp, gap 20-400, indel 0.2%, "
                        "plain form"),
}


def make_workload(wl, args, chroms):
    """Reads (single-end ReadSet) or pairs ((ReadSet, ReadSet)) of a workload."""
    from fame_amd import synth
    n = args.reads or wl["n"]
    if wl["paired"]:
        return synth.make_pairs(chroms, n, wl["readlen"], gap=wl["gap"], seed=args.seed + 2,
                                indel_rate=wl["indel"])
    if "near_copies" not in wl:
        return synth.make_reads(chroms, n, wl["readlen"], seed=args.seed + 2, len_range=wl.get("len_range"))
    regions = wl["regions"]
    nc = int(n * wl["copied_frac"])
    a = synth.make_reads(chroms, n - nc, wl["readlen"], seed=args.seed + 2, len_range=wl["len_range"])
    b = synth.make_reads(chroms, nc, wl["readlen"], seed=args.seed + 3, len_range=wl["len_range"],
                         regions=regions)
    order = np.random.default_rng(args.seed + 4).permutation(n)
    seqs = np.concatenate([a.seqs, b.seqs])[order]
    lens = np.concatenate([a.lens, b.lens])[order]
    return synth.ReadSet(seqs, lens)


def write_fastq(path, seqs, lens):
    with open(path, "wb") as f:
        q = b"I" * int(lens.max() if lens.size else 0)
        for i in range(lens.size):
            L = int(lens[i])
            f.write(b"@r%d\n" % i + seqs[i, :L].tobytes() + b"\n+\n" + q[:L] + b"\n")


def parse_tsv_counts(path):
    a = np.loadtxt(path, dtype=np.int64, usecols=(2, 3, 4, 5), ndmin=2)
    return a.astype(np.uint16)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="genome_se")
    ap.add_argument("--reads", type=int, default=0, help="reads (pairs) per step; 0 = the workload's")
    ap.add_argument("--genome-scale", type=float, default=1.0, help="1.0 = 300 Mbp")
    ap.add_argument("--htab", type=int, default=1 << 30)
    ap.add_argument("--batch", type=int, default=2_000_000, help="reads (pairs) per call after the first chunk")
    ap.add_argument("--cpu-sample", type=int, default=0, help="reads (pairs) of the CPU leg; 0 = default")
    ap.add_argument("--cpu-cores", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per load point")
    ap.add_argument("--serve-form", default="batched", help="pipeline form of the server's batches")
    ap.add_argument("--sub-batch", type=int, default=0, help="batched form: reads per pipelined sub-batch (0 = library default)")
    ap.add_argument("--forms", default="plain,batched",
                    help="single-end forms to time (plain: one kernel per stage; batched: one graph per batch)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--gather-probe", type=int, default=0,
                    help="only run a random 8-byte gather of this many lookups from a 2^30-word table "
                         "(HBM counter calibration for random reads)")
    ap.add_argument("--pmc-summary", nargs="+", default=None, metavar=("OUT.md", "CSV"),
                    help="only write OUT.md: per-launch rocprofv3 counters of the last dispatch of each "
                         "kernel named by --pmc-kernels, from the counter_collection CSVs given")
    ap.add_argument("--pmc-kernels", default="k_verify<2>;k_seed_light<unsigned int, 128;k_meth_dp",
                    help="';'-separated parts of kernel names")
    ap.add_argument("--pmc-command", default="", help="the profiled command, for the summaries' header")
    ap.add_argument("--traffic-summary", nargs=6, default=None,
                    metavar=("OUT.md", "KERNEL", "FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="only write OUT.md: FETCH_SIZE / WRITE_SIZE of KERNEL's last dispatch, scaled by "
                         "what the counters read for the --copy-probe-mib copy (size --copy-mib)")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--gather-summary", nargs=3, default=None, metavar=("OUT.md", "GATHER_FETCH_CSV", "COPY_FETCH_CSV"),
                    help="only write OUT.md: FETCH_SIZE per lookup of the --gather-probe run (lookups: "
                         "--gather-lookups) against the copy's calibration")
    ap.add_argument("--gather-lookups", type=int, default=138_000_000)
    ap.add_argument("--stats-summary", nargs=2, default=None, metavar=("OUT.md", "DIR"),
                    help="only write OUT.md: the kernel table of a rocprofv3 --kernel-trace --stats run "
                         "(its CSVs in DIR) and the per-launch durations of the kernels --pmc-kernels names")
    ap.add_argument("--extras", default="uneven_se,paired,serve",
                    help="genome_se only: other workloads run after it (1 step each, no CPU leg) and "
                         "summarised under other_workloads; empty = none")
    ap.add_argument("--no-host-io", action="store_true", help="skip the host-buffer (fame_match_reads) timing")
    args = ap.parse_args()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gather_probe:
        gather_probe(args.gather_probe, 1 << 30, reps=2)
        return
    if args.gather_summary:
        gather_summary(*args.gather_summary, args.copy_mib, args.gather_lookups, args.pmc_command)
        return
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary, args.copy_mib, args.pmc_command)
        return
    if args.stats_summary:
        stats_summary(args.stats_summary[0], args.stats_summary[1], args.pmc_kernels.split(";"), args.pmc_command)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1:], args.pmc_kernels.split(";"), args.pmc_command)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by read "
              "chunks with a replicated index, which this build does not launch", file=sys.stderr)
        raise SystemExit(2)
    wl = dict(WORKLOADS[args.workload])
    if not args.cpu_sample:
        args.cpu_sample = 300_000 if wl["paired"] else 2_400_000

    import torch  # device memory and synchronisation only

    from fame_amd import _lib, index, synth

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    dev = torch.device("cuda:0")
    t_all = time.time()
    # ---------------- workload ----------------
    lens_c = synth.config_genome("genome", scale=args.genome_scale, seed=args.seed)
    log(f"genome: {len(lens_c)} chromosomes, {sum(lens_c) / 1e6:.1f} Mbp")
    chroms = synth.make_genome(lens_c, seed=args.seed)
    if wl.get("repeat_frac"):
        k = synth.add_repeat_families(chroms, frac=wl["repeat_frac"], seed=args.seed + 5)
        log(f"{k} repeat-family copies placed")
    if wl.get("near_copies"):
        wl["regions"] = synth.add_near_copies(chroms, segments=wl["near_copies"], seed=args.seed + 6)
        log(f"{len(wl['regions'][0])} near-copy regions")
    est = args.htab * 8 + sum(lens_c) * 10 + (1 << 30)
    wd = workdir_for(est, args.workdir)
    idx_path = os.path.join(wd, "genome.idx")
    RL = wl["readlen"]
    try:
        t0 = time.time()
        index.build_index(chroms, idx_path, readlen=RL, htab=args.htab, kmercutoff=1500, noloss=wl["noloss"])
        t_build = time.time() - t0
        log(f"index built in {t_build:.1f}s ({os.path.getsize(idx_path) / 1e9:.2f} GB) in {wd}")
        t0 = time.time()
        reads = make_workload(wl, args, chroms)
        t_reads = time.time() - t0
        paired = wl["paired"]
        if paired:
            r1, r2 = reads
            s1, o1, l1 = r1.flat()
            s2, o2, l2 = r2.flat()
            seqs = np.concatenate([s1, s2])
            offs = np.concatenate([o1, o2 + np.uint64(s1.size)])
            lens = np.concatenate([l1, l2])
            n = len(l1)
        else:
            seqs, offs, lens = reads.flat()
            n = len(lens)
        log(f"{n} {'pairs' if paired else 'reads'} generated in {t_reads:.1f}s")
        if wl.get("serve"):
            out = serve_bench(args, wl, idx_path, reads, lens_c, t_build, t_reads)
            out["wall_s"] = round(time.time() - t_all, 1)
            print(json.dumps(out), flush=True)
            return
        # ---------------- device ----------------
        t0 = time.time()
        ctx = _lib.Context(idx_path)
        if paired:
            ctx.set_pair_dist(wl["minpd"], wl["maxpd"])
        t_load = time.time() - t0
        log(f"index resident in HBM after {t_load:.1f}s: {ctx.info}")
        d_seqs = torch.from_numpy(seqs).to(dev)
        d_offs = torch.from_numpy(offs.view(np.int64)).to(dev)
        d_lens = torch.from_numpy(lens.view(np.int32)).to(dev)
        d_valid = torch.zeros(len(lens), dtype=torch.uint8, device=dev)
        d_mat = torch.zeros(len(lens), dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        calls = []  # per call: reads, stage timings (device ms), work counts
        ncnt = 5 if paired else 3

        def call(s, e, stranded, cnt):
            if paired:
                ctx.match_pairs_device(d_seqs.data_ptr(), d_offs.data_ptr() + 8 * s, d_lens.data_ptr() + 4 * s,
                                       d_offs.data_ptr() + 8 * (n + s), d_lens.data_ptr() + 4 * (n + s), e - s,
                                       stranded, cnt, d_valid.data_ptr() + s, d_mat.data_ptr() + 8 * s,
                                       d_valid.data_ptr() + n + s, d_mat.data_ptr() + 8 * (n + s))
            else:
                ctx.match_device(d_seqs.data_ptr(), d_offs.data_ptr() + 8 * s, d_lens.data_ptr() + 4 * s,
                                 e - s, stranded, cnt, d_valid.data_ptr() + s, d_mat.data_ptr() + 8 * s)
            calls.append((e - s, ctx.timings(), ctx.stats()))

        cur = {"form": "plain"}

        def step():
            # plain: FAME's loop with batches of --batch reads; batched: the
            # rest in one call, which the library splits into pipelined
            # sub-batches on two lanes
            ctx.reset()
            cnt = np.zeros(ncnt, np.uint64)
            e = min(n, CHUNKSIZE)
            call(0, e, True, cnt)
            ctx.decide_strand()
            s = e
            B = n if cur["form"] == "batched" else args.batch
            while s < n:
                e = min(n, s + B)
                call(s, e, False, cnt)
                s = e
            return cnt

        forms = [f for f in args.forms.split(",") if f]
        results = {}
        for form in forms:
            cur["form"] = form
            ctx.set_form(form)
            if args.sub_batch:
                ctx.set_pipeline(args.sub_batch)
            for _ in range(args.warmup):
                step()
            calls.clear()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                cnt = step()
            torch.cuda.synchronize()
            t_run = time.perf_counter() - t0
            ms_f = 1e3 * t_run / max(1, args.steps)
            results[form] = (ms_f, cnt.copy(), list(calls))
            log(f"[{form}] step {ms_f:.1f} ms; counters {cnt.tolist()} over {n} {'pairs' if paired else 'reads'}"
                + (f"; {ctx.form_info()}" if form == "batched" else ""))
        # headline: the fastest form; the per-stage figures come from the plain form
        best = min(results, key=lambda f: results[f][0])
        ms_step, cnt, _ = results[best]
        calls = results[forms[0]][2] if "plain" not in results else results["plain"][2]
        succ = int(cnt[0])
        value = succ / (ms_step / 1e3)
        for f in results:
            assert results[f][1].tolist() == results[best][1].tolist(), "forms disagree"
        K = max(1, args.steps)
        per_stage, per_work = {}, {}
        for _, tm, st in calls:
            for k, v in tm.items():
                per_stage[k] = per_stage.get(k, 0.0) + v / K
            for k, v in st.items():
                per_work[k] = v if k == "tab_bytes" else per_work.get(k, 0) + v
        per_work = {k: (v // K if k != "tab_bytes" else v) for k, v in per_work.items()}
        log("stage ms per step:", {k: round(v, 2) for k, v in per_stage.items()})
        log("work per step:", per_work)
        names = ["succ", "nonunique", "unsucc"] + (["succ_paired", "too_short"] if paired else [])
        out = {
            "metric": METRIC, "value": round(value, 1), "unit": "reads/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic",
            "config": {"workload": wl["desc"], "genome_bp": int(sum(lens_c)), ("pairs" if paired else "reads"): n,
                       "read_len": RL, "batch": args.batch, "first_chunk": CHUNKSIZE, "form": best},
            "forms": {f: {"ms_per_step": round(results[f][0], 3),
                          "aligned_reads_per_s": round(int(results[f][1][0]) / (results[f][0] / 1e3), 1)}
                      for f in results},
            "input_reads_per_s": round((2 if paired else 1) * n / (ms_step / 1e3), 1),
            "counters": dict(zip(names, (int(x) for x in cnt))),
            "stage_ms": {k: round(v, 3) for k, v in per_stage.items()},
            "work": per_work,
            "setup_s": {"index_build": round(t_build, 1), "reads": round(t_reads, 1),
                        "index_load": round(t_load, 1)},
        }
        if "plain" in results and "batched" in results:
            out["batched_vs_plain"] = round(results["plain"][0] / results["batched"][0], 3)
        if not paired and "plain" in results:
            # ---------------- roofline ----------------
            # over the full-size calls (one launch of each kernel per call)
            full = [c for c in calls if c[0] == args.batch] or calls
            dev_stages = {k: per_stage[k] for k in ("prep", "seed", "verify", "select", "meth")}
            # the counter summaries under profiles/ were taken on the default
            # workload's 2 M-read launches
            same = args.workload == "genome_se" and args.batch == 2_000_000 and args.genome_scale == 1.0
            r_seed = roofline_seed(full, same)
            r_verify = roofline_verify(full, same)
            r_dp = roofline_dp(full, same)
            # the seed kernel's bucket lookups set against a plain random gather
            # of as many 8-byte words from a table of the slot table's size
            lookups = int(sum(c[2]["kmers"] for c in full) / len(full))
            seed_rate = lookups / (r_seed["avg_launch_ms"] / 1e3)
            probe_rate = gather_probe(lookups, args.htab)
            r_seed["lookups_per_launch"] = lookups
            r_seed["lookups_per_s"] = round(seed_rate, 1)
            r_seed["gather_probe_lookups_per_s"] = round(probe_rate, 1)
            r_seed["frac_of_gather_probe"] = round(seed_rate / probe_rate, 3)
            # `roofline` (the contract's object) is the HBM roofline of the
            # seed kernel, the largest stage that moves memory; verify and the
            # banded DP are integer-ALU bound and carry their own objects
            roof = dict(r_seed)
            roof["dominant_stage"] = max(dev_stages, key=dev_stages.get)
            verify_rate = per_work["window_bytes"] / (per_stage["verify"] / 1e3) if per_stage.get("verify") else 0
            out["verify_window_GBps"] = round(verify_rate / 1e9, 2)
            out["roofline"] = roof
            out["roofline_seed"] = r_seed
            out["roofline_verify"] = r_verify
            out["roofline_dp"] = r_dp
        if not paired and not args.no_host_io:
            out["host_buffers"] = host_leg(ctx, seqs, offs, lens, n, forms)
        # ---------------- CPU baseline + parity on a sample ----------------
        if not args.no_cpu:
            if paired:
                out["cpu_baseline"], out["parity"] = cpu_leg_pairs(args, wl, ctx, reads, idx_path, wd,
                                                                   len(lens_c))
            else:
                out["cpu_baseline"], out["parity"] = cpu_leg(args, wl, ctx, reads, idx_path, wd, len(lens_c))
        ctx.close()
        del d_seqs, d_offs, d_lens, d_valid, d_mat
        torch.cuda.empty_cache()
        if args.extras and args.workload == "genome_se":
            shutil.rmtree(wd, ignore_errors=True)
            out["other_workloads"] = run_extras(args)
        out["wall_s"] = round(time.time() - t_all, 1)
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


# MI355X figures the rooflines are set against: HBM3E bandwidth; 256 CUs x 4
# SIMDs of 16 lanes at 2.4 GHz, so 39.3 T 32-bit integer lane-ops/s, and one
# wave64 VALU instruction per SIMD every 4 cycles (6.14e11 wave-instructions/s)
CUS, SIMDS, CLOCK_HZ = 256, 1024, 2.4e9
INT_OPS_PEAK = SIMDS * 16 * CLOCK_HZ
VALU_ISSUE_PEAK = SIMDS * CLOCK_HZ / 4

# kernel, its HBM traffic summary, the VALU counter summary (all three kernels)
VALU_PROFILE = "seed_verify_dp_genome300M_2Mreads_valu_counters.md"
PROFILES = {
    # the variant for reads of at most 128 letters (the other returns at once)
    "seed": ("k_seed_light<unsigned int, 128", "k_seed_light_genome300M_2Mreads_traffic.md"),
    "verify": ("k_verify<2>", "k_verify2_genome300M_2Mreads_traffic.md"),
    "dp": ("k_meth_dp", None),
}


def profile_value(fname, key):
    """A `key: value` line of a committed profile summary (JSON value), or None."""
    if not fname:
        return None
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return json.loads(line.split(":", 1)[1])
    return None


def mean_over(calls, f):
    return sum(f(c) for c in calls) / len(calls)


def roofline_seed(calls, same):
    """HBM roofline of k_seed_light (the seed stage of a plain call): the
    algorithmic bytes per (read, orientation) are FAME's own lookups, 16 B
    of tabIndex per k-mer, 8.125 B per index entry of every bucket named
    (kmerTableSmall entry + strand bit) and the read's letters; `design_bytes`
    is what this design must read (one 8-byte slot per k-mer read at random,
    4 B of key and 8 B of slot per k-mer found in the genome-ordered copy,
    4 B per key-position lookup, the k-mer table only for buckets of three
    or more, the read, the candidate records written).  traffic: HBM bytes
    per launch from the committed counter summary (null off the default
    workload)."""
    ms = mean_over(calls, lambda c: c[1]["seed"])
    alg = mean_over(calls, lambda c: 16 * c[2]["kmers"] + 8.125 * c[2]["bucket_entries"] + c[2]["read_bytes"])
    design = mean_over(calls, lambda c: c[2]["tab_bytes"] * c[2]["random_slots"] + 12 * c[2]["ordered_slots"]
                       + 4 * c[2]["entries"] + c[2]["read_bytes"] + 8 * c[2]["jobs"])
    seeded = mean_over(calls, lambda c: c[2]["seeded"]) or 1
    gbs = alg / (ms / 1e3) / 1e9
    traffic = profile_value(PROFILES["seed"][1], "traffic_bytes_per_launch") if same else None
    return {"kernel": "k_seed_light", "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_GBps": round(traffic / (ms / 1e3) / 1e9, 1) if traffic else None,
            "algorithmic_bytes_per_launch": int(alg), "design_bytes_per_launch": int(design),
            "design_GBps": round(design / (ms / 1e3) / 1e9, 2),
            "bucket_entries_per_launch": int(mean_over(calls, lambda c: c[2]["bucket_entries"])),
            "random_slot_reads_per_read_orientation": round(mean_over(calls, lambda c: c[2]["random_slots"]) / seeded, 2),
            "ordered_slot_reads_per_read_orientation": round(mean_over(calls, lambda c: c[2]["ordered_slots"]) / seeded, 2),
            "avg_launch_ms": round(ms, 3), "launches": len(calls),
            "reads_per_launch": int(mean_over(calls, lambda c: c[0]))}


def valu_fields(stage, ms, same):
    """VALU issue fraction, lane utilisation and integer-op rate of a kernel
    from its committed counter summary (per-launch counts, which do not
    depend on timing) over the launch time measured here."""
    cnt = profile_value(VALU_PROFILE, f"counters_per_launch[{PROFILES[stage][0]}]") if same else None
    if not cnt:
        return {"valu_counters": None}
    s = ms / 1e3
    out = {"valu_insts_per_launch": int(cnt["SQ_INSTS_VALU"]),
           "valu_issue_frac": round(cnt["SQ_INSTS_VALU"] / s / VALU_ISSUE_PEAK, 3)}
    if "SQ_THREAD_CYCLES_VALU" in cnt and cnt.get("SQ_ACTIVE_INST_VALU"):
        # AMD's VALUUtilization: active lanes per issued VALU instruction
        util = cnt["SQ_THREAD_CYCLES_VALU"] / (64 * cnt["SQ_ACTIVE_INST_VALU"])
        out["lane_utilisation"] = round(util, 3)
        # active-lane operations per second against 39.3 T lane-ops/s
        out["lane_ops_per_s"] = round(cnt["SQ_INSTS_VALU"] * 64 * util / s, 1)
        out["lane_ops_frac"] = round(cnt["SQ_INSTS_VALU"] * 64 * util / s / INT_OPS_PEAK, 3)
    if "SQ_INSTS_VALU_INT32" in cnt:
        out["int_valu_insts_per_launch"] = int(cnt["SQ_INSTS_VALU_INT32"] + cnt.get("SQ_INSTS_VALU_INT64", 0))
    return out


def roofline_verify(calls, same):
    """k_verify<2> is bound by integer VALU issue.  achieved = the bit-parallel
    scan work FAME's ShiftAnd would do over every candidate window (7 64-bit
    word updates per letter and pattern word, a 64-bit op counted as two
    32-bit ops) per second of this kernel, against 39.3 T 32-bit lane-ops/s;
    the filter lets this kernel skip most of that scan, so frac can pass 1.
    The counter fields give what the kernel itself issues."""
    ms = mean_over(calls, lambda c: c[1]["verify"])
    ops = mean_over(calls, lambda c: 7 * 2 * c[2]["word_letters"])
    r = {"kernel": "k_verify<2>", "bound": "valu", "achieved": round(ops / (ms / 1e3) / 1e12, 3),
         "peak": round(INT_OPS_PEAK / 1e12, 1), "unit": "T int32 ops/s",
         "frac": round(ops / (ms / 1e3) / INT_OPS_PEAK, 4), "algorithmic_ops_per_launch": int(ops),
         "window_bytes_per_launch": int(mean_over(calls, lambda c: c[2]["window_bytes"])),
         "windows_per_launch": int(mean_over(calls, lambda c: c[2]["jobs"])),
         "traffic": profile_value(PROFILES["verify"][1], "traffic_bytes_per_launch") if same else None,
         "avg_launch_ms": round(ms, 3), "launches": len(calls),
         "reads_per_launch": int(mean_over(calls, lambda c: c[0]))}
    r.update(valu_fields("verify", ms, same))
    return r


def roofline_dp(calls, same):
    """k_meth_dp (banded DP of the reads matched with errors): (L + 1) x 17
    band cells per read, as GCUPS over this kernel's own time."""
    ms = mean_over(calls, lambda c: c[1]["dp"])
    cells = mean_over(calls, lambda c: 17 * (c[2]["dp_rows"] + c[2]["dp"]))
    r = {"kernel": "k_meth_dp", "bound": "valu", "gcups": round(cells / (ms / 1e3) / 1e9, 2) if ms else None,
         "cells_per_launch": int(cells), "alignments_per_launch": int(mean_over(calls, lambda c: c[2]["dp"])),
         "avg_launch_ms": round(ms, 3), "launches": len(calls)}
    r.update(valu_fields("dp", ms, same))
    return r


def stats_summary(out_md, d, kernels, command):
    """Markdown of a rocprofv3 --kernel-trace --stats run: every kernel's
    calls and durations, then each named kernel's launches in dispatch order."""
    import csv
    import glob
    st = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    tr = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if not st or not tr:
        raise SystemExit(f"no kernel_stats / kernel_trace CSV under {d}")
    rows = list(csv.DictReader(open(st[0])))
    with open(out_md, "w") as f:
        f.write(f"# Kernel times (rocprofv3 --kernel-trace --stats)\n\n")
        if command:
            f.write("Command (from /tmp with TMPDIR=/tmp):\n\n```\n" + command + "\n```\n\n")
        f.write("| Kernel | Calls | Total ms | Avg ms | Min ms | Max ms | % |\n|---|---:|---:|---:|---:|---:|---:|\n")
        for r in rows:
            name = r["Name"][5:] if r["Name"].startswith("void ") else r["Name"]
            name = name if len(name) <= 100 else name[:97] + "..."
            f.write(f"| `{name}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                    f"{float(r['AverageNs']) / 1e6:.3f} | {float(r['MinNs']) / 1e6:.3f} | "
                    f"{float(r['MaxNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |\n")
        launches = {}
        for r in csv.DictReader(open(tr[0])):
            for k in kernels:
                if k in r["Kernel_Name"]:
                    launches.setdefault(k, []).append(
                        (int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6))
        f.write("\nPer-launch durations (ms), in dispatch order:\n\n")
        for k, v in launches.items():
            v.sort()
            f.write(f"- `{k}`: " + ", ".join(f"{ms:.3f}" for _, ms in v) + "\n")
    log(f"wrote {out_md}")


def _counter_by_dispatch(path, counter, match):
    import csv
    import collections
    d = collections.defaultdict(float)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter and match in r["Kernel_Name"]:
            d[int(r["Dispatch_Id"])] += float(r["Counter_Value"])
    return [d[k] for k in sorted(d)]


def traffic_summary(out_md, kernel, fetch_csv, write_csv, cfetch_csv, cwrite_csv, copy_mib, command):
    """HBM bytes per launch of `kernel` (its last dispatch: the 2 M-read
    launch) from FETCH_SIZE and WRITE_SIZE (KiB), each scaled by the ratio of
    the bytes a plain copy moves to what the same counter reads for it."""
    kib = copy_mib * 1024
    cf = _counter_by_dispatch(cfetch_csv, "FETCH_SIZE", "neg")
    cw = _counter_by_dispatch(cwrite_csv, "WRITE_SIZE", "neg")
    if not cf or not cw:
        raise SystemExit("no copy-probe dispatches in the calibration CSVs")
    f_scale = kib / (sum(cf) / len(cf))
    w_scale = kib / (sum(cw) / len(cw))
    kf = _counter_by_dispatch(fetch_csv, "FETCH_SIZE", kernel)
    kw = _counter_by_dispatch(write_csv, "WRITE_SIZE", kernel)
    if not kf or not kw:
        raise SystemExit(f"no dispatch of {kernel} in the kernel CSVs")
    fb, wb = kf[-1] * 1024 * f_scale, kw[-1] * 1024 * w_scale
    with open(out_md, "w") as f:
        f.write(f"# HBM traffic of `{kernel}`: its last launch in the profiled run\n\n")
        if command:
            f.write("Commands (from /tmp with TMPDIR=/tmp; each counter in its own run):\n\n```\n" + command
                    + "\n```\n\n")
        f.write(f"Calibration: a plain copy (`torch.neg` over {copy_mib} MiB of int32: {kib} KiB read and {kib} "
                f"KiB written) reads FETCH_SIZE = {sum(cf) / len(cf):.0f} KiB and WRITE_SIZE = "
                f"{sum(cw) / len(cw):.0f} KiB per launch (mean of {len(cf)} / {len(cw)} launches): FETCH_SIZE "
                f"is scaled by {f_scale:.3f} and WRITE_SIZE by {w_scale:.3f}.\n\n")
        f.write("| Counter | KiB per launch | scaled bytes |\n|---|---:|---:|\n")
        f.write(f"| FETCH_SIZE | {kf[-1]:.0f} | {fb:.0f} |\n| WRITE_SIZE | {kw[-1]:.0f} | {wb:.0f} |\n\n")
        f.write(f"traffic_bytes_per_launch: {int(fb + wb)}\n")
    log(f"wrote {out_md}")


def gather_summary(out_md, gather_csv, cfetch_csv, copy_mib, lookups, command):
    """FETCH_SIZE of the random-gather probe (torch gather kernel: `lookups`
    8-byte indexes read in order, as many random 8-byte table words, 8-byte
    results written), per lookup, beside the copy's calibration."""
    kib = copy_mib * 1024
    cf = _counter_by_dispatch(cfetch_csv, "FETCH_SIZE", "neg")
    f_scale = kib / (sum(cf) / len(cf))
    g = _counter_by_dispatch(gather_csv, "FETCH_SIZE", "scatter_gather")
    if not g:
        raise SystemExit("no gather dispatches")
    g_kib = sum(g) / len(g)
    idx_kib = 8 * lookups / 1024 / f_scale  # the sequential index stream, as the counter reads it
    per = (g_kib - idx_kib) * 1024 / lookups
    with open(out_md, "w") as f:
        f.write("# FETCH_SIZE of random 8-byte gathers (calibration check)\n\n")
        if command:
            f.write("Commands (from /tmp with TMPDIR=/tmp):\n\n```\n" + command + "\n```\n\n")
        f.write(f"The probe gathers {lookups} words at uniform random positions of a 2^30-word (8 GiB) "
                f"table with torch's gather kernel, reading the {8 * lookups} bytes of indexes in order "
                f"and writing as many bytes of results.\n\n")
        f.write(f"| | KiB per launch |\n|---|---:|\n| FETCH_SIZE (mean of {len(g)} launches) | {g_kib:.0f} |\n"
                f"| index stream at the copy's factor {f_scale:.3f} | {idx_kib:.0f} |\n\n")
        f.write(f"counter_bytes_per_random_lookup: {per:.2f}\n\n")
        f.write(f"The counter reads {per:.1f} bytes per random lookup: one 64-byte unit, so one read "
                f"request per lookup.  The copy reads at 1/{f_scale:.0f} of the bytes it moves, one unit per "
                f"{64 * f_scale:.0f} bytes.  If the read requests to HBM are all of one size, that size is "
                f"{64 * f_scale:.0f} bytes, and the factor {f_scale:.3f} holds for random gathers too "
                f"({64 * f_scale:.0f} bytes of HBM traffic per independent lookup); the traffic figures "
                f"here use it.\n")
    log(f"wrote {out_md}")


def pmc_summary(out_md, csvs, kernels, command):
    """Per-launch counters of the last dispatch of each kernel (the 2 M-read
    launch of `--reads 2300000 --steps 1 --warmup 0 --forms plain`), summed
    over XCDs, shader engines and SIMDs as rocprofv3 reports them."""
    import csv
    import collections
    per = collections.defaultdict(lambda: collections.defaultdict(float))  # (kernel, dispatch) -> counter
    for path in csvs:
        for row in csv.DictReader(open(path)):
            name = row["Kernel_Name"]
            for k in kernels:
                if k in name:
                    per[(k, int(row["Dispatch_Id"]))][row["Counter_Name"]] += float(row["Counter_Value"])
    with open(out_md, "w") as f:
        f.write("# VALU counters per launch (rocprofv3 --pmc, one pass per counter group)\n\n")
        if command:
            f.write("Command (from /tmp with TMPDIR=/tmp):\n\n```\n" + command + "\n```\n\n")
        for k in kernels:
            if not any(kk == k for (kk, d) in per):
                continue
            # counters from several runs: the last dispatch of the kernel in each
            last = {}
            for (kk, d), cs in per.items():
                if kk != k:
                    continue
                for c, v in cs.items():
                    if c not in last or d > last[c][0]:
                        last[c] = (d, v)
            cnt = {c: v for c, (d, v) in sorted(last.items())}
            f.write(f"## `{k}`\n\n| Counter | per launch |\n|---|---:|\n")
            for c, v in cnt.items():
                f.write(f"| {c} | {v:.6g} |\n")
            f.write(f"\ncounters_per_launch[{k}]: " + json.dumps({c: round(v) for c, v in cnt.items()}) + "\n\n")
    log(f"wrote {out_md}")


def copy_probe(mib):
    """A plain elementwise copy (torch.neg on int32: reads and writes `mib`
    MiB once each), three launches; used to scale the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {mib} MiB read + {mib} MiB written")


def gather_probe(lookups, words, reps=3):
    """torch.index_select of `lookups` 8-byte words at uniform random positions
    of a `words`-word device table: the rate (lookups/s, best of `reps` after
    one warm-up) at which a plain gather kernel reads independent random words
    of a table that size, for comparison with the seed stage's bucket lookups."""
    import torch
    tab = torch.zeros(words, dtype=torch.int64, device="cuda:0")
    idx = torch.randint(0, words, (lookups,), dtype=torch.int64, device="cuda:0")
    out = torch.empty(lookups, dtype=torch.int64, device="cuda:0")
    best = None
    for i in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.index_select(tab, 0, idx, out=out)
        e1.record()
        e1.synchronize()
        if i:
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
    del tab, idx, out
    torch.cuda.empty_cache()
    log(f"gather probe: {lookups} random 8 B reads of a {words}-word table in {best:.3f} ms")
    return lookups / (best / 1e3)


def host_leg(ctx, seqs, offs, lens, n, forms):
    """The whole job through fame_match_reads with the reads in pinned host
    memory and FAME's CHUNKSIZE of 300 000 reads per call (what FAME's loop
    gets through the drop-in): host-to-device copies of the reads and the
    copies of the per-read results back are inside the timed region.  One
    untimed pass, then one timed pass per form."""
    import torch
    pin = torch.empty(seqs.size, dtype=torch.uint8, pin_memory=True)
    pin.numpy()[:] = seqs
    hs = pin.numpy()
    res = {"reads_per_call": CHUNKSIZE, "reads": n}
    for form in forms:
        ctx.set_form(form)
        for rep in range(2):
            ctx.reset()
            cnt = np.zeros(3, np.uint64)
            t0 = time.perf_counter()
            e = min(n, CHUNKSIZE)
            ctx.match(hs, offs[:e], lens[:e], True, cnt)
            ctx.decide_strand()
            s = e
            while s < n:
                e = min(n, s + CHUNKSIZE)
                ctx.match(hs, offs[s:e], lens[s:e], False, cnt)
                s = e
            dt = time.perf_counter() - t0
        res[form] = {"ms_per_step": round(dt * 1e3, 3), "aligned_reads_per_s": round(int(cnt[0]) / dt, 1),
                     "input_reads_per_s": round(n / dt, 1)}
        log(f"[host buffers, {form}] {dt * 1e3:.1f} ms for {n} reads in calls of {CHUNKSIZE}")
    del pin
    return res


def host_facts():
    """CPU model, threads visible and the C++ runtime of this host."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    libstdcxx = None
    for d in ("/usr/lib/x86_64-linux-gnu", "/usr/lib64", "/lib/x86_64-linux-gnu"):
        p = os.path.join(d, "libstdc++.so.6")
        if os.path.exists(p):
            libstdcxx = os.path.basename(os.path.realpath(p))
            break
    gxx = None
    try:
        gxx = subprocess.run(["g++", "-dumpfullversion"], capture_output=True, text=True).stdout.strip() or None
    except OSError:
        pass
    return {"cpu_model": model, "host_threads": os.cpu_count(), "libstdcxx": libstdcxx, "gxx": gxx}


def chunk_rates(st):
    """FAME's per-chunk times (the harness's steady_clock around each
    matchReads call): the first chunk (both orientations) and the stranded
    ones separately, in input reads/s."""
    ch = st.get("chunks", [])
    first = [c for c in ch if c["stranded"]]
    rest = [c for c in ch if not c["stranded"] and c["n"]]
    out = {}
    if first:
        out["first_chunk_reads_per_s"] = round(sum(c["n"] for c in first) / sum(c["s"] for c in first), 1)
    if rest:
        out["stranded_reads_per_s"] = round(sum(c["n"] for c in rest) / sum(c["s"] for c in rest), 1)
    return out


def run_extras(args):
    """The other workloads, each in a child process (1 warm-up + 1 timed step,
    no CPU leg), summarised."""
    out = {}
    for w in [x for x in args.extras.split(",") if x]:
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", w, "--steps", "1", "--warmup", "1",
               "--no-cpu", "--no-host-io", "--extras", "", "--seed", str(args.seed)]
        log(f"other workload {w}: {' '.join(cmd[1:])}")
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True)
        lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
        if r.returncode != 0 or not lines:
            log(f"{w} failed (rc {r.returncode}):", r.stderr[-800:])
            out[w] = {"error": f"rc {r.returncode}"}
            continue
        d = json.loads(lines[-1])
        keep = ("value", "unit", "ms_per_step", "forms", "batched_vs_plain", "counters", "stage_ms",
                "p50_ms", "p99_ms", "batched_peak_reads_per_s", "plain_peak_reads_per_s", "host_buffers")
        out[w] = {"workload": d["config"]["workload"], **{k: d[k] for k in keep if k in d}}
        if "sweep" in d:
            out[w]["sweep"] = [{"load": row["load"], "batched_p50_ms": row["batched"]["p50_ms"],
                                "batched_p99_ms": row["batched"]["p99_ms"], "plain_p50_ms": row["plain"]["p50_ms"],
                                "plain_p99_ms": row["plain"]["p99_ms"]} for row in d["sweep"]]
        out[w]["wall_s"] = round(time.time() - t0, 1)
        log(f"{w}: {out[w].get('value')} {out[w].get('unit')}")
    return out


def cpu_leg(args, wl, ctx, reads, idx_path, wd, nchrom):
    """FAME's own CPU path on the first `cpu_sample` reads of the workload, and
    the HIP path on the same sample for a per-read / report comparison."""
    S = min(args.cpu_sample, len(reads))
    binp = ref_binary(args.htab, wl["readlen"], args.cpu_cores, nchrom)
    sample = f"first {S} reads of the workload (chunk 1 of {CHUNKSIZE} in both orientations, the rest stranded)"
    if binp is None:
        log("reference CPU binary not built; cpu_baseline skipped")
        return None, None
    fq = os.path.join(wd, "sample.fq")
    write_fastq(fq, reads.seqs[:S], reads.lens[:S])
    log(f"CPU reference on {S} reads with {args.cpu_cores} threads ...")
    t0 = time.time()
    r = subprocess.run([binp, "match", idx_path, fq, os.path.join(wd, "cpu")], cwd=wd,
                       capture_output=True, text=True)
    if r.returncode != 0:
        log("reference run failed:", r.stderr[-500:])
        return None, None
    st = json.load(open(os.path.join(wd, "cpu_stats.json")))
    log(f"CPU reference: {st['succ']} matched in {st['match_s']:.2f}s (load {st['load_s']:.1f}s, "
        f"wall {time.time() - t0:.1f}s)")
    cpu = {"value": round(st["succ"] / st["match_s"], 1), "unit": "reads/s", "cores": args.cpu_cores,
           "kind": "reference", "sample": sample, "match_s": round(st["match_s"], 3),
           "reads": int(st["reads"]), "succ": int(st["succ"]), **chunk_rates(st), **host_facts()}
    if "stranded_reads_per_s" in cpu and "first_chunk_reads_per_s" in cpu:
        # the whole job's mix (one first chunk, the rest stranded) at the
        # sample's rates and aligned fraction
        N = args.reads or wl["n"]
        t = CHUNKSIZE / cpu["first_chunk_reads_per_s"] + (N - CHUNKSIZE) / cpu["stranded_reads_per_s"]
        cpu["whole_job_estimate_aligned_reads_per_s"] = round(N * st["succ"] / st["reads"] / t, 1)
    # HIP path on the same sample, fresh state
    raw = open(os.path.join(wd, "cpu_reads.bin"), "rb").read()
    nr = int(np.frombuffer(raw[8:16], np.uint64)[0])
    rvalid = np.frombuffer(raw[16:16 + nr], np.uint8)
    rmat = np.frombuffer(raw[16 + nr:16 + 9 * nr], np.uint64)
    ctx.reset()
    sub = reads.subset(np.arange(S))
    seqs, offs, lens = sub.flat()
    cnt = np.zeros(3, np.uint64)
    valid = np.zeros(S, np.uint8)
    mat = np.zeros(S, np.uint64)
    e = min(S, CHUNKSIZE)
    valid[:e], mat[:e] = ctx.match(seqs, offs[:e], lens[:e], True, cnt)
    ctx.decide_strand()
    s = e
    while s < S:
        e = min(S, s + CHUNKSIZE)
        valid[s:e], mat[s:e] = ctx.match(seqs, offs[s:e], lens[s:e], False, cnt)
        s = e
    base = os.path.join(wd, "hip")
    ctx.print_methylation_levels(base)
    same_tsv = open(base + "_cpg.tsv", "rb").read() == open(os.path.join(wd, "cpu_cpg.tsv"), "rb").read()
    mism = int(np.count_nonzero((valid != rvalid) | (mat != rmat)))
    parity = {"reads": S, "per_read_mismatches": mism, "cpg_report_identical": bool(same_tsv),
              "counters_equal": cnt.tolist() == [st["succ"], st["nonunique"], st["unsucc"]]}
    log("parity on the sample:", parity)
    return cpu, parity


def chunk_sizes(rng, k):
    """log-normal chunk sizes, median 20k reads, sigma 1, clipped to [1k, 300k]"""
    return np.clip(np.round(rng.lognormal(np.log(20000), 1.0, size=k)), 1000, 300000).astype(np.int64)


def run_stream(submit, wait, arrivals, sizes, starts):
    """Submits chunk i at arrivals[i] (s from now) from this thread and waits
    for completions in another; returns per-chunk latency (ms, from the
    scheduled arrival) and results."""
    import threading
    import queue
    q = queue.Queue()
    res = [None] * len(sizes)

    def collector():
        while True:
            item = q.get()
            if item is None:
                return
            i, tk, t_arr = item
            res[i] = wait(tk, t_arr)

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    for i, (a, k, s0) in enumerate(zip(arrivals, sizes, starts)):
        now = time.perf_counter() - t0
        if a > now:
            time.sleep(a - now)
        q.put((i, submit(int(s0), int(k)), t0 + a))
    q.put(None)
    th.join()
    return res, time.perf_counter() - t0


def serve_bench(args, wl, idx_path, reads, lens_c, t_build, t_reads):
    """Continuous-batching server vs the plain form on a Poisson stream."""
    import torch
    from fame_amd import _lib
    seqs, offs, lens = reads.flat()
    # the read pool in pinned host memory: chunks are copied to the device
    # straight from it
    pin = torch.empty(seqs.size, dtype=torch.uint8, pin_memory=True)
    pin.numpy()[:] = seqs
    seqs = pin.numpy()
    n = len(lens)
    rng = np.random.default_rng(args.seed + 9)

    def new_ctx():
        ctx = _lib.Context(idx_path)
        cnt = np.zeros(3, np.uint64)
        ctx.match(seqs, offs[:CHUNKSIZE], lens[:CHUNKSIZE], True, cnt)  # the strand decision chunk
        ctx.decide_strand()
        return ctx

    def plan(k_req, rate_chunks):
        sz = chunk_sizes(rng, k_req)
        st = rng.integers(CHUNKSIZE, n - sz, size=k_req)
        arr = np.cumsum(rng.exponential(1.0 / rate_chunks, size=k_req)) if rate_chunks else np.zeros(k_req)
        return arr, sz, st

    mean_chunk = float(np.mean(chunk_sizes(np.random.default_rng(1), 100000)))
    # batched peak: every chunk queued at once
    ctx = new_ctx()
    ctx.set_form(args.serve_form)
    srv = _lib.Server(ctx)
    tick = {}

    def s_submit(s0, k):
        return srv.submit(seqs, offs[s0:s0 + k], lens[s0:s0 + k])

    def s_wait(tk, t_arr):
        v, m, cnt, tm = srv.wait(tk)
        return (time.perf_counter() - t_arr) * 1e3, int(cnt[0]), tm["batch_reads"]

    arr, sz, st = plan(300, 0)
    srv.wait(s_submit(int(st[0]), int(sz[0])))  # warm-up
    res, wall = run_stream(s_submit, s_wait, arr, sz, st)
    peak_reads = float(sz.sum()) / wall
    peak_aligned = sum(r[1] for r in res) / wall
    log(f"batched peak: {peak_reads / 1e6:.1f} M reads/s ({peak_aligned / 1e6:.1f} M aligned), "
        f"{srv.stats()}")
    # plain form: one call per chunk, in arrival order
    pctx = new_ctx()

    def p_submit(s0, k):
        return (s0, k)

    def p_wait(tk, t_arr):
        s0, k = tk
        cnt = np.zeros(3, np.uint64)
        pctx.match(seqs, offs[s0:s0 + k], lens[s0:s0 + k], False, cnt)
        return (time.perf_counter() - t_arr) * 1e3, int(cnt[0]), k

    arr, sz, st = plan(100, 0)
    res, wall = run_stream(p_submit, p_wait, arr, sz, st)
    plain_peak = float(sz.sum()) / wall
    log(f"plain peak: {plain_peak / 1e6:.1f} M reads/s")
    sweep = []
    for frac in (0.2, 0.4, 0.6, 0.8, 1.0):
        rate_reads = frac * peak_reads
        rate_chunks = rate_reads / mean_chunk
        k_req = int(min(3000, max(60, rate_chunks * args.serve_seconds)))
        arr, sz, st = plan(k_req, rate_chunks)
        row = {"load": frac, "offered_reads_per_s": round(rate_reads, 1), "chunks": k_req}
        for form, sub, wt in (("batched", s_submit, s_wait), ("plain", p_submit, p_wait)):
            res, wall = run_stream(sub, wt, arr, sz, st)
            lat = np.array([r[0] for r in res])
            row[form] = {"served_reads_per_s": round(float(sz.sum()) / wall, 1),
                         "aligned_reads_per_s": round(sum(r[1] for r in res) / wall, 1),
                         "p50_ms": round(float(np.percentile(lat, 50)), 2),
                         "p99_ms": round(float(np.percentile(lat, 99)), 2),
                         "mean_batch_reads": round(float(np.mean([r[2] for r in res])), 1)}
        log(f"load {frac:.1f}: batched p50 {row['batched']['p50_ms']} p99 {row['batched']['p99_ms']} ms, "
            f"plain p50 {row['plain']['p50_ms']} p99 {row['plain']['p99_ms']} ms")
        sweep.append(row)
    srv.close()
    ctx.close()
    pctx.close()
    top = sweep[-1]["batched"]
    return {
        "metric": METRIC, "value": top["aligned_reads_per_s"], "unit": "reads/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": wl["desc"], "genome_bp": int(sum(lens_c)), "read_pool": n, "read_len": 100,
                   "form": f"serving (continuous batching, {args.serve_form} batches) vs plain"},
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "batched_peak_reads_per_s": round(peak_reads, 1), "plain_peak_reads_per_s": round(plain_peak, 1),
        "sweep": sweep,
        "setup_s": {"index_build": round(t_build, 1), "reads": round(t_reads, 1)},
    }


def read_pairs_bin(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"FAMEPE01"
    n = int(np.frombuffer(raw[8:16], np.uint64)[0])
    v1 = np.frombuffer(raw[16:16 + n], np.uint8)
    v2 = np.frombuffer(raw[16 + n:16 + 2 * n], np.uint8)
    m1 = np.frombuffer(raw[16 + 2 * n:16 + 10 * n], np.uint64)
    m2 = np.frombuffer(raw[16 + 10 * n:16 + 18 * n], np.uint64)
    return v1, m1, v2, m2


def cpu_leg_pairs(args, wl, ctx, pairs, idx_path, wd, nchrom):
    """FAME's paired CPU path (queryRoutinePaired) on the first `cpu_sample`
    pairs, and the HIP path on the same sample for a per-pair comparison."""
    r1, r2 = pairs
    S = min(args.cpu_sample, len(r1))
    binp = ref_binary(args.htab, wl["readlen"], args.cpu_cores, nchrom, wl["minpd"], wl["maxpd"])
    sample = (f"first {S} pairs of the workload (chunk 1 of {CHUNKSIZE} in both combinations, the rest "
              "stranded)")
    if binp is None:
        log("reference CPU binary not built; cpu_baseline skipped")
        return None, None
    fq1, fq2 = os.path.join(wd, "s1.fq"), os.path.join(wd, "s2.fq")
    write_fastq(fq1, r1.seqs[:S], r1.lens[:S])
    write_fastq(fq2, r2.seqs[:S], r2.lens[:S])
    log(f"CPU reference on {S} pairs with {args.cpu_cores} threads ...")
    t0 = time.time()
    r = subprocess.run([binp, "match_pe", idx_path, fq1, fq2, os.path.join(wd, "cpu")], cwd=wd,
                       capture_output=True, text=True)
    if r.returncode != 0:
        log("reference run failed:", r.stderr[-500:])
        return None, None
    st = json.load(open(os.path.join(wd, "cpu_stats.json")))
    log(f"CPU reference: {st['succ']} matched reads in {st['match_s']:.2f}s (load {st['load_s']:.1f}s, "
        f"wall {time.time() - t0:.1f}s)")
    cpu = {"value": round(st["succ"] / st["match_s"], 1), "unit": "reads/s", "cores": args.cpu_cores,
           "kind": "reference", "sample": sample, "match_s": round(st["match_s"], 3),
           "pairs": int(st["reads"]), "succ": int(st["succ"]), **chunk_rates(st), **host_facts()}
    rv1, rm1, rv2, rm2 = read_pairs_bin(os.path.join(wd, "cpu_pairs.bin"))
    ctx.reset()
    a = r1.subset(np.arange(S)).flat()
    b = r2.subset(np.arange(S)).flat()
    cnt = np.zeros(5, np.uint64)
    v1 = np.zeros(S, np.uint8)
    v2 = np.zeros(S, np.uint8)
    m1 = np.zeros(S, np.uint64)
    m2 = np.zeros(S, np.uint64)
    s = 0
    stranded = True
    while s < S:
        e = min(S, s + CHUNKSIZE)
        v1[s:e], m1[s:e], v2[s:e], m2[s:e] = ctx.match_pairs((a[0], a[1][s:e], a[2][s:e]),
                                                             (b[0], b[1][s:e], b[2][s:e]), stranded, cnt)
        if stranded:
            ctx.decide_strand()
            stranded = False
        s = e
    base = os.path.join(wd, "hip")
    ctx.print_methylation_levels(base)
    same_tsv = open(base + "_cpg.tsv", "rb").read() == open(os.path.join(wd, "cpu_cpg.tsv"), "rb").read()
    mism = int(np.count_nonzero((v1 != rv1) | (v2 != rv2) | (m1 != rm1) | (m2 != rm2)))
    parity = {"pairs": S, "per_pair_mismatches": mism, "cpg_report_identical": bool(same_tsv),
              "counters_equal": cnt.tolist() == [st["succ"], st["nonunique"], st["unsucc"], st["succ_paired"],
                                                 st["too_short"]]}
    log("parity on the sample:", parity)
    return cpu, parity


if __name__ == "__main__":
    main()
