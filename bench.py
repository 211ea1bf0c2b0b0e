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
    args = ap.parse_args()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by read "
              "chunks with a replicated index, which this build does not launch", file=sys.stderr)
        raise SystemExit(2)
    wl = dict(WORKLOADS[args.workload])
    if not args.cpu_sample:
        args.cpu_sample = 300_000 if wl["paired"] else 600_000

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

        forms = ["plain"] if paired else [f for f in args.forms.split(",") if f]
        results = {}
        for form in forms:
            cur["form"] = form
            if not paired:
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
            # `roofline` is the kernel of the larger of the two dominant stages
            # (seed, verify); both are also given under their own names
            r_seed, r_verify = roofline("seed", full, same), roofline("verify", full, same)
            # the seed kernel's bucket lookups set against a plain random gather
            # of as many 8-byte words from a table of the slot table's size
            lookups = int(sum(c[2]["kmers"] for c in full) / len(full))
            seed_rate = lookups / (r_seed["avg_launch_ms"] / 1e3)
            probe_rate = gather_probe(lookups, args.htab)
            r_seed["lookups_per_launch"] = lookups
            r_seed["lookups_per_s"] = round(seed_rate, 1)
            r_seed["gather_probe_lookups_per_s"] = round(probe_rate, 1)
            r_seed["frac_of_gather_probe"] = round(seed_rate / probe_rate, 3)
            dominant = max(dev_stages, key=dev_stages.get)
            roof = dict(r_verify if dev_stages["verify"] > dev_stages["seed"] else r_seed)
            roof["dominant_stage"] = dominant
            verify_rate = per_work["window_bytes"] / (per_stage["verify"] / 1e3) if per_stage.get("verify") else 0
            out["verify_window_GBps"] = round(verify_rate / 1e9, 2)
            out["roofline"] = roof
            out["roofline_seed"] = r_seed
            out["roofline_verify"] = r_verify
        # ---------------- CPU baseline + parity on a sample ----------------
        if not args.no_cpu:
            if paired:
                out["cpu_baseline"], out["parity"] = cpu_leg_pairs(args, wl, ctx, reads, idx_path, wd,
                                                                   len(lens_c))
            else:
                out["cpu_baseline"], out["parity"] = cpu_leg(args, wl, ctx, reads, idx_path, wd, len(lens_c))
        ctx.close()
        out["wall_s"] = round(time.time() - t_all, 1)
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def algorithmic_bytes(stage, st):
    """Bytes the stage's algorithm must move for one call, from the kernels'
    own work counters: seed = the bucket lookup of every k-mer (one 8-byte
    slot, which holds buckets of one or two entries), the index entries read
    from the k-mer table (buckets of three or more), the read's letters and
    the candidate records written; verify = the window text of every
    candidate (WINLEN + MISCOUNT + ADDMIS - 1 letters, as FAME's ShiftAnd
    scans it)."""
    if stage == "seed":
        return st["tab_bytes"] * st["kmers"] + 4 * st["entries"] + st["read_bytes"] + 8 * st["jobs"]
    return st["window_bytes"]


TRAFFIC = {"seed": ("k_seed_light", "k_seed_light_genome300M_2Mreads_traffic.md"),
           "verify": ("k_verify<2>", "k_verify2_genome300M_2Mreads_traffic.md")}


def measured_traffic(stage):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", TRAFFIC[stage][1])
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1].strip())
    return None


def roofline(stage, calls, traffic_known=True):
    """traffic_known: the committed counter summary was taken on this
    workload (the default one); otherwise traffic is null."""
    n = len(calls)
    ms = sum(c[1][stage] for c in calls) / n
    nbytes = sum(algorithmic_bytes(stage, c[2]) for c in calls) / n
    gbs = nbytes / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    traffic = measured_traffic(stage) if traffic_known else None
    return {"kernel": TRAFFIC[stage][0], "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_GBps": round(traffic / (ms / 1e3) / 1e9, 1) if traffic and ms > 0 else None,
            "algorithmic_bytes_per_launch": int(nbytes), "avg_launch_ms": round(ms, 3),
            "launches": n, "reads_per_launch": int(sum(c[0] for c in calls) / n)}


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
           "reads": int(st["reads"]), "succ": int(st["succ"])}
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
           "pairs": int(st["reads"]), "succ": int(st["succ"])}
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
