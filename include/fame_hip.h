/* fame_hip.h -- C ABI of the MI355X read-matching library (libfamehip.so).
 *
 * The reference's matching boundary is a C++ member API (FischerJo/FAME
 * ReadQueue.h:58-129, RefGenome.h:55).  Each entry point below replaces one
 * of its members; the C++ drop-in (fame_amd/dropin/ReadQueue.h and
 * RefGenome.h, linked with FAME's own main.cpp) and the Python mirror in
 * fame_amd/readqueue.py call only these functions.
 *
 * Conventions: integer status codes (FAME_OK = 0, negative on error, message
 * in fame_last_error); caller-owned host buffers; device memory owned by the
 * context; one host thread per context.  Counters are accumulated (+=) into
 * the caller's array like the reference's uint64_t& arguments
 * (ReadQueue.cpp:926-931).
 */
#ifndef FAME_HIP_H
#define FAME_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fame_ctx fame_ctx;

enum {
    FAME_OK = 0,
    FAME_ERR_IO = -1,       /* index file missing / truncated */
    FAME_ERR_FORMAT = -2,   /* header values this library cannot serve */
    FAME_ERR_HIP = -3,      /* a HIP runtime call failed */
    FAME_ERR_ARG = -4,      /* bad arguments (e.g. read longer than FAME_MAX_READ) */
    FAME_ERR_CAPACITY = -5  /* a per-read structure outgrew the device limits */
};

/* Single-end reads of any length are matched as FAME matches them (it has no
 * limit): up to 255 letters by the short-read kernels, up to FAME_MAX_READ by
 * the long-read path (plain form; the batched form re-runs a batch holding
 * one on the plain path), and longer ones are reported unmatched without
 * being seeded when they exceed the index's window text plus the errors
 * allowed (WINLEN + 12 letters: no window can hold a match, as FAME's
 * ShiftAnd finds none).  A longer read that could match (an index with
 * WINLEN over FAME_MAX_READ - 12) makes the call fail with FAME_ERR_ARG.
 * Paired mates are limited to FAME_MAX_MATE letters (FAME_ERR_ARG beyond),
 * and so is fame_match_reads_noalloc. */
#define FAME_MAX_READ 4096
#define FAME_MAX_MATE 255

/* index header / table sizes, in this order:
 * htab, readlen, winlen, kmerlen, kmercutoff, seedbits, ncpg, nchrom, nkmer,
 * nwin, nfiltered */
#define FAME_INFO_FIELDS 11

/* RefGenome(std::string filePath) RefGenome.cpp:77-80 (load :230-387 plus the
 * <index>_strands companion :414-433) followed by the single-end ReadQueue
 * constructor ReadQueue.cpp:25-78: loads the --load_index file, validates
 * the header and makes the index resident in HBM on the current device.
 * both_strands is FAME's --unord_reads flag (bsFlag).  On failure *out
 * still receives a context that only carries the message for
 * fame_last_error; it must be released with fame_close. */
int fame_open(const char* index_path, int both_strands, fame_ctx** out);

/* releases device and host memory */
int fame_close(fame_ctx* ctx);

/* message of the last failing call (never NULL) */
const char* fame_last_error(const fame_ctx* ctx);

int fame_index_info(const fame_ctx* ctx, uint64_t out[FAME_INFO_FIELDS]);

/* ReadQueue::matchReads ReadQueue.cpp:404-933.  Reads are n byte strings
 * seqs[offs[i] .. offs[i]+lens[i]).  counters[0..2] += succMatch,
 * nonUniqueMatch, unSuccMatch.  valid[i] = !Read::isInvalid and mat[i] =
 * Read::mat (0 for invalid reads).  valid / mat may be NULL. */
int fame_match_reads(fame_ctx* ctx, const uint8_t* seqs, const uint64_t* offs,
                     const uint32_t* lens, uint32_t n, int get_stranded, uint64_t counters[3],
                     uint8_t* valid, uint64_t* mat);

/* The same with every buffer already in device memory (HBM); d_valid /
 * d_mat receive the per-read results on the device.  `stream` is a
 * hipStream_t (NULL = the context's own stream); the call returns after the
 * work completed.  No counterpart in the reference: it is the resident-input
 * form of matchReads used by benchmarks and serving loops. */
int fame_match_reads_device(fame_ctx* ctx, const uint8_t* d_seqs, const uint64_t* d_offs,
                            const uint32_t* d_lens, uint32_t n, int get_stranded,
                            uint64_t counters[3], uint8_t* d_valid, uint64_t* d_mat,
                            void* stream);

/* No counterpart in the reference: fame_match_reads_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or other allocating routine: its per-batch buffers are carved
 * from one file-scope device array of FAME_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  Limits: at most FAME_NOALLOC_MAX_READS reads
 * per call, every read ending within FAME_NOALLOC_MAX_BYTES of d_seqs
 * (FAME_ERR_ARG otherwise, nothing done).  A batch that outgrows one of the
 * array's capacities (verify pools, deferred reads, match overflow) is
 * detected before anything is counted and matched again in halves; a
 * single read that still does not fit (more than 2^18 index hits in one
 * orientation) returns FAME_ERR_CAPACITY, with the reads before it
 * matched and counted.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other contexts or threads) race and must
 * be serialised by the caller.  The context's device must be current.
 * Results are those of fame_match_reads_device. */
#define FAME_NOALLOC_MAX_READS 32768u
#define FAME_NOALLOC_MAX_BYTES (8ull << 20)
#define FAME_NOALLOC_SCRATCH_BYTES (112ull << 20)
int fame_match_reads_noalloc(fame_ctx* ctx, const uint8_t* d_seqs, const uint64_t* d_offs,
                             const uint32_t* d_lens, uint32_t n, int get_stranded,
                             uint64_t counters[3], uint8_t* d_valid, uint64_t* d_mat, void* stream);

/* MINPDIST / MAXPDIST (CONST.h:41-42, compile-time constants in the
 * reference; defaults 50 / 550 here).  Only MAXPDIST enters the matching:
 * contextWLen = MAXPDIST / WINLEN + 1 (ReadQueue.cpp:2854, 4540, 4607) and the
 * mate distance bound MAXPDIST + READLEN of extractPairedMatch (:3270).
 * FAME_ERR_ARG when MAXPDIST spans 30 windows or more. */
int fame_set_pair_dist(fame_ctx* ctx, uint32_t minpd, uint32_t maxpd);

/* ReadQueue::matchPairedReads ReadQueue.cpp:935-1655.  Pair i is mate 1
 * seqs1[offs1[i] .. +lens1[i]) and mate 2 seqs2[offs2[i] .. +lens2[i]).
 * counters[0..4] += succMatch, nonUniqueMatch, unSuccMatch, succPairedMatch,
 * tooShortCount (main.cpp:405-409).  valid1/2[i] = !isInvalid of the mates and
 * mat1/2[i] = their Read::mat, set for matched pairs and 0 otherwise (the
 * reference leaves mat unset there).  Needs ceil(0.75 READLEN) >= 32
 * (FAME_ERR_FORMAT otherwise).  The output pointers may be NULL. */
int fame_match_pairs(fame_ctx* ctx, const uint8_t* seqs1, const uint64_t* offs1, const uint32_t* lens1,
                     const uint8_t* seqs2, const uint64_t* offs2, const uint32_t* lens2, uint32_t n,
                     int get_stranded, uint64_t counters[5], uint8_t* valid1, uint64_t* mat1,
                     uint8_t* valid2, uint64_t* mat2);

/* The same with every buffer in device memory: both mates index d_seqs
 * (d_seqs[d_offs1[i] ..] is mate 1 of pair i, d_seqs[d_offs2[i] ..] its mate
 * 2); results go to the device arrays (any may be NULL).  `stream` as for
 * fame_match_reads_device.  No counterpart in the reference: the
 * resident-input form of matchPairedReads used by benchmarks. */
int fame_match_pairs_device(fame_ctx* ctx, const uint8_t* d_seqs, const uint64_t* d_offs1,
                            const uint32_t* d_lens1, const uint64_t* d_offs2, const uint32_t* d_lens2,
                            uint32_t n, int get_stranded, uint64_t counters[5], uint8_t* d_valid1,
                            uint64_t* d_mat1, uint8_t* d_valid2, uint64_t* d_mat2, void* stream);

/* ReadQueue::decideStrand ReadQueue.cpp:376-400 (without the console
 * report): matchR1Fwd = r1FwdMatches > r1RevMatches.  Any pointer may be
 * NULL. */
int fame_decide_strand(fame_ctx* ctx, uint64_t* r1_fwd, uint64_t* r1_rev, int* match_r1_fwd);

/* methLevels (ReadQueue.h:354-371): 4 x u16 per CpG in cpgTable order
 * (methFwd, unmethFwd, methRev, unmethRev), counts modulo 2^16 */
int fame_get_meth(fame_ctx* ctx, uint16_t* out, uint64_t n_cpg);

/* zeroes the CpG counts and strand tallies (a fresh ReadQueue) */
int fame_reset(fame_ctx* ctx);

/* ReadQueue::printMethylationLevels ReadQueue.cpp:1860-1893: writes
 * <basename>_cpg.tsv, byte-identical to the reference's report */
int fame_print_methylation_levels(fame_ctx* ctx, const char* basename);

/* ---- single cell ----------------------------------------------------------
 * The count matrix of FAME's single-cell driver: fame_sc_open writes the
 * header of the single-cell ReadQueue (ReadQueue.cpp:181-216: per CpG its
 * chromosome, then its position + READLEN - 2); fame_sc_write_cell is
 * printSCMethylationLevels (:1894-1921): four rows for the cell from the
 * context's counts (the fourth, unmethRev, carries the label "unmethFwd" as
 * in the reference), after which the counts are zeroed.  A cell's reads are
 * matched on the same context between fame_reset and fame_sc_write_cell
 * (matchSCBatch / matchSCBatchPaired :1660-1854). */
typedef struct fame_sc fame_sc;
int fame_sc_open(fame_ctx* ctx, const char* path, fame_sc** out);
int fame_sc_write_cell(fame_sc* sc, const char* sc_id);
int fame_sc_close(fame_sc* sc);

/* ---- batched form (no counterpart in the reference) ----------------------
 * FAME_FORM_PLAIN (default): one kernel per stage, the host reads the job
 * counts between stages to size the next launches.  FAME_FORM_BATCHED: the
 * whole single-end batch is one hipGraph launch (captured per input buffers
 * and capacity class of batch sizes, replayed afterwards): the counts stay on
 * the device, the verify / heavy-seed / overflow kernels are persistent over
 * them, and the only host wait is at the end.  A batch that outgrows the
 * graph's fixed capacities (a pool, 16384 index hits for one read, the match
 * overflow buffer) is detected on the device before anything is counted and
 * re-run on the plain path, so results are the same in either form.  The
 * persistent verify grid is picked per capacity class by timing candidate
 * grids on the first batches (online autotuning on real work).  Paired
 * matching always uses the plain form. */
#define FAME_FORM_PLAIN 0
#define FAME_FORM_BATCHED 1
int fame_set_form(fame_ctx* ctx, int form);

/* In the batched form a call of more than sub_batch_reads reads (default
 * 3 * 2^20), and the stranded first chunk from 2^16 reads on, is split into
 * an even number of equal sub-batches of at most sub_batch_reads that
 * alternate between two lanes (graph buffers and HIP streams of their own),
 * so that one lane's seeding overlaps the other's verification and DP.
 * FAME_ERR_ARG below 256. */
int fame_set_pipeline(fame_ctx* ctx, uint32_t sub_batch_reads);

/* form, batched graph launches, of which re-run plain, graphs captured,
 * capacity classes, the autotuned verify grid (blocks) of the largest tuned
 * class, pipelined calls, sub-batch size, graphs held now (each capacity
 * class keeps at most 64, the least recently used destroyed first) */
#define FAME_FORM_FIELDS 9
int fame_form_info(const fame_ctx* ctx, uint64_t out[FAME_FORM_FIELDS]);

/* ---- serving loop (no counterpart in the reference) ----------------------
 * A request loop over the context's resident index for callers that submit
 * chunks of single-end reads as they arrive (the strand decided beforehand,
 * i.e. matchReads with getStranded = false; with --unord_reads both).
 * Waiting requests are merged into one device batch while it fits the HBM
 * budget; the next batch is copied to the device while the current one
 * computes.  Every read gets the result fame_match_reads gives it, whatever
 * it is batched with; CpG counts accumulate in the context.  The context
 * must not be used for anything else while a server is open on it. */
typedef struct fame_server fame_server;

/* hbm_budget: device bytes the two in-flight batches may use (0 = half the
 * free HBM); max_batch_reads: reads per batch (0 = 2M) */
int fame_serve_open(fame_ctx* ctx, uint64_t hbm_budget, uint32_t max_batch_reads, fame_server** out);

/* queues a chunk; the read buffers must stay valid, and valid / mat (n
 * entries each, may be NULL) are written, before fame_serve_wait returns */
int fame_serve_submit(fame_server* srv, const uint8_t* seqs, const uint64_t* offs, const uint32_t* lens,
                      uint32_t n, uint8_t* valid, uint64_t* mat, uint64_t* ticket);

/* waits for a chunk: counters[0..2] += succ, nonunique, unsucc of its reads;
 * times_ms = queued, on the GPU, arrival to completion, reads of its batch */
int fame_serve_wait(fame_server* srv, uint64_t ticket, uint64_t counters[3], double times_ms[4]);

/* batches, requests, reads served, reads per batch at most; busy_ms = time
 * the device spent on batches */
int fame_serve_stats(fame_server* srv, uint64_t out[4], double* busy_ms);

int fame_serve_close(fame_server* srv);

/* ---- FASTQ ingestion ------------------------------------------------------
 * ReadQueue::parseChunk / parseChunkGZ (ReadQueue.cpp:218-373): records are
 * every fourth line (the reference's std::getline loop: '\n' ends a line, a
 * last unterminated line counts, a record's sequence is its second line or
 * empty at end of input).  fame_fastq_next parses up to max_reads records
 * with the host's threads into pinned buffers (reads back to back, u64
 * offsets, u32 lengths), valid until the next call on the handle; *n = 0 at
 * the end of the file.  gz selects gzip input (parseChunkGZ). */
typedef struct fame_fastq fame_fastq;
int fame_fastq_open(const char* path, int gz, fame_fastq** out);
int fame_fastq_next(fame_fastq* fq, uint32_t max_reads, uint32_t* n, const uint8_t** seqs,
                    const uint64_t** offs, const uint32_t** lens);
int fame_fastq_close(fame_fastq* fq);

/* per-stage device time of the last fame_match_reads* call, milliseconds:
 * [0] prep, [1] seed, [2] verify, [3] select, [4] meth, [5] the total
 * including host waits, [6] the banded-DP part of meth.  The batched form
 * reports its whole graph as seed, and the total.
 * No counterpart in the reference (instrumentation). */
#define FAME_TIMING_FIELDS 7
int fame_last_timings(const fame_ctx* ctx, double out[FAME_TIMING_FIELDS]);

/* work done by the last fame_match_reads* call, for roofline accounting:
 * [0] k-mers hashed, [1] index entries read from the k-mer table (with the
 * slot table: those of buckets of three or more), [2] (read, orientation)
 * pairs seeded, [3] window verifications, [4] window bytes scanned, [5]
 * banded alignments, [6] read bytes seeded, [7] bytes read per k-mer bucket
 * lookup, [8] the bucket lengths of every k-mer looked up, summed (index
 * entries the seeds name), [9] read letters aligned by the banded DP (rows
 * of 17 band cells each), [10] window letters scanned times the pattern's
 * 64-bit words (bit-parallel word updates of a full scan), [11] k-mer
 * slots read at random (the bucket lookups, plus the key-position lookups
 * of the genome-ordered slot copy), [12] k-mer slots read from the
 * genome-ordered copy (contiguous per read).
 * No counterpart in the reference (instrumentation). */
#define FAME_STATS_FIELDS 13
int fame_last_stats(const fame_ctx* ctx, uint64_t out[FAME_STATS_FIELDS]);

#ifdef __cplusplus
}
#endif
#endif
