/* CPU restatement of FAME's read-matching path (single-end and paired-end).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product (libfamehip.so) never links or calls it.
 *
 * Every function restates the reference algorithm (file:line in
 * FischerJo/FAME) with the same integer types, so that its outputs are
 * bit-identical to the reference CPU path on the same index and reads.
 */
#ifndef FAME_ORACLE_H
#define FAME_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_index orc_index;
typedef struct orc_ctx orc_ctx;

/* RefGenome::load RefGenome.cpp:230-387 + read_strands :414-433 */
orc_index* orc_load_index(const char* path, char* err, size_t errlen);
void orc_free_index(orc_index* ix);
/* header and table sizes: htab, readlen, winlen, kmerlen, kmercutoff,
 * seedbits, ncpg, nchrom, nkmer, nwin, nfiltered */
void orc_index_info(const orc_index* ix, uint64_t out[11]);

/* ReadQueue ctor ReadQueue.cpp:25-78 (single-end) */
orc_ctx* orc_ctx_new(orc_index* ix, int both_strands);
void orc_ctx_free(orc_ctx* c);

/* ReadQueue::matchReads ReadQueue.cpp:404-933 over n reads given as a byte
 * buffer + per-read offsets / lengths.  Counters are accumulated (+=) into
 * counters[0..2] = succ, nonunique, unsucc.  valid[i] = !isInvalid and
 * mat[i] = Read::mat (0 when invalid).  nthreads <= 0 uses OpenMP's default. */
int orc_match_reads(orc_ctx* c, const uint8_t* seqs, const uint64_t* offs, const uint32_t* lens,
                    uint32_t n, int get_stranded, uint64_t counters[3], uint8_t* valid,
                    uint64_t* mat, int nthreads);
/* ReadQueue::decideStrand ReadQueue.cpp:376-400; returns matchR1Fwd */
int orc_decide_strand(orc_ctx* c, uint64_t* fwd, uint64_t* rev);
/* 4 x u16 per CpG (methFwd, unmethFwd, methRev, unmethRev) in cpgTable order */
void orc_get_meth(const orc_ctx* c, uint16_t* out);
/* ReadQueue::printMethylationLevels ReadQueue.cpp:1860-1893 */
int orc_write_tsv(const orc_ctx* c, const char* path);
/* number of reads whose DP read reference bytes outside their chromosome
 * (undefined behaviour in the reference; see orc_dp_align) */
uint64_t orc_ub_reads(const orc_ctx* c);

/* ---- unit-level entry points (pinned individually by the tests) ---- */

/* ntHash::NTPS64 spaced_nthash/nthash.hpp:480-498; returns the spaced hash of
 * the 32-mer at seq[0..32) under the seed given as SEEDBITS (MSB = pos 0). */
uint64_t orc_ntps64(const uint8_t* kmer, uint32_t seedbits);
/* same value computed with the rolling update (nthash.hpp:492-498) for every
 * k-mer of seq; out[k] = hash of seq[k..k+32) */
void orc_ntps64_rolling(const uint8_t* seq, uint32_t len, uint32_t seedbits, uint64_t* out);

/* ShiftAnd<6>::querySeq / queryRevSeq (ShiftAnd.h:130-255).  text[0..tlen) is
 * scanned forward (rev=0) or from text[tlen-1] down to text[0] with
 * complementing (rev=1).  Positions are reported exactly as the reference
 * does relative to its iterator pair; see the implementation.  Returns the
 * number of matches written (<= cap); *total receives the full count. */
int orc_shiftand(const uint8_t* pat, uint32_t plen, const uint8_t* text, uint32_t tlen, int rev,
                 uint64_t* pos, uint8_t* err, int cap, int* total);
/* the same for ShiftAnd<E>, E <= 6 (the known-answer tests use E = 0..2) */
int orc_shiftand_e(const uint8_t* pat, uint32_t plen, const uint8_t* text, uint32_t tlen, int rev,
                   unsigned E, uint64_t* pos, uint8_t* err, int cap, int* total);
/* ShiftAnd<E>'s state after construction and reset(): masks[4][W] (A, C, G,
 * T), accepted[W], active[E+1][W]; returns W = plen / 64 + 1 */
int orc_shiftand_state(const uint8_t* pat, uint32_t plen, unsigned E, uint64_t* masks, uint64_t* acc,
                       uint64_t* act);

/* LevenshtDP<uint16_t,8> fill + backtrace (LevenshtDP.h:207-544) of read
 * `row` (len L) against reference bytes ending at ref[ref_end] read
 * backwards; ref[0..ref_len) is the chromosome.  rev selects runDPFillRev /
 * backtrackDPRev with CompiRev.  Writes the ERROR_T trace (0 MATCHING,
 * 1 MISMATCH, 2 INSERTION, 3 DELETION) and returns its length; *ub is set if
 * the reference was read outside [0, ref_len). */
int orc_dp_align(const uint8_t* row, uint32_t L, const uint8_t* ref, uint64_t ref_len,
                 int64_t ref_end, int rev, uint8_t* trace, int cap, int* ub);

/* LevenshtDP<uint16_t, band> with a chosen comparator: CompiFwd / CompiRev
 * (ReadQueue.h:295-338) as on the matching path, or the equality and
 * complement comparators of LevenshtDP_unittest.cpp.  rev selects
 * runDPFillRev / backtrackDPRev.  *dist receives getEditDist(). */
enum { ORC_CMP_FWD = 0, ORC_CMP_REV = 1, ORC_CMP_EQ = 2, ORC_CMP_WC = 3 };
int orc_dp_align_e(const uint8_t* row, uint32_t L, const uint8_t* ref, uint64_t ref_len, int64_t ref_end,
                   int rev, unsigned band, int cmp, uint8_t* trace, int cap, int* ub, int* dist);

/* tsl::hopscotch_map<uint32_t,uint16_t,identity> emulation: ++map[key] for
 * the keys in the given order on a map constructed with `bucket_count`;
 * writes the keys in iteration order, their counts (vals, may be NULL) and
 * the final bucket_count() (may be NULL).  Returns the number of keys. */
int orc_hopscotch_order(const uint32_t* keys, uint32_t n, uint64_t bucket_count, uint32_t* out, uint32_t* vals,
                        uint64_t* final_bucket_count);

/* getSeedRefs ReadQueue.cpp:2329-2461 for one sequence: writes the fwd and
 * rev maps in iteration order as (metaId, count) pairs. */
int orc_seed_maps(orc_ctx* c, const uint8_t* seq, uint32_t len, uint32_t* fwd_keys,
                  uint16_t* fwd_cnt, int* nfwd, uint32_t* rev_keys, uint16_t* rev_cnt, int* nrev,
                  int cap, uint32_t* bucket_count);

/* ---- paired-end path ---- */

/* MINPDIST / MAXPDIST (CONST.h:41-42; compile-time in FAME), default 50 / 550 */
void orc_set_pdist(orc_ctx* c, uint32_t minpd, uint32_t maxpd);
/* ReadQueue::matchPairedReads ReadQueue.cpp:935-1655 over n pairs.
 * counters[0..4] += succ, nonunique, unsucc, succPaired, tooShort; valid =
 * !Read::isInvalid of each mate, mat = Read::mat when the pair matched */
int orc_match_pairs(orc_ctx* c, const uint8_t* seqs1, const uint64_t* offs1, const uint32_t* lens1,
                    const uint8_t* seqs2, const uint64_t* offs2, const uint32_t* lens2, uint32_t n,
                    int get_stranded, uint64_t counters[5], uint8_t* valid1, uint64_t* mat1,
                    uint8_t* valid2, uint64_t* mat2, int nthreads);
/* std::sort (libstdc++ 11) of (key, packed value) pairs by count greater:
 * c0 (cmpMetaFirst) or c1 (cmpMetaSecond) */
void orc_std_sort_counts(uint32_t* keys, uint32_t* vals, uint32_t n, int second);

#ifdef __cplusplus
}
#endif
#endif
