// Internal device-side data structures shared by fame_device.hip and
// fame_host.cpp.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

struct fame_ctx;

// reads of at most FAME_DEV_MAX_READ letters take the short-read kernels;
// longer ones (up to FAME_LONG_MAX_READ) the long-read path: seeded by the
// heavy kernel from rows of their own, verified one wave per job (pool
// FAME_LONG_POOL), aligned by the DP kernel on scratch sized for them
#define FAME_DEV_MAX_READ 255
#define FAME_LONG_MAX_READ 4096
#define FAME_LONG_POOL 4
#define FAME_DEV_TRACE_CAP (2 * FAME_DEV_MAX_READ + 16)
#define FAME_PE_MAP_CHUNK 256 // map pool entries a paired seed wave reserves at a time
#define FAME_PE_WCAP 8         // windows of an item k_pe_walk1 walks in LDS, one lane per item
#define FAME_PE_WBIG 1360      // windows of an item k_pe_walk1_big walks in LDS, one wave per item
#define FAME_NPOOL 5 // verify pools by 64-bit words of the read: 1..4, then reads over FAME_DEV_MAX_READ
#define FAME_NSTATS 13
#define FAME_SEEDBITS 0xDEFDFFFDu // CONST.h:58 (0b11011110111111011111111111111101)
#define FAME_SEED_WAVES 16384      // persistent waves of the seed kernel
#ifndef FAME_SEED_WPB
#define FAME_SEED_WPB 2            // independent waves per block of the seed kernel
#endif
#define FAME_DPBAND 8 // 2*MISCOUNT + ADDMIS (ReadQueue.cpp:3640)

namespace fame {

enum : uint32_t {
    FLAG_TOO_LONG = 1u,
    FLAG_POOL_FULL = 2u,
    FLAG_HEAVY_FULL = 4u,
    FLAG_CAPACITY = 8u,
    FLAG_OVF_FULL = 16u,
    // any of these makes the batched form's later stages skip the batch,
    // which the host then runs on the plain path
    FLAG_ABORT = 1u | 2u | 4u | 8u | 16u,
    // a banded alignment whose traceback left the band or its trace buffer:
    // raised by the last stage, after other reads were counted, so it is a
    // hard error of the call and never a reason to run the batch again
    FLAG_DP_BROKEN = 32u,
};

// The index as resident in HBM.
//   tab      tabIndex, u32 when every prefix fits (nkmer < 2^32), else u64
//   kent     one u32 per k-mer table entry: isStart << 31 | metaId << 1 |
//            forward-strand bit (KMER_S::getMetaCpG / isStartCpG +
//            strandTable, structs.h:118-151); metaId < 2^30
//   kpe      the paired-end path's copy: kent | KMER_S::tmask << 32 (one
//            gather per entry)
//   filt32   the filteredKmers values below 2^32, sorted (the paired path
//            looks up a 32-bit k-mer code there, ReadQueue.cpp:2527)
//   win      metaWindow as {startPos, chrom, startInd, endInd}
//   seq      all chromosomes back to back, coff / clen per chromosome
//   pk       the same packed two bits per letter with an N mask (verify)
//   cpg_pos  cpgTable[i].pos, cpg_chrom cpgTable[i].chrom
struct DevIndex {
    const void* tab;
    int tab_is64;
    uint64_t htab, hmask;
    int hpow2;
    const uint32_t* kent;
    // per hash key, one u64 that holds a bucket of one or two entries itself
    // (so that the single-end seed stage reads one line per k-mer), or its
    // offset and length in kent; NULL when some bucket does not fit
    // (see k_build_slots)
    const uint64_t* slot;
    // the slot table in genome order, so that the k-mers of a read that
    // occur where its other k-mers put it are read from a few contiguous
    // lines instead of one random line each.  Positions p < gn/2 are the
    // forward k-mers (letters seq[p..p+31] reduced C->T), positions
    // gn/2 + r those of the reverse complement (letters comp(seq[gn/2-1-r-i])
    // reduced C->T).  gkey[p] = the hash key of the k-mer at p (~0u when it
    // holds a letter other than A/C/G/T or runs past the end), gslot[p] =
    // slot[gkey[p]], kpos[key] = some p with gkey[p] == key (~0u: none).
    // Since a slot is a function of its key alone, a k-mer whose key equals
    // gkey[p] takes gslot[p] whatever p is: the prediction only decides how
    // many lines are read, never what is found.  NULL when off.
    const uint32_t* gkey;
    const uint64_t* gslot;
    const uint32_t* kpos;
    uint64_t gn;
    const uint64_t* kpe;
    const uint32_t* filt32;
    uint32_t nfilt32;
    const uint4* win;
    const uint8_t* seq;
    const uint64_t* coff;
    const uint64_t* clen;
    // the verify stage's copy of the chromosomes: per 16 letters one u64,
    // letter j's code (A 0, C 1, G 2, T 3) at bits 2j, bit 32 + j set when
    // letter j is not A, C, G or T; chromosome c from word pkoff[c], two
    // words of padding after each
    const uint64_t* pk;
    const uint64_t* pkoff;
    // per window: its chromosome's first word in pk and length (so that the
    // verify kernels find a window's text with one load after the job's)
    const uint2* wtext;
    const uint32_t* cpg_pos;
    const uint8_t* cpg_chrom;
    uint32_t readlen, winlen, seedbits;
};

// Per-batch device buffers.  A (read, orientation) pair is ro = 2*read + ori
// (ori 0 = the read as given, 1 = its reverse complement).
struct BatchDev {
    const uint32_t* n_dev; // batched form: the batch size in device memory (NULL: plain form)
    const uint8_t* seqs;
    const uint64_t* offs;
    const uint32_t* lens;
    uint8_t* rev;
    uint8_t* pre;      // 0 = to match, 3 = invalid (short, N, too long)
    uint8_t* codes;    // [2n][256]: reduced letters of each (read, orientation), byte 255 = length, 0 = skip
    uint32_t* max_len; // longest read to be seeded (k_prep)
    uint32_t ori_mask; // bit0: query ori 0, bit1: query ori 1
    // candidate windows per ro, in hopscotch iteration order (fwd map then rev map)
    uint32_t* cand_off;
    uint32_t* cand_nf;
    uint32_t* cand_nr;
    // verify job pools, by 64-bit words of the pattern
    uint32_t* pool_n; // [FAME_NPOOL] device counters
    uint32_t pool_cap[FAME_NPOOL];
    uint32_t* job_meta[FAME_NPOOL]; // metaId | rev-map << 31
    uint32_t* job_ro[FAME_NPOOL];
    uint32_t* job_nm[FAME_NPOOL];    // matches found
    uint32_t* job_match[FAME_NPOOL]; // first match_cap matches: pos << 8 | err
    uint32_t* job_ovf_off[FAME_NPOOL];
    uint32_t match_cap;
    uint32_t verify_runs_cap; // verify's run queue: runs per block round (0: what its LDS holds)
    uint32_t* ovf_n;
    uint32_t* ovf_pool;
    uint32_t* ovf_job;
    uint32_t ovf_list_cap;
    uint32_t* ovf_buf;
    uint32_t ovf_buf_cap; // batched form: entries of ovf_buf (0: sized by the host)
    // reads over FAME_DEV_MAX_READ letters (plain form; NULL: k_prep flags
    // them FLAG_TOO_LONG): their list, and per long item q = 2 * (index in
    // the list) + orientation a row of reduced letters (long_stride bytes)
    uint32_t* long_n;
    uint32_t* long_list;
    uint8_t* long_codes;
    uint32_t long_stride;
    // DP reads over FAME_DEV_MAX_READ letters (k_select; NULL: in dp_list)
    uint32_t* dp_long_list;
    uint32_t* dp_long_n;
    // heavy (read, orientation) pairs deferred from the LDS seed kernel, and
    // long items (heavy_ro = 0x80000000 | q)
    uint32_t* heavy_n;
    uint32_t* heavy_ro;
    uint32_t* heavy_e;
    uint32_t* heavy_bc;
    uint32_t heavy_cap;
    // results
    uint8_t* status; // 0 matched, 1 nonunique, 2 unmatched, 3 invalid
    uint8_t* use;    // 1 = r.seq, 2 = revSeq carries the methylation call
    uint64_t* mat;
    uint8_t* valid;
    uint64_t* counters; // succ, nonunique, unsucc, r1fwd, r1rev (accumulated)
    uint32_t* meth;     // 4 x u32 per CpG
    // matched reads whose window has CpGs: errNum == 0 / errNum > 0
    uint32_t* direct_list;
    uint32_t* direct_n;
    uint32_t* dp_list;
    uint32_t* dp_n;
    uint32_t* flags;
    // work counters: k-mers hashed, bucket entries gathered, (read,
    // orientation) pairs seeded, verify jobs, window bytes scanned, DP
    // alignments, read bytes seeded
    unsigned long long* stats;
};

// Paired-end batch (matchPairedReads ReadQueue.cpp:935-1655).  The batch's
// 2n mates are laid out as 2n reads of a BatchDev: read i is mate 1 of pair
// i, read n + i its mate 2.  An item t = 2i + c is combination c of pair i:
// c = 0 matches (mate 1, revcomp mate 2), c = 1 (revcomp mate 1, mate 2);
// the first of the two is "the first read" of FAME's paired seeding.
struct PeItem {
    uint32_t k, v, j; // window, value (c0 | c1 << 8 | flags), verify slot
};

struct PairDev {
    uint32_t n;      // pairs
    uint32_t combos; // bit c: combination c is queried
    uint32_t cw;     // contextWLen: neighbouring windows that pair a mate
    uint32_t maxpd;  // MAXPDIST
    // per item: the item's slice of the map pool (both strand maps, each in
    // hopscotch iteration order) and its first-read match count
    uint32_t* map_off;
    uint32_t* map_nf;
    uint32_t* map_nr;
    uint32_t* m1_n;
    uint32_t* map_n; // pool fill (device counter)
    uint32_t* next;  // next item to take (device counter, LDS seed kernel)
    // items of many windows, walked one wave each (device list, its length
    // and the next entry to take)
    uint32_t* wlist;
    uint32_t* wlist_n;
    uint32_t* wnext;
    uint32_t wcap; // items of more windows go to k_pe_walk1_big (<= FAME_PE_WCAP)
    uint32_t wbig; // which keeps at most this many in LDS (<= FAME_PE_WBIG)
    uint32_t map_cap;
    PeItem* map;     // [map_cap] maps as built by the seed stage
    PeItem* srt;     // [map_cap] the same slices sorted by a count
    uint8_t* acc1[FAME_NPOOL]; // per first-read verify slot: 1 walked | 2 first match dropped
    uint8_t* acc2[FAME_NPOOL]; // per second-read verify slot
    uint64_t* counters;        // succ, nonunique, unsucc, succPaired, tooShort
    // items that outgrow the LDS seed kernel are listed and re-run with one
    // global scratch slice per block
    uint8_t* scratch;
    size_t stride;
    uint32_t nslices;
    uint32_t cap_e, hsize, tb_cap, cap_t, mcap, mhsz;
    const uint32_t* list; // NULL: every item; else the items to run
    uint32_t list_n;
    uint32_t* big_n;
    uint32_t* big_item;
    uint32_t* big_e;
    uint32_t big_cap;
};

struct HeavyScratch {
    uint8_t* base;
    size_t stride;
    uint32_t first, count;
    uint32_t cap_e, hsize, tb_cap;
    uint32_t kmax = 0; // k-mers per item at most (0: FAME_DEV_MAX_READ; more for long items)
    uint32_t e_lo, e_hi; // device-list mode: only the items with e_lo <= E <= e_hi
    uint32_t handed = 0;  // device-list mode: only the items the LDS tier handed on (bit 31 of heavy_e)
};

// bytes of one block's slice of heavy scratch: the arrays k_seed_heavy takes
// from its slice, in the same order, each rounded up to 256 bytes
inline size_t heavy_scratch_stride(const HeavyScratch& hs)
{
    size_t stride = 0;
    auto add = [&](size_t b) { stride += (b + 255) & ~(size_t)255; };
    const size_t e = hs.cap_e;
    const size_t km = hs.kmax > FAME_DEV_MAX_READ ? hs.kmax : FAME_DEV_MAX_READ;
    add(8 * km);                // klo
    add(4 * km);                // klen
    add(4 * km);                // koff
    add(4 * km);                // k2
    add(4 * e);                 // ent
    add(2 * e);                 // ekm (u16)
    add(4 * e);                 // eaux
    for (int i = 0; i < 3; ++i) add(4ull * hs.hsize); // hkey, hfirst, hcnt
    add(8 * e);                 // ins_key
    add(4 * e);                 // ins_cnt
    add(4ull * hs.tb_cap);      // tb
    add(4 * e);                 // tmp
    add(4 * e);                 // over
    add(8 * e);                 // cand
    return stride;
}

void launch_prep(const BatchDev& bd, uint32_t n, uint32_t readlen, hipStream_t st);
// batched form: counters and lists of one batch cleared on the device
void launch_batch_reset(const BatchDev& bd, hipStream_t st);
// batched form: the deferred (read, orientation) pairs from the device list,
// `blocks` scratch slices of hs
void launch_seed_heavy_dev(const DevIndex& ix, const BatchDev& bd, const HeavyScratch& hs, uint32_t blocks,
                           hipStream_t st);
// max_waves: persistent waves at most (0 = FAME_SEED_WAVES); each wave may
// leave up to 63 verify slots unused per pool it writes
void launch_seed_light(const DevIndex& ix, const BatchDev& bd, uint32_t n_ro, hipStream_t st, uint32_t lmax_hint = 0,
                       uint32_t max_waves = 0);
// *p = v on the stream (a kernel: no host staging)
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t st);
// p[0 .. n) = 0 on the stream (a kernel)
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st);
// shortest and longest of n lengths (out[0] must start at ~0, out[1] at 0)
void launch_len_range(const uint32_t* lens, uint32_t n, uint32_t* out, hipStream_t st);
// plain form: blocks (at most hs.count) stride over items first .. first + count - 1
void launch_seed_heavy(const DevIndex& ix, const BatchDev& bd, const HeavyScratch& hs, uint32_t blocks,
                       hipStream_t st);
uint32_t seed_heavy_blocks(bool tab64);
// njobs = FAME_JOBS_ON_DEVICE: as many as the device counter holds, on
// `blocks` persistent blocks (batched form)
#define FAME_JOBS_ON_DEVICE 0xFFFFFFFFu
void launch_verify(const DevIndex& ix, const BatchDev& bd, uint32_t pool, uint32_t njobs,
                   const uint32_t* only, hipStream_t st, uint32_t blocks = 0);
void launch_select(const DevIndex& ix, const BatchDev& bd, uint32_t n, int get_stranded, hipStream_t st);
// DP scratch per lane: FAME_DP_SROWS * 19 u16 values, FAME_DEV_MAX_READ + 1
// u64 choice words, FAME_DP_TRACE_WORDS u64 trace words (lane-strided)
#define FAME_DP_SROWS 12
#define FAME_DP_TRACE_WORDS ((FAME_DEV_TRACE_CAP + 31) / 32)
// between (optional): recorded after the direct kernel, before the banded DP
void launch_meth(const DevIndex& ix, const BatchDev& bd, uint32_t n, uint16_t* dp, uint64_t* dec, uint64_t* trace,
                 uint32_t grid_threads, hipStream_t st, hipEvent_t between = nullptr);
uint32_t meth_dp_lanes();
// the banded DP of the long reads' list (bd.dp_long_list / dp_long_n) on
// `lanes` lanes whose scratch holds L + 1 choice words and trace_cap / 32 + 1
// trace words for the longest of them
void launch_meth_dp_long(const DevIndex& ix, const BatchDev& bd, uint16_t* dp, uint64_t* dec, uint64_t* trace,
                         uint32_t lanes, uint32_t trace_cap, hipStream_t st);
// rows, bucket sums and heavy-list entries of the 2 * long_n long items
void launch_long_measure(const DevIndex& ix, const BatchDev& bd, uint32_t items, hipStream_t st);
void launch_prep_pairs(const BatchDev& bd, uint32_t n_reads, uint32_t nsplit, uint32_t minlen, hipStream_t st);
void launch_pe_seed(const DevIndex& ix, const BatchDev& bd, const PairDev& pd, uint32_t grid, hipStream_t st);
void launch_pe_walk1(const DevIndex& ix, const BatchDev& bd1, const BatchDev& bd2, const PairDev& pd, hipStream_t st);
void launch_pe_final(const DevIndex& ix, const BatchDev& bd1, const BatchDev& bd2, const PairDev& pd,
                     int get_stranded, hipStream_t st);
// serving layer hooks (fame_host.cpp): one batch on stream st with the
// strand already decided; the per-read outputs stay in the context's buffers
int serve_run(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
              uint64_t total_bytes, hipStream_t st, const uint8_t** d_status, const uint8_t** d_valid,
              const uint64_t** d_mat);
hipStream_t serve_stream(fame_ctx* c);
int serve_device(fame_ctx* c);
uint32_t serve_readlen(fame_ctx* c);
uint32_t serve_winlen(fame_ctx* c);
// a second context for concurrent batches: own buffers, stream and counters,
// the parent's index and CpG accumulators, its strand decision and form as
// of now (NULL on failure)
fame_ctx* serve_worker_open(fame_ctx* c);
void serve_worker_close(fame_ctx* w);
void serve_set_err(fame_ctx* c, const std::string& m);
void launch_pack_tmask(const uint32_t* raw, const uint32_t* kent, uint64_t n, uint64_t* out, hipStream_t st);
void launch_pack_kmers(const uint32_t* raw, const uint8_t* sbits, uint64_t first, uint64_t n, uint32_t* out,
                       hipStream_t st);
void launch_narrow_tab(const uint64_t* in, uint64_t n, uint32_t* out, hipStream_t st);
// nwords packed words (16 letters each) of one chromosome of len letters
void launch_pack_seq(const uint8_t* seq, uint64_t len, uint64_t nwords, uint64_t* out, hipStream_t st);
void launch_seed_mid(const DevIndex& ix, const BatchDev& bd, uint32_t first, uint32_t count, uint32_t* ovl,
                     uint32_t blocks, hipStream_t st);
uint32_t seed_mid_max_e();
uint32_t seed_mid_blocks();
void launch_gorder(const uint8_t* seq, uint64_t ng, uint64_t htab, const uint64_t* slot, uint32_t* gkey,
                   uint64_t* gslot, uint32_t* kpos, hipStream_t st);
void launch_build_slots(const uint32_t* tab, const uint32_t* kent, uint64_t htab, uint64_t* slot, uint32_t* bad,
                        hipStream_t st);
void launch_meth_u16(const uint32_t* in, uint64_t n, uint16_t* out, hipStream_t st);
void launch_ovf_offsets(const BatchDev& bd, uint32_t n_ovf, uint32_t* total, hipStream_t st);
void launch_max_end(const uint64_t* offs, const uint32_t* lens, uint32_t n, unsigned long long* out, hipStream_t st);

} // namespace fame
