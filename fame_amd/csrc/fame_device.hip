// HIP kernels of the MI355X read-matching path (gfx950).
//
// One batch of reads flows through five kernels (the "plain" one kernel per
// stage form):
//   k_prep    per read: length / N validation and the reverse complement
//             (ReadQueue.cpp:428-490)
//   k_seed    one wavefront per (read, orientation): spaced-seed ntHash of
//             every k-mer, tabIndex / kmer-table gathers, per-window q-gram
//             counting with the reference's insertion rules, emulation of
//             tsl::hopscotch_map's iteration order, and emission of the
//             windows with count >= QTHRESH in that order
//             (getSeedRefs ReadQueue.cpp:2329-2461, saQuerySeedSetRef's
//             iteration :1937-1950, :2035-2044)
//   k_verify  one lane per (read, orientation, window): bit-parallel
//             approximate matching of the read over the 2053-byte window
//             (ShiftAnd<6>::querySeq / queryRevSeq ShiftAnd.h:130-255) in
//             Myers' bit-vector form, which yields the same lowest error layer
//             per text position with ~4x fewer word operations than the
//             7-layer automaton
//   k_select  one lane per read: the sequential uniqueness logic of
//             saQuerySeedSetRef (:1967-2159) and the decision block
//             (:630-805)
//   k_meth    one lane per matched read: CpG lookup or banded DP with
//             backtrace, then per-CpG counting into u32 accumulators
//             (computeMethLvl :3278-3947; LevenshtDP.h; BandedMatrix.h)
//
// All arithmetic is integer except two float expressions of the hopscotch
// map's load factors, which are reproduced with the reference's float
// rounding (see hs_ceil_div09).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fame_internal.h"

namespace fame {

// CONST.h:80-97 and ReadQueue.cpp:532, 3640
constexpr uint32_t KMERLEN = 32;
constexpr uint32_t QTHRESH = 5;
constexpr uint32_t MISCOUNT = 2;
constexpr uint32_t ADDMIS = 4;
constexpr uint32_t NERR = MISCOUNT + ADDMIS;
constexpr int DPBAND = FAME_DPBAND;
static_assert(DPBAND == 2 * MISCOUNT + ADDMIS, "DP band");
constexpr uint32_t CPGDUMMY = 0xFFFFFFFFu;
constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr uint32_t NBH = 62;             // hopscotch NeighborhoodSize
constexpr uint32_t MAX_PROBES = 12 * NBH; // MAX_PROBES_FOR_EMPTY_BUCKET

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t rol64(uint64_t v, int s)
{
    s &= 63;
    return s ? (v << s) | (v >> (64 - s)) : v;
}

// code of a byte in ntHash's reduced (C->T) table msTab (nthash.hpp:88-121):
// 0 = seedA, 1 = seedG, 2 = seedT, 3 = zero row
__device__ __forceinline__ uint32_t red_code(uint8_t c)
{
    switch (c) {
    case 'A': case 'a': case 4:
        return 0;
    case 'G': case 'g': case 3:
        return 1;
    case 'C': case 'c': case 'T': case 't': case 1: case 7:
        return 2;
    default:
        return 3;
    }
}

__device__ __forceinline__ uint64_t red_seed(uint32_t code)
{
    // nthash.hpp:24-27 (seedC is never used by the reduced table)
    return code == 0 ? 0x3c8bfbb395c60474ULL
         : code == 1 ? 0x20323ed082572324ULL
         : code == 2 ? 0x295549f54be24456ULL
                     : 0ULL;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// inclusive scan over the 64 lanes of a wave
__device__ __forceinline__ uint32_t wave_iscan_u32(uint32_t v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t rup_pow2(uint64_t v)
{
    if (v && !(v & (v - 1))) return v;
    if (!v) return 1;
    --v;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    v |= v >> 32;
    return v + 1;
}

// size_type(float(x) * f) with f = 0.9f or 0.1f (hopscotch_hash.h:1175-1176)
__device__ __forceinline__ uint32_t hs_thr(uint64_t b, float f)
{
    float x = __fmul_rn((float)b, f);
    return (uint32_t)x;
}

// size_type(std::ceil(float(n) / 0.9f)) (hopscotch_hash.h:1180); the float
// quotient is obtained exactly as the correctly rounded double quotient
// rounded to float (checked exhaustively for n < 2^24 on the host)
__device__ __forceinline__ uint64_t hs_ceil_div09(uint64_t n)
{
    double q = (double)n / (double)0.9f;
    float f = __double2float_rn(q);
    return (uint64_t)ceilf(f);
}

// ---------------------------------------------------------------------------
// tsl::hopscotch_map<uint32_t, uint16_t, identity> placement emulation.
// Only the order of iteration is reproduced: keys are inserted in the order
// in which operator[] first created them (counts never move an element).
// tb[] holds, per physical bucket, the index of the key stored there.
// ---------------------------------------------------------------------------

struct Hs {
    uint32_t* tb;
    uint32_t tb_cap;
    uint32_t* over;
    uint32_t over_cap;
    uint32_t* tmp;
    uint32_t tmp_cap;
    const uint32_t* key;
    uint64_t B, mask, arr;
    uint32_t nelem, nover, max_thr, min_thr;
};

enum { HS_OK = 0, HS_FAIL = 1 };

// hopscotch_hash ctor (hopscotch_hash.h:599-628) on a fresh bucket array;
// the overflow list is kept (rehash_impl moves it to the new table)
__device__ bool hs_setup(Hs& h, uint64_t bucket_count)
{
    uint64_t B = bucket_count ? rup_pow2(bucket_count) : 0;
    uint64_t arr = B ? B + NBH - 1 : 0;
    if (arr > h.tb_cap) return false;
    h.B = B;
    h.mask = B ? B - 1 : 0;
    h.arr = arr;
    for (uint64_t i = 0; i < arr; ++i) h.tb[i] = EMPTY;
    h.max_thr = hs_thr(B, 0.9f);
    h.min_thr = hs_thr(B, 0.1f);
    return true;
}

// swap_empty_bucket_closer hopscotch_hash.h:1545-1578; the neighbourhood
// bitmap of bucket tc is "the buckets tc+i holding a key whose home is tc"
__device__ bool hs_swap_closer(Hs& h, uint64_t& e)
{
    const uint64_t start = e - NBH + 1;
    for (uint64_t tc = start; tc < e; ++tc) {
        for (uint64_t ts = tc; ts < e; ++ts) {
            uint32_t i = h.tb[ts];
            if (i != EMPTY && (h.key[i] & h.mask) == tc) {
                h.tb[e] = i;
                h.tb[ts] = EMPTY;
                e = ts;
                return true;
            }
        }
    }
    return false;
}

// will_neighborhood_change_on_rehash hopscotch_hash.h:1466-1486
__device__ bool hs_will_change(const Hs& h, uint64_t ib)
{
    const uint64_t mask2 = rup_pow2((h.mask + 1) * 2) - 1;
    for (uint64_t b = ib; b < h.arr && (b - ib) < NBH; ++b) {
        uint32_t i = h.tb[b];
        if (i == EMPTY) continue;
        uint64_t k = h.key[i];
        if ((k & h.mask) != (k & mask2)) return true;
    }
    return false;
}

enum { HS_PLACED = 0, HS_NEED_REHASH = 2 };

// one pass of insert_impl hopscotch_hash.h:1426-1458 without its rehash
// calls: returns HS_NEED_REHASH where the reference rehashes (load
// threshold reached first, or no bucket within the neighbourhood and the
// element may not go to the overflow list)
__device__ int hs_place(Hs& h, uint32_t idx)
{
    if (h.nelem - h.nover >= h.max_thr) return HS_NEED_REHASH;
    const uint64_t ib = h.key[idx] & h.mask;
    uint64_t e = h.arr;
    {
        const uint64_t lim = ib + MAX_PROBES < h.arr ? ib + MAX_PROBES : h.arr;
        for (uint64_t b = ib; b < lim; ++b)
            if (h.tb[b] == EMPTY) {
                e = b;
                break;
            }
    }
    if (e < h.arr) {
        for (;;) {
            if (e - ib < NBH) {
                h.tb[e] = idx;
                h.nelem++;
                return HS_PLACED;
            }
            if (!hs_swap_closer(h, e)) break;
        }
    }
    if (h.nelem < h.min_thr || !hs_will_change(h, ib)) {
        if (h.nover >= h.over_cap) return HS_FAIL;
        h.over[h.nover++] = idx;
        h.nelem++;
        return HS_PLACED;
    }
    return HS_NEED_REHASH;
}

// rehash_impl hopscotch_hash.h:1244-1290: a new table, the overflow list
// carried over, then the old buckets re-inserted in physical order.  A
// rehash that would be needed inside this re-insertion is reported as a
// capacity failure.
__device__ int hs_rehash(Hs& h, uint64_t count)
{
    const uint64_t need = hs_ceil_div09(h.nelem);
    if (need > count) count = need;
    uint32_t n = 0;
    for (uint64_t b = 0; b < h.arr; ++b) {
        uint32_t i = h.tb[b];
        if (i != EMPTY) {
            if (n >= h.tmp_cap) return HS_FAIL;
            h.tmp[n++] = i;
        }
    }
    if (!hs_setup(h, count)) return HS_FAIL;
    h.nelem = h.nover;
    for (uint32_t t = 0; t < n; ++t)
        if (hs_place(h, h.tmp[t]) != HS_PLACED) return HS_FAIL;
    return HS_OK;
}

// operator[] creating a key: insert_impl with its rehash-and-retry loop
__device__ int hs_insert(Hs& h, uint32_t idx)
{
    for (;;) {
        const int r = hs_place(h, idx);
        if (r == HS_PLACED) return HS_OK;
        if (r == HS_FAIL) return HS_FAIL;
        if (hs_rehash(h, (h.mask + 1) * 2) != HS_OK) return HS_FAIL;
    }
}

// ---------------------------------------------------------------------------
// k_prep
// ---------------------------------------------------------------------------

__global__ void k_prep(const uint8_t* __restrict__ seqs, const uint64_t* __restrict__ offs,
                       const uint32_t* __restrict__ lens, uint32_t n, uint32_t readlen,
                       uint8_t* __restrict__ rev, uint8_t* __restrict__ pre,
                       uint32_t* __restrict__ err_flag)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t L = lens[r];
    const uint8_t* s = seqs + offs[r];
    uint8_t* q = rev + offs[r];
    uint8_t st = 0;
    // readSize < READLEN - 20 (ReadQueue.cpp:430): the unsigned subtraction
    // of the reference is kept
    if ((uint64_t)L < (uint64_t)(uint32_t)(readlen - 20u)) st = 3;
    if (L > FAME_DEV_MAX_READ) {
        atomicOr(err_flag, 1u);
        st = 3;
    }
    if (L < KMERLEN) st = 3;
    if (st == 0) {
        for (uint32_t p = 0; p < L; ++p) {
            uint8_t c = s[p], o = 0;
            switch (c) {
            case 'A': o = 'T'; break;
            case 'C': o = 'G'; break;
            case 'G': o = 'C'; break;
            case 'T': o = 'A'; break;
            case 'N': st = 3; break;
            default: break; // unknown characters stay '\0' in revSeq
            }
            q[L - 1 - p] = o;
        }
    }
    pre[r] = st;
}

// ---------------------------------------------------------------------------
// k_seed: seed + filter + map order for one (read, orientation) per wave
// ---------------------------------------------------------------------------

struct SeedWork {
    // phase A (gather / count)
    uint64_t* klo;     // [K]
    uint32_t* klen;    // [K]
    uint32_t* koff;    // [K] exclusive prefix of klen
    uint32_t* ent;     // [cap_e] packed metaId<<1 | fwd
    uint8_t* ekm;      // [cap_e] k-mer index of each entry
    uint32_t* eaux;    // [cap_e] flags (bit0 kept, bit1 first) or hash slot (heavy)
    uint32_t cap_e;
    // heavy path only: hash table keyed by packed entry
    uint32_t* hkey;
    uint32_t* hfirst;
    uint32_t* hcnt;
    uint32_t hsize; // power of two, 0 = use pairwise compare
    // insertion lists (alive across phases)
    uint32_t* ins_key[2];
    uint32_t* ins_cnt[2];
    uint32_t cap_ins;
    // phase B (hopscotch)
    uint32_t* tb[2];
    uint32_t tb_cap;
    uint32_t* tmp[2];
    uint32_t* over[2];
    uint32_t cap_list;
};

struct SeedShared {
    uint8_t seq[FAME_DEV_MAX_READ + 1];
    uint8_t code[FAME_DEV_MAX_READ + 1];
    uint32_t nins[2];
    uint32_t ncand[2];
    uint32_t base;
    uint32_t fail;
};

template <typename TabT>
__device__ __forceinline__ uint64_t tab_at(const void* tab, uint64_t i)
{
    return (uint64_t)((const TabT*)tab)[i];
}

// Returns 0 done, 1 = needs the heavy path (capacity), with E / bucketCount
// reported through the out pointers.
template <typename TabT>
__device__ int seed_one(const DevIndex& ix, const BatchDev& bd, uint32_t ro, SeedShared& sh,
                        SeedWork& w, uint64_t* E_out, uint32_t* bc_out)
{
    const int lane = threadIdx.x;
    const uint32_t r = ro >> 1, ori = ro & 1;
    const uint32_t L = bd.lens[r];
    const uint8_t* s = (ori ? bd.rev : bd.seqs) + bd.offs[r];
    for (uint32_t i = lane; i < L; i += 64) {
        uint8_t c = s[i];
        sh.seq[i] = c;
        sh.code[i] = (uint8_t)red_code(c);
    }
    __syncthreads();
    const uint32_t K = L - KMERLEN + 1;
    const uint32_t sb = ix.seedbits;
    // hashes + tabIndex gathers (NTPS64 closed form; key = h % HTABSIZE)
    uint32_t bsum = 0;
    for (uint32_t k = lane; k < K; k += 64) {
        uint64_t h = 0;
#pragma unroll
        for (uint32_t i = 0; i < KMERLEN; ++i)
            if ((sb >> (KMERLEN - 1 - i)) & 1) h ^= rol64(red_seed(sh.code[k + i]), (int)(KMERLEN - 1 - i));
        const uint64_t key = ix.hpow2 ? (h & ix.hmask) : (h % ix.htab);
        const uint64_t lo = tab_at<TabT>(ix.tab, key);
        const uint64_t hi = tab_at<TabT>(ix.tab, key + 1);
        w.klo[k] = lo;
        const uint64_t len = hi - lo;
        w.klen[k] = len > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)len;
        bsum += (uint32_t)len; // bucketCount += (u32 truncation of the u64 difference)
    }
    bsum = wave_sum_u32(bsum);
    const uint32_t bucketCount = bsum / 5;
    // exclusive prefix over k-mers in index order
    uint64_t carry = 0;
    bool big = false;
    for (uint32_t k0 = 0; k0 < K; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t v = k < K ? w.klen[k] : 0;
        // buckets beyond 2^24 entries cannot be served (reported as capacity)
        if (v > 0xFFFFFFu) {
            v = 0xFFFFFFu;
            big = true;
        }
        uint32_t inc = wave_iscan_u32(v, lane);
        uint32_t tot = __shfl(inc, 63, 64);
        if (k < K) w.koff[k] = (uint32_t)(carry + inc - v);
        carry += tot;
        if (carry > 0x7FFFFFFFull) big = true;
    }
    big = __ballot(big) != 0;
    const uint64_t E = carry;
    *E_out = E;
    *bc_out = bucketCount;
    if (lane == 0) {
        atomicAdd(&bd.stats[0], (unsigned long long)K);
        atomicAdd(&bd.stats[1], (unsigned long long)E);
        atomicAdd(&bd.stats[2], 1ull);
        atomicAdd(&bd.stats[6], (unsigned long long)L);
    }
    if (big || E > w.cap_e) return 1;
    // entry gather: ent[koff[k] + i] = kmer[lo + i]
    for (uint32_t k = lane; k < K; k += 64) {
        const uint64_t lo = w.klo[k];
        const uint32_t n = w.klen[k], o = w.koff[k];
        for (uint32_t i = 0; i < n; ++i) {
            w.ent[o + i] = ix.kent[lo + i];
            w.ekm[o + i] = (uint8_t)k;
        }
    }
    __syncthreads();
    // consecutive dedup inside a bucket (ReadQueue.cpp:2381-2391, 2412-2422)
    for (uint32_t j = lane; j < E; j += 64) {
        const bool skip = j > 0 && w.ekm[j] == w.ekm[j - 1] && w.ent[j] == w.ent[j - 1];
        w.eaux[j] = skip ? 0u : 1u;
    }
    __syncthreads();
    // first occurrence and total count of every kept key
    if (w.hsize == 0) {
        for (uint32_t j = lane; j < E; j += 64) {
            if (!(w.eaux[j] & 1)) continue;
            const uint32_t key = w.ent[j];
            bool first = true;
            uint32_t cnt = 0;
            for (uint32_t i = 0; i < E; ++i) {
                if ((w.eaux[i] & 1) && w.ent[i] == key) {
                    ++cnt;
                    if (i < j) first = false;
                }
            }
            // count is kept in bits 16..31 (u16 as in the reference's map)
            w.eaux[j] = 1u | (first ? 2u : 0u) | ((cnt & 0xFFFFu) << 16);
        }
    } else {
        const uint32_t hm = w.hsize - 1;
        for (uint32_t j = lane; j < E; j += 64) {
            if (!(w.eaux[j] & 1)) continue;
            const uint32_t key = w.ent[j];
            uint32_t slot = (key * 2654435761u) & hm;
            for (;;) {
                uint32_t old = atomicCAS(&w.hkey[slot], EMPTY, key);
                if (old == EMPTY || old == key) break;
                slot = (slot + 1) & hm;
            }
            atomicMin(&w.hfirst[slot], j);
            atomicAdd(&w.hcnt[slot], 1u);
            w.eaux[j] = 1u | (slot << 8);
        }
        __threadfence_block();
        __syncthreads();
        for (uint32_t j = lane; j < E; j += 64) {
            uint32_t a = w.eaux[j];
            if (!(a & 1)) continue;
            const uint32_t slot = a >> 8;
            const bool first = w.hfirst[slot] == j;
            w.eaux[j] = 1u | (first ? 2u : 0u) | ((w.hcnt[slot] & 0xFFFFu) << 16);
        }
    }
    __syncthreads();
    // insertion sequences: a key enters its map at its first occurrence if
    // that occurrence may insert (k-mer 0, or k-mer cIdx+1 with cIdx < maxQPos)
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    uint32_t nins0 = 0, nins1 = 0;
    for (uint32_t j0 = 0; j0 < E; j0 += 64) {
        const uint32_t j = j0 + lane;
        bool take0 = false, take1 = false;
        uint32_t key = 0, cnt = 0;
        if (j < E) {
            const uint32_t a = w.eaux[j];
            const uint32_t km = w.ekm[j];
            const bool ins = km == 0 || (km - 1) < maxQPos;
            if ((a & 2) && ins) {
                const uint32_t e = w.ent[j];
                key = e >> 1;
                cnt = a >> 16;
                take0 = (e & 1) != 0; // forward-strand windows -> fwdMetaIDs
                take1 = !take0;
            }
        }
        const uint64_t b0 = __ballot(take0), b1 = __ballot(take1);
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (take0) {
            uint32_t p = nins0 + __popcll(b0 & below);
            if (p < w.cap_ins) {
                w.ins_key[0][p] = key;
                w.ins_cnt[0][p] = cnt;
            }
        }
        if (take1) {
            uint32_t p = nins1 + __popcll(b1 & below);
            if (p < w.cap_ins) {
                w.ins_key[1][p] = key;
                w.ins_cnt[1][p] = cnt;
            }
        }
        nins0 += __popcll(b0);
        nins1 += __popcll(b1);
    }
    if (nins0 > w.cap_ins || nins1 > w.cap_ins) return 1;
    __syncthreads();
    // hopscotch placement, one lane per map
    if (lane == 0) sh.fail = 0;
    __syncthreads();
    if (lane < 2) {
        Hs h;
        h.tb = w.tb[lane];
        h.tb_cap = w.tb_cap;
        h.over = w.over[lane];
        h.over_cap = w.cap_list;
        h.tmp = w.tmp[lane];
        h.tmp_cap = w.cap_list;
        h.key = w.ins_key[lane];
        h.nelem = h.nover = 0;
        const uint32_t ni = lane ? nins1 : nins0;
        bool ok = hs_setup(h, bucketCount);
        for (uint32_t i = 0; ok && i < ni; ++i) ok = hs_insert(h, i) == HS_OK;
        uint32_t nc = 0;
        if (ok) {
            // iteration order: buckets, then the overflow list; keep the
            // windows passing the q-gram threshold (count is a u16)
            for (uint64_t b = 0; b < h.arr; ++b) {
                uint32_t i = h.tb[b];
                if (i != EMPTY && w.ins_cnt[lane][i] >= QTHRESH) h.tmp[nc++] = w.ins_key[lane][i];
            }
            for (uint32_t t = 0; t < h.nover; ++t) {
                uint32_t i = h.over[t];
                if (w.ins_cnt[lane][i] >= QTHRESH) h.tmp[nc++] = w.ins_key[lane][i];
            }
        } else {
            atomicOr(&sh.fail, 1u);
        }
        sh.ncand[lane] = nc;
    }
    __syncthreads();
    if (sh.fail) return 1;
    // emit candidates: fwd-map windows then rev-map windows
    const uint32_t nf = sh.ncand[0], nr = sh.ncand[1];
    const uint32_t pool = (L + 63) / 64 - 1;
    if (lane == 0) {
        uint32_t base = (nf + nr) ? atomicAdd(&bd.pool_n[pool], nf + nr) : 0;
        if (base + nf + nr > bd.pool_cap[pool]) atomicOr(bd.flags, FLAG_POOL_FULL);
        sh.base = base;
        bd.cand_off[ro] = base;
        bd.cand_nf[ro] = nf;
        bd.cand_nr[ro] = nr;
    }
    __syncthreads();
    const uint32_t base = sh.base;
    if (base + nf + nr <= bd.pool_cap[pool]) {
        for (uint32_t t = lane; t < nf + nr; t += 64) {
            const bool isrev = t >= nf;
            const uint32_t id = isrev ? w.tmp[1][t - nf] : w.tmp[0][t];
            bd.job_meta[pool][base + t] = id | (isrev ? 0x80000000u : 0u);
            bd.job_ro[pool][base + t] = ro;
        }
    }
    return 0;
}

constexpr uint32_t L_CAP_E = 256;   // entries per (read, orientation) in LDS
constexpr uint32_t L_CAP_INS = 256; // keys per map in LDS
constexpr uint32_t L_TB = 512 + NBH - 1;

template <typename TabT>
__global__ __launch_bounds__(64) void k_seed_light(DevIndex ix, BatchDev bd, uint32_t n_ro)
{
    __shared__ SeedShared sh;
    __shared__ uint64_t s_klo[FAME_DEV_MAX_READ];
    __shared__ uint32_t s_klen[FAME_DEV_MAX_READ];
    __shared__ uint32_t s_koff[FAME_DEV_MAX_READ];
    // phase A arrays and phase B arrays share storage
    __shared__ union {
        struct {
            uint32_t ent[L_CAP_E];
            uint8_t ekm[L_CAP_E];
            uint32_t eaux[L_CAP_E];
        } a;
        struct {
            uint32_t tb[2][L_TB];
            uint32_t tmp[2][L_CAP_INS];
            uint32_t over[2][L_CAP_INS];
        } b;
    } u;
    __shared__ uint32_t s_ins_key[2][L_CAP_INS];
    __shared__ uint32_t s_ins_cnt[2][L_CAP_INS];

    const uint32_t ro = blockIdx.x;
    if (ro >= n_ro) return;
    const uint32_t r = ro >> 1, ori = ro & 1;
    if (!((bd.ori_mask >> ori) & 1) || bd.pre[r] != 0) {
        if (threadIdx.x == 0) {
            bd.cand_nf[ro] = 0;
            bd.cand_nr[ro] = 0;
            bd.cand_off[ro] = 0;
        }
        return;
    }
    SeedWork w;
    w.klo = s_klo;
    w.klen = s_klen;
    w.koff = s_koff;
    w.ent = u.a.ent;
    w.ekm = u.a.ekm;
    w.eaux = u.a.eaux;
    w.cap_e = L_CAP_E;
    w.hsize = 0;
    w.hkey = w.hfirst = w.hcnt = nullptr;
    w.ins_key[0] = s_ins_key[0];
    w.ins_key[1] = s_ins_key[1];
    w.ins_cnt[0] = s_ins_cnt[0];
    w.ins_cnt[1] = s_ins_cnt[1];
    w.cap_ins = L_CAP_INS;
    w.tb[0] = u.b.tb[0];
    w.tb[1] = u.b.tb[1];
    w.tb_cap = L_TB;
    w.tmp[0] = u.b.tmp[0];
    w.tmp[1] = u.b.tmp[1];
    w.over[0] = u.b.over[0];
    w.over[1] = u.b.over[1];
    w.cap_list = L_CAP_INS;
    uint64_t E;
    uint32_t bc;
    int rc = seed_one<TabT>(ix, bd, ro, sh, w, &E, &bc);
    if (rc != 0 && threadIdx.x == 0) {
        uint32_t h = atomicAdd(bd.heavy_n, 1u);
        if (h < bd.heavy_cap) {
            bd.heavy_ro[h] = ro;
            bd.heavy_e[h] = E > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)E;
            bd.heavy_bc[h] = bc;
        } else {
            atomicOr(bd.flags, FLAG_HEAVY_FULL);
        }
        bd.cand_nf[ro] = 0;
        bd.cand_nr[ro] = 0;
        bd.cand_off[ro] = 0;
    }
}

// heavy path: one block per deferred (read, orientation), global scratch
template <typename TabT>
__global__ __launch_bounds__(64) void k_seed_heavy(DevIndex ix, BatchDev bd, HeavyScratch hs)
{
    __shared__ SeedShared sh;
    const uint32_t hi = blockIdx.x;
    if (hi >= hs.count) return;
    const uint32_t ro = bd.heavy_ro[hs.first + hi];
    uint8_t* base = hs.base + (size_t)hi * hs.stride;
    SeedWork w;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t* p = base + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const uint32_t ce = hs.cap_e;
    w.klo = (uint64_t*)take(8 * FAME_DEV_MAX_READ);
    w.klen = (uint32_t*)take(4 * FAME_DEV_MAX_READ);
    w.koff = (uint32_t*)take(4 * FAME_DEV_MAX_READ);
    w.ent = (uint32_t*)take(4ull * ce);
    w.ekm = (uint8_t*)take(ce);
    w.eaux = (uint32_t*)take(4ull * ce);
    w.cap_e = ce;
    w.hsize = hs.hsize;
    w.hkey = (uint32_t*)take(4ull * hs.hsize);
    w.hfirst = (uint32_t*)take(4ull * hs.hsize);
    w.hcnt = (uint32_t*)take(4ull * hs.hsize);
    w.cap_ins = ce;
    w.ins_key[0] = (uint32_t*)take(4ull * ce);
    w.ins_key[1] = (uint32_t*)take(4ull * ce);
    w.ins_cnt[0] = (uint32_t*)take(4ull * ce);
    w.ins_cnt[1] = (uint32_t*)take(4ull * ce);
    w.tb_cap = hs.tb_cap;
    w.tb[0] = (uint32_t*)take(4ull * hs.tb_cap);
    w.tb[1] = (uint32_t*)take(4ull * hs.tb_cap);
    w.cap_list = ce;
    w.tmp[0] = (uint32_t*)take(4ull * ce);
    w.tmp[1] = (uint32_t*)take(4ull * ce);
    w.over[0] = (uint32_t*)take(4ull * ce);
    w.over[1] = (uint32_t*)take(4ull * ce);
    for (uint32_t i = threadIdx.x; i < hs.hsize; i += 64) {
        w.hkey[i] = EMPTY;
        w.hfirst[i] = 0xFFFFFFFFu;
        w.hcnt[i] = 0;
    }
    __threadfence_block();
    __syncthreads();
    uint64_t E;
    uint32_t bc;
    int rc = seed_one<TabT>(ix, bd, ro, sh, w, &E, &bc);
    if (rc != 0 && threadIdx.x == 0) {
        atomicOr(bd.flags, FLAG_CAPACITY);
        bd.cand_nf[ro] = 0;
        bd.cand_nr[ro] = 0;
    }
}

// ---------------------------------------------------------------------------
// k_verify: approximate matching of one read over one window per lane
// ---------------------------------------------------------------------------

// lmap ReadQueue.cpp:73-76 (A0 C1 G2 T3), 4 = N / other
__device__ __forceinline__ uint32_t base_code(uint8_t c)
{
    return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 3u : 4u;
}

// returns the number of window bytes scanned (0 when the lane has no job)
template <int W>
__device__ __forceinline__ uint64_t verify_one(const DevIndex& ix, const BatchDev& bd, uint32_t pool,
                                               const uint32_t* __restrict__ only, uint32_t n_only,
                                               uint64_t* peq)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t njobs = only ? n_only : bd.pool_n[pool];
    if (tid >= njobs) return 0;
    if (only && bd.ovf_pool[tid] != pool) return 0;
    const uint32_t j = only ? only[tid] : tid;
    const uint32_t meta = bd.job_meta[pool][j];
    const uint32_t ro = bd.job_ro[pool][j];
    const uint32_t r = ro >> 1, ori = ro & 1;
    const bool rev = (meta >> 31) != 0;
    const uint32_t id = meta & 0x7FFFFFFFu;
    const uint32_t m = bd.lens[r];
    const uint8_t* pat = (ori ? bd.rev : bd.seqs) + bd.offs[r];
    // Peq masks of the pattern (ShiftAnd loadBitmasks ShiftAnd.h:345-432):
    // pattern T also matches text C (bisulfite asymmetry)
#pragma unroll
    for (int i = 0; i < 4 * W; ++i) peq[i] = 0;
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t c = base_code(pat[i]);
        const uint64_t bit = 1ull << (i & 63);
        const uint32_t wi = i >> 6;
        if (c < 4) peq[c * W + wi] |= bit;
        if (c == 3) peq[1 * W + wi] |= bit;
    }
    const uint4 win = ix.win[id];
    const uint32_t startPos = win.x, chrom = win.y;
    const uint64_t clen = ix.clen[chrom];
    const uint8_t* cs = ix.seq + ix.coff[chrom];
    const uint64_t span = (uint64_t)ix.winlen + MISCOUNT + ADDMIS - 1;
    const uint8_t* text;
    uint64_t tlen;
    if (!rev) {
        uint64_t end = (uint64_t)startPos + span;
        if (end > clen) end = clen;
        text = cs + startPos;
        tlen = end - startPos;
    } else {
        const uint64_t e0 = startPos > 0 ? (uint64_t)startPos - 1 : 0;
        uint64_t s0 = (uint64_t)startPos + span;
        if (s0 >= clen) s0 = clen - 1;
        text = cs + e0 + 1;
        tlen = s0 > e0 ? s0 - e0 : 0;
    }
    uint64_t Pv[W], Mv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        Pv[w] = ~0ull;
        Mv[w] = 0;
    }
    int score = (int)m;
    const int lastw = (int)((m - 1) >> 6);
    const uint64_t lastbit = 1ull << ((m - 1) & 63);
    bool was = false;
    uint32_t prevErrs = MISCOUNT + 1;
    uint64_t ncomp = 0;
    uint32_t nm = 0;
    const uint32_t cap = bd.match_cap;
    uint32_t* outm = bd.job_match[pool] + (size_t)j * cap;
    uint32_t* ovf = nullptr;
    if (only) ovf = bd.ovf_buf + bd.job_ovf_off[pool][j];
    for (uint64_t q = 0; q < tlen; ++q) {
        const uint64_t p = rev ? tlen - 1 - q : q;
        uint32_t c = base_code(text[p]);
        if (c == 4) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                Pv[w] = ~0ull;
                Mv[w] = 0;
            }
            score = (int)m;
            continue;
        }
        if (rev) c = 3 - c; // complement (ShiftAnd.h:197-211)
        uint64_t carry = 0, phc = 0, mhc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t eq = peq[c * W + w];
            const uint64_t pv = Pv[w], mv = Mv[w];
            const uint64_t xv = eq | mv;
            const uint64_t t = eq & pv;
            uint64_t s1 = t + pv;
            const uint64_t c1 = s1 < t;
            const uint64_t s2 = s1 + carry;
            carry = c1 | (uint64_t)(s2 < s1);
            const uint64_t xh = (s2 ^ pv) | eq;
            const uint64_t ph = mv | ~(xh | pv);
            const uint64_t mh = pv & xh;
            if (w == lastw) score += (ph & lastbit) ? 1 : ((mh & lastbit) ? -1 : 0);
            const uint64_t ph2 = (ph << 1) | phc;
            const uint64_t mh2 = (mh << 1) | mhc;
            phc = ph >> 63;
            mhc = mh >> 63;
            Pv[w] = mh2 | ~(xv | ph2);
            Mv[w] = ph2 & xv;
        }
        ++ncomp;
        if (ncomp < (uint64_t)m - NERR) continue;
        if (score <= (int)NERR) {
            const uint32_t e = (uint32_t)score;
            // reported end: querySeq it-start+MISCOUNT; queryRevSeq it-end+pLen-1+MISCOUNT
            const uint32_t pos = rev ? (uint32_t)(p + 1 + m - 1 + MISCOUNT) : (uint32_t)(p + MISCOUNT);
            const uint32_t packed = (pos << 8) | e;
            if (was) {
                if (e <= prevErrs) {
                    const uint32_t k = nm - 1;
                    if (ovf) ovf[k] = packed;
                    else if (k < cap) outm[k] = packed;
                    prevErrs = e;
                }
            } else {
                if (ovf) ovf[nm] = packed;
                else if (nm < cap) outm[nm] = packed;
                ++nm;
                was = true;
                prevErrs = e;
            }
        } else {
            was = false;
        }
    }
    if (!only) {
        bd.job_nm[pool][j] = nm;
        if (nm > cap) {
            uint32_t o = atomicAdd(bd.ovf_n, 1u);
            if (o < bd.ovf_list_cap) {
                bd.ovf_pool[o] = pool;
                bd.ovf_job[o] = j;
            } else {
                atomicOr(bd.flags, FLAG_OVF_FULL);
            }
        }
    }
    return tlen ? tlen : 1;
}

template <int W>
__global__ __launch_bounds__(256) void k_verify(DevIndex ix, BatchDev bd, uint32_t pool,
                                                const uint32_t* __restrict__ only, uint32_t n_only)
{
    __shared__ uint64_t s_peq[256][4 * W];
    __shared__ unsigned long long s_work[2];
    if (threadIdx.x < 2) s_work[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t t = verify_one<W>(ix, bd, pool, only, n_only, s_peq[threadIdx.x]);
    if (t && !only) {
        atomicAdd(&s_work[0], 1ull);
        atomicAdd(&s_work[1], (unsigned long long)t);
    }
    __syncthreads();
    if (threadIdx.x < 2 && s_work[threadIdx.x]) atomicAdd(&bd.stats[3 + threadIdx.x], s_work[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// k_select: saQuerySeedSetRef uniqueness logic + decision block, per read
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mk_match(uint32_t off, uint32_t err, uint64_t fwd, uint64_t meta)
{
    // MATCH::constructMatch structs.h:185-188 (isStart is always 0)
    return (uint64_t)(off & 0xFFFFu) | ((uint64_t)(err & 0xFFu) << 16) | (fwd << 24) | (meta << 32);
}

__device__ __forceinline__ uint32_t job_match_at(const BatchDev& bd, uint32_t pool, uint32_t j,
                                                 uint32_t k)
{
    if (bd.job_nm[pool][j] > bd.match_cap) return bd.ovf_buf[bd.job_ovf_off[pool][j] + k];
    return bd.job_match[pool][(size_t)j * bd.match_cap + k];
}

__device__ int query_seed_set(const DevIndex& ix, const BatchDev& bd, uint32_t ro, uint32_t pool,
                              uint64_t* mat)
{
    uint8_t multi[NERR + 1];
    uint64_t uniq[NERR + 1];
#pragma unroll
    for (int i = 0; i <= (int)NERR; ++i) {
        multi[i] = 0;
        uniq[i] = 0;
    }
    const uint32_t off = bd.cand_off[ro], nf = bd.cand_nf[ro], nr = bd.cand_nr[ro];
    uint32_t prevChr = 0;
    uint64_t prevOff = ~0ull;
    for (uint32_t c = 0; c < nf + nr; ++c) {
        if (c == nf) {
            prevChr = 0;
            prevOff = ~0ull;
        }
        const bool isfwd = c < nf;
        const uint32_t j = off + c;
        const uint32_t id = bd.job_meta[pool][j] & 0x7FFFFFFFu;
        const uint4 w = ix.win[id];
        const uint32_t nm = bd.job_nm[pool][j];
        uint32_t i = 0;
        if (nm > 0) {
            const uint32_t m0 = job_match_at(bd, pool, j, 0);
            if ((uint64_t)(m0 >> 8) + w.x == prevOff && w.y == prevChr) ++i;
        }
        for (; i < nm; ++i) {
            const uint32_t mm = job_match_at(bd, pool, j, i);
            const uint32_t e = mm & 0xFF, pos = mm >> 8;
            if (multi[e]) {
                const uint64_t m2 = uniq[e];
                const bool s_fwd = (m2 >> 24) & 1;
                const bool same_strand = isfwd ? s_fwd : !s_fwd;
                if (same_strand && (uint64_t)ix.win[(uint32_t)(m2 >> 32)].x + (m2 & 0xFFFF) ==
                                       (uint64_t)w.x + pos)
                    continue;
                if (e == 0) return -1; // mat is not written (ReadQueue.cpp:2001)
                multi[e] = 2;
            } else {
                uniq[e] = mk_match(pos, e, isfwd ? 1 : 0, id);
                multi[e] = 1;
            }
        }
        if (nm > 0) {
            prevChr = w.y;
            prevOff = (uint64_t)w.x + (job_match_at(bd, pool, j, nm - 1) >> 8);
        } else {
            prevChr = 0;
            prevOff = ~0ull;
        }
    }
    for (uint32_t e = 0; e <= NERR; ++e) {
        if (!multi[e]) continue;
        *mat = uniq[e];
        return multi[e] > 1 ? -1 : 1;
    }
    return 0;
}

__global__ __launch_bounds__(256) void k_select(DevIndex ix, BatchDev bd, uint32_t n,
                                                int get_stranded)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ uint32_t s_cnt[5];
    if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    if (r < n) {
        uint8_t st = 3; // invalid, no counter
        uint64_t out = 0;
        uint8_t use = 0;
        if (bd.pre[r] == 0) {
            const uint32_t pool = (bd.lens[r] + 63) / 64 - 1;
            uint64_t mF = 0, mR = 0;
            int sF = 0, sR = 0;
            if (bd.ori_mask & 1) sF = query_seed_set(ix, bd, 2 * r, pool, &mF);
            if (bd.ori_mask & 2) sR = query_seed_set(ix, bd, 2 * r + 1, pool, &mR);
            const uint32_t eF = (uint32_t)((mF >> 16) & 0xFF), eR = (uint32_t)((mR >> 16) & 0xFF);
            int useFwd = -1, tally = 0;
            if (sF == 1 && sR == 1) {
                if (eF < eR) {
                    useFwd = 1;
                    tally = 1;
                } else if (eF > eR) {
                    useFwd = 0;
                    tally = -1;
                } else {
                    const uint32_t p1 = ix.win[(uint32_t)(mF >> 32)].x + (uint32_t)(mF & 0xFFFF);
                    const uint32_t p2 = ix.win[(uint32_t)(mR >> 32)].x + (uint32_t)(mR & 0xFFFF);
                    const bool f1 = (mF >> 24) & 1, f2 = (mR >> 24) & 1;
                    if (f1 == f2 && p1 == p2) {
                        useFwd = 1;
                        tally = 1;
                    } else {
                        st = 1;
                    }
                }
            } else if (sF == 1) {
                if (sR == -1 && !(eF < eR))
                    st = 1;
                else {
                    useFwd = 1;
                    tally = 1;
                }
            } else if (sR == 1) {
                if (sF == -1 && !(eR < eF))
                    st = 1;
                else {
                    useFwd = 0;
                    tally = -1;
                }
            } else {
                st = (sF == -1 || sR == -1) ? 1 : 2;
            }
            if (useFwd >= 0) {
                st = 0;
                out = useFwd ? mF : mR;
                use = useFwd ? 1 : 2;
                if (get_stranded) atomicAdd(&s_cnt[tally > 0 ? 3 : 4], 1u);
            }
            atomicAdd(&s_cnt[st], 1u);
        }
        bd.status[r] = st;
        bd.use[r] = use;
        bd.mat[r] = out;
        bd.valid[r] = st == 0;
    }
    __syncthreads();
    if (threadIdx.x < 5 && s_cnt[threadIdx.x])
        atomicAdd((unsigned long long*)&bd.counters[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// k_meth: computeMethLvl
// ---------------------------------------------------------------------------

// CompiFwd / CompiRev ReadQueue.h:295-338
__device__ __forceinline__ int cmp_fwd(uint8_t r, uint8_t g)
{
    if (r == 'T' && g == 'C') return 0;
    return r == g ? 0 : 1;
}
__device__ __forceinline__ int cmp_rev(uint8_t r, uint8_t g)
{
    switch (r) {
    case 'T': return (g == 'A' || g == 'G') ? 0 : 1;
    case 'G': return g == 'C' ? 0 : 1;
    case 'A': return g == 'T' ? 0 : 1;
    case 'C': return g == 'G' ? 0 : 1;
    default: return 1;
    }
}

// BandedMatrix<uint16_t, 9>::operator() BandedMatrix.h:67-75, stored per
// lane with a stride so that the lanes of a wave touch consecutive words
struct DpMat {
    uint16_t* base;
    size_t stride;
    long rows;
    __device__ __forceinline__ long idx(long r, long c) const
    {
        const long bnd = DPBAND + 1;
        return r * (2 * bnd + 1) + ((bnd - r > 0) ? bnd - r : 0) + (c - r + bnd);
    }
    __device__ __forceinline__ uint16_t get(long r, long c) const
    {
        long i = idx(r, c);
        if (i < 0 || i >= rows * (2 * (DPBAND + 1) + 1)) return 0xFFFF;
        return base[(size_t)i * stride];
    }
    __device__ __forceinline__ void put(long r, long c, uint16_t v) const
    {
        long i = idx(r, c);
        if (i < 0 || i >= rows * (2 * (DPBAND + 1) + 1)) return;
        base[(size_t)i * stride] = v;
    }
};

// *(colPat - j + 1) with colPat = chromosome + ref_end; bytes outside the
// chromosome read as 0
__device__ __forceinline__ uint8_t col_at(const uint8_t* cs, uint64_t clen, int64_t ref_end, long j)
{
    const int64_t p = ref_end - j + 1;
    return (p < 0 || (uint64_t)p >= clen) ? 0 : cs[p];
}

__global__ __launch_bounds__(256) void k_meth(DevIndex ix, BatchDev bd, uint32_t n, uint16_t* dp_scratch,
                                              size_t dp_stride, uint8_t* trace_scratch)
{
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gsz = gridDim.x * blockDim.x;
    const uint32_t RL = ix.readlen;
    const uint32_t RL2 = RL - 2;
    for (uint32_t r = gid; r < n; r += gsz) {
        if (bd.status[r] != 0) continue;
        const uint64_t mat = bd.mat[r];
        const bool isFwd = (mat >> 24) & 1;
        const uint32_t metaID = (uint32_t)(mat >> 32);
        const uint32_t offset = (uint32_t)(mat & 0xFFFF);
        const uint32_t errNum = (uint32_t)((mat >> 16) & 0xFF);
        const uint4 m = ix.win[metaID];
        if (m.z == CPGDUMMY) continue;
        const uint32_t L = bd.lens[r];
        const uint8_t* seq = (bd.use[r] == 1 ? bd.seqs : bd.rev) + bd.offs[r];
        const uint32_t metaPos = m.x;
        const uint32_t minPos = (uint32_t)((uint64_t)metaPos + offset - ((uint64_t)L - 1));
        const uint32_t maxPos = metaPos + offset;
        uint32_t* meth = bd.meth;
        if (errNum == 0) {
            for (uint32_t id = m.z; id <= m.w; ++id) {
                const uint32_t pos = ix.cpg_pos[id];
                if (pos + RL2 < minPos) continue;
                if (isFwd) {
                    if (pos + RL2 > maxPos) break;
                } else {
                    if (pos + RL - 1 > maxPos) break;
                }
                const uint32_t trueCpGPos = pos - 1 + RL - 1;
                const uint32_t readCpGPos =
                    (uint32_t)((uint64_t)L - (uint64_t)(uint32_t)(metaPos + offset - trueCpGPos) - 1);
                if (isFwd) {
                    const uint32_t k = readCpGPos + MISCOUNT;
                    if ((uint64_t)k >= L) continue;
                    const uint8_t ch = seq[k];
                    if (ch == 'T')
                        atomicAdd(&meth[4ull * id + 1], 1u);
                    else if (ch == 'C')
                        atomicAdd(&meth[4ull * id + 0], 1u);
                } else {
                    if ((uint64_t)(uint32_t)(readCpGPos + 1 + MISCOUNT + 2) > L) continue;
                    const uint8_t ch = seq[(uint64_t)L - readCpGPos - 1 - MISCOUNT - 2];
                    if (ch == 'T')
                        atomicAdd(&meth[4ull * id + 3], 1u);
                    else if (ch == 'C')
                        atomicAdd(&meth[4ull * id + 2], 1u);
                }
            }
            continue;
        }
        // DP branch ReadQueue.cpp:3625-3912
        atomicAdd(&bd.stats[5], 1ull);
        const uint32_t chrom = ix.cpg_chrom[m.z];
        const uint8_t* cs = ix.seq + ix.coff[chrom];
        const uint64_t clen = ix.clen[chrom];
        const int64_t ref_end = (int64_t)metaPos + offset;
        int32_t minIndex = (int32_t)m.z, maxIndex = (int32_t)m.w;
        for (int32_t id = (int32_t)m.z; (uint32_t)id <= m.w; ++id) {
            const uint32_t pos = ix.cpg_pos[id];
            if (pos + RL2 < minPos) {
                ++minIndex;
            } else if (pos + RL2 > maxPos) {
                maxIndex = id - 1;
                break;
            }
        }
        DpMat M{dp_scratch + gid, dp_stride, (long)L + 1};
        const long band = DPBAND;
        const long ncell = M.rows * (2 * (band + 1) + 1);
        for (long i = 0; i < ncell; ++i) M.base[(size_t)i * dp_stride] = 0xFFFF;
        // border init in the reference's write order (LevenshtDP.h:212-233)
        M.put(0, 0, 0);
        for (long c = 1; c <= band; ++c) M.put(0, c, (uint16_t)c);
        for (long rr = 1; rr <= band; ++rr) M.put(rr, 0, (uint16_t)rr);
        for (long c = 0; c < (long)(L - band); ++c) M.put(c + band + 1, c, (uint16_t)(0xFFFF - c - band - 1));
        for (long rr = 0; rr < (long)L; ++rr) M.put(rr, rr + band + 1, (uint16_t)(0xFFFF - rr - band - 1));
        const bool rv = !isFwd;
        for (long rr = 1; rr <= (long)L; ++rr)
            for (long o = -band; o <= band; ++o) {
                if (rr + o <= 0) continue;
                const long j = rr + o;
                const uint8_t rc = rv ? seq[rr - 1] : seq[L - rr];
                const uint8_t gc = col_at(cs, clen, ref_end, j);
                const int mm = rv ? cmp_rev(rc, gc) : cmp_fwd(rc, gc);
                int v = (int)M.get(rr - 1, j) + 1;
                const int v2 = (int)M.get(rr, j - 1) + 1;
                const int v3 = (int)M.get(rr - 1, j - 1) + mm;
                v = v2 < v ? v2 : v;
                v = v3 < v ? v3 : v;
                M.put(rr, j, (uint16_t)v);
            }
        // backtrace (backtrackDP / backtrackDPRev LevenshtDP.h:329-544)
        uint16_t minimum = 0xFFFF;
        unsigned long col = 0;
        for (long o = -band; o <= band; ++o) {
            const uint16_t v = M.get((long)L, (long)L + o);
            if (v < minimum) {
                col = (unsigned long)((long)L + o);
                minimum = v;
            }
        }
        uint8_t* tr = trace_scratch + gid;
        const size_t tstride = gsz;
        int tn = 0;
        for (long rr = (long)L; rr > 0;) {
            const uint8_t rc = rv ? seq[rr - 1] : seq[L - rr];
            const uint8_t gc = col_at(cs, clen, ref_end, (long)col);
            const int mm = rv ? cmp_rev(rc, gc) : cmp_fwd(rc, gc);
            const long c = (long)col;
            uint8_t t;
            const uint16_t up = M.get(rr - 1, c), left = M.get(rr, c - 1), dg = M.get(rr - 1, c - 1);
            if (up < left) {
                if ((int)up + 1 < (int)dg + mm) {
                    t = 2;
                    --rr;
                } else {
                    t = mm ? 1 : 0;
                    --rr;
                    --col;
                }
            } else {
                if ((int)left + 1 < (int)dg + mm) {
                    t = 3;
                    --col;
                } else {
                    t = mm ? 1 : 0;
                    --rr;
                    --col;
                }
            }
            if (tn < FAME_DEV_TRACE_CAP) tr[(size_t)tn * tstride] = t;
            ++tn;
        }
        while (col > 0) {
            if (tn < FAME_DEV_TRACE_CAP) tr[(size_t)tn * tstride] = 3;
            ++tn;
            --col;
        }
        if (tn > FAME_DEV_TRACE_CAP) {
            atomicOr(bd.flags, FLAG_CAPACITY);
            continue;
        }
        uint32_t refSeqPos = metaPos + offset;
        int32_t readSeqPos = (int32_t)L - 1;
        int32_t alignPos = tn - 1;
        for (int32_t id = maxIndex; id >= minIndex; --id) {
            const uint32_t pos = ix.cpg_pos[id];
            while (alignPos >= 0 && pos + RL2 < refSeqPos) {
                switch (tr[(size_t)alignPos * tstride]) {
                case 0:
                case 1:
                    --readSeqPos;
                    --refSeqPos;
                    break;
                case 3:
                    --refSeqPos;
                    break;
                default:
                    --readSeqPos;
                    break;
                }
                if (readSeqPos < 0) break;
                --alignPos;
            }
            if (readSeqPos < 0) break;
            if (isFwd) {
                const uint8_t ch = seq[readSeqPos];
                if (ch == 'C')
                    atomicAdd(&meth[4ull * id + 0], 1u);
                else if (ch == 'T')
                    atomicAdd(&meth[4ull * id + 1], 1u);
            } else {
                // std::reverse(seq) (ReadQueue.cpp:3796), then seq[readSeqPos+1];
                // position L of the reversed string is its terminating '\0'
                const int32_t k = readSeqPos + 1;
                const uint8_t ch = (k >= (int32_t)L) ? 0 : seq[L - 1 - k];
                if (ch == 'C')
                    atomicAdd(&meth[4ull * id + 2], 1u);
                else if (ch == 'T')
                    atomicAdd(&meth[4ull * id + 3], 1u);
            }
        }
    }
}

// exclusive scan of the match counts of overflowed jobs (one thread; the
// list is empty in practice and short when not)
__global__ void k_ovf_offsets(BatchDev bd, uint32_t n_ovf, uint32_t* total)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t o = 0;
    for (uint32_t i = 0; i < n_ovf; ++i) {
        const uint32_t p = bd.ovf_pool[i], j = bd.ovf_job[i];
        bd.job_ovf_off[p][j] = o;
        o += bd.job_nm[p][j];
    }
    *total = o;
}

// ---------------------------------------------------------------------------
// index preparation kernels
// ---------------------------------------------------------------------------

__global__ void k_pack_kmers(const uint32_t* __restrict__ raw, const uint8_t* __restrict__ strand_bits,
                             uint64_t first, uint64_t n, uint32_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t g = first + i;
    const uint32_t meta = raw[2 * i] & 0x7FFFFFFFu; // KMER_S::getMetaCpG
    const uint32_t fwd = (strand_bits[g >> 3] >> (g & 7)) & 1;
    out[g] = (meta << 1) | fwd;
}

__global__ void k_narrow_tab(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)in[i];
}

__global__ void k_meth_u16(const uint32_t* __restrict__ in, uint64_t n, uint16_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint16_t)in[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_prep(const BatchDev& bd, uint32_t n, uint32_t readlen, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_prep, dim3((n + 255) / 256), dim3(256), 0, st, bd.seqs, bd.offs, bd.lens, n,
                       readlen, bd.rev, bd.pre, bd.flags);
}

void launch_seed_light(const DevIndex& ix, const BatchDev& bd, uint32_t n_ro, hipStream_t st)
{
    if (!n_ro) return;
    if (ix.tab_is64)
        hipLaunchKernelGGL(k_seed_light<uint64_t>, dim3(n_ro), dim3(64), 0, st, ix, bd, n_ro);
    else
        hipLaunchKernelGGL(k_seed_light<uint32_t>, dim3(n_ro), dim3(64), 0, st, ix, bd, n_ro);
}

void launch_seed_heavy(const DevIndex& ix, const BatchDev& bd, const HeavyScratch& hs, hipStream_t st)
{
    if (!hs.count) return;
    if (ix.tab_is64)
        hipLaunchKernelGGL(k_seed_heavy<uint64_t>, dim3(hs.count), dim3(64), 0, st, ix, bd, hs);
    else
        hipLaunchKernelGGL(k_seed_heavy<uint32_t>, dim3(hs.count), dim3(64), 0, st, ix, bd, hs);
}

void launch_verify(const DevIndex& ix, const BatchDev& bd, uint32_t pool, uint32_t njobs,
                   const uint32_t* only, hipStream_t st)
{
    if (!njobs) return;
    const dim3 g((njobs + 255) / 256), b(256);
    switch (pool) {
    case 0: hipLaunchKernelGGL(k_verify<1>, g, b, 0, st, ix, bd, 0u, only, only ? njobs : 0u); break;
    case 1: hipLaunchKernelGGL(k_verify<2>, g, b, 0, st, ix, bd, 1u, only, only ? njobs : 0u); break;
    case 2: hipLaunchKernelGGL(k_verify<3>, g, b, 0, st, ix, bd, 2u, only, only ? njobs : 0u); break;
    default: hipLaunchKernelGGL(k_verify<4>, g, b, 0, st, ix, bd, 3u, only, only ? njobs : 0u); break;
    }
}

void launch_select(const DevIndex& ix, const BatchDev& bd, uint32_t n, int get_stranded, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_select, dim3((n + 255) / 256), dim3(256), 0, st, ix, bd, n, get_stranded);
}

void launch_meth(const DevIndex& ix, const BatchDev& bd, uint32_t n, uint16_t* dp, size_t dp_stride,
                 uint8_t* trace, uint32_t grid_threads, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_meth, dim3(grid_threads / 256), dim3(256), 0, st, ix, bd, n, dp, dp_stride, trace);
}

void launch_ovf_offsets(const BatchDev& bd, uint32_t n_ovf, uint32_t* total, hipStream_t st)
{
    hipLaunchKernelGGL(k_ovf_offsets, dim3(1), dim3(1), 0, st, bd, n_ovf, total);
}

__global__ void k_max_end(const uint64_t* __restrict__ offs, const uint32_t* __restrict__ lens, uint32_t n,
                          unsigned long long* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, (unsigned long long)(offs[i] + lens[i]));
}

void launch_max_end(const uint64_t* offs, const uint32_t* lens, uint32_t n, unsigned long long* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_max_end, dim3((n + 255) / 256), dim3(256), 0, st, offs, lens, n, out);
}

void launch_pack_kmers(const uint32_t* raw, const uint8_t* sbits, uint64_t first, uint64_t n, uint32_t* out,
                       hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_pack_kmers, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, raw, sbits, first,
                       n, out);
}

void launch_narrow_tab(const uint64_t* in, uint64_t n, uint32_t* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_narrow_tab, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out);
}

void launch_meth_u16(const uint32_t* in, uint64_t n, uint16_t* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_meth_u16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out);
}

} // namespace fame
