// Device helpers shared by the single-end (fame_device.hip) and paired-end
// (fame_paired.hip) kernels: FAME's constants, the reduced spaced-seed hash
// tables, wavefront scans, the tsl::hopscotch_map placement emulation, the
// verify job pools and MATCH packing.  Device code only.
#pragma once

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
constexpr uint32_t JOB_GAP = 0xFFFFFFFFu; // unused verify slot
// spaced seed (CONST.h:57-58); RefGenome::load rejects other seeds
// (RefGenome.cpp:275-280), and so does fame_open
constexpr uint32_t SEEDBITS = FAME_SEEDBITS;
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
// verify pool of a read of L letters: 64-bit words - 1 for the short-read
// kernels, FAME_LONG_POOL beyond FAME_DEV_MAX_READ
__device__ __forceinline__ uint32_t pool_of_len(uint32_t L)
{
    return L > FAME_DEV_MAX_READ ? (uint32_t)FAME_LONG_POOL : (L + 63) / 64 - 1;
}

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

// inclusive scan over the 64 lanes of a wave with DPP moves: shifts by 1,
// 2, 4, 8 inside each row of 16 lanes, then the row totals are broadcast
// into the rows above (row_bcast:15 into rows 1 and 3, row_bcast:31 into
// rows 2 and 3).  Lanes without a source take 0.
__device__ __forceinline__ uint32_t wave_iscan_u32(uint32_t v, int lane)
{
    (void)lane;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false); // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false); // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false); // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false); // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false); // row_bcast:31
    return v;
}

// inclusive max-scan over the 64 lanes of a wave (values >= -1; lanes
// without a source take -1), the same DPP pattern as wave_iscan_u32
__device__ __forceinline__ int wave_imaxscan_i32(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xC, 0xF, false));
    return v;
}

// each lane receives the value of the lane below it (lane 0: -1), wave_shr:1
__device__ __forceinline__ int wave_shr1_i32(int v) { return __builtin_amdgcn_update_dpp(-1, v, 0x138, 0xF, 0xF, false); }

// lane i's value of a 32-bit register, to every lane (a scalar read)
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int i) { return (uint32_t)__builtin_amdgcn_readlane((int)v, i); }

// sum over the 64 lanes of a wave (every lane receives it)
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_iscan_u32(v, 0), 63);
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

// Placement of keys 0..n-1 (key t held by lane t) into a fresh hopscotch table
// of bucket_count buckets without the table, where the insertions need no
// rehash (n at most 0.9 B) and no displacement beyond the neighbourhood: the
// table is then linear probing from each key's home bucket (key & (B - 1)) in
// insertion order (find_empty_bucket, hopscotch_hash.h:1491-1500), kept as a
// bitmap of at most 128 buckets (B <= 64, n <= 64) in scalar registers.
// Lane t receives key t's bucket in mypos.  Returns false where that does not
// hold (the caller emulates the table).
__device__ __forceinline__ bool lp_place(uint32_t mykey, uint32_t n, uint64_t bucket_count, uint64_t tb_cap,
                                         uint32_t& mypos)
{
    const uint64_t B = bucket_count ? rup_pow2(bucket_count) : 0;
    const uint64_t arr = B ? B + NBH - 1 : 0;
    if (B == 0 || B > 64 || n > 64 || arr > tb_cap || n > hs_thr(B, 0.9f)) return false;
    const uint32_t lane = threadIdx.x & 63;
    uint64_t occ0 = 0, occ1 = 0; // buckets 0..63, 64..127
    mypos = 0;
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t ib = lane_u32(mykey, (int)t) & (uint32_t)(B - 1); // < 64
        const uint64_t f0 = ~occ0 & (~0ull << ib);
        const uint32_t e = f0 ? (uint32_t)__builtin_ctzll(f0) : (~occ1 ? 64u + (uint32_t)__builtin_ctzll(~occ1) : 128u);
        if (e >= arr || e - ib >= NBH) return false;
        if (e < 64) occ0 |= 1ull << e;
        else occ1 |= 1ull << (e - 64);
        if (lane == t) mypos = e;
    }
    return true;
}

// bucket-order rank of each selected lane's key (bucket mypos < 128) among the
// selected keys, and their number (the table's iteration order, by bucket)
__device__ __forceinline__ uint32_t lp_rank(bool sel, uint32_t mypos, uint32_t& rank)
{
    uint64_t k0 = 0, k1 = 0;
    for (uint64_t b = __ballot(sel); b; b &= b - 1) {
        const uint32_t p = lane_u32(mypos, (int)__builtin_ctzll(b));
        if (p < 64) k0 |= 1ull << p;
        else k1 |= 1ull << (p - 64);
    }
    rank = mypos < 64 ? (uint32_t)__popcll(k0 & ((1ull << mypos) - 1))
                      : (uint32_t)(__popcll(k0) + __popcll(k1 & ((1ull << (mypos - 64)) - 1)));
    return (uint32_t)(__popcll(k0) + __popcll(k1));
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
// tsl::hopscotch_map<uint32_t, uint16_t, identity> placement emulation,
// executed by a whole wavefront: the probes for an empty bucket, the
// neighbourhood checks, the rehash collection and the final iteration are
// ballots over 64 buckets at a time; the decisions (and so the control flow)
// are wave-uniform.  Only the order of iteration is reproduced: keys are
// inserted in the order in which operator[] first created them (count
// updates never move an element).  tb[] holds, per physical bucket, the
// index of the key stored there.
// ---------------------------------------------------------------------------

template <typename IdxT>
struct Hs {
    IdxT* tb;
    uint32_t tb_cap;
    IdxT* over;
    uint32_t over_cap;
    IdxT* tmp;
    uint32_t tmp_cap;
    const uint32_t* key;
    uint64_t B, mask, arr;
    uint32_t nelem, nover, max_thr, min_thr;
    static constexpr IdxT NONE = (IdxT)~(IdxT)0;
};

enum { HS_OK = 0, HS_FAIL = 1, HS_NEED_REHASH = 2 };

// the lanes of one wavefront see each other's LDS / global writes: a
// wavefront-scope release/acquire pair (no instruction on CDNA, an ordering
// point for the compiler); the waves of a block run independently
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

// Prefix form of the spaced-seed hash.  With Q[t] = rol(seed(c_t), -t) and
// P[j] = Q[0] ^ ... ^ Q[j-1] over a read's reduced letters c_t, the hash of
// k-mer k,
//   XOR over care positions i of rol(seed(c_{k+i}), 31 - i),
// equals rol(XOR over care i of Q[k+i], 31 + k); a run [a, b] of care
// positions contributes P[k+b+1] ^ P[k+a], so the hash is rol of the XOR of
// P[k + e] over the ends e of the runs: ten LDS reads per k-mer for the
// shipped seed instead of 28 letters and 28 table terms.
struct SeedEnds {
    int n;
    int e[KMERLEN + 1];
};
constexpr SeedEnds seed_ends()
{
    SeedEnds s{};
    bool prev = false;
    for (int i = 0; i <= (int)KMERLEN; ++i) {
        const bool cur = i < (int)KMERLEN && ((SEEDBITS >> (KMERLEN - 1 - i)) & 1);
        if (cur != prev) s.e[s.n++] = i;
        prev = cur;
    }
    return s;
}

// inclusive XOR scan of a u64 over the lanes (the DPP pattern of
// wave_iscan_u32 on both halves)
__device__ __forceinline__ uint32_t wave_ixor_u32(uint32_t v)
{
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false); // row_shr:1
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false); // row_shr:2
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false); // row_shr:4
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false); // row_shr:8
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false); // row_bcast:15
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false); // row_bcast:31
    return v;
}
__device__ __forceinline__ uint64_t wave_ixor_u64(uint64_t v)
{
    return ((uint64_t)wave_ixor_u32((uint32_t)(v >> 32)) << 32) | wave_ixor_u32((uint32_t)v);
}

// hopscotch_hash ctor (hopscotch_hash.h:599-628) on a fresh bucket array;
// the overflow list is kept (rehash_impl moves it to the new table)
template <typename IdxT>
__device__ bool hs_setup(Hs<IdxT>& h, uint64_t bucket_count, int lane)
{
    const uint64_t B = bucket_count ? rup_pow2(bucket_count) : 0;
    const uint64_t arr = B ? B + NBH - 1 : 0;
    if (arr > h.tb_cap) return false;
    h.B = B;
    h.mask = B ? B - 1 : 0;
    h.arr = arr;
    for (uint64_t i = lane; i < arr; i += 64) h.tb[i] = Hs<IdxT>::NONE;
    h.max_thr = hs_thr(B, 0.9f);
    h.min_thr = hs_thr(B, 0.1f);
    wave_sync();
    return true;
}

// find_empty_bucket hopscotch_hash.h:1491-1500
template <typename IdxT>
__device__ uint64_t hs_find_empty(const Hs<IdxT>& h, uint64_t ib, int lane)
{
    const uint64_t lim = ib + MAX_PROBES < h.arr ? ib + MAX_PROBES : h.arr;
    for (uint64_t b0 = ib; b0 < lim; b0 += 64) {
        const uint64_t b = b0 + lane;
        const uint64_t bal = __ballot(b < lim && h.tb[b] == Hs<IdxT>::NONE);
        if (bal) return b0 + __ffsll((long long)bal) - 1;
    }
    return h.arr;
}

// swap_empty_bucket_closer hopscotch_hash.h:1545-1578 (rare: the
// neighbourhood of a home bucket is full); computed by lane 0, where the
// neighbourhood bitmap of bucket tc is "the buckets tc+i holding a key whose
// home is tc"
template <typename IdxT>
__device__ bool hs_swap_closer(Hs<IdxT>& h, uint64_t& e, int lane)
{
    uint64_t ne = e;
    int found = 0;
    if (lane == 0) {
        const uint64_t start = e - NBH + 1;
        for (uint64_t tc = start; tc < e && !found; ++tc) {
            for (uint64_t ts = tc; ts < e; ++ts) {
                const IdxT i = h.tb[ts];
                if (i != Hs<IdxT>::NONE && (h.key[i] & h.mask) == tc) {
                    h.tb[e] = i;
                    h.tb[ts] = Hs<IdxT>::NONE;
                    ne = ts;
                    found = 1;
                    break;
                }
            }
        }
    }
    wave_sync();
    found = __shfl(found, 0, 64);
    e = __shfl((unsigned long long)ne, 0, 64);
    return found != 0;
}

// will_neighborhood_change_on_rehash hopscotch_hash.h:1466-1486
template <typename IdxT>
__device__ bool hs_will_change(const Hs<IdxT>& h, uint64_t ib, int lane)
{
    const uint64_t mask2 = rup_pow2((h.mask + 1) * 2) - 1;
    const uint64_t b = ib + lane;
    bool ch = false;
    if (lane < (int)NBH && b < h.arr) {
        const IdxT i = h.tb[b];
        if (i != Hs<IdxT>::NONE) {
            const uint64_t k = h.key[i];
            ch = (k & h.mask) != (k & mask2);
        }
    }
    return __ballot(ch) != 0;
}

// one pass of insert_impl hopscotch_hash.h:1426-1458 without its rehash
// calls: HS_NEED_REHASH where the reference rehashes (load threshold reached
// first, or no bucket in the neighbourhood and no overflow allowed)
template <typename IdxT>
__device__ int hs_place(Hs<IdxT>& h, uint32_t idx, int lane)
{
    if (h.nelem - h.nover >= h.max_thr) return HS_NEED_REHASH;
    const uint64_t ib = h.key[idx] & h.mask;
    uint64_t e = hs_find_empty(h, ib, lane);
    if (e < h.arr) {
        for (;;) {
            if (e - ib < NBH) {
                if (lane == 0) h.tb[e] = (IdxT)idx;
                wave_sync();
                h.nelem++;
                return HS_OK;
            }
            if (!hs_swap_closer(h, e, lane)) break;
        }
    }
    if (h.nelem < h.min_thr || !hs_will_change(h, ib, lane)) {
        if (h.nover >= h.over_cap) return HS_FAIL;
        if (lane == 0) h.over[h.nover] = (IdxT)idx;
        wave_sync();
        h.nover++;
        h.nelem++;
        return HS_OK;
    }
    return HS_NEED_REHASH;
}

// rehash_impl hopscotch_hash.h:1244-1290: a new table, the overflow list
// carried over, then the old buckets re-inserted in physical order.  A
// rehash needed inside this re-insertion is reported as a capacity failure.
template <typename IdxT>
__device__ int hs_rehash(Hs<IdxT>& h, uint64_t count, int lane)
{
    const uint64_t need = hs_ceil_div09(h.nelem);
    if (need > count) count = need;
    uint32_t n = 0;
    for (uint64_t b0 = 0; b0 < h.arr; b0 += 64) {
        const uint64_t b = b0 + lane;
        const IdxT i = b < h.arr ? h.tb[b] : Hs<IdxT>::NONE;
        const bool occ = i != Hs<IdxT>::NONE;
        const uint64_t bal = __ballot(occ);
        const uint32_t p = n + __popcll(bal & lanes_below(lane));
        if (occ && p < h.tmp_cap) h.tmp[p] = i;
        n += __popcll(bal);
    }
    if (n > h.tmp_cap) return HS_FAIL;
    wave_sync();
    if (!hs_setup(h, count, lane)) return HS_FAIL;
    h.nelem = h.nover;
    for (uint32_t t = 0; t < n; ++t)
        if (hs_place(h, (uint32_t)h.tmp[t], lane) != HS_OK) return HS_FAIL;
    return HS_OK;
}

// operator[] creating a key: insert_impl with its rehash-and-retry loop
template <typename IdxT>
__device__ int hs_insert(Hs<IdxT>& h, uint32_t idx, int lane)
{
    for (;;) {
        const int r = hs_place(h, idx, lane);
        if (r == HS_OK) return HS_OK;
        if (r == HS_FAIL) return HS_FAIL;
        if (hs_rehash(h, (h.mask + 1) * 2, lane) != HS_OK) return HS_FAIL;
    }
}

struct SeedShared {
    // rot[i][c] = rol(seed_c, 31 - i): the term of position i of a spaced
    // k-mer (nthash.hpp:480-489); code 3 (N and others) contributes 0
    uint64_t rot[KMERLEN][4];
};

// filled by the whole block at kernel start
__device__ __forceinline__ void seed_table_init(SeedShared& sh)
{
    for (int t = threadIdx.x; t < (int)KMERLEN * 4; t += blockDim.x) {
        const int i = t >> 2, c = t & 3;
        sh.rot[i][c] = rol64(red_seed((uint32_t)c), (int)(KMERLEN - 1 - i));
    }
    __syncthreads();
}

// Returns 0 when done, 1 when the (read, orientation) needs the heavy path
// (capacity); E / bucketCount are reported through the out pointers.
// Candidate slots of the verify pools.  Persistent seed waves take pool
// space in chunks (one atomic per chunk instead of one per read); the slots
// a wave leaves unused are marked JOB_GAP.  `exact` takes exactly what a
// read needs (heavy path).
struct PoolChunk {
    uint32_t pool, base, left; // one open chunk at a time (a batch rarely mixes pools)
};
constexpr uint32_t POOL_CHUNK = 64;

__device__ __forceinline__ uint32_t pc_get_cap(const BatchDev& bd, uint32_t p)
{
    uint32_t v = bd.pool_cap[0];
#pragma unroll
    for (uint32_t q = 1; q < FAME_NPOOL; ++q) v = p == q ? bd.pool_cap[q] : v;
    return v;
}
__device__ __forceinline__ uint32_t* pc_meta(const BatchDev& bd, uint32_t p)
{
    uint32_t* v = bd.job_meta[0];
#pragma unroll
    for (uint32_t q = 1; q < FAME_NPOOL; ++q) v = p == q ? bd.job_meta[q] : v;
    return v;
}

__device__ inline void pool_gap_fill(const BatchDev& bd, uint32_t p, uint32_t base, uint32_t n, int lane)
{
    if (base + n > pc_get_cap(bd, p)) return;
    uint32_t* meta = pc_meta(bd, p);
    for (uint32_t t = lane; t < n; t += 64) meta[base + t] = JOB_GAP;
}

// returns the first slot of `need` consecutive slots, or ~0u if the pool is full
__device__ inline uint32_t pool_take(const BatchDev& bd, PoolChunk& pc, uint32_t p, uint32_t need, bool exact, int lane)
{
    if (exact) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&bd.pool_n[p], need);
        b = lane_u32(b, 0);
        return b + need <= bd.pool_cap[p] ? b : ~0u;
    }
    if (pc.pool != p || pc.left < need) {
        if (pc.left) pool_gap_fill(bd, pc.pool, pc.base, pc.left, lane);
        const uint32_t sz = need > POOL_CHUNK ? need : POOL_CHUNK;
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&bd.pool_n[p], sz);
        b = lane_u32(b, 0);
        pc.pool = p;
        pc.base = b;
        pc.left = sz;
        if (b + sz > pc_get_cap(bd, p)) {
            pc.left = 0;
            return ~0u;
        }
    }
    const uint32_t b = pc.base;
    pc.base += need;
    pc.left -= need;
    return b;
}

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


// slot table entries (k_build_slots): bucket length, first entry or kent
// offset, second entry
__device__ inline uint32_t slot_len(uint64_t s)
{
    const uint32_t t = (uint32_t)(s >> 62);
    return t < 3 ? t : ((uint32_t)(s >> 32) & 0x3FFFFFFFu);
}
__device__ inline uint32_t slot_lo(uint64_t s) { return (s >> 62) == 3 ? (uint32_t)s : ((uint32_t)s & 0x7FFFFFFFu); }
__device__ inline uint32_t slot_e1(uint64_t s) { return (uint32_t)(s >> 31) & 0x7FFFFFFFu; }

} // namespace fame
