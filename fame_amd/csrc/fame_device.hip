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
#include "fame_dev_common.h"

namespace fame {

// ---------------------------------------------------------------------------
// k_prep
// ---------------------------------------------------------------------------

// one wavefront per read (persistent, grid-stride over groups of 64 reads):
// coalesced loads of the read, the reverse complement stored reversed, N
// found by a ballot
// Also writes, per (read, orientation) to be seeded, a 256-byte row of the
// reduced letters (red_code) with the length in byte 255 (0 for a pair the
// seed stage skips), so that the seed stage reads one coalesced row per pair
// instead of chasing offsets.
__global__ __launch_bounds__(256) void k_prep(const uint8_t* __restrict__ seqs, const uint64_t* __restrict__ offs,
                                              const uint32_t* __restrict__ lens, uint32_t n, uint32_t readlen,
                                              uint8_t* __restrict__ rev, uint8_t* __restrict__ pre,
                                              uint8_t* __restrict__ codes, uint32_t ori_mask,
                                              uint32_t* __restrict__ err_flag, uint32_t* __restrict__ max_len,
                                              uint32_t nsplit, uint32_t minlen, int paired,
                                              const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ long_n,
                                              uint32_t* __restrict__ long_list)
{
    // per byte value c: complement (bits 0-7), red_code(c) (8-9),
    // red_code(complement) (10-11), c == 'N' (12), c in A C G T N (13)
    __shared__ uint16_t s_lt[256];
    for (uint32_t v = threadIdx.x; v < 256; v += blockDim.x) {
        uint32_t o = 0;
        switch (v) {
        case 'A': o = 'T'; break;
        case 'C': o = 'G'; break;
        case 'G': o = 'C'; break;
        case 'T': o = 'A'; break;
        default: break;
        }
        const bool known = v == 'A' || v == 'C' || v == 'G' || v == 'T' || v == 'N';
        s_lt[v] = (uint16_t)(o | (red_code((uint8_t)v) << 8) | (red_code((uint8_t)o) << 10) | ((v == 'N') << 12) |
                             (known << 13));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    if (n_dev) n = *n_dev;
    uint32_t lmax = 0; // longest read this wave prepared for seeding
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwave = (gridDim.x * blockDim.x) >> 6;
    // groups of 64 reads per wave: one coalesced load of their lengths and
    // offsets, then read by read, with the bytes of the next reads in flight
    // while one is worked on (lane p holds letters p, p + 64, p + 128, p + 192)
    for (uint32_t r0 = wave * 64; r0 < n; r0 += nwave * 64) {
        const uint32_t cnt = n - r0 < 64 ? n - r0 : 64;
        const uint32_t myL = (uint32_t)lane < cnt ? lens[r0 + lane] : 0u;
        const uint64_t myO = (uint32_t)lane < cnt ? offs[r0 + lane] : 0ull;
        auto len_of = [&](uint32_t i) { return (uint32_t)__builtin_amdgcn_readlane((int)myL, (int)i); };
        auto off_of = [&](uint32_t i) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)myO, (int)i);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(myO >> 32), (int)i);
            return ((uint64_t)hi << 32) | lo;
        };
        // read i of the group, its letters in cb (lane p: letters p, p + 64,
        // p + 128, p + 192)
        auto prep_one = [&](uint32_t i, const uint8_t (&cb)[4]) {
            const uint32_t r = r0 + i;
            const uint32_t L = len_of(i);
            uint8_t* qrev = rev + off_of(i);
            uint8_t st = 0;
            // ori 0 / 1 of mate 2 (reads from nsplit on) pair with ori 1 / 0 of mate 1
            const uint32_t om = r < nsplit ? ori_mask : (((ori_mask & 1) << 1) | ((ori_mask >> 1) & 1));
            if (!paired) {
                // readSize < READLEN - 20 (ReadQueue.cpp:430): the unsigned
                // subtraction of the reference is kept
                if ((uint64_t)L < (uint64_t)(uint32_t)(readlen - 20u)) st = 3;
                if (L < KMERLEN) st = 3;
            } else if (L < minlen) {
                st = 1; // readSize < ceil(0.75 READLEN) (ReadQueue.cpp:971-980); minlen >= KMERLEN
            }
            uint8_t* c0 = codes + (size_t)(2 * r) * 256;
            uint8_t* c1 = c0 + 256;
            if (L > FAME_DEV_MAX_READ) {
                // a long read (single-end, plain form): validated and
                // reverse-complemented here, listed for the long-read path;
                // no row (its seed-stage rows get length 0).  Otherwise (the
                // batched form, paired mates) FLAG_TOO_LONG.
                if (!paired && long_list && st == 0) {
                    const uint8_t* sp = seqs + off_of(i);
                    bool nflag = false;
                    for (uint32_t p = (uint32_t)lane; p < L; p += 64) {
                        const uint32_t t = s_lt[sp[p]];
                        nflag = nflag || ((t >> 12) & 1u);
                        qrev[L - 1 - p] = (uint8_t)t;
                    }
                    if (__ballot(nflag)) st = 3;
                    if (st == 0 && lane == 0) {
                        long_list[atomicAdd(long_n, 1u)] = r;
                        atomicMax(long_n + 1, L); // the longest of them
                    }
                } else {
                    if (lane == 0) atomicOr(err_flag, 1u);
                    st = 3;
                }
                if (lane == 0) {
                    pre[r] = st;
                    c0[255] = 0;
                    c1[255] = 0;
                }
                return;
            }
            if (st == 0 || (paired && st == 1)) {
                bool nflag = false;
                // unknown characters stay '\0' in revSeq (branch-free per letter)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t p = (uint32_t)lane + 64u * q;
                    const uint32_t t = s_lt[cb[q]];
                    const bool in = p < L;
                    nflag = nflag || (in && (((t >> 12) & 1u) || (paired && !((t >> 13) & 1u))));
                    if (in) {
                        qrev[L - 1 - p] = (uint8_t)t;
                        if (om & 1) c0[p] = (uint8_t)((t >> 8) & 3u);
                        if (om & 2) c1[L - 1 - p] = (uint8_t)((t >> 10) & 3u);
                    }
                }
                // single-end: N invalidates (ReadQueue.cpp:456-460); paired: N and
                // unknown letters invalidate the mate (:1016-1027)
                if (__ballot(nflag)) st = paired ? (uint8_t)(st | 2) : (uint8_t)3;
            }
            if (st == 0 && L > lmax) lmax = L;
            if (lane == 0) {
                pre[r] = st;
                c0[255] = (st == 0 && (om & 1)) ? (uint8_t)L : 0;
                c1[255] = (st == 0 && (om & 2)) ? (uint8_t)L : 0;
            }
        };
        // the letters of read i are loaded U reads ahead into slot i % U, a
        // register set the unrolled loop below names statically
        constexpr int U = 4;
        uint8_t nb[U][4];
        // (always issued: lanes past the read, or reads too long for the
        // rows, load a byte of this group's lengths instead, which prep_one
        // does not use)
        const uint8_t* dummy = (const uint8_t*)(lens + r0);
        auto fetch = [&](uint32_t i, uint8_t (&d)[4]) {
            const uint32_t L = len_of(i);
            const uint32_t Lm = L <= FAME_DEV_MAX_READ ? L : 0u; // longer reads are flagged, not read
            const uint8_t* sp = seqs + off_of(i);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t p = (uint32_t)lane + 64u * q;
                d[q] = *(p < Lm ? sp + p : dummy);
            }
        };
#pragma unroll
        for (int u = 0; u < U; ++u) fetch(min((uint32_t)u, cnt - 1), nb[u]);
        for (uint32_t i0 = 0; i0 < cnt; i0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u;
                if (i >= cnt) break;
                prep_one(i, nb[u]);
                fetch(min(i + U, cnt - 1), nb[u]);
            }
        }
    }
    if (lane == 0 && lmax) atomicMax(max_len, lmax);
}

// ---------------------------------------------------------------------------
// k_seed: seed + filter + map order for one (read, orientation) per wave
// ---------------------------------------------------------------------------

template <typename TabT, typename IdxT, typename EkmT = uint8_t>
struct SeedWork {
    TabT* klo;      // [K] bucket start
    uint32_t* klen; // [K] bucket length
    uint32_t* koff; // [K] exclusive prefix of klen
    uint32_t* k2;   // [K] second entry of a two-entry bucket read from ix.slot (NULL: no slots)
    uint32_t* ent;  // [cap_e] packed metaId<<1 | fwd
    EkmT* ekm;      // [cap_e] k-mer index of each entry
    uint32_t* eaux; // [cap_e] kept / first / count (or hash slot, heavy)
    uint32_t cap_e;
    uint32_t* hkey; // heavy path: hash table keyed by packed entry
    uint32_t* hfirst;
    uint32_t* hcnt;
    uint32_t hsize; // power of two, >= cap_e
    uint32_t* ins_key; // [2][cap_ins] keys in first-insertion order per map
    uint16_t* ins_cnt; // [2][cap_ins] their u16 counts
    uint32_t cap_ins;
    IdxT* tb;          // hopscotch buckets (one map at a time)
    uint32_t tb_cap;
    IdxT* tmp;         // rehash / output list
    IdxT* over;        // overflow list
    uint32_t* cand;    // [2][cap_ins] candidate windows per map
    uint8_t* code;     // [L] reduced base codes of the read
    uint64_t* srt = nullptr; // [hsize] the streaming counter's sort buffer (seed_tail_stream)
};

struct SeedStats {
    uint64_t kmers, entries, seeded, bytes, buckets;
};

// NTPS64 of k-mer k of the reduced-letter row `code` (closed form), reduced
// to a hash key (h % HTABSIZE)
__device__ __forceinline__ uint64_t seed_key(const DevIndex& ix, const SeedShared& sh, const uint8_t* code,
                                             uint32_t k)
{
    uint64_t h = 0;
#pragma unroll
    for (uint32_t i = 0; i < KMERLEN; ++i)
        if ((SEEDBITS >> (KMERLEN - 1 - i)) & 1) h ^= sh.rot[i][code[k + i]];
    return ix.hpow2 ? (h & ix.hmask) : (h % ix.htab);
}

// slot-table mode: the hashes of one (read, orientation) and the loads of
// its K slots, issued without waiting for them (lanes past K load slot 0,
// which the consumer ignores); seed_one takes the values with staged = true
template <int KPL>
__device__ __forceinline__ void seed_stage(const DevIndex& ix, const SeedShared& sh, const uint8_t* code,
                                           uint32_t K, uint64_t (&sl)[KPL])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < KPL; ++q) {
        const uint32_t k = lane + 64 * q;
        const uint64_t key = k < K ? seed_key(ix, sh, code, k) : 0;
        sl[q] = ix.slot[key];
    }
}

// exclusive prefix of the bucket lengths over k-mers in index order (koff),
// E = all index hits of the (read, orientation); false when it needs the
// heavy path (E beyond the capacity, or a bucket beyond what the device
// path serves); else the work counters
template <typename TabT, typename IdxT, typename EkmT>
__device__ __forceinline__ bool seed_prefix(const SeedWork<TabT, IdxT, EkmT>& w, uint32_t K, uint32_t L, bool inl,
                                            SeedStats& ss, uint64_t& E)
{
    const int lane = threadIdx.x & 63;
    uint64_t carry = 0;
    bool big = false;
    uint32_t gathered = 0; // entries read from the k-mer table (slot mode: buckets of three or more)
    for (uint32_t k0 = 0; k0 < K; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t v = k < K ? w.klen[k] : 0;
        if (v > 0xFFFFFFu) { // beyond what the device path serves
            v = 0xFFFFFFu;
            big = true;
        }
        gathered += (!inl || v >= 3) ? v : 0u;
        const uint32_t inc = wave_iscan_u32(v, lane);
        const uint32_t tot = lane_u32(inc, 63);
        if (k < K) w.koff[k] = (uint32_t)(carry + inc - v);
        carry += tot;
        if (carry > 0x7FFFFFFFull) big = true;
    }
    big = __ballot(big) != 0;
    E = carry;
    if (big || E > w.cap_e) return false;
    ss.kmers += K;
    ss.entries += wave_sum_u32(gathered);
    ss.seeded += 1;
    ss.bytes += L;
    ss.buckets += E;
    wave_sync();
    return true;
}

// The windows of a map at or above QTHRESH in hopscotch iteration order,
// without the table, where lp_place applies (returns false elsewhere: the
// caller emulates the table).
template <typename TabT, typename IdxT, typename EkmT>
__device__ __forceinline__ bool seed_order_small(const SeedWork<TabT, IdxT, EkmT>& w, int mp, uint32_t nins,
                                                 uint64_t bucket_count, uint32_t& nc)
{
    const int lane = threadIdx.x & 63;
    if (nins > 64) return false;
    const uint32_t* key = w.ins_key + mp * w.cap_ins;
    const uint16_t* cnt = w.ins_cnt + mp * w.cap_ins;
    const bool in = (uint32_t)lane < nins;
    const uint32_t mykey = in ? key[lane] : 0u;
    const uint32_t mycnt = in ? cnt[lane] : 0u;
    uint32_t mypos;
    if (!lp_place(mykey, nins, bucket_count, w.tb_cap, mypos)) return false;
    const bool keep = in && mycnt >= QTHRESH;
    uint32_t rank;
    nc = lp_rank(keep, mypos, rank);
    if (keep) w.cand[mp * w.cap_ins + rank] = mykey;
    wave_sync();
    return true;
}

// the maps of one (read, orientation) once their insertion sequences are in
// w.ins_key / w.ins_cnt (nins keys each; nq of them at or above QTHRESH, the
// first such in w.cand when nq == 1): hopscotch order where a map holds two or
// more such windows, and the candidates into the verify pools (returns 1 when
// a capacity of the caller's arrays is exceeded)
template <typename TabT, typename IdxT, typename EkmT>
__device__ __forceinline__ int seed_maps(const DevIndex& ix, const BatchDev& bd, uint32_t ro,
                                         const SeedWork<TabT, IdxT, EkmT>& w, PoolChunk& pc, bool exact,
                                         const uint32_t (&nins)[2], const uint32_t (&nq)[2], uint32_t bucketCount,
                                         uint32_t L)
{
    const int lane = threadIdx.x & 63;
    if ((nins[0] > w.cap_ins && nq[0] > 1) || (nins[1] > w.cap_ins && nq[1] > 1)) return 1;
    wave_sync();
    // maps with two or more windows at or above QTHRESH: hopscotch placement,
    // then those windows in iteration order (buckets, then the overflow list)
    uint32_t ncand[2];
#pragma unroll
    for (int mp = 0; mp < 2; ++mp) {
        if (nq[mp] <= 1) {
            ncand[mp] = nq[mp];
            continue;
        }
        if (seed_order_small<TabT, IdxT, EkmT>(w, mp, nins[mp], bucketCount, ncand[mp])) continue;
        Hs<IdxT> h;
        h.tb = w.tb;
        h.tb_cap = w.tb_cap;
        h.over = w.over;
        h.over_cap = w.cap_ins;
        h.tmp = w.tmp;
        h.tmp_cap = w.cap_ins;
        h.key = w.ins_key + mp * w.cap_ins;
        h.nelem = h.nover = 0;
        if (!hs_setup(h, bucketCount, lane)) return 1;
        for (uint32_t i = 0; i < nins[mp]; ++i)
            if (hs_insert(h, i, lane) != HS_OK) return 1;
        const uint16_t* cnt = w.ins_cnt + mp * w.cap_ins;
        uint32_t* cand = w.cand + mp * w.cap_ins;
        uint32_t nc = 0;
        for (uint64_t b0 = 0; b0 < h.arr + h.nover; b0 += 64) {
            const uint64_t b = b0 + lane;
            IdxT i = Hs<IdxT>::NONE;
            if (b < h.arr) i = h.tb[b];
            else if (b < h.arr + h.nover) i = h.over[b - h.arr];
            const bool keep = i != Hs<IdxT>::NONE && cnt[i] >= QTHRESH;
            const uint64_t bal = __ballot(keep);
            if (keep) cand[nc + __popcll(bal & lanes_below(lane))] = h.key[i];
            nc += __popcll(bal);
        }
        ncand[mp] = nc;
        wave_sync();
    }
    // emit candidates: fwd-map windows then rev-map windows
    const uint32_t nf = ncand[0], nr = ncand[1];
    const uint32_t pool = pool_of_len(L);
    uint32_t base = 0;
    if (nf + nr) {
        base = pool_take(bd, pc, pool, nf + nr, exact, lane);
        if (base == ~0u) {
            if (lane == 0) atomicOr(bd.flags, FLAG_POOL_FULL);
            base = 0;
        } else {
            uint32_t* jm = pc_meta(bd, pool);
            uint32_t* jr = bd.job_ro[0];
#pragma unroll
            for (uint32_t q = 1; q < FAME_NPOOL; ++q) jr = pool == q ? bd.job_ro[q] : jr;
            for (uint32_t t = lane; t < nf + nr; t += 64) {
                const bool isrev = t >= nf;
                const uint32_t id = isrev ? w.cand[w.cap_ins + (t - nf)] : w.cand[t];
                jm[base + t] = id | (isrev ? 0x80000000u : 0u);
                jr[base + t] = ro;
            }
        }
    }
    if (lane == 0) {
        bd.cand_off[ro] = base;
        bd.cand_nf[ro] = nf;
        bd.cand_nr[ro] = nr;
    }
    return 0;
}

// getSeedRefs' counting without per-hit arrays (the LDS tier): the hits are
// visited in index order, 64 per round and four rounds' k-mer table loads in
// flight, each lane finding its hit's k-mer by walking the bucket offsets from
// the k-mer where the round starts.  A hit is dropped when it repeats the
// previous hit of the same bucket (ReadQueue.cpp:2381-2391, 2412-2422); a kept
// hit counts for its window in an LDS table keyed by the packed entry, whose
// first occurrence is the least (hit << 8 | k-mer) seen.  The windows whose
// first occurrence may insert (k-mer 0, or k-mer cIdx+1 with cIdx < maxQPos)
// enter their map in first-occurrence order: the table's windows are sorted by
// it (bitonic, in LDS).  Returns 1 when the distinct windows outgrow the table
// or the maps outgrow their arrays (the heavy tier then takes the pair).
template <typename TabT, typename IdxT, typename EkmT>
__device__ __forceinline__ int seed_tail_stream(const DevIndex& ix, const BatchDev& bd, uint32_t ro,
                                                const SeedWork<TabT, IdxT, EkmT>& w, PoolChunk& pc, bool exact,
                                                uint64_t E64, uint32_t bucketCount, uint32_t L, uint32_t K, bool inl)
{
    const int lane = threadIdx.x & 63;
    if (E64 >= (1ull << 24)) return 1; // hit indexes are packed in 24 bits
    const uint32_t E = (uint32_t)E64;
    const uint32_t hs = w.hsize, hm = hs - 1;
    const int hshift = __clz(hs) + 1;
    for (uint32_t i = lane; i < hs; i += 64) {
        w.hkey[i] = EMPTY;
        w.hfirst[i] = 0xFFFFFFFFu;
        w.hcnt[i] = 0;
    }
    wave_sync();
    constexpr int U = 4;
    uint32_t kb = 0;                  // k-mer of the round's first hit (no later than any of its hits)
    uint32_t last_v = 0, last_k = ~0u; // the previous round's last hit
    bool full = false;
    for (uint32_t t0 = 0; t0 < E && !full; t0 += 64 * U) {
        uint32_t kk[U], vv[U], inl_v[U];
        uint64_t at[U];
        bool isinl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = t0 + 64u * u + (uint32_t)lane;
            uint32_t k = kb;
            at[u] = 0;
            isinl[u] = false;
            inl_v[u] = 0;
            if (j < E) {
                while (w.koff[k] + w.klen[k] <= j) ++k; // also steps over empty buckets
                const uint32_t d = j - w.koff[k];
                if (inl && w.klen[k] <= 2) {
                    isinl[u] = true;
                    inl_v[u] = d == 0 ? (uint32_t)w.klo[k] : w.k2[k];
                } else {
                    at[u] = (uint64_t)w.klo[k] + d;
                }
            }
            kk[u] = k;
            // the next round starts at the k-mer of this round's last hit
            const uint32_t jl = t0 + 64u * u + 63u;
            kb = lane_u32(k, jl < E ? 63 : (int)((E - 1) & 63));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) vv[u] = ix.kent[at[u]]; // always issued (entry 0 past the hits)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = t0 + 64u * u + (uint32_t)lane;
            const bool live = j < E;
            const uint32_t v = isinl[u] ? inl_v[u] : (vv[u] & 0x7FFFFFFFu); // without isStart
            const uint32_t k = kk[u];
            uint32_t pv = (uint32_t)wave_shr1_i32((int)v), pk = (uint32_t)wave_shr1_i32((int)k);
            if (lane == 0) {
                pv = last_v;
                pk = last_k;
            }
            last_v = lane_u32(v, 63);
            last_k = lane_u32(k, 63);
            const bool kept = live && !(j > 0 && pk == k && pv == v);
            if (kept && !full) {
                uint32_t slot = (v * 2654435761u) >> hshift;
                for (uint32_t probes = 0;; ++probes) {
                    if (probes == hs) {
                        full = true;
                        break;
                    }
                    const uint32_t old = atomicCAS(&w.hkey[slot], EMPTY, v);
                    if (old == EMPTY || old == v) break;
                    slot = (slot + 1) & hm;
                }
                if (!full) {
                    atomicMin(&w.hfirst[slot], (j << 8) | k);
                    atomicAdd(&w.hcnt[slot], 1u);
                }
            }
        }
        full = __ballot(full) != 0;
    }
    if (full) return 1;
    wave_sync();
    // windows whose first occurrence may insert, as (first hit << 32 | slot)
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    uint32_t nd = 0;
    for (uint32_t i0 = 0; i0 < hs; i0 += 64) {
        const uint32_t i = i0 + lane;
        bool take = false;
        uint32_t fj = 0;
        if (i < hs && w.hkey[i] != EMPTY) {
            fj = w.hfirst[i];
            const uint32_t km = fj & 0xFFu;
            take = km == 0 || (km - 1) < maxQPos;
        }
        const uint64_t b = __ballot(take);
        if (take) w.srt[nd + __popcll(b & lanes_below(lane))] = ((uint64_t)(fj >> 8) << 32) | i;
        nd += __popcll(b);
    }
    uint32_t n2 = 64;
    while (n2 < nd) n2 <<= 1;
    for (uint32_t i = nd + lane; i < n2; i += 64) w.srt[i] = ~0ull;
    wave_sync();
    // bitonic sort of n2 (a power of two, at most hs) by first hit
    for (uint32_t size = 2; size <= n2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t p = lane; p < n2 / 2; p += 64) {
                const uint32_t lo = 2 * p - (p & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = w.srt[lo], c = w.srt[hi];
                if ((a > c) == up) {
                    w.srt[lo] = c;
                    w.srt[hi] = a;
                }
            }
            wave_sync();
        }
    }
    // the two maps' insertion sequences in first-occurrence order
    uint32_t nins[2] = {0, 0}, nq[2] = {0, 0};
    for (uint32_t j0 = 0; j0 < nd; j0 += 64) {
        const uint32_t j = j0 + lane;
        bool take0 = false, take1 = false;
        uint32_t key = 0, cnt = 0;
        if (j < nd) {
            const uint32_t slot = (uint32_t)w.srt[j];
            const uint32_t e = w.hkey[slot];
            key = e >> 1;
            cnt = w.hcnt[slot] & 0xFFFFu;
            take0 = (e & 1) != 0; // forward-strand windows -> fwdMetaIDs
            take1 = !take0;
        }
        const uint64_t b0 = __ballot(take0), b1 = __ballot(take1);
        const uint64_t below = lanes_below(lane);
        if (take0) {
            const uint32_t p = nins[0] + __popcll(b0 & below);
            if (p < w.cap_ins) {
                w.ins_key[p] = key;
                w.ins_cnt[p] = (uint16_t)cnt;
            }
        }
        if (take1) {
            const uint32_t p = nins[1] + __popcll(b1 & below);
            if (p < w.cap_ins) {
                w.ins_key[w.cap_ins + p] = key;
                w.ins_cnt[w.cap_ins + p] = (uint16_t)cnt;
            }
        }
        nins[0] += __popcll(b0);
        nins[1] += __popcll(b1);
        const bool q0 = take0 && cnt >= QTHRESH, q1 = take1 && cnt >= QTHRESH;
        const uint64_t c0 = __ballot(q0), c1 = __ballot(q1);
        if (q0 && nq[0] == 0 && (c0 & below) == 0) w.cand[0] = key;
        if (q1 && nq[1] == 0 && (c1 & below) == 0) w.cand[w.cap_ins] = key;
        nq[0] += __popcll(c0);
        nq[1] += __popcll(c1);
    }
    wave_sync();
    return seed_maps<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, nins, nq, bucketCount, L);
}

// the rest of getSeedRefs for one (read, orientation) once every index hit
// is in w.ent / w.ekm (E of them, in k-mer order): consecutive dedup, counts
// and first occurrences, the two maps' insertion sequences, hopscotch order
// where it matters, and the candidate windows into the verify pools.
// Returns 1 when a capacity of the caller's arrays is exceeded (heavy path).
template <typename TabT, typename IdxT, typename EkmT>
__device__ __forceinline__ int seed_tail(const DevIndex& ix, const BatchDev& bd, uint32_t ro,
                                          const SeedWork<TabT, IdxT, EkmT>& w,
                         PoolChunk& pc, bool exact, uint64_t E, uint32_t bucketCount, uint32_t L)
{
    const int lane = threadIdx.x & 63;
    // consecutive dedup inside a bucket (ReadQueue.cpp:2381-2391, 2412-2422)
    for (uint32_t j = lane; j < E; j += 64) {
        const bool skip = j > 0 && w.ekm[j] == w.ekm[j - 1] && w.ent[j] == w.ent[j - 1];
        w.eaux[j] = skip ? 0u : 1u;
    }
    wave_sync();
    // first occurrence and total count (u16) of every kept key: an open
    // addressing table keyed by the packed entry (a power of two of at least
    // twice the entries, at most hsize; large scratch tables are only cleared
    // as far as the read needs).  With hsize >= 2 cap_e every probe sequence
    // ends at a free slot; a smaller table (the LDS tier, sized for the
    // distinct windows of a read) gives up after hsize probes and the read
    // goes to the next tier
    {
        uint32_t hs = 64;
        while (hs < 2 * (uint32_t)E && hs < w.hsize) hs <<= 1;
        if (hs > w.hsize) hs = w.hsize;
        const uint32_t hm = hs - 1;
        const int hshift = __clz(hs) + 1;
        for (uint32_t i = lane; i < hs; i += 64) {
            w.hkey[i] = EMPTY;
            w.hfirst[i] = 0xFFFFFFFFu;
            w.hcnt[i] = 0;
        }
        wave_sync();
        bool full = false;
        for (uint32_t j = lane; j < E && !full; j += 64) {
            if (!(w.eaux[j] & 1)) continue;
            const uint32_t key = w.ent[j];
            uint32_t slot = (key * 2654435761u) >> hshift;
            for (uint32_t probes = 0;; ++probes) {
                if (probes == hs) {
                    full = true;
                    break;
                }
                const uint32_t old = atomicCAS(&w.hkey[slot], EMPTY, key);
                if (old == EMPTY || old == key) break;
                slot = (slot + 1) & hm;
            }
            if (full) break;
            atomicMin(&w.hfirst[slot], j);
            atomicAdd(&w.hcnt[slot], 1u);
            w.eaux[j] = 1u | (slot << 8);
        }
        if (__ballot(full)) return 1;
        wave_sync();
        for (uint32_t j = lane; j < E; j += 64) {
            const uint32_t a = w.eaux[j];
            if (!(a & 1)) continue;
            const uint32_t slot = a >> 8;
            const bool first = w.hfirst[slot] == j;
            w.eaux[j] = 1u | (first ? 2u : 0u) | ((w.hcnt[slot] & 0xFFFFu) << 16);
        }
    }
    wave_sync();
    // insertion sequences: a key enters its map at its first occurrence if
    // that occurrence may insert (k-mer 0, or k-mer cIdx+1 with cIdx < maxQPos)
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    uint32_t nins[2] = {0, 0}, nq[2] = {0, 0};
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
        const uint64_t below = lanes_below(lane);
        if (take0) {
            const uint32_t p = nins[0] + __popcll(b0 & below);
            if (p < w.cap_ins) {
                w.ins_key[p] = key;
                w.ins_cnt[p] = (uint16_t)cnt;
            }
        }
        if (take1) {
            const uint32_t p = nins[1] + __popcll(b1 & below);
            if (p < w.cap_ins) {
                w.ins_key[w.cap_ins + p] = key;
                w.ins_cnt[w.cap_ins + p] = (uint16_t)cnt;
            }
        }
        nins[0] += __popcll(b0);
        nins[1] += __popcll(b1);
        // windows at or above QTHRESH; the first one of each map is kept
        // directly, since a map holding at most one such window yields it
        // whatever the iteration order
        const bool q0 = take0 && cnt >= QTHRESH, q1 = take1 && cnt >= QTHRESH;
        const uint64_t c0 = __ballot(q0), c1 = __ballot(q1);
        if (q0 && nq[0] == 0 && (c0 & below) == 0) w.cand[0] = key;
        if (q1 && nq[1] == 0 && (c1 & below) == 0) w.cand[w.cap_ins] = key;
        nq[0] += __popcll(c0);
        nq[1] += __popcll(c1);
    }
    return seed_maps<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, nins, nq, bucketCount, L);
}

// KPL > 0: the read's row of reduced letters arrives as this lane's word cw
// (a 256-byte row, the length in its byte 255); KPL == 0 (long reads): the
// row is already in w.code, L letters long, and the k-mers are hashed and
// looked up 64 at a time (no register arrays)
struct SeedNoMid {
    __device__ void operator()() const {}
};

// mid() runs once on every path, after the loads of the entry gather are
// issued and before their values are used: a caller's loads issued there do
// not delay the gather's (the memory counter is waited on in issue order)
template <typename TabT, typename IdxT, int KPL, typename EkmT, bool STREAM = false, typename Mid = SeedNoMid>
__device__ __forceinline__ int seed_one(const DevIndex& ix, const BatchDev& bd, uint32_t ro, SeedShared& sh,
                        const SeedWork<TabT, IdxT, EkmT>& w, PoolChunk& pc, bool exact, SeedStats& ss,
                        uint64_t* E_out, uint32_t* bc_out, uint32_t cw, bool staged = false,
                        const uint64_t* pre = nullptr, uint32_t Llong = 0, const Mid& mid = Mid{})
{
    const int lane = threadIdx.x & 63;
    // cw: this lane's word of the row of reduced letters (the length is in
    // byte 255, the top byte of lane 63's word)
    const uint32_t L = KPL == 0 ? Llong : lane_u32(cw, 63) >> 24;
    if (KPL > 0 && !staged) {
        ((uint32_t*)w.code)[lane] = cw;
        wave_sync();
    }
    const uint32_t K = L - KMERLEN + 1;
    const TabT* tab = (const TabT*)ix.tab;
    const bool inl = ix.slot != nullptr && w.k2 != nullptr;
    // hashes + bucket gathers (NTPS64 closed form; key = h % HTABSIZE): one
    // slot per k-mer, or tabIndex[key], tabIndex[key + 1]; all gathers of a
    // lane are issued before the first is consumed.  staged: the slots were
    // loaded by seed_stage (pre)
    uint32_t bsum = 0;
    if constexpr (KPL == 0) {
        for (uint32_t k0 = 0; k0 < K; k0 += 64) {
            const uint32_t k = k0 + lane;
            if (k < K) {
                const uint64_t key = seed_key(ix, sh, w.code, k);
                if (inl) {
                    const uint64_t sv = ix.slot[key];
                    const uint32_t len = slot_len(sv);
                    w.klo[k] = (TabT)slot_lo(sv);
                    w.k2[k] = slot_e1(sv);
                    w.klen[k] = len;
                    bsum += len;
                } else {
                    const TabT lo = tab[key], hi = tab[key + 1];
                    w.klo[k] = lo;
                    const uint64_t len = (uint64_t)(hi - lo);
                    w.klen[k] = len > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)len;
                    bsum += (uint32_t)len; // bucketCount += (u32 truncation of the u64 difference)
                }
            }
        }
    } else {
        TabT lo[KPL], hi[KPL];
        uint64_t sl[KPL];
    #pragma unroll
        for (int q = 0; q < KPL; ++q) {
            const uint32_t k = lane + 64 * q;
            lo[q] = hi[q] = 0;
            sl[q] = staged ? pre[q] : 0;
            if (k < K && !staged) {
                const uint64_t key = seed_key(ix, sh, w.code, k);
                if (inl) {
                    sl[q] = ix.slot[key];
                } else {
                    lo[q] = tab[key];
                    hi[q] = tab[key + 1];
                }
            }
        }
    #pragma unroll
        for (int q = 0; q < KPL; ++q) {
            const uint32_t k = lane + 64 * q;
            if (k < K) {
                if (inl) {
                    const uint32_t len = slot_len(sl[q]);
                    w.klo[k] = (TabT)slot_lo(sl[q]);
                    w.k2[k] = slot_e1(sl[q]);
                    w.klen[k] = len;
                    bsum += len;
                } else {
                    w.klo[k] = lo[q];
                    const uint64_t len = (uint64_t)(hi[q] - lo[q]);
                    w.klen[k] = len > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)len;
                    bsum += (uint32_t)len; // bucketCount += (u32 truncation of the u64 difference)
                }
            }
        }
    }
    const uint32_t bucketCount = wave_sum_u32(bsum) / 5;
    uint64_t E;
    const bool fits = seed_prefix(w, K, L, inl, ss, E);
    *E_out = E;
    *bc_out = bucketCount;
    if (!fits) {
        mid();
        return 1;
    }
    if constexpr (STREAM) {
        mid();
        return seed_tail_stream<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, E, bucketCount, L, K, inl);
    }
    // entry gather, one entry per lane: entry j belongs to the k-mer whose
    // [koff, koff + klen) contains j.  Up to 256 entries each k-mer writes
    // its index over its entries first; beyond, each k-mer also writes the
    // k-mer table position of its entries into eaux (free until the dedup),
    // or the entries themselves for a bucket held in its slot, and the
    // gather is one load per entry (a 64-bit tabIndex: a binary search over
    // koff instead)
    if constexpr (sizeof(TabT) == 4) {
        if (E > 256) {
            mid();
            for (uint32_t k = lane; k < K; k += 64) {
                const uint32_t o = w.koff[k], n = w.klen[k];
                const uint32_t lo = (uint32_t)w.klo[k];
                if (inl && n <= 2) {
                    for (uint32_t i = 0; i < n; ++i) {
                        w.ekm[o + i] = (EkmT)k;
                        w.ent[o + i] = i == 0 ? lo : w.k2[k];
                        w.eaux[o + i] = ~0u;
                    }
                } else {
                    for (uint32_t i = 0; i < n; ++i) {
                        w.ekm[o + i] = (EkmT)k;
                        w.eaux[o + i] = lo + i;
                    }
                }
            }
            wave_sync();
            for (uint32_t j = lane; j < E; j += 64) {
                const uint32_t at = w.eaux[j];
                if (at != ~0u) w.ent[j] = ix.kent[at] & 0x7FFFFFFFu; // without isStart
            }
            wave_sync();
            return seed_tail<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, E, bucketCount, L);
        }
    }
    if constexpr (sizeof(TabT) == 4) {
        if (E <= 256) {
            // eaux[j] = the k-mer table position of entry j, or ~0u for an
            // entry held in its slot (written to ent at once); ekm[j] = its
            // k-mer.  The table entries are then loaded, mid(), and written.
            for (uint32_t k = lane; k < K; k += 64) {
                const uint32_t o = w.koff[k], n = w.klen[k], lo = (uint32_t)w.klo[k];
                const bool held = inl && n <= 2; // the bucket is in its slot
                const uint32_t e1 = held && n == 2 ? w.k2[k] : 0u;
                for (uint32_t i = 0; i < n; ++i) {
                    w.ekm[o + i] = (EkmT)k;
                    if (held) w.ent[o + i] = i == 0 ? lo : e1;
                    w.eaux[o + i] = held ? ~0u : lo + i;
                }
            }
            wave_sync();
            uint32_t gv[4];
            bool need[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t j = lane + 64u * q;
                const uint32_t at = j < E ? w.eaux[j] : ~0u;
                need[q] = at != ~0u;
                gv[q] = ix.kent[need[q] ? at : 0u];
            }
            mid();
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (need[q]) w.ent[lane + 64u * q] = gv[q] & 0x7FFFFFFFu; // without isStart
            wave_sync();
            return seed_tail<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, E, bucketCount, L);
        }
    }
    const bool scatter = E <= 256;
    if (scatter) {
        for (uint32_t k = lane; k < K; k += 64) {
            const uint32_t o = w.koff[k], n = w.klen[k];
            for (uint32_t i = 0; i < n; ++i) w.ekm[o + i] = (EkmT)k;
        }
        wave_sync();
    }
    mid();
    for (uint32_t j = lane; j < E; j += 64) {
        uint32_t k;
        if (scatter) {
            k = w.ekm[j];
        } else {
            uint32_t lo = 0, hi = K; // last k with koff[k] <= j and klen[k] > 0
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (w.koff[mid] <= j) lo = mid;
                else hi = mid;
            }
            k = lo;
            while (w.klen[k] == 0 || w.koff[k] + w.klen[k] <= j) ++k; // skip empty buckets sharing koff
        }
        const uint32_t d = j - w.koff[k];
        if (inl && w.klen[k] <= 2) // the bucket is in its slot
            w.ent[j] = d == 0 ? (uint32_t)w.klo[k] : w.k2[k];
        else
            w.ent[j] = ix.kent[(uint64_t)w.klo[k] + d] & 0x7FFFFFFFu; // without isStart
        if (!scatter) w.ekm[j] = (EkmT)k;
    }
    wave_sync();
    return seed_tail<TabT, IdxT, EkmT>(ix, bd, ro, w, pc, exact, E, bucketCount, L);
}

// ---------------------------------------------------------------------------
// Slot lookups through the genome-ordered copy (DevIndex::gkey)
// ---------------------------------------------------------------------------

// A (read, orientation) passes three stages before seed_one takes its
// slots, one per iteration of the seed loop, so that each stage's loads have
// a whole iteration to arrive:
//   hashed     keys known; the key positions (kpos) of up to three anchor
//              k-mers in flight
//   predicted  the anchors give where the read sits in genome order (one or
//              two offsets); gkey of every k-mer there in flight
//   slots      k-mers whose key equals a predicted gkey load gslot there
//              (contiguous), the others slot[key] at random
template <int KPL>
struct GoHashed {
    uint32_t L; // letters, 0 = nothing to seed
    uint32_t ky[KPL];
    uint32_t anc; // kpos of this lane's anchor k-mer, ~0u: none
};
template <int KPL>
struct GoPredicted {
    uint32_t L;
    uint32_t ky[KPL];
    uint32_t gk1[KPL], gk2[KPL]; // as loaded: valid where the offset puts k-mer k inside the copy
    int64_t d1, d2;
    int nd; // offsets found (0-2)
};
template <int KPL>
struct GoSlots {
    uint32_t L;
    uint64_t sl[KPL];
};

// anchors of K k-mers: k-mers 0, (K-1)/2 and K-1, at most one per lane (an
// anchor on a lane already taken is dropped)
struct GoAnchors {
    int a[3];
    bool v[3];
};
__device__ __forceinline__ GoAnchors go_anchors(uint32_t K)
{
    GoAnchors g;
    g.a[0] = 0;
    g.a[1] = (int)(K - 1) / 2;
    g.a[2] = (int)K - 1;
    g.v[0] = true;
    g.v[2] = (g.a[2] & 63) != 0;
    g.v[1] = (g.a[1] & 63) != 0 && (!g.v[2] || (g.a[1] & 63) != (g.a[2] & 63));
    return g;
}

// keys of the k-mers of the row cw (lane l holds letters 4l .. 4l+3) through
// the prefix array pf (this wave's LDS, L + 1 entries), anchor loads issued
// rq[t][c] = rol(seed_c, -t mod 64): the term of letter t of a read in the
// prefix form of the hash (filled by the block at kernel start)
struct GoShared {
    uint64_t rq[64][4];
};
__device__ __forceinline__ void go_table_init(GoShared& g)
{
    for (int t = threadIdx.x; t < 64 * 4; t += blockDim.x) {
        const int i = t >> 2, c = t & 3;
        g.rq[i][c] = rol64(red_seed((uint32_t)c), (64 - i) & 63);
    }
    __syncthreads();
}

template <int KPL>
__device__ __forceinline__ void go_hash(const DevIndex& ix, const GoShared& sh, uint64_t* pf, uint32_t cw,
                                        uint32_t L, GoHashed<KPL>& h, uint32_t& rnd)
{
    const int lane = threadIdx.x & 63;
    h.L = L;
    h.anc = ~0u;
#pragma unroll
    for (int q = 0; q < KPL; ++q) h.ky[q] = 0;
    if (L < KMERLEN) return; // 0: not seeded (k_prep never writes 1 .. KMERLEN-1)
    uint64_t x[4], acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t t = 4u * (uint32_t)lane + (uint32_t)j;
        acc ^= t < L ? sh.rq[t & 63][(cw >> (8 * j)) & 3u] : 0ull;
        x[j] = acc;
    }
    const uint64_t ex = wave_ixor_u64(acc) ^ acc; // letters below this lane's
    wave_sync();
    if (lane == 0) pf[0] = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t t = 4u * (uint32_t)lane + (uint32_t)j;
        if (t < L) pf[t + 1] = ex ^ x[j];
    }
    wave_sync();
    constexpr SeedEnds se = seed_ends();
    const uint32_t K = L - KMERLEN + 1;
#pragma unroll
    for (int q = 0; q < KPL; ++q) {
        const uint32_t k = (uint32_t)lane + 64u * q;
        uint32_t key = 0;
        if (k < K) {
            uint64_t X = 0;
#pragma unroll
            for (int j = 0; j < se.n; ++j) X ^= pf[k + se.e[j]];
            const uint64_t hv = rol64(X, (int)((KMERLEN - 1 + k) & 63));
            key = (uint32_t)(ix.hpow2 ? (hv & ix.hmask) : (hv % ix.htab));
        }
        h.ky[q] = key;
    }
    wave_sync();
}

// the anchors' key positions: one load per lane, always issued (lanes
// without an anchor read entry 0 and drop it), so that the waits of the
// seed loop count loads exactly
template <int KPL>
__device__ __forceinline__ void go_anchor_load(const DevIndex& ix, GoHashed<KPL>& h, uint32_t& rnd)
{
    const int lane = threadIdx.x & 63;
    const bool go = h.L >= KMERLEN;
    const GoAnchors g = go_anchors(go ? h.L - KMERLEN + 1 : 1u);
    int my = -1;
    if (lane == 0) my = g.a[0];
    else if (g.v[2] && lane == (g.a[2] & 63)) my = g.a[2];
    else if (g.v[1] && lane == (g.a[1] & 63)) my = g.a[1];
    if (!go) my = -1;
    uint32_t key = h.ky[0];
#pragma unroll
    for (int q = 1; q < KPL; ++q)
        if ((my >> 6) == q) key = h.ky[q];
    // only the anchor lanes' values are read (go_predict); the loaded value
    // is not touched here, which would wait for it
    h.anc = ix.kpos[my >= 0 ? key : 0u];
    rnd += my >= 0 ? 1u : 0u;
}

// offsets from the anchors (one two of them agree on, else the first two
// distinct ones) and the genome-ordered loads at them
template <int KPL>
__device__ __forceinline__ void go_predict(const DevIndex& ix, const GoHashed<KPL>& h, GoPredicted<KPL>& p)
{
    const int lane = threadIdx.x & 63;
    p.L = h.L;
    const bool go = h.L >= KMERLEN;
    const uint32_t K = go ? h.L - KMERLEN + 1 : 0u;
    const GoAnchors g = go_anchors(go ? K : 1u);
    int64_t d[3];
    bool ok[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const uint32_t pos = g.v[i] ? lane_u32(h.anc, g.a[i] & 63) : ~0u;
        ok[i] = pos != ~0u;
        d[i] = (int64_t)pos - g.a[i];
    }
    int64_t d1 = 0, d2 = 0;
    int nd = 0;
    if (!go) {
    } else if (ok[0] && ok[1] && d[0] == d[1]) {
        d1 = d[0];
        nd = 1;
    } else if (ok[0] && ok[2] && d[0] == d[2]) {
        d1 = d[0];
        nd = 1;
    } else if (ok[1] && ok[2] && d[1] == d[2]) {
        d1 = d[1];
        nd = 1;
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (!ok[i]) continue;
            if (nd == 0) {
                d1 = d[i];
                nd = 1;
            } else if (nd == 1 && d[i] != d1) {
                d2 = d[i];
                nd = 2;
            }
        }
    }
    p.d1 = d1;
    p.d2 = d2;
    p.nd = nd;
    const int64_t gn = (int64_t)ix.gn;
#pragma unroll
    for (int q = 0; q < KPL; ++q) {
        const uint32_t k = (uint32_t)lane + 64u * q;
        p.ky[q] = h.ky[q];
        const int64_t p1 = d1 + (int64_t)k, p2 = d2 + (int64_t)k;
        const bool in1 = nd >= 1 && k < K && p1 >= 0 && p1 < gn;
        const bool in2 = nd >= 2 && k < K && p2 >= 0 && p2 < gn;
        // always issued (entry 0 for lanes without a prediction), and not
        // touched before go_fetch, which checks the offsets again
        p.gk1[q] = ix.gkey[in1 ? p1 : 0];
        p.gk2[q] = ix.gkey[in2 ? p2 : 0];
    }
}

// each k-mer's slot: gslot where the predicted key is its own, else the
// slot table at random
template <int KPL>
__device__ __forceinline__ void go_fetch(const DevIndex& ix, const GoPredicted<KPL>& p, GoSlots<KPL>& f,
                                         uint32_t& rnd, uint32_t& ord)
{
    const int lane = threadIdx.x & 63;
    f.L = p.L;
    const uint32_t K = p.L >= KMERLEN ? p.L - KMERLEN + 1 : 0u;
    const int64_t gn = (int64_t)ix.gn;
#pragma unroll
    for (int q = 0; q < KPL; ++q) {
        const uint32_t k = (uint32_t)lane + 64u * q;
        const bool in = k < K;
        const int64_t p1 = p.d1 + (int64_t)k, p2 = p.d2 + (int64_t)k;
        const bool in1 = in && p.nd >= 1 && p1 >= 0 && p1 < gn;
        const bool in2 = in && p.nd >= 2 && p2 >= 0 && p2 < gn;
        const bool o1 = in1 && p.gk1[q] == p.ky[q], o2 = !o1 && in2 && p.gk2[q] == p.ky[q];
        // always issued (slot 0 for lanes past the k-mers); go_gather reads
        // only the slots of k-mers below K
        const uint64_t* src = o1   ? ix.gslot + p1
                            : o2 ? ix.gslot + p2
                                 : ix.slot + (in ? p.ky[q] : 0u);
        f.sl[q] = *src;
        ord += (o1 || o2) ? 1u : 0u;
        rnd += (in && !(o1 || o2)) ? 1u : 0u;
    }
}

// Capacities of the light seed kernel for reads of at most ML letters
// (ML = 128 or 256): index hits ML, keys per map ML/2, counting slots ML.
// Larger (read, orientation) pairs go to the heavy kernel.
// (LB: the longest read of the variant, which sets the k-mers per lane;
// at most ML)
template <int ML, int LB = ML>
struct SeedCap {
    static constexpr uint32_t E = ML;
    static constexpr uint32_t INS = ML / 2;
    static constexpr uint32_t TB = ML + NBH - 1;
    static constexpr uint32_t HASH = ML;
    static constexpr int KPL = (LB - (int)KMERLEN + 63) / 64; // k-mers per lane
    static_assert((HASH & (HASH - 1)) == 0 && HASH >= E, "counting table");
    static_assert(LB <= ML, "read bound");
};
constexpr int SEED_SHORT = 128; // read length bound of the short variant
constexpr int SEED_GO_MID = 159; // the longest read with at most 128 k-mers (two per lane)

// per-wave LDS of the light seed kernel; the gather / count arrays and the
// hopscotch arrays share storage
template <typename TabT, int ML>
struct SeedLds {
    using C = SeedCap<ML>;
    TabT klo[ML];
    uint32_t klen[ML];
    uint32_t koff[ML];
    uint32_t k2[ML];
    // the gather / count arrays; between pairs also the prefix array of the
    // next pair's hash (go_hash)
    union alignas(8) {
        struct {
            uint32_t ent[C::E];
            uint32_t eaux[C::E];
            uint8_t ekm[C::E];
        } a;
        struct {
            uint16_t tb[C::TB];
            uint16_t tmp[C::INS];
            uint16_t over[C::INS];
        } b;
    } u;
    // the counting table is dead before the insertion lists are written
    union {
        struct {
            uint32_t key[C::HASH];
            uint32_t first[C::HASH];
            uint32_t cnt[C::HASH];
        } h;
        struct {
            uint32_t ins_key[2 * C::INS];
            uint16_t ins_cnt[2 * C::INS];
            uint32_t cand[2 * C::INS];
        } c;
    } v;
    alignas(4) uint8_t code[256];
};

// FAME_SEED_WPB independent waves per block (one block slot of a CU then
// holds several waves); each wave works through (read, orientation) pairs
// in grid-stride order.  Both variants are launched; the one that does not
// fit the batch's longest read (from k_prep) returns at once.
// GO: the slots through the genome-ordered copy (ix.gkey set)
// Variants by the batch's longest read to seed (k_prep): the one whose
// range (len_lo, len_hi] holds it runs, the others return at once.
template <typename TabT, int ML, bool GO, int LB = ML>
__global__ __launch_bounds__(64 * FAME_SEED_WPB) __attribute__((amdgpu_waves_per_eu(4))) void k_seed_light(DevIndex ix, BatchDev bd, uint32_t n_ro, uint32_t len_lo, uint32_t len_hi)
{
    using C = SeedCap<ML, LB>;
    if (bd.n_dev) n_ro = 2 * *bd.n_dev;
    {
        const uint32_t ml = *bd.max_len;
        if (ml <= len_lo || ml > len_hi) return;
    }
    __shared__ SeedShared sh;
    __shared__ SeedLds<TabT, ML> lds[FAME_SEED_WPB];
    const int lane = threadIdx.x & 63;
    // wave-uniform (a scalar): the LDS pointers below then live in SGPRs
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    SeedLds<TabT, ML>& m = lds[wv];
    SeedWork<TabT, uint16_t> w;
    w.klo = m.klo;
    w.klen = m.klen;
    w.koff = m.koff;
    w.k2 = m.k2;
    w.ent = m.u.a.ent;
    w.ekm = m.u.a.ekm;
    w.eaux = m.u.a.eaux;
    w.cap_e = C::E;
    w.hsize = C::HASH;
    w.hkey = m.v.h.key;
    w.hfirst = m.v.h.first;
    w.hcnt = m.v.h.cnt;
    w.ins_key = m.v.c.ins_key;
    w.ins_cnt = m.v.c.ins_cnt;
    w.cap_ins = C::INS;
    w.tb = m.u.b.tb;
    w.tb_cap = C::TB;
    w.tmp = m.u.b.tmp;
    w.over = m.u.b.over;
    w.cand = m.v.c.cand;
    w.code = m.code;
    seed_table_init(sh);
    PoolChunk pc = {0, 0, 0};
    SeedStats ss{};
    const uint32_t nw = gridDim.x * FAME_SEED_WPB;
    // the (read, orientation) pairs of the queried orientations only: pair q
    // is row 2q + o when one orientation o is queried (k_select never reads
    // the other's rows), row q when both are; an even wave stride over all
    // rows would leave every other wave without work
    const uint32_t om = bd.ori_mask;
    const bool both = om == 3u;
    const uint32_t nq = both ? n_ro : (om ? n_ro >> 1 : 0u);
    const uint32_t o1 = om == 2u ? 1u : 0u;
    auto ro_of = [&](uint32_t q) { return both ? q : 2 * q + o1; };
    auto defer = [&](uint32_t ro, uint64_t E, uint32_t bc) {
        if (lane == 0) {
            const uint32_t h = atomicAdd(bd.heavy_n, 1u);
            if (h < bd.heavy_cap) {
                bd.heavy_ro[h] = ro;
                bd.heavy_e[h] = E > 0x7FFFFFFFull ? 0x7FFFFFFFu : (uint32_t)E; // bit 31: handed on by the LDS tier
                bd.heavy_bc[h] = bc;
            } else {
                atomicOr(bd.flags, FLAG_HEAVY_FULL);
            }
            bd.cand_nf[ro] = 0;
            bd.cand_nr[ro] = 0;
            bd.cand_off[ro] = 0;
        }
    };
    static_assert(sizeof(m.u) >= 8 * (ML + 1), "prefix array of the hash");
    uint32_t go_rnd = 0, go_ord = 0;
    if constexpr (GO) {
        // genome-ordered slots, four pairs in flight per wave: pair qC is
        // matched while qC + nw has its slot loads in flight, qC + 2nw its
        // genome-ordered loads, qC + 3nw its anchor loads, and the row of
        // qC + 4nw loads (the first three iterations only fill the stages)
        uint64_t* pf = (uint64_t*)&m.u;
        __shared__ GoShared gsh;
        go_table_init(gsh);
        auto row = [&](uint32_t q) {
            return ((const uint32_t*)(bd.codes + (size_t)(q < nq ? ro_of(q) : 0u) * 256))[lane];
        };
        GoHashed<C::KPL> sH;
        GoPredicted<C::KPL> sP;
        GoSlots<C::KPL> sF;
        sH.L = sP.L = sF.L = 0;
        sH.anc = ~0u;
        const uint32_t q0 = blockIdx.x * FAME_SEED_WPB + wv;
        uint32_t cwH = row(q0);
        for (int64_t qC = (int64_t)q0 - 3 * (int64_t)nw; qC < (int64_t)nq; qC += nw) {
            const uint32_t qH = (uint32_t)(qC + 3 * (int64_t)nw);
            uint32_t cwN;
            GoHashed<C::KPL> nH;
            GoPredicted<C::KPL> nP;
            GoSlots<C::KPL> nF;
            go_hash<C::KPL>(ix, gsh, pf, cwH, qH < nq ? (lane_u32(cwH, 63) >> 24) : 0u, nH, go_rnd);
            // the stages' loads, issued inside seed_one after its entry
            // gather's loads (or here for a pair without one)
            auto stages = [&]() {
                cwN = row(qH + nw);
                go_anchor_load<C::KPL>(ix, nH, go_rnd);
                go_predict<C::KPL>(ix, sH, nP);
                go_fetch<C::KPL>(ix, sP, nF, go_rnd, go_ord);
            };
            if (qC >= 0 && sF.L != 0) {
                const uint32_t ro = ro_of((uint32_t)qC);
                uint64_t E;
                uint32_t bc;
                const int rc = seed_one<TabT, uint16_t, C::KPL, uint8_t, false>(
                    ix, bd, ro, sh, w, pc, false, ss, &E, &bc, sF.L << 24, true, sF.sl, 0u, stages);
                if (rc != 0) defer(ro, E, bc);
            } else {
                stages();
                if (qC >= 0 && lane == 0) { // not seeded: invalid read
                    const uint32_t ro = ro_of((uint32_t)qC);
                    bd.cand_nf[ro] = 0;
                    bd.cand_nr[ro] = 0;
                    bd.cand_off[ro] = 0;
                }
            }
            sH = nH;
            sP = nP;
            sF = nF;
            cwH = cwN;
        }
    } else if (ix.slot) {
        // slot-table mode, software-pipelined over this wave's pairs: while
        // pair j is worked on, the slot loads of pair j+1 and the code row of
        // pair j+2 are in flight (rows past the batch read row 0, unused)
        auto row = [&](uint32_t q) {
            return ((const uint32_t*)(bd.codes + (size_t)(q < nq ? ro_of(q) : 0u) * 256))[lane];
        };
        uint32_t q = blockIdx.x * FAME_SEED_WPB + wv;
        uint32_t cwc = row(q), cwn = row(q + nw);
        uint32_t Lc = q < nq ? (lane_u32(cwc, 63) >> 24) : 0;
        uint64_t slc[C::KPL], sln[C::KPL];
#pragma unroll
        for (int q = 0; q < C::KPL; ++q) slc[q] = sln[q] = 0;
        if (Lc) {
            ((uint32_t*)w.code)[lane] = cwc;
            wave_sync();
            seed_stage<C::KPL>(ix, sh, w.code, Lc - KMERLEN + 1, slc);
        }
        for (; q < nq; q += nw) {
            const uint32_t ro = ro_of(q);
            const uint32_t qn = q + nw;
            const uint32_t cwnn = row(qn + nw);
            const uint32_t Ln = qn < nq ? (lane_u32(cwn, 63) >> 24) : 0;
            if (Ln) {
                wave_sync(); // the previous pair's hashing has read w.code
                ((uint32_t*)w.code)[lane] = cwn;
                wave_sync();
                seed_stage<C::KPL>(ix, sh, w.code, Ln - KMERLEN + 1, sln);
            }
            if (Lc == 0) { // not seeded: invalid read
                if (lane == 0) {
                    bd.cand_nf[ro] = 0;
                    bd.cand_nr[ro] = 0;
                    bd.cand_off[ro] = 0;
                }
            } else {
                uint64_t E;
                uint32_t bc;
                const int rc = seed_one<TabT, uint16_t, C::KPL>(ix, bd, ro, sh, w, pc, false, ss, &E, &bc, cwc,
                                                                  true, slc);
                if (rc != 0) defer(ro, E, bc);
            }
#pragma unroll
            for (int q = 0; q < C::KPL; ++q) slc[q] = sln[q];
            cwc = cwn;
            cwn = cwnn;
            Lc = Ln;
        }
    }
    for (uint32_t q = blockIdx.x * FAME_SEED_WPB + wv; !GO && !ix.slot && q < nq; q += nw) {
        const uint32_t ro = ro_of(q);
        const uint32_t cw = ((const uint32_t*)(bd.codes + (size_t)ro * 256))[lane];
        if ((lane_u32(cw, 63) >> 24) == 0) { // not seeded: invalid read
            if (lane == 0) {
                bd.cand_nf[ro] = 0;
                bd.cand_nr[ro] = 0;
                bd.cand_off[ro] = 0;
            }
            continue;
        }
        wave_sync();
        uint64_t E;
        uint32_t bc;
        const int rc = seed_one<TabT, uint16_t, C::KPL>(ix, bd, ro, sh, w, pc, false, ss, &E, &bc, cw);
        if (rc != 0) defer(ro, E, bc);
    }
    if (pc.left) pool_gap_fill(bd, pc.pool, pc.base, pc.left, lane);
    if constexpr (GO) {
        go_rnd = wave_sum_u32(go_rnd);
        go_ord = wave_sum_u32(go_ord);
    } else {
        go_rnd = (uint32_t)ss.kmers;
    }
    if (lane == 0 && ss.seeded) {
        atomicAdd(&bd.stats[0], (unsigned long long)ss.kmers);
        atomicAdd(&bd.stats[1], (unsigned long long)ss.entries);
        atomicAdd(&bd.stats[2], (unsigned long long)ss.seeded);
        atomicAdd(&bd.stats[6], (unsigned long long)ss.bytes);
        atomicAdd(&bd.stats[8], (unsigned long long)ss.buckets);
    }
    // slot reads, also those of pairs deferred to the heavy kernel
    if (lane == 0 && (go_rnd || go_ord)) {
        atomicAdd(&bd.stats[11], (unsigned long long)go_rnd);
        atomicAdd(&bd.stats[12], (unsigned long long)go_ord);
    }
}

// heavy path: one block per deferred (read, orientation), global scratch
template <typename TabT>
__global__ __launch_bounds__(64) void k_seed_heavy(DevIndex ix, BatchDev bd, HeavyScratch hs)
{
    __shared__ SeedShared sh;
    __shared__ alignas(4) uint8_t s_code[256];
    // plain form: block hi serves item first + hi; batched form (count =
    // FAME_JOBS_ON_DEVICE): persistent blocks over the device list
    uint32_t count = hs.count;
    if (count == FAME_JOBS_ON_DEVICE) {
        if (*bd.flags & FLAG_ABORT) return;
        count = min(*bd.heavy_n, bd.heavy_cap);
    }
    if (blockIdx.x >= count) return;
    // (a tier's blocks skip the items of the other tier)
    uint8_t* base = hs.base + (size_t)blockIdx.x * hs.stride;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t* p = base + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const uint32_t ce = hs.cap_e;
    const size_t km = hs.kmax > FAME_DEV_MAX_READ ? hs.kmax : FAME_DEV_MAX_READ;
    SeedWork<TabT, uint32_t, uint16_t> w;
    w.klo = (TabT*)take(8 * km);
    w.klen = (uint32_t*)take(4 * km);
    w.koff = (uint32_t*)take(4 * km);
    w.k2 = (uint32_t*)take(4 * km);
    w.ent = (uint32_t*)take(4ull * ce);
    w.ekm = (uint16_t*)take(2ull * ce);
    w.eaux = (uint32_t*)take(4ull * ce);
    w.cap_e = ce;
    w.hsize = hs.hsize;
    w.hkey = (uint32_t*)take(4ull * hs.hsize);
    w.hfirst = (uint32_t*)take(4ull * hs.hsize);
    w.hcnt = (uint32_t*)take(4ull * hs.hsize);
    w.cap_ins = ce;
    w.ins_key = (uint32_t*)take(8ull * ce);
    w.ins_cnt = (uint16_t*)take(4ull * ce);
    w.tb_cap = hs.tb_cap;
    w.tb = (uint32_t*)take(4ull * hs.tb_cap);
    w.tmp = (uint32_t*)take(4ull * ce);
    w.over = (uint32_t*)take(4ull * ce);
    w.cand = (uint32_t*)take(8ull * ce);
    w.code = s_code;
    seed_table_init(sh);
    SeedStats ss{};
    const bool tiered = hs.count == FAME_JOBS_ON_DEVICE;
    for (uint32_t hi = blockIdx.x; hi < count; hi += gridDim.x) {
        if (tiered) {
            uint32_t e = bd.heavy_e[hi];
            if (hs.handed && !(e >> 31)) continue; // the LDS tier counted it
            e &= 0x7FFFFFFFu;
            if (e < hs.e_lo || e > hs.e_hi) continue;
        }
        const uint32_t hro = bd.heavy_ro[hs.first + hi];
        uint64_t E;
        uint32_t bc;
        PoolChunk pc = {0, 0, 0};
        uint32_t ro;
        int rc;
        if (hro >> 31) {
            // a long item: its row of reduced letters from k_long_measure
            const uint32_t q = hro & 0x7FFFFFFFu;
            const uint32_t r = bd.long_list[q >> 1];
            ro = 2 * r + (q & 1);
            w.code = bd.long_codes + (size_t)q * bd.long_stride;
            rc = seed_one<TabT, uint32_t, 0>(ix, bd, ro, sh, w, pc, true, ss, &E, &bc, 0u, false, nullptr,
                                             bd.lens[r]);
        } else {
            ro = hro;
            w.code = s_code;
            const uint32_t cw = ((const uint32_t*)(bd.codes + (size_t)ro * 256))[threadIdx.x];
            rc = seed_one<TabT, uint32_t, SeedCap<256>::KPL>(ix, bd, ro, sh, w, pc, true, ss, &E, &bc, cw);
        }
        if (rc != 0 && threadIdx.x == 0) {
            atomicOr(bd.flags, FLAG_CAPACITY);
            bd.cand_nf[ro] = 0;
            bd.cand_nr[ro] = 0;
        }
    }
    if (threadIdx.x == 0 && ss.seeded) {
        atomicAdd(&bd.stats[0], (unsigned long long)ss.kmers);
        atomicAdd(&bd.stats[1], (unsigned long long)ss.entries);
        atomicAdd(&bd.stats[2], (unsigned long long)ss.seeded);
        atomicAdd(&bd.stats[6], (unsigned long long)ss.bytes);
        atomicAdd(&bd.stats[8], (unsigned long long)ss.buckets);
        atomicAdd(&bd.stats[11], (unsigned long long)ss.kmers); // random slot lookups
    }
}

// Long reads (more than FAME_DEV_MAX_READ letters, listed by k_prep): one wave
// per (read, orientation) item q = 2 * (list index) + orientation writes the
// item's row of reduced letters (the short path's 256-byte rows cannot hold
// it), sums its bucket lengths as seed_one does (E, and bucketCount from the
// u32 sum) and lists it for the heavy kernel (heavy_ro = 0x80000000 | q),
// which the host sizes by E.  A read longer than any window text plus the
// errors allowed (winlen + MISCOUNT + ADDMIS + NERR letters) can match no
// window, so it is not seeded: k_select finds no candidate and counts it
// unmatched, as FAME's ShiftAnd finds nothing.  A read over
// FAME_LONG_MAX_READ that could match raises FLAG_TOO_LONG.
// The LDS tier between the light kernel (at most 256 index hits) and the
// heavy kernel (global scratch): (read, orientation) pairs of reads of at
// most FAME_DEV_MAX_READ letters, one wave per pair, every array in LDS and
// none per index hit (seed_tail_stream).  The counting table, the maps' key
// lists and the hopscotch arrays are sized for the distinct windows a read
// names (MID_H / MID_INS), not for its hits: repeat reads name each of a few
// dozen windows hundreds of times.  A pair whose windows do not fit goes to
// the heavy kernel: plain form, listed in ovl (ovl[0] = count, then positions
// in the heavy list); device form, bit 31 of its hit count set, which the
// heavy tiers that follow on the stream select on.
constexpr uint32_t MID_E = 0xFFFFFFu; // hits (24-bit hit indexes)
constexpr uint32_t MID_H = 512;
constexpr uint32_t MID_INS = 256;
constexpr uint32_t MID_TB = 1152; // two hopscotch bucket arrays' worth for MID_INS keys

struct SeedMidLds {
    uint32_t klo[FAME_DEV_MAX_READ + 1];
    uint32_t klen[FAME_DEV_MAX_READ + 1];
    uint32_t koff[FAME_DEV_MAX_READ + 1];
    uint32_t k2[FAME_DEV_MAX_READ + 1];
    // the counting table, dead once the maps' key lists are written; then
    // the hopscotch arrays
    union {
        struct {
            uint32_t key[MID_H];
            uint32_t first[MID_H];
            uint32_t cnt[MID_H];
        } h;
        struct {
            uint16_t tb[MID_TB];
            uint16_t tmp[MID_INS];
            uint16_t over[MID_INS];
        } b;
    } u;
    uint64_t srt[MID_H];
    uint32_t ins_key[2 * MID_INS];
    uint16_t ins_cnt[2 * MID_INS];
    uint32_t cand[2 * MID_INS];
    alignas(4) uint8_t code[256];
};

__global__ __launch_bounds__(64) void k_seed_mid(DevIndex ix, BatchDev bd, uint32_t first, uint32_t count,
                                                 uint32_t* __restrict__ ovl, uint32_t hsize)
{
    __shared__ SeedShared sh;
    __shared__ SeedMidLds m;
    const int lane = threadIdx.x;
    const bool dev = count == FAME_JOBS_ON_DEVICE;
    if (dev) {
        if (*bd.flags & FLAG_ABORT) return;
        count = min(*bd.heavy_n, bd.heavy_cap);
    }
    SeedWork<uint32_t, uint16_t> w;
    w.klo = m.klo;
    w.klen = m.klen;
    w.koff = m.koff;
    w.k2 = m.k2;
    w.ent = nullptr; // no per-hit arrays
    w.ekm = nullptr;
    w.eaux = nullptr;
    w.cap_e = MID_E;
    w.hsize = hsize; // a power of two, 64 .. MID_H
    w.hkey = m.u.h.key;
    w.hfirst = m.u.h.first;
    w.hcnt = m.u.h.cnt;
    w.srt = m.srt;
    w.ins_key = m.ins_key;
    w.ins_cnt = m.ins_cnt;
    w.cap_ins = MID_INS;
    w.tb = m.u.b.tb;
    w.tb_cap = MID_TB;
    w.tmp = m.u.b.tmp;
    w.over = m.u.b.over;
    w.cand = m.cand;
    w.code = m.code;
    seed_table_init(sh);
    SeedStats ss{};
    for (uint32_t hi = blockIdx.x; hi < count; hi += gridDim.x) {
        const uint32_t h = first + hi;
        if (dev && (bd.heavy_e[h] >> 31)) continue; // already handed on
        const uint32_t hro = bd.heavy_ro[h];
        int rc = 1;
        if (!(hro >> 31)) { // long items (own rows) always go to the heavy kernel
            const uint32_t cw = ((const uint32_t*)(bd.codes + (size_t)hro * 256))[lane];
            uint64_t E;
            uint32_t bc;
            PoolChunk pc = {0, 0, 0};
            wave_sync();
            rc = seed_one<uint32_t, uint16_t, SeedCap<256>::KPL, uint8_t, true>(ix, bd, hro, sh, w, pc, true, ss, &E,
                                                                                &bc, cw);
        }
        if (rc != 0 && lane == 0) {
            if (dev) bd.heavy_e[h] |= 0x80000000u;
            else ovl[1 + atomicAdd(ovl, 1u)] = h;
        }
    }
    if (lane == 0 && ss.seeded) {
        atomicAdd(&bd.stats[0], (unsigned long long)ss.kmers);
        atomicAdd(&bd.stats[1], (unsigned long long)ss.entries);
        atomicAdd(&bd.stats[2], (unsigned long long)ss.seeded);
        atomicAdd(&bd.stats[6], (unsigned long long)ss.bytes);
        atomicAdd(&bd.stats[8], (unsigned long long)ss.buckets);
        atomicAdd(&bd.stats[11], (unsigned long long)ss.kmers); // random slot lookups
    }
}

template <typename TabT>
__global__ __launch_bounds__(64) void k_long_measure(DevIndex ix, BatchDev bd)
{
    __shared__ SeedShared sh;
    seed_table_init(sh);
    const int lane = threadIdx.x;
    const uint32_t nitems = 2 * *bd.long_n;
    const TabT* tab = (const TabT*)ix.tab;
    for (uint32_t q = blockIdx.x; q < nitems; q += gridDim.x) {
        const uint32_t ori = q & 1;
        if (!((bd.ori_mask >> ori) & 1)) continue;
        const uint32_t r = bd.long_list[q >> 1];
        const uint32_t L = bd.lens[r];
        if (L > ix.winlen + MISCOUNT + ADDMIS + NERR) continue;
        if (L > FAME_LONG_MAX_READ) {
            if (lane == 0) atomicOr(bd.flags, FLAG_TOO_LONG);
            continue;
        }
        const uint8_t* src = (ori ? bd.rev : bd.seqs) + bd.offs[r];
        uint8_t* row = bd.long_codes + (size_t)q * bd.long_stride;
        for (uint32_t j = lane; j < L; j += 64) row[j] = (uint8_t)red_code(src[j]);
        wave_sync();
        const uint32_t K = L - KMERLEN + 1;
        uint64_t E = 0;
        uint32_t bsum = 0;
        for (uint32_t k = lane; k < K; k += 64) {
            const uint64_t key = seed_key(ix, sh, row, k);
            uint32_t len;
            if (ix.slot) {
                len = slot_len(ix.slot[key]);
            } else {
                const uint64_t d = (uint64_t)(tab[key + 1] - tab[key]);
                len = (uint32_t)d; // bucketCount's u32 truncation
                E += d > 0xFFFFFFull ? 0xFFFFFFull : d;
                bsum += len;
                continue;
            }
            E += len > 0xFFFFFFu ? 0xFFFFFFu : len;
            bsum += len;
        }
        for (int o = 32; o > 0; o >>= 1) E += __shfl_xor(E, o, 64);
        bsum = wave_sum_u32(bsum);
        if (lane == 0) {
            const uint32_t h = atomicAdd(bd.heavy_n, 1u);
            if (h < bd.heavy_cap) {
                bd.heavy_ro[h] = 0x80000000u | q;
                bd.heavy_e[h] = E > 0x7FFFFFFFull ? 0x7FFFFFFFu : (uint32_t)E; // bit 31: handed on by the LDS tier
                bd.heavy_bc[h] = bsum / 5;
            } else {
                atomicOr(bd.flags, FLAG_HEAVY_FULL);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_verify: approximate matching of one read over one window per lane
// ---------------------------------------------------------------------------

// Myers / Hyyro bit-vector step for one text letter.  The pattern occupies
// bits [z, 64W) of the W words (z = 64W - m) and the z low bits never
// match, which adds exactly z to every distance of the last row: the score
// kept here is D[m][j] + z, the edit distance of the whole pattern to the
// best substring ending at this text position, plus z.  The last pattern
// bit is bit 63 of the top word.  D[m][j] equals the lowest layer of
// ShiftAnd<6>'s automaton (ShiftAnd.h:276-341) holding the accepting state,
// whenever that distance is <= 6.
template <int W>
__device__ __forceinline__ void myers_step(uint64_t (&Pv)[W], uint64_t (&Mv)[W], const uint64_t* __restrict__ peq,
                                           int& score)
{
    unsigned long long carry = 0;
    uint64_t phc = 0, mhc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t eq = peq[w * 256];
        const uint64_t pv = Pv[w], mv = Mv[w];
        const uint64_t xv = eq | mv;
        unsigned long long co;
        const uint64_t sum = __builtin_addcll(eq & pv, pv, carry, &co);
        carry = co;
        const uint64_t xh = (sum ^ pv) | eq;
        const uint64_t ph = mv | ~(xh | pv);
        const uint64_t mh = pv & xh;
        if (w == W - 1) score += (int)(ph >> 63) - (int)(mh >> 63);
        const uint64_t ph2 = (ph << 1) | phc;
        const uint64_t mh2 = (mh << 1) | mhc;
        phc = ph >> 63;
        mhc = mh >> 63;
        Pv[w] = mh2 | ~(xv | ph2);
        Mv[w] = ph2 & xv;
    }
}

// Piece filter of the window scan.  A match with at most NERR errors aligns
// one of NERR+1 disjoint pieces of the pattern (length FP each, pattern
// positions [i*FP, (i+1)*FP)) to the text without an error.  Letters are
// compared in the three-letter alphabet A, G, {C,T}: a pattern letter
// matching a text letter (Peq) implies equal reduced letters.  A piece found
// at scan position q (its first letter) bounds the matches it can belong
// to: the matched substring starts at >= q - i*FP - NERR and ends at
// <= q - i*FP + m - 1 + NERR, at >= q - i*FP + m - 1 - NERR.  The bit-vector
// scan restarted at the start of that region therefore reproduces every
// end position with at most NERR errors there, with its error count, and
// no position outside the regions is one.  Regions are kept as a mask of
// 16-byte scan chunks: a chunk in which pieces ilo..ihi end (lowest and
// highest index among its hits) marks the union of the regions of those
// pieces ending anywhere in it, and a run of consecutive marked chunks is
// scanned from a fresh automaton.  Windows holding an N, patterns shorter than
// (NERR+1)*FILT_MIN_P and windows wider than the mask are scanned whole.
constexpr int FILT_MIN_P = 10;
constexpr int FILT_MAX_P = 14;
constexpr int FILT_MASKW = 4; // mask words: up to 256 chunks of 16 bytes

__device__ __forceinline__ uint32_t spread16(uint32_t x)
{
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// returns the number of window bytes scanned (0 when the lane has no job)
// multipliers of the piece hash (odd, 24 bits), tried in turn per pattern
__constant__ uint32_t c_piece_mul[16] = {0x9E3779u, 0x7F4A7Du, 0x85EBCBu, 0xC2B2AFu, 0x27D4EBu, 0x165667u,
                                         0xD6E8FDu, 0xA0761Du, 0xE70371u, 0x8EBC6Bu, 0x5851F5u, 0x2545F5u,
                                         0xB5297Bu, 0x68E31Du, 0xF1357Bu, 0x1B8735u};

// the low 32 bits of the product of the low 24 bits of a and b (HIP's
// __umul24 returns them as a signed int, which a right shift would extend)
__device__ __forceinline__ uint32_t mul24u(uint32_t a, uint32_t b) { return (uint32_t)__umul24(a, b); }

// bucket of a top-aligned 16-letter code: the top 3 bits of the 24-bit
// product of its last 10 letters (bits 12-31) with the multiplier
__device__ __forceinline__ uint32_t piece_bucket(uint32_t x, uint32_t mul) { return mul24u(x >> 12, mul) >> 29; }

// A job whose filter marked chunks can hand its runs of marked chunks to the
// block's run queue (k_verify): verify_one then stops after the filter and
// returns the job's text context and chunk mask here.  ctx.x = index of the
// text's memory chunk 0 in ix.pk; ctx.y = T | skip << 13 | rev << 17 |
// m << 18.
struct VerifyDefer {
    bool queue;        // hand the runs over (false: scan in place)
    bool deferred;     // the filter marked nr runs: ctx and cm hold the job
    uint32_t nr;
    uint2 ctx;
    uint64_t cm[4];
};

template <int W>
__device__ __forceinline__ uint64_t verify_one(const DevIndex& ix, const BatchDev& bd, uint32_t pool,
                                               const uint32_t* __restrict__ only, uint32_t tid,
                                               uint64_t* __restrict__ s_peq, uint2* __restrict__ s_ptab,
                                               VerifyDefer& df)
{
    if (only && bd.ovf_pool[tid] != pool) return 0;
    const uint32_t j = only ? only[tid] : tid;
    const uint32_t meta = bd.job_meta[pool][j];
    if (meta == JOB_GAP) return 0;
    const uint32_t ro = bd.job_ro[pool][j];
    const uint32_t r = ro >> 1, ori = ro & 1;
    const bool rev = (meta >> 31) != 0;
    const uint32_t id = meta & 0x3FFFFFFFu;
    // bit 30: matchRevSecond's scan, which has no guard for a window at
    // the chromosome start (ReadQueue.cpp:4736) and then also reads letter 0
    const bool unguarded = ((meta >> 30) & 1) != 0;
    const uint32_t m = bd.lens[r]; // 64(W-1) < m <= 64W: the pool holds reads of W words
    const int z = 64 * W - (int)m;
    const uint8_t* pat = (ori ? bd.rev : bd.seqs) + bd.offs[r];
    // the window's text first: its loads lead the pattern work
    const uint4 win = ix.win[id];
    const uint2 wtx = ix.wtext[id]; // the chromosome's first word in ix.pk, its length
    const uint32_t startPos = win.x;
    const uint64_t clen = wtx.y;
    const uint64_t span = (uint64_t)ix.winlen + MISCOUNT + ADDMIS - 1;
    // the text: chromosome positions [tstart, tstart + tlen), scanned
    // forward, or backward and complemented for the reverse map
    uint64_t tstart, tlen;
    if (!rev) {
        uint64_t end = (uint64_t)startPos + span;
        if (end > clen) end = clen;
        tstart = startPos;
        tlen = end - startPos;
    } else {
        const uint64_t e0 = startPos > 0 ? (uint64_t)startPos - 1 : 0;
        uint64_t s0 = (uint64_t)startPos + span;
        if (s0 >= clen) s0 = clen - 1;
        tstart = e0 + 1;
        tlen = s0 > e0 ? s0 - e0 : 0;
        if (unguarded && startPos == 0) {
            tstart = 0;
            tlen = s0 + 1;
        }
    }
    const int T = (int)tlen;
    const int skip = (int)(tstart & 15);
    const uint64_t* __restrict__ pk = ix.pk + wtx.x + (tstart >> 4);
    const int nchunk = (skip + T + 15) >> 4;
    // scan chunk 0's word (the last memory word of the reverse map's text)
    const uint64_t v0 = nchunk > 0 ? pk[rev ? nchunk - 1 : 0] : 0ull;
    // Peq masks of the pattern, laid out [letter*W + word][lane] so that the
    // lanes of a wave read consecutive LDS words (ShiftAnd loadBitmasks
    // ShiftAnd.h:345-432: pattern T also matches text C), shifted up by z;
    // the reduced piece codes of the filter (two bits per letter, A 0, G 2,
    // C/T 3, the piece's first letter in the low bits)
    uint64_t* peq = s_peq + threadIdx.x;
    constexpr int NP = (int)NERR + 1;
    const int FP = (int)m / NP < FILT_MAX_P ? (int)m / NP : FILT_MAX_P;
    uint32_t pc[NP];
    bool filt = W >= 2 && FP >= FILT_MIN_P;
    {
        uint64_t pm[4][W];
        {
            // The pattern arrives in aligned 16-byte chunks, each holding at
            // least one pattern byte (so no load leaves the buffer's pages).
            // The letters of each class (A, C, G, T, exact byte compares:
            // other bytes match nothing) become one bit per byte at raw
            // position 16 * chunk + byte; the raw masks are then shifted down
            // by the pattern's offset in its first chunk and cut at m.
            const uint32_t psk = (uint32_t)((uintptr_t)pat & 15);
            const uint8_t* p0 = pat - psk;
            const uint32_t nch = (psk + m + 15) >> 4; // chunks holding pattern bytes
            constexpr int RW = W + 1;                 // psk + m <= 15 + 64W bits
            uint64_t raw[4][RW];
            // bit 7 of each byte of x equal to ch (exact)
            auto eqb = [](uint32_t x, uint32_t ch) -> uint32_t {
                const uint32_t t = x ^ (ch * 0x01010101u);
                return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
            };
            // those four bits as a nibble (byte 0 -> bit 0)
            auto nib = [](uint32_t z) -> uint64_t { return (uint64_t)((((z >> 7) * 0x204081u) >> 21) & 15u); };
#pragma unroll
            for (int rw = 0; rw < RW; ++rw) {
                uint4 v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t c = (uint32_t)(4 * rw + q);
                    v[q] = *(const uint4*)(p0 + 16 * (c < nch ? c : nch - 1));
                }
                uint64_t wa = 0, wc = 0, wg = 0, wt = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool live = (uint32_t)(4 * rw + q) < nch;
                    const uint32_t d4[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                    for (int dd = 0; dd < 4; ++dd) {
                        const int sh = 16 * q + 4 * dd;
                        const uint32_t x = live ? d4[dd] : 0u;
                        wa |= nib(eqb(x, 'A')) << sh;
                        wc |= nib(eqb(x, 'C')) << sh;
                        wg |= nib(eqb(x, 'G')) << sh;
                        wt |= nib(eqb(x, 'T')) << sh;
                    }
                }
                raw[0][rw] = wa;
                raw[1][rw] = wc | wt; // pattern T also matches text C
                raw[2][rw] = wg;
                raw[3][rw] = wt;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint64_t x = raw[c][w] >> psk;
                    if (psk) x |= raw[c][w + 1] << (64 - psk);
                    const int lim = (int)m - 64 * w; // pattern bits in this word
                    if (lim < 64) x &= lim > 0 ? (~0ull >> (64 - lim)) : 0ull;
                    pm[c][w] = x;
                }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int w = W - 1; w >= 0; --w) {
                uint64_t x = pm[c][w] << z;
                if (w > 0 && z > 0) x |= pm[c][w - 1] >> (64 - z);
                peq[(c * W + w) * 256] = x;
            }
        if (W >= 2) {
            const uint32_t fm = (1u << FP) - 1;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                // pieces lie in pattern positions < NP * FILT_MAX_P <= 128
                const uint32_t off = (uint32_t)(i * FP);
                auto ext = [&](int c) -> uint32_t {
                    const uint64_t lo = pm[c][0], hi = pm[c][W >= 2 ? 1 : 0];
                    const uint64_t v = off == 0 ? lo : off < 64 ? ((lo >> off) | (hi << (64 - off))) : (hi >> (off - 64));
                    return (uint32_t)v & fm;
                };
                const uint32_t ba = ext(0), bct = ext(1), bg = ext(2);
                const uint32_t sct = spread16(bct), sg = spread16(bg);
                pc[i] = ((ba | bct | bg) == fm) ? (sct | (sct << 1) | (sg << 1)) : 0xFFFFFFFFu;
            }
        }
    }
    // The pieces in a per-lane hash table in LDS (8 buckets of two entries):
    // each piece top-aligned in 32 bits (its last letter in bits 30-31, the
    // bits below 32 - 2*FP masked off in compares).  The multiplier is the
    // first of c_piece_mul under which no bucket holds more than two
    // distinct pieces; free entries repeat a piece (never a spurious hit).
    // Pieces with a letter other than A, C, G, T never match and are left
    // out.  No multiplier found (or no piece): the window is scanned whole.
    // An entry's bits below 32 - 2*FP (at least 4) hold its piece's index,
    // so that a hit tells which piece ended there; pieces whose code repeats
    // an earlier piece's are left out of the table, and then a hit may be
    // any of them (pdup: the chunk's regions are those of every piece).
    uint32_t hmul = 0, tmask = 0;
    bool pdup = false;
    if (filt) {
        const uint32_t tsh = 32u - 2u * (uint32_t)FP;
        tmask = ~0u << tsh;
        uint32_t pt[NP];
        bool pv[NP];
        uint32_t fill = 0;
        bool any = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const bool valid = pc[i] != 0xFFFFFFFFu;
            pv[i] = valid;
            pt[i] = pv[i] ? pc[i] << tsh : 0u;
#pragma unroll
            for (int j = 0; j < i; ++j) pv[i] = pv[i] && !(pv[j] && pt[j] == pt[i]);
            pdup = pdup || (valid && !pv[i]);
            fill = (!any && pv[i]) ? (pt[i] | (uint32_t)i) : fill;
            any = any || pv[i];
        }
        bool ok = false;
        for (int t = 0; t < 16 && any && !ok; ++t) {
            const uint32_t mu = c_piece_mul[t];
            uint32_t occ = 0;
            bool good = true;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                if (!pv[i]) continue;
                const uint32_t b = piece_bucket(pt[i], mu);
                good = good && ((occ >> (2 * b)) & 3u) < 2u;
                occ += 1u << (2 * b);
            }
            if (good) {
                ok = true;
                hmul = mu;
            }
        }
        if (!ok) {
            filt = false;
        } else {
#pragma unroll
            for (uint32_t b = 0; b < 8; ++b) {
                uint32_t e0 = fill, e1 = fill, n = 0;
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const bool in = pv[i] && piece_bucket(pt[i], hmul) == b;
                    e0 = (in && n == 0) ? (pt[i] | (uint32_t)i) : e0;
                    e1 = (in && n == 1) ? (pt[i] | (uint32_t)i) : e1;
                    n += in ? 1u : 0u;
                }
                s_ptab[b * 256 + threadIdx.x] = make_uint2(e0, e1);
            }
        }
    }
    uint64_t Pv[W], Mv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        Pv[w] = ~0ull;
        Mv[w] = 0;
    }
    const int score0 = 64 * W; // m + z
    const int thr = (int)NERR + z;
    int score = score0;
    bool was = false;
    uint32_t prevErrs = MISCOUNT + 1;
    uint32_t ncomp = 0;
    const uint32_t minc = m - NERR;
    uint32_t nm = 0;
    const uint32_t cap = bd.match_cap;
    uint32_t* outm = bd.job_match[pool] + (size_t)j * cap;
    uint32_t* ovf = only ? bd.ovf_buf + bd.job_ovf_off[pool][j] : nullptr;
    // FAME's run logic: a run of consecutive accepting positions reports one
    // match, its first position, replaced by later ones with no more errors.
    // The run's current entry is kept in runbest and stored when the run
    // ends (the stored values equal those of storing at every replacement).
    uint32_t runbest = 0;
    auto flush = [&]() {
        const uint32_t k = nm - 1;
        if (ovf) ovf[k] = runbest;
        else if (k < cap) outm[k] = runbest;
    };
    auto step = [&](uint32_t c, int p) {
        myers_step<W>(Pv, Mv, peq + c * (W * 256), score);
        ++ncomp;
        const bool valid = ncomp >= minc; // earlier positions leave the run state alone
        const bool acc = valid && score <= thr;
        const uint32_t e = (uint32_t)(score - z);
        // reported end: querySeq it-start+MISCOUNT,
        // queryRevSeq it-end+pLen-1+MISCOUNT
        const uint32_t pos = rev ? (uint32_t)(p + 1 + m - 1 + MISCOUNT) : (uint32_t)(p + MISCOUNT);
        const uint32_t packed = (pos << 8) | e;
        const bool newrun = acc && !was;
        const bool upd = acc && (newrun || e <= prevErrs);
        if (valid && !acc && was) flush();
        nm += newrun ? 1u : 0u;
        runbest = upd ? packed : runbest;
        prevErrs = upd ? e : prevErrs;
        was = valid ? acc : was;
    };
    // The text comes from the packed copy of the genome: one 64-bit word per
    // 16 letters of a chromosome (ix.pk, from ix.wtext[window].x), two bits per
    // letter (A 0, C 1, G 2, T 3; letter j at bits 2j) and an N mask in bits
    // 32-47 (every byte other than A, C, G, T, as FAME's reader writes N for
    // them).  Scan chunk ci is memory chunk ci (forward) or nchunk-1-ci
    // (reverse, whose letters are reversed and complemented: A<->T, C<->G,
    // ShiftAnd.h:197-211).  Positions are 32-bit (texts are at most a few
    // windows long).
    struct Chunk {
        uint32_t x, nm; // codes in scan order (letter ii at bits 2ii), N bits of the text letters
        int lo, hi;     // scan-order range of text letters in the chunk
        int pb, pd;     // text position of scan-order letter 0, and its step
    };
    // the packed word of scan chunk ci (zero beyond the last chunk)
    auto raw_chunk = [&](int ci) -> uint64_t {
        uint64_t v = 0;
        if (ci >= 0 && ci < nchunk) v = pk[rev ? nchunk - 1 - ci : ci];
        return v;
    };
    // the same, for 0 <= ci (indexes past the end read the last chunk)
    auto raw_chunk_c = [&](int ci) -> uint64_t {
        const int c = ci < nchunk ? ci : nchunk - 1;
        return pk[rev ? nchunk - 1 - c : c];
    };
    auto load_chunk = [&](int ci, const uint64_t v, Chunk& k) {
        const int chunk = rev ? nchunk - 1 - ci : ci;
        uint32_t x = (uint32_t)v, nmk = (uint32_t)(v >> 32) & 0xFFFFu;
        if (rev) {
            // letter order reversed (bit reversal, then the two bits of each
            // letter swapped back) and complemented
            const uint32_t y = __brev(x);
            x = ~(((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1));
            nmk = __brev(nmk) >> 16;
        }
        const int base = 16 * chunk - skip; // text position of the chunk's first memory letter
        const int mlo = base < 0 ? -base : 0;
        const int rem = T - base;
        const int mhi = rem < 16 ? rem : 16;
        k.lo = rev ? 16 - mhi : mlo;
        k.hi = rev ? 16 - mlo : mhi;
        k.x = x;
        k.nm = nmk & ((0xFFFFu >> (16 - k.hi)) & (0xFFFFu << k.lo)); // N among the text letters only
        k.pb = rev ? base + 15 : base;
        k.pd = rev ? -1 : 1;
    };
    auto reset = [&]() {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            Pv[w] = ~0ull;
            Mv[w] = 0;
        }
        score = score0;
    };
    // the bit-vector scan over one chunk
    auto scan_chunk = [&](const Chunk& k) {
        if (k.lo == 0 && k.hi == 16 && k.nm == 0) {
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) step((k.x >> (2 * ii)) & 3u, k.pb + k.pd * ii);
        } else {
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                if (ii < k.lo || ii >= k.hi) continue;
                if ((k.nm >> ii) & 1u) { // an N restarts the automaton (ShiftAnd.h:142-146, 212-214)
                    reset();
                    continue;
                }
                step((k.x >> (2 * ii)) & 3u, k.pb + k.pd * ii);
            }
        }
    };
    // filter pass: the mask of scan chunks to scan
    uint64_t cmask[FILT_MASKW];
#pragma unroll
    for (int q = 0; q < FILT_MASKW; ++q) cmask[q] = 0;
    if (filt && nchunk > 64 * FILT_MASKW) filt = false;
    if (filt) {
        int x = 0; // scan index of the next text letter
        // marks the scan chunks holding scan positions [s0, s1]
        auto mark_range = [&](int s0, int s1) {
            if (s0 < 0) s0 = 0;
            if (s1 > T - 1) s1 = T - 1;
            if (s0 > s1) return;
            int c0, c1;
            if (!rev) {
                c0 = (s0 + skip) >> 4;
                c1 = (s1 + skip) >> 4;
            } else {
                c0 = nchunk - 1 - ((T - 1 - s0 + skip) >> 4);
                c1 = nchunk - 1 - ((T - 1 - s1 + skip) >> 4);
            }
#pragma unroll
            for (int q2 = 0; q2 < FILT_MASKW; ++q2) {
                const int lo = c0 > 64 * q2 ? c0 : 64 * q2;
                const int hi = c1 < 64 * q2 + 63 ? c1 : 64 * q2 + 63;
                if (lo <= hi) cmask[q2] |= (~0ull >> (63 - (hi - lo))) << (lo - 64 * q2);
            }
        };
        // Per text letter q: the 16 letters ending at q, top-aligned, from
        // this chunk's and the previous chunk's reduced codes (2 bits each,
        // A 0, G 2, C/T 3, scan order from bit 0), looked up in the piece
        // table.  Letters outside the window only ever add spurious hits.
        // Chunks go in blocks of two, the next block loaded while this one is
        // worked on (the empty asm keeps the loads where they are issued).
        uint32_t pprev = 0;
        const uint32_t tbase = threadIdx.x * 8u;
        // byte offset of letter ii's bucket in the table: bucket * 2048
        // (= bits 29-31 of the product moved to bits 11-13) | lane * 8.
        // The masked code XOR an entry is its piece index (< 16) on a hit, at
        // least 16 otherwise.
        auto probe = [&](uint32_t P, int ii) -> uint32_t {
            const uint32_t xc = ii < 15 ? __builtin_amdgcn_alignbit(P, pprev, 2 * (ii + 1)) : P;
            const uint32_t off = ((mul24u(xc >> 12, hmul) >> 18) & 0x3800u) | tbase;
            const uint2 e = *(const uint2*)((const char*)s_ptab + off);
            const uint32_t xm = xc & tmask;
            return min(xm ^ e.x, xm ^ e.y);
        };
        // the chunk's 16 positions (scan-order letters lo .. hi-1 are text);
        // x is the scan index of letter lo.  Any hit (the common case is
        // none) leads to the piece indexes and the regions they mark.
        auto probe_chunk = [&](uint32_t P, int lo, int hi) {
            uint32_t any = 0xFFFFFFFFu;
            uint32_t r[16];
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                r[ii] = probe(P, ii);
                any = min(any, r[ii]);
            }
            pprev = P;
            if (any < 16u) {
                // lowest piece index among the hits at text letters, and 15 -
                // the highest
                uint32_t alo = 0xFFFFFFFFu, ahi = 0xFFFFFFFFu;
#pragma unroll
                for (int ii = 0; ii < 16; ++ii) {
                    const bool in = ii >= lo && ii < hi;
                    alo = in ? min(alo, r[ii]) : alo;
                    ahi = in ? min(ahi, r[ii] ^ 15u) : ahi;
                }
                if (alo < 16u) {
                    // pieces ilo..ihi end in this chunk (every piece when codes
                    // repeat): the union of their regions.  Piece i ending at
                    // scan position e starts the match at >= e - (i+1)*FP + 1 -
                    // NERR and ends it at <= e - (i+1)*FP + m + NERR.
                    const int ilo = pdup ? 0 : (int)alo;
                    const int ihi = pdup ? NP - 1 : 15 - (int)ahi;
                    mark_range(x - (ihi + 1) * FP + 1 - (int)NERR,
                               x + (hi - lo) - 1 - (ilo + 1) * FP + (int)m + (int)NERR);
                }
            }
            x += hi - lo;
        };
        // reduced codes (C and T both 3) of a chunk's letters in scan order
        auto reduced = [](uint32_t c) { return c | ((c & 0x55555555u) << 1); };
        // a chunk with text letters outside [lo, hi) or an N: the general
        // form; false when the window holds an N
        auto filter_chunk = [&](int ci, const uint64_t v) -> bool {
            Chunk k;
            load_chunk(ci, v, k);
            if (k.nm) return false; // an N restarts the automaton: scan the window whole
            probe_chunk(reduced(k.x), k.lo, k.hi);
            return true;
        };
        // a chunk's letters in scan order (reverse map: letter order
        // reversed, complemented; any N bit means the same whichever way
        // round)
        auto scan_codes = [&](uint64_t v) -> uint32_t {
            const uint32_t xw = (uint32_t)v;
            const uint32_t y = __brev(xw);
            const uint32_t xr = ~(((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1));
            return rev ? xr : xw;
        };
        if (nchunk > 0 && filter_chunk(0, v0)) {
            // interior chunks are whole: their words follow one another in
            // memory (backwards for the reverse map), two loaded ahead.  A
            // chunk only records whether any of its positions hit (bit ci of
            // hm); the hits' pieces and regions are worked out afterwards,
            // chunk by chunk, from the same probes.  (Hits are rare per lane
            // but some lane of a wave has one at almost every chunk, and
            // working them out in the loop cost every lane the work.)
            const int last = nchunk - 1;
            const int x1 = x; // scan index of chunk 1's first letter
            const uint64_t* wp = pk + (rev ? last - 1 : 1);
            const int dw = rev ? -1 : 1;
            uint64_t c1 = last > 1 ? wp[0] : 0, c2 = last > 2 ? wp[dw] : 0;
            uint64_t hm[FILT_MASKW];
#pragma unroll
            for (int q = 0; q < FILT_MASKW; ++q) hm[q] = 0;
            uint64_t hacc = 0;
            for (int ci = 1; ci < last; ++ci) {
                const uint64_t v = c1;
                c1 = c2;
                if (ci + 2 < last) c2 = wp[(ci + 1) * dw];
                asm volatile("" ::: "memory");
                if ((uint32_t)(v >> 32) & 0xFFFFu) {
                    filt = false;
                    break;
                }
                const uint32_t P = reduced(scan_codes(v));
                uint32_t any = 0xFFFFFFFFu;
#pragma unroll
                for (int ii = 0; ii < 16; ++ii) any = min(any, probe(P, ii));
                pprev = P;
                hacc |= (uint64_t)(any < 16u) << (ci & 63);
                if ((ci & 63) == 63) {
#pragma unroll
                    for (int q = 0; q < FILT_MASKW; ++q)
                        if (q == (ci >> 6)) hm[q] = hacc;
                    hacc = 0;
                }
            }
            x = x1 + 16 * (last > 1 ? last - 1 : 0);
            if (filt) {
                const int cl = last - 1; // the last interior chunk
#pragma unroll
                for (int q = 0; q < FILT_MASKW; ++q)
                    if (cl >= 1 && q == (cl >> 6)) hm[q] |= hacc;
                const uint32_t pprev_last = pprev;
                // the interior chunks with hits: lowest piece index among
                // the hits, and 15 - the highest; pieces ilo..ihi end in the
                // chunk (every piece when codes repeat): the union of their
                // regions.  Piece i ending at scan position e starts the match
                // at >= e - (i+1)*FP + 1 - NERR and ends it at <= e - (i+1)*FP +
                // m + NERR.
#pragma unroll
                for (int q = 0; q < FILT_MASKW; ++q) {
                    for (uint64_t b = hm[q]; b; b &= b - 1) {
                        const int ci = 64 * q + (int)__builtin_ctzll(b);
                        pprev = reduced(scan_codes(raw_chunk(ci - 1)));
                        const uint32_t P = reduced(scan_codes(raw_chunk(ci)));
                        uint32_t alo = 0xFFFFFFFFu, ahi = 0xFFFFFFFFu;
#pragma unroll
                        for (int ii = 0; ii < 16; ++ii) {
                            const uint32_t r = probe(P, ii);
                            alo = min(alo, r);
                            ahi = min(ahi, r ^ 15u);
                        }
                        const int xc = x1 + 16 * (ci - 1);
                        const int ilo = pdup ? 0 : (int)alo;
                        const int ihi = pdup ? NP - 1 : 15 - (int)ahi;
                        mark_range(xc - (ihi + 1) * FP + 1 - (int)NERR, xc + 15 - (ilo + 1) * FP + (int)m + (int)NERR);
                    }
                }
                pprev = pprev_last;
            }
            if (filt && last > 0 && !filter_chunk(last, raw_chunk_c(last))) filt = false;
        } else {
            filt = false;
        }
    }
    if (filt && df.queue) {
        // the runs of marked chunks go to the block's queue (the caller
        // scans them, and stores the job's matches)
        uint32_t nr = 0;
#pragma unroll
        for (int q = 0; q < FILT_MASKW; ++q) {
            const uint64_t prevw = q ? cmask[q - 1] : 0ull;
            nr += (uint32_t)__popcll(cmask[q] & ~((cmask[q] << 1) | (prevw >> 63)));
        }
        df.deferred = true;
        df.nr = nr;
        df.ctx = make_uint2((uint32_t)(pk - ix.pk),
                            (uint32_t)T | ((uint32_t)skip << 13) | ((rev ? 1u : 0u) << 17) | (m << 18));
#pragma unroll
        for (int q = 0; q < FILT_MASKW; ++q) df.cm[q] = cmask[q];
        return tlen ? tlen : 1;
    }
    if (filt) {
        // the marked chunks in scan order, the next one loaded ahead; a fresh
        // automaton at the start of every run of consecutive chunks
        auto next_marked = [&]() -> int {
            int ci = -1;
#pragma unroll
            for (int q = 0; q < FILT_MASKW; ++q)
                if (ci < 0 && cmask[q]) {
                    ci = 64 * q + __ffsll((long long)cmask[q]) - 1;
                    cmask[q] &= cmask[q] - 1;
                }
            return ci;
        };
        int prevci = -2;
        int ci = next_marked();
        uint64_t v1 = raw_chunk(ci);
        while (ci >= 0) {
            const int cn = next_marked();
            const uint64_t v = v1;
            v1 = raw_chunk(cn);
            asm volatile("" ::: "memory");
            if (ci != prevci + 1) {
                reset();
                if (was) flush();
                was = false;
            }
            prevci = ci;
            Chunk k;
            load_chunk(ci, v, k);
            // scan index of the chunk's first text letter (no N in the window)
            const int p0 = k.pb + k.pd * k.lo;
            ncomp = (uint32_t)(rev ? T - 1 - p0 : p0);
            scan_chunk(k);
            ci = cn;
        }
    } else {
        uint64_t v1 = raw_chunk(0);
        for (int ci = 0; ci < nchunk; ++ci) {
            const uint64_t v = v1;
            v1 = raw_chunk(ci + 1);
            asm volatile("" ::: "memory");
            Chunk k;
            load_chunk(ci, v, k);
            scan_chunk(k);
        }
    }
    if (was) flush();
    if (!only) {
        bd.job_nm[pool][j] = nm;
        if (nm > cap) {
            const uint32_t o = atomicAdd(bd.ovf_n, 1u);
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

// One run of marked chunks (scan chunks c0 .. c0+len-1) of another lane's job,
// from the job's context (VerifyDefer::ctx) and its Peq masks (column owner of
// s_peq): the same bit-vector scan and run logic as verify_one, from a fresh
// automaton.  Returns the matches found; res[0 .. cap) holds the first of them.
template <int W>
__device__ __forceinline__ uint32_t scan_run(const DevIndex& ix, uint2 ctx, const uint64_t* __restrict__ s_peq,
                                         uint32_t owner, int c0, int len, uint32_t cap, uint32_t* __restrict__ res)
{
    const uint64_t* __restrict__ pk = ix.pk + ctx.x;
    const int T = (int)(ctx.y & 0x1FFFu);
    const int skip = (int)((ctx.y >> 13) & 15u);
    const bool rev = ((ctx.y >> 17) & 1u) != 0;
    const uint32_t m = ctx.y >> 18;
    const int nchunk = (skip + T + 15) >> 4;
    const int z = 64 * W - (int)m;
    const uint64_t* peq = s_peq + owner;
    uint64_t Pv[W], Mv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        Pv[w] = ~0ull;
        Mv[w] = 0;
    }
    const int thr = (int)NERR + z;
    int score = 64 * W;
    bool was = false;
    uint32_t prevErrs = MISCOUNT + 1;
    uint32_t ncomp = 0;
    const uint32_t minc = m - NERR;
    uint32_t nm = 0;
    uint32_t runbest = 0;
    auto flush = [&]() {
        const uint32_t k = nm - 1;
        if (k < cap) res[k] = runbest;
    };
    auto step = [&](uint32_t c, int p) {
        myers_step<W>(Pv, Mv, peq + c * (W * 256), score);
        ++ncomp;
        const bool valid = ncomp >= minc;
        const bool acc = valid && score <= thr;
        const uint32_t e = (uint32_t)(score - z);
        const uint32_t pos = rev ? (uint32_t)(p + 1 + m - 1 + MISCOUNT) : (uint32_t)(p + MISCOUNT);
        const uint32_t packed = (pos << 8) | e;
        const bool newrun = acc && !was;
        const bool upd = acc && (newrun || e <= prevErrs);
        if (valid && !acc && was) flush();
        nm += newrun ? 1u : 0u;
        runbest = upd ? packed : runbest;
        prevErrs = upd ? e : prevErrs;
        was = valid ? acc : was;
    };
    auto word = [&](int ci) -> uint64_t { return pk[rev ? nchunk - 1 - ci : ci]; };
    const int c1 = c0 + len;
    uint64_t vn = word(c0);
    for (int ci = c0; ci < c1; ++ci) {
        const uint64_t v = vn;
        if (ci + 1 < c1) vn = word(ci + 1);
        asm volatile("" ::: "memory");
        const int chunk = rev ? nchunk - 1 - ci : ci;
        uint32_t x = (uint32_t)v;
        if (rev) {
            const uint32_t y = __brev(x);
            x = ~(((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1));
        }
        const int base = 16 * chunk - skip;
        const int mlo = base < 0 ? -base : 0;
        const int rem = T - base;
        const int mhi = rem < 16 ? rem : 16;
        const int lo = rev ? 16 - mhi : mlo, hi = rev ? 16 - mlo : mhi;
        const int pb = rev ? base + 15 : base, pd = rev ? -1 : 1;
        const int p0 = pb + pd * lo;
        ncomp = (uint32_t)(rev ? T - 1 - p0 : p0);
        if (lo == 0 && hi == 16) {
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) step((x >> (2 * ii)) & 3u, pb + pd * ii);
        } else {
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                if (ii < lo || ii >= hi) continue;
                step((x >> (2 * ii)) & 3u, pb + pd * ii);
            }
        }
    }
    if (was) flush();
    return nm;
}

// The first chunk at or after `from` whose bit in the chunk mask is `set`
// (64 * FILT_MASKW if none); indexes the mask words with constants only.
__device__ __forceinline__ int cmask_first(const uint64_t (&cm)[4], bool set, int from)
{
    int r = 64 * FILT_MASKW;
    const int fq = from >> 6;
#pragma unroll
    for (int q = FILT_MASKW - 1; q >= 0; --q) {
        uint64_t w = set ? cm[q] : ~cm[q];
        if (q == fq) w &= ~0ull << (from & 63);
        if (q >= fq && w) r = 64 * q + __ffsll((long long)w) - 1;
    }
    return r;
}

// Jobs of at least two pattern words run in block rounds of 256: every lane
// prepares and filters its job; the runs of marked chunks then go to a
// queue in the LDS the filters' tables used (long runs from its top, short
// ones from its bottom), which the block's waves take 64 runs at a time,
// long ones first, so that a wave's scan no longer lasts as long as its
// lane with the most marked chunks; each owner finally gathers its runs'
// matches in scan order.  A run's matches do not depend on the others' (each
// starts from a fresh automaton), so the results equal verify_one's.
// One-word patterns scan in place, one job per lane; so do the jobs of the
// overflow re-verification (`only`), in the same block rounds (one copy of
// verify_one in the kernel keeps its code within the instruction cache).
template <int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_verify(DevIndex ix, BatchDev bd, uint32_t pool,
                                                const uint32_t* __restrict__ only, uint32_t n_only)
{
    // exactly 32 KiB of LDS for two pattern words: five blocks per CU
    __shared__ uint64_t s_peq[4 * W * 256];
    __shared__ uint2 s_ptab[W >= 2 ? 8 * 256 : 1];
    // plain form: one thread per job (the host sizes the grid); batched form:
    // persistent blocks over the device's job count, nothing after an abort
    uint32_t njobs;
    if (only) njobs = n_only != FAME_JOBS_ON_DEVICE ? n_only : min(*bd.ovf_n, bd.ovf_list_cap);
    else njobs = min(bd.pool_n[pool], pc_get_cap(bd, pool));
    if (bd.n_dev && (*bd.flags & FLAG_ABORT)) njobs = 0;
    uint64_t jobs = 0, bytes = 0;
    if constexpr (W < 2) {
        VerifyDefer df{};
        for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < njobs; t += gridDim.x * blockDim.x) {
            const uint64_t b = verify_one<W>(ix, bd, pool, only, t, s_peq, s_ptab, df);
            jobs += b ? 1 : 0;
            bytes += b;
        }
    } else {
        // the queue in s_ptab: owners' contexts [256] (uint2), their first
        // result slot [256] (u16), run descriptors [Q] (u32: owner | first
        // chunk << 8 | (length - 1) << 16 | index of the run in its job << 24),
        // results [Q][1 + cap] (u32: count, matches), and in its last 16
        // bytes (bucket 7 of lanes 254 and 255 while the tables live) the
        // counters: result slots taken, short runs, long runs, runs handed out
        const uint32_t cap = bd.match_cap;
        uint2* s_ctx = s_ptab;
        uint16_t* s_rb = (uint16_t*)(s_ptab + 256);
        uint32_t* s_desc = (uint32_t*)(s_ptab + 320);
        uint32_t* s_q = (uint32_t*)(s_ptab + 8 * 256 - 2);
        constexpr uint32_t QBYTES = 8 * 256 * 8 - 320 * 8 - 16;
        const uint32_t Qmax = QBYTES / (4 + 4 * (1 + cap));
        uint32_t* s_resv = s_desc + Qmax;
        uint32_t Q = Qmax < 4096 ? Qmax : 4096;
        if (bd.verify_runs_cap && bd.verify_runs_cap < Q) Q = bd.verify_runs_cap;
        constexpr int LONG_RUN = 12; // chunks
        const int lane = threadIdx.x & 63;
        // a job whose runs did not fit the queue is verified again, in
        // place, in a repeat of its round in which only such jobs run
        bool redo = false, again = false;
        for (uint32_t base = blockIdx.x * 256; base < njobs; base += again ? 0u : gridDim.x * 256) {
            const uint32_t t = base + threadIdx.x;
            VerifyDefer df{};
            df.queue = !redo && !only; // overflow re-verification scans in place
            if (t < njobs && redo == again) {
                const uint64_t b = verify_one<W>(ix, bd, pool, only, t, s_peq, s_ptab, df);
                jobs += redo ? 0 : (b ? 1 : 0);
                bytes += redo ? 0 : b;
            }
            // this lane's table is dead: lanes 254 and 255 clear the counters
            if (threadIdx.x >= 254) *(uint2*)(s_q + 2 * (threadIdx.x - 254)) = make_uint2(0u, 0u);
            __syncthreads(); // the tables are dead: the queue is written
            uint32_t rb = 0;
            if (df.deferred) {
                rb = atomicAdd(&s_q[0], df.nr);
                redo = rb + df.nr > Q;
            } else {
                redo = false;
            }
            if (df.deferred && !redo) {
                s_ctx[threadIdx.x] = df.ctx;
                s_rb[threadIdx.x] = (uint16_t)rb;
                // runs in scan order: first set bit, first clear bit after it
                uint32_t k = 0;
                for (int c0 = cmask_first(df.cm, true, 0); c0 < 64 * FILT_MASKW; ++k) {
                    const int c1 = cmask_first(df.cm, false, c0);
                    const int len = c1 - c0;
                    const uint32_t pos = len > LONG_RUN ? Q - 1 - atomicAdd(&s_q[2], 1u) : atomicAdd(&s_q[1], 1u);
                    s_desc[pos] = threadIdx.x | ((uint32_t)c0 << 8) | ((uint32_t)(len - 1) << 16) | (k << 24);
                    c0 = cmask_first(df.cm, true, c1);
                }
            }
            __syncthreads(); // the queue is complete
            again = s_q[0] > Q; // some job's runs did not fit
            {
                const uint32_t ns = s_q[1], nl = s_q[2], nq = ns + nl;
                for (;;) {
                    uint32_t g = 0;
                    if (lane == 0) g = atomicAdd(&s_q[3], 64u);
                    g = (uint32_t)__builtin_amdgcn_readfirstlane((int)lane_u32(g, 0));
                    if (g >= nq) break;
                    const uint32_t v = g + lane;
                    if (v < nq) {
                        const uint32_t pos = v < nl ? Q - 1 - v : v - nl;
                        const uint32_t d = s_desc[pos];
                        const uint32_t owner = d & 0xFFu;
                        uint32_t* r = s_resv + (size_t)((uint32_t)s_rb[owner] + (d >> 24)) * (1 + cap);
                        r[0] = scan_run<W>(ix, s_ctx[owner], s_peq, owner, (int)((d >> 8) & 0xFFu),
                                           (int)((d >> 16) & 0xFFu) + 1, cap, r + 1);
                    }
                }
            }
            __syncthreads(); // every run is scanned
            if (df.deferred && !redo) {
                // the owner's matches in scan order, then as verify_one ends
                const uint32_t j = t;
                uint32_t* outm = bd.job_match[pool] + (size_t)j * cap;
                uint32_t nm = 0;
                for (uint32_t k = 0; k < df.nr; ++k) {
                    const uint32_t* r = s_resv + (size_t)(rb + k) * (1 + cap);
                    const uint32_t cnt = r[0];
                    for (uint32_t q = 0; q < cnt && q < cap; ++q)
                        if (nm + q < cap) outm[nm + q] = r[1 + q];
                    nm += cnt;
                }
                bd.job_nm[pool][j] = nm;
                if (nm > cap) {
                    const uint32_t o = atomicAdd(bd.ovf_n, 1u);
                    if (o < bd.ovf_list_cap) {
                        bd.ovf_pool[o] = pool;
                        bd.ovf_job[o] = j;
                    } else {
                        atomicOr(bd.flags, FLAG_OVF_FULL);
                    }
                }
            }
            __syncthreads(); // the queue's LDS is free for the next round's tables
        }
    }
    if (!only) {
        // one atomic per wave
        for (int o = 32; o > 0; o >>= 1) {
            jobs += __shfl_xor(jobs, o, 64);
            bytes += __shfl_xor(bytes, o, 64);
        }
        if ((threadIdx.x & 63) == 0 && jobs) {
            atomicAdd(&bd.stats[3], (unsigned long long)jobs);
            atomicAdd(&bd.stats[4], (unsigned long long)bytes);
            atomicAdd(&bd.stats[10], (unsigned long long)bytes * W);
        }
    }
}

// Verify jobs of reads over FAME_DEV_MAX_READ letters (pool FAME_LONG_POOL):
// one wave per job, lane w holding word w of the pattern's bit vectors (at
// most 64 words: FAME_LONG_MAX_READ letters), the pattern top-aligned in its
// W words as in verify_one (pattern position i at bit i + z).  The Myers
// step's addition carries across words through the wave: each word's sum
// without carry-in generates (overflow) or propagates (all ones) a carry,
// and with G and P the masks of the two over the words, the carry into each
// word is bit w of (G + (G | P)) ^ G ^ (G | P); the one-bit shifts take the
// top bit of the word below.  The whole window is scanned (no piece
// filter); N letters, run logic and reported positions as in verify_one.
__global__ __launch_bounds__(64) void k_verify_long(DevIndex ix, BatchDev bd, const uint32_t* __restrict__ only,
                                                    uint32_t n_only)
{
    constexpr uint32_t pool = FAME_LONG_POOL;
    const int lane = threadIdx.x;
    uint32_t njobs;
    if (only) njobs = n_only != FAME_JOBS_ON_DEVICE ? n_only : min(*bd.ovf_n, bd.ovf_list_cap);
    else njobs = min(bd.pool_n[pool], pc_get_cap(bd, pool));
    if (bd.n_dev && (*bd.flags & FLAG_ABORT)) njobs = 0;
    uint64_t jobs = 0, bytes = 0, wletters = 0;
    const uint32_t cap = bd.match_cap;
    for (uint32_t tid = blockIdx.x; tid < njobs; tid += gridDim.x) {
        if (only && bd.ovf_pool[tid] != pool) continue;
        const uint32_t j = only ? only[tid] : tid;
        const uint32_t meta = bd.job_meta[pool][j];
        if (meta == JOB_GAP) continue;
        const uint32_t ro = bd.job_ro[pool][j];
        const uint32_t r = ro >> 1, ori = ro & 1;
        const bool rev = (meta >> 31) != 0;
        const uint32_t id = meta & 0x3FFFFFFFu;
        const bool unguarded = ((meta >> 30) & 1) != 0;
        const uint32_t m = bd.lens[r];
        const int W = (int)((m + 63) >> 6);
        const int z = 64 * W - (int)m;
        const uint8_t* pat = (ori ? bd.rev : bd.seqs) + bd.offs[r];
        // Peq of this lane's word (ShiftAnd loadBitmasks ShiftAnd.h:345-432:
        // pattern T also matches text C)
        uint64_t pe0 = 0, pe1 = 0, pe2 = 0, pe3 = 0;
        if (lane < W) {
            for (int b = 0; b < 64; ++b) {
                const int i = 64 * lane + b - z;
                if (i < 0 || i >= (int)m) continue;
                const uint8_t ch = pat[i];
                const uint64_t bit = 1ull << b;
                pe0 |= ch == 'A' ? bit : 0ull;
                pe1 |= (ch == 'C' || ch == 'T') ? bit : 0ull;
                pe2 |= ch == 'G' ? bit : 0ull;
                pe3 |= ch == 'T' ? bit : 0ull;
            }
        }
        // the text (verify_one's bounds)
        const uint4 win = ix.win[id];
        const uint32_t startPos = win.x, chrom = win.y;
        const uint64_t clen = ix.clen[chrom];
        const uint64_t span = (uint64_t)ix.winlen + MISCOUNT + ADDMIS - 1;
        uint64_t tstart, tlen;
        if (!rev) {
            uint64_t end = (uint64_t)startPos + span;
            if (end > clen) end = clen;
            tstart = startPos;
            tlen = end - startPos;
        } else {
            const uint64_t e0 = startPos > 0 ? (uint64_t)startPos - 1 : 0;
            uint64_t s0 = (uint64_t)startPos + span;
            if (s0 >= clen) s0 = clen - 1;
            tstart = e0 + 1;
            tlen = s0 > e0 ? s0 - e0 : 0;
            if (unguarded && startPos == 0) {
                tstart = 0;
                tlen = s0 + 1;
            }
        }
        const int T = (int)tlen;
        const uint64_t* __restrict__ pk = ix.pk + ix.pkoff[chrom];
        const uint64_t wmask = W >= 64 ? ~0ull : ((1ull << W) - 1);
        uint64_t pv = lane < W ? ~0ull : 0ull, mv = 0;
        const int score0 = 64 * W;
        const int thr = (int)NERR + z;
        int score = score0;
        bool was = false;
        uint32_t prevErrs = MISCOUNT + 1;
        uint32_t ncomp = 0;
        const uint32_t minc = m - NERR;
        uint32_t nm = 0;
        uint32_t runbest = 0;
        uint32_t* outm = bd.job_match[pool] + (size_t)j * cap;
        uint32_t* ovf = only ? bd.ovf_buf + bd.job_ovf_off[pool][j] : nullptr;
        auto flush = [&]() {
            const uint32_t k = nm - 1;
            if (lane == 0) {
                if (ovf) ovf[k] = runbest;
                else if (k < cap) outm[k] = runbest;
            }
        };
        uint64_t wv = 0;
        int64_t widx = -1;
        for (int sidx = 0; sidx < T; ++sidx) {
            const int p = rev ? T - 1 - sidx : sidx;
            const uint64_t x = tstart + (uint64_t)p;
            if ((int64_t)(x >> 4) != widx) {
                widx = (int64_t)(x >> 4);
                wv = pk[widx];
            }
            const uint32_t sh = (uint32_t)(x & 15);
            if ((wv >> (32 + sh)) & 1) { // an N restarts the automaton (ShiftAnd.h:142-146, 212-214)
                pv = lane < W ? ~0ull : 0ull;
                mv = 0;
                score = score0;
                continue;
            }
            uint32_t c = (uint32_t)(wv >> (2 * sh)) & 3u;
            if (rev) c ^= 3u; // the reverse map scans the complement
            const uint64_t eq = c == 0 ? pe0 : c == 1 ? pe1 : c == 2 ? pe2 : pe3;
            const uint64_t xv = eq | mv;
            const uint64_t s0 = (eq & pv) + pv;
            const uint64_t G = __ballot(s0 < pv) & wmask;
            const uint64_t P = __ballot(s0 == ~0ull) & wmask;
            const uint64_t B = G | P;
            const uint64_t C = (G + B) ^ G ^ B;
            const uint64_t sum = s0 + ((C >> lane) & 1ull);
            const uint64_t xh = (sum ^ pv) | eq;
            const uint64_t ph = mv | ~(xh | pv);
            const uint64_t mh = pv & xh;
            const uint32_t phb = (uint32_t)(ph >> 63), mhb = (uint32_t)(mh >> 63);
            score += (int)lane_u32(phb, W - 1) - (int)lane_u32(mhb, W - 1);
            const uint32_t phc = __shfl_up(phb, 1, 64), mhc = __shfl_up(mhb, 1, 64);
            const uint64_t ph2 = (ph << 1) | (lane ? (uint64_t)phc : 0ull);
            const uint64_t mh2 = (mh << 1) | (lane ? (uint64_t)mhc : 0ull);
            pv = lane < W ? (mh2 | ~(xv | ph2)) : 0ull;
            mv = lane < W ? (ph2 & xv) : 0ull;
            ++ncomp;
            const bool valid = ncomp >= minc;
            const bool acc = valid && score <= thr;
            const uint32_t e = (uint32_t)(score - z);
            const uint32_t pos = rev ? (uint32_t)(p + 1 + m - 1 + MISCOUNT) : (uint32_t)(p + MISCOUNT);
            const uint32_t packed = (pos << 8) | e;
            const bool newrun = acc && !was;
            const bool upd = acc && (newrun || e <= prevErrs);
            if (valid && !acc && was) flush();
            nm += newrun ? 1u : 0u;
            runbest = upd ? packed : runbest;
            prevErrs = upd ? e : prevErrs;
            was = valid ? acc : was;
        }
        if (was) flush();
        if (!only && lane == 0) {
            bd.job_nm[pool][j] = nm;
            if (nm > cap) {
                const uint32_t o = atomicAdd(bd.ovf_n, 1u);
                if (o < bd.ovf_list_cap) {
                    bd.ovf_pool[o] = pool;
                    bd.ovf_job[o] = j;
                } else {
                    atomicOr(bd.flags, FLAG_OVF_FULL);
                }
            }
        }
        const uint64_t b = tlen ? tlen : 1;
        jobs += 1;
        bytes += b;
        wletters += b * (uint64_t)W;
    }
    if (!only && lane == 0 && jobs) {
        atomicAdd(&bd.stats[3], (unsigned long long)jobs);
        atomicAdd(&bd.stats[4], (unsigned long long)bytes);
        atomicAdd(&bd.stats[10], (unsigned long long)wletters);
    }
}

// ---------------------------------------------------------------------------
// k_select: saQuerySeedSetRef uniqueness logic + decision block, per read
// ---------------------------------------------------------------------------

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
        const uint32_t id = bd.job_meta[pool][j] & 0x3FFFFFFFu;
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
    if (bd.n_dev) {
        if (*bd.flags & FLAG_ABORT) return; // the host re-runs the batch on the plain path
        n = *bd.n_dev;
    }
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    // per block: the counters, then the three lists' appends (one global
    // atomic per list and block: the lists' heads are the same address for
    // every block)
    __shared__ uint32_t s_cnt[5];
    __shared__ uint32_t s_app[3], s_base[3];
    if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x < 3) s_app[threadIdx.x] = 0;
    __syncthreads();
    int want = 0;
    int cnt_at = -1;   // the counter of the read's status (0..2; invalid reads none)
    int tally_at = -1; // counter 3 / 4 of get_stranded
    if (r < n) {
        uint8_t st = 3; // invalid, no counter
        uint64_t out = 0;
        uint8_t use = 0;
        if (bd.pre[r] == 0) {
            const uint32_t pool = pool_of_len(bd.lens[r]);
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
                if (get_stranded) tally_at = tally > 0 ? 3 : 4;
            }
            cnt_at = st;
        }
        bd.status[r] = st;
        bd.use[r] = use;
        bd.mat[r] = out;
        bd.valid[r] = st == 0;
        // computeMethLvl returns at once for windows without CpGs (:3287);
        // the others are queued by branch (wave-aggregated appends)
        if (st == 0 && ix.win[(uint32_t)(out >> 32)].z != CPGDUMMY) {
            if (((out >> 16) & 0xFF) == 0)
                want = 1;
            else if (bd.dp_long_list && bd.lens[r] > FAME_DEV_MAX_READ)
                want = 3; // the DP of a long read runs on scratch sized for it
            else
                want = 2;
        }
    }
    const int lane = threadIdx.x & 63;
    const uint64_t below = lanes_below(lane);
    // the counters: one LDS atomic per wave and counter
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const uint64_t b = __ballot(cnt_at == i || tally_at == i);
        if (lane == 0 && b) atomicAdd(&s_cnt[i], (uint32_t)__popcll(b));
    }
    const uint64_t bdir = __ballot(want == 1), bdp = __ballot(want == 2), blong = __ballot(want == 3);
    uint32_t wofs[3] = {0, 0, 0};
    if (lane == 0) {
        if (bdir) wofs[0] = atomicAdd(&s_app[0], (uint32_t)__popcll(bdir));
        if (bdp) wofs[1] = atomicAdd(&s_app[1], (uint32_t)__popcll(bdp));
        if (blong) wofs[2] = atomicAdd(&s_app[2], (uint32_t)__popcll(blong));
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t* head = threadIdx.x == 0 ? bd.direct_n : threadIdx.x == 1 ? bd.dp_n : bd.dp_long_n;
        s_base[threadIdx.x] = s_app[threadIdx.x] ? atomicAdd(head, s_app[threadIdx.x]) : 0u;
    }
    if (threadIdx.x < 5 && s_cnt[threadIdx.x])
        atomicAdd((unsigned long long*)&bd.counters[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
    __syncthreads();
    if (want == 1) bd.direct_list[s_base[0] + lane_u32(wofs[0], 0) + __popcll(bdir & below)] = r;
    if (want == 2) bd.dp_list[s_base[1] + lane_u32(wofs[1], 0) + __popcll(bdp & below)] = r;
    if (want == 3) bd.dp_long_list[s_base[2] + lane_u32(wofs[2], 0) + __popcll(blong & below)] = r;
}

// ---------------------------------------------------------------------------
// k_meth: computeMethLvl (ReadQueue.cpp:3278-3947), split by branch:
// k_meth_direct serves reads matched without errors (direct CpG lookup),
// k_meth_dp the others (banded DP + backtrace + CpG walk).  k_select queues
// both lists.
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

// *(colPat - j + 1) with colPat = chromosome + ref_end; bytes outside the
// chromosome read as 0
__device__ __forceinline__ uint8_t col_at(const uint8_t* cs, uint64_t clen, int64_t ref_end, long j)
{
    const int64_t p = ref_end - j + 1;
    return (p < 0 || (uint64_t)p >= clen) ? 0 : cs[p];
}

// The first id in [lo, hi) whose CpG coordinate pos + RL2 is above thr (at
// or above it when !strict), else hi; the coordinates of a window's CpGs
// ascend with id (cpgTable is in genome order), so this is where the
// reference's forward scans over a window's CpGs (ReadQueue.cpp:3355-3484,
// 3690-3912) stop skipping or stop.  Eight probes per round, loaded
// together, narrow the range about ninefold; `near` probes lo .. lo + 7
// first, for an answer expected close to lo.
__device__ __forceinline__ uint32_t cpg_search(const uint32_t* __restrict__ cpos, uint32_t lo, uint32_t hi,
                                               uint32_t RL2, uint32_t thr, bool strict, bool near)
{
    while (lo < hi) {
        const uint32_t n = hi - lo;
        const bool dense = near || n <= 8;
        uint32_t pr[8];
        bool t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            pr[i] = dense ? lo + (uint32_t)i : lo + ((uint32_t)(i + 1) * n) / 9u;
            const bool ok = pr[i] < hi;
            const uint32_t v = cpos[ok ? pr[i] : lo] + RL2;
            t[i] = !ok || (strict ? v > thr : v >= thr);
        }
        near = false;
        int c = 8; // the first probe at or past the answer
#pragma unroll
        for (int i = 7; i >= 0; --i)
            if (t[i]) c = i;
        uint32_t nlo = lo, nhi = hi;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i == c - 1) nlo = pr[i] + 1;
            if (i == c && pr[i] < hi) nhi = pr[i];
        }
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_meth_direct(DevIndex ix, BatchDev bd)
{
    const uint32_t n = *bd.direct_n;
    const uint32_t RL = ix.readlen, RL2 = RL - 2;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t r = bd.direct_list[t];
        const uint64_t mat = bd.mat[r];
        const bool isFwd = (mat >> 24) & 1;
        const uint32_t metaID = (uint32_t)(mat >> 32);
        const uint32_t offset = (uint32_t)(mat & 0xFFFF);
        const uint4 m = ix.win[metaID];
        const uint32_t L = bd.lens[r];
        const uint8_t* seq = (bd.use[r] == 1 ? bd.seqs : bd.rev) + bd.offs[r];
        const uint32_t metaPos = m.x;
        const uint32_t minPos = (uint32_t)((uint64_t)metaPos + offset - ((uint64_t)L - 1));
        const uint32_t maxPos = metaPos + offset;
        uint32_t* meth = bd.meth;
        // errNum == 0 branch :3355-3484 (the CpGs before the read's span
        // found by cpg_search instead of being stepped over)
        for (uint32_t id = cpg_search(ix.cpg_pos, m.z, m.w + 1, RL2, minPos, false, false); id <= m.w; ++id) {
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
    }
}

// The DP matrix of LevenshtDP<uint16_t, 8> lives in a BandedMatrix<uint16_t, 9>
// (BandedMatrix.h:67-75): row r, column c at flat index
// r*19 + max(9-r, 0) + (c-r+9).  Rows 0..8 then occupy [17r+18, 18r+27] with
// unused slots between them, and cell (8,17) shares its slot with (9,0).
// Here each lane stores its cells by (row, offset c-r+9) and the backtrace
// maps the reference's flat index back to the cell holding that slot last
// (unused slots read as the matrix's fill value 0xFFFF).
// Stored value of cell (r, c) of the banded matrix: the flat index of
// BandedMatrix<uint16_t, 9> (r*19 + max(9-r, 0) + c-r+9, so that the slots of
// neighbouring rows alias at the band edges) mapped to the rows kept in S
// (rows 0..DP_SROWS-1, slot kk of row rr at S[(rr*19 + kk) * stride]).
// Flat positions no row was written to read as 0xFFFF; *bad is set when a
// position beyond the kept rows is asked for.
constexpr int DP_SROWS = FAME_DP_SROWS;

__device__ __forceinline__ uint16_t dp_get(const uint16_t* S, size_t stride, long L, long r, long c, bool* bad)
{
    const long idx = r * 19 + ((9 - r) > 0 ? 9 - r : 0) + (c - r + 9);
    if (idx < 0 || idx >= (L + 1) * 19) return 0xFFFF;
    long rr, kk;
    if (idx >= 171) {
        rr = idx / 19;
        kk = idx % 19;
    } else {
        rr = -1;
#pragma unroll
        for (long q = 0; q <= 8; ++q)
            if (idx >= 17 * q + 18 && idx <= 18 * q + 27) rr = q;
        if (rr < 0) return 0xFFFF;
        kk = (idx - 17 * rr - 18) - rr + 9;
    }
    if (rr >= DP_SROWS) {
        *bad = true;
        return 0xFFFF;
    }
    return S[(size_t)(rr * 19 + kk) * stride];
}

// a reference letter g matches the read letter when it equals t1 or t2
// (cmp_fwd / cmp_rev above): forward t1 = r and t2 = C for r = T; reverse
// the complement, and G for r = T; no target for other reverse letters
__device__ __forceinline__ void dp_targets(uint32_t rc, bool rv, uint32_t& t1, uint32_t& t2)
{
    // without branches: the complements of A, C, T, G sit in byte
    // (letter >> 1) & 3 (0, 1, 2, 3) of 0x43414754, and a letter is one of
    // them exactly when the complement of its complement is itself
    constexpr uint32_t comp = 0x43414754u;
    const uint32_t c1 = (comp >> (8 * ((rc >> 1) & 3u))) & 0xFFu;
    const uint32_t c2 = (comp >> (8 * ((c1 >> 1) & 3u))) & 0xFFu;
    t1 = rv ? (c2 == rc ? c1 : 0x100u) : rc;
    t2 = rc == 'T' ? (rv ? (uint32_t)'G' : (uint32_t)'C') : t1;
}

// backtrackDP's choice at slot k of row rr (LevenshtDP.h:329-544) from the
// stored values up = (rr-1, c), left = (rr, c-1), diagonal = (rr-1, c-1):
// 2 insertion, 3 deletion, else the diagonal step (0 match, 1 mismatch)
__device__ __forceinline__ uint32_t dp_choice(int up, int left, int dg, int mm)
{
    if (up < left) return (up + 1 < dg + mm) ? 2u : (uint32_t)mm;
    return (left + 1 < dg + mm) ? 3u : (uint32_t)mm;
}

// One lane per read.  Rows 0..9 are filled by the general recurrence (the
// first column and the cells left of it lie in the band); from row 10 on
// every band slot but the two edges holds a computed distance, the edges
// hold 0xFFFF - rr (left) and 0xFFFF - rr - 9 (right), so a traceback never
// leaves slots 1..17 there.  Rows 0..DP_SROWS-1 are kept as values for the
// last part of the traceback; for rows >= 11 only the traceback choice of
// slots 1..17 is kept (2 bits each, one word per row), computed during the
// fill from the same stored values the traceback would read.
// dec_scratch holds L + 1 choice words per lane, trace_scratch
// (trace_cap + 31) / 32 trace words (lane-strided): launched for the reads of
// at most FAME_DEV_MAX_READ letters with FAME_DEV_TRACE_CAP, and for the long
// reads' list with scratch and cap sized for the longest of them
__global__ __launch_bounds__(256) void k_meth_dp(DevIndex ix, BatchDev bd, uint16_t* __restrict__ dp_scratch,
                                                 uint64_t* __restrict__ dec_scratch,
                                                 uint64_t* __restrict__ trace_scratch, uint32_t trace_cap)
{
    const uint32_t n = *bd.dp_n;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gsz = gridDim.x * blockDim.x;
    const size_t stride = gsz;
    const uint32_t RL = ix.readlen, RL2 = RL - 2;
    uint16_t* S = dp_scratch + gid;
    uint64_t* D = dec_scratch + gid;
    uint64_t* TR = trace_scratch + gid;
    if (gid == 0) atomicAdd(&bd.stats[5], (unsigned long long)n);
    uint32_t rows = 0; // read letters aligned, for the cell count
    for (uint32_t t = gid; t < n; t += gsz) {
        const uint32_t r = bd.dp_list[t];
        const uint64_t mat = bd.mat[r];
        rows += bd.lens[r];
        const bool isFwd = (mat >> 24) & 1;
        const bool rv = !isFwd;
        const uint32_t metaID = (uint32_t)(mat >> 32);
        const uint32_t offset = (uint32_t)(mat & 0xFFFF);
        const uint4 m = ix.win[metaID];
        const uint32_t L = bd.lens[r];
        const uint8_t* seq = (bd.use[r] == 1 ? bd.seqs : bd.rev) + bd.offs[r];
        const uint32_t metaPos = m.x;
        const uint32_t minPos = (uint32_t)((uint64_t)metaPos + offset - ((uint64_t)L - 1));
        const uint32_t maxPos = metaPos + offset;
        const uint32_t chrom = ix.cpg_chrom[m.z];
        const uint8_t* cs = ix.seq + ix.coff[chrom];
        const uint64_t clen = ix.clen[chrom];
        const int64_t ref_end = (int64_t)metaPos + offset;
        // the CpGs of the read's span (the reference steps over the window's
        // CpGs from its first: minIndex counts those before minPos, maxIndex
        // ends before the first one past maxPos, m.w if none)
        const uint32_t lo_id = cpg_search(ix.cpg_pos, m.z, m.w + 1, RL2, minPos, false, false);
        const int32_t minIndex = (int32_t)lo_id;
        const int32_t maxIndex = (int32_t)cpg_search(ix.cpg_pos, lo_id, m.w + 1, RL2, maxPos, true, true) - 1;
        // fill (runDPFill / runDPFillRev LevenshtDP.h:207-323): one row at a
        // time in registers; values are computed in int and kept as u16.
        // gw holds the reference bytes of columns r-8 .. r+8 (byte i = column
        // r-8+i), shifted by one byte per row.  Row r's letters (reference
        // column r+8 and its read letter) are loaded PD rows ahead into slot
        // r % PD, which the unrolled loops below name statically (no register
        // rotation, so a row waits only for its own letters); the loads are
        // always issued, at positions clamped into the chromosome and the
        // read, and col_at's bounds (0 outside the chromosome) are applied
        // where a letter is used.
        // A letter is loaded as the aligned 32-bit word that holds it (an
        // aligned word never crosses a page, so the word of a byte inside an
        // allocation is readable) and taken out of the word where it is used.
        constexpr int PD = 8;
        const int64_t cmax = (int64_t)clen - 1;
        const uint32_t cs_mis = (uint32_t)(uintptr_t)cs & 3u, seq_mis = (uint32_t)(uintptr_t)seq & 3u;
        auto word_at = [](const uint8_t* a) -> uint32_t {
            return *(const uint32_t*)(a - ((uintptr_t)a & 3u));
        };
        auto gload = [&](long j) -> uint32_t { // the word of column j, clamped
            int64_t p = ref_end - j + 1;
            p = p < 0 ? 0 : (p > cmax ? cmax : p);
            return word_at(cs + p);
        };
        auto gkeep = [&](long j, uint32_t w) -> uint32_t { // column j's letter (col_at)
            const int64_t p = ref_end - j + 1;
            const uint32_t b = (w >> (8 * ((cs_mis + (uint32_t)p) & 3u))) & 0xFFu;
            return (p < 0 || (uint64_t)p >= clen) ? 0u : b;
        };
        auto ridx = [&](long rq) -> long { return rv ? rq - 1 : (long)L - rq; };
        auto rload = [&](long rq) -> uint32_t { // the word of row rq's read letter, clamped
            long i = ridx(rq);
            i = i < 0 ? 0 : (i > (long)L - 1 ? (long)L - 1 : i);
            return word_at(seq + i);
        };
        auto rkeep = [&](long rq, uint32_t w) -> uint32_t {
            return (w >> (8 * ((seq_mis + (uint32_t)ridx(rq)) & 3u))) & 0xFFu;
        };
        uint32_t bq[PD], cq[PD], g0[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) g0[i] = gload(i + 1); // row 0: columns 1..8
#pragma unroll
        for (int rq = 1; rq <= PD; ++rq) {
            bq[rq % PD] = gload(rq + 8);
            cq[rq % PD] = rload(rq);
        }
        int prev[19], cur[19];
        uint32_t gw[5];
#pragma unroll
        for (int k = 0; k < 19; ++k) prev[k] = 0xFFFF;
        prev[9] = 0;
#pragma unroll
        for (int o = 1; o <= 8; ++o) prev[9 + o] = o;
        prev[18] = 0xFFFF - 9; // right band of row 0
#pragma unroll
        for (int k = 9; k < 19; ++k) S[(size_t)k * stride] = (uint16_t)prev[k];
#pragma unroll
        for (int i = 0; i < 5; ++i) gw[i] = 0;
#pragma unroll
        for (int k = 9; k < 17; ++k) gw[k >> 2] |= gkeep(k - 8, g0[k - 9]) << (8 * (k & 3)); // row 0: col k-8
        auto shift_in = [&](uint32_t b) {
#pragma unroll
            for (int i = 0; i < 4; ++i) gw[i] = (gw[i] >> 8) | (gw[i + 1] << 24);
            gw[4] = b;
        };
        auto gbyte = [&](int i) -> uint32_t { return (gw[i >> 2] >> (8 * (i & 3))) & 0xFFu; };
        // The rows are stepped to the longest read of the wave's lanes (Lw,
        // the same for all of them, so that leaving the unrolled loops is a
        // scalar branch and the letters in flight need no copies); a lane
        // past its own L computes rows that it neither stores nor uses.
        uint32_t Lw = 0;
        for (uint64_t m = __ballot(true); m;) {
            Lw = lane_u32(L, (int)__builtin_ctzll(m));
            m &= ~__ballot(L <= Lw);
        }
        // rows 1..9: the general recurrence (the first column and the cells
        // left of it lie in the band)
#pragma unroll
        for (int rr = 1; rr <= 9; ++rr) {
            if (rr > (long)Lw) break;
            const int s = rr % PD;
            const uint32_t gb = gkeep(rr + 8, bq[s]);
            const uint32_t rc = rkeep(rr, cq[s]);
            bq[s] = gload(rr + PD + 8);
            cq[s] = rload(rr + PD);
            shift_in(gb);
            uint32_t t1, t2;
            dp_targets(rc, rv, t1, t2);
            // left band (offset -9): the first column for rows 1..8 is not at
            // offset -9; row 9's slot holds the right band of row 8 (65518)
            cur[0] = rr == 9 ? 0xFFFF - 8 - 9 : 0xFFFF;
#pragma unroll
            for (int k = 1; k < 18; ++k) {
                const int c = rr + k - 9;
                int v;
                if (c < 0) {
                    v = 0xFFFF;
                } else if (c == 0) {
                    v = rr; // first column (rows 1..8)
                } else {
                    const uint32_t gc = gbyte(k - 1);
                    const int mm = (gc != t1 && gc != t2) ? 1 : 0;
                    int a = prev[k + 1] + 1;
                    const int b = cur[k - 1] + 1;
                    const int d = prev[k] + mm;
                    a = b < a ? b : a;
                    a = d < a ? d : a;
                    v = a & 0xFFFF;
                }
                cur[k] = v;
            }
            cur[18] = rr < (long)L ? (int)(0xFFFF - rr - 9) : 0xFFFF;
            const int k0 = rr >= 9 ? 0 : 9 - rr;
            if (rr <= (long)L) {
                uint16_t* row = S + (size_t)(rr * 19) * stride;
#pragma unroll
                for (int k = k0; k < 19; ++k) row[(size_t)k * stride] = (uint16_t)cur[k];
            }
#pragma unroll
            for (int k = 0; k < 19; ++k) prev[k] = cur[k];
        }
        // rows 10..L: every slot 1..17 is a computed cell; row r uses slot
        // r % PD, i.e. (10 + q) % PD in step q of a block of PD rows.  The
        // backtrace starts at the first minimum of row L over offsets -8..8
        // (backtrackDP / backtrackDPRev LevenshtDP.h:329-544), taken here
        // when row L is computed.
        uint16_t minimum = 0xFFFF;
        unsigned long col = 0;
        for (long rb = 10; rb <= (long)Lw; rb += PD) {
#pragma unroll
            for (int q = 0; q < PD; ++q) {
                const long rr = rb + q;
                if (rr > (long)Lw) break;
                const int s = (10 + q) % PD;
                const uint32_t gb = gkeep(rr + 8, bq[s]);
                const uint32_t rc = rkeep(rr, cq[s]);
                bq[s] = gload(rr + PD + 8);
                cq[s] = rload(rr + PD);
                shift_in(gb);
                uint32_t t1, t2;
                dp_targets(rc, rv, t1, t2);
                uint64_t dw = 0;
                cur[0] = (int)(0xFFFF - rr);
#pragma unroll
                for (int k = 1; k < 18; ++k) {
                    const uint32_t gc = gbyte(k - 1);
                    const int mm = (gc != t1 && gc != t2) ? 1 : 0;
                    const int up = prev[k + 1], left = cur[k - 1], dg = prev[k];
                    dw |= (uint64_t)dp_choice(up, left, dg, mm) << (2 * (k - 1));
                    int a = up + 1;
                    const int b = left + 1;
                    const int d = dg + mm;
                    a = b < a ? b : a;
                    a = d < a ? d : a;
                    cur[k] = a & 0xFFFF;
                }
                cur[18] = rr < (long)L ? (int)(0xFFFF - rr - 9) : 0xFFFF;
                if (rr <= (long)L) {
                    if (rr < DP_SROWS) {
                        uint16_t* row = S + (size_t)(rr * 19) * stride;
#pragma unroll
                        for (int k = 0; k < 19; ++k) row[(size_t)k * stride] = (uint16_t)cur[k];
                    }
                    D[(size_t)rr * stride] = dw; // row 10's word is never read
                }
                if (rr == (long)L) {
#pragma unroll
                    for (int k = 1; k < 18; ++k)
                        if (cur[k] < (int)minimum) {
                            col = (unsigned long)((long)L + k - 9);
                            minimum = (uint16_t)cur[k];
                        }
                }
#pragma unroll
                for (int k = 0; k < 19; ++k) prev[k] = cur[k];
            }
        }
        // backtrace (backtrackDP / backtrackDPRev LevenshtDP.h:329-544): the
        // end column is the first minimum of row L over offsets -8..8 (for
        // L >= 10 taken during the fill)
        bool bad = false;
        if (L < 10) {
            for (long o = -DPBAND; o <= DPBAND; ++o) {
                const uint16_t v = dp_get(S, stride, L, (long)L, (long)L + o, &bad);
                if (v < minimum) {
                    col = (unsigned long)((long)L + o);
                    minimum = v;
                }
            }
        }
        // the trace: 2-bit steps, 32 to a word, gathered in a ring of two
        // words (a row of the traceback adds at most 18 steps, so it completes
        // at most one word) and stored outside the step loops
        int tn = 0;
        uint64_t tw[2] = {0, 0};
        auto put = [&](uint32_t tt) {
            if (tn < (int)trace_cap) tw[(tn >> 5) & 1] |= (uint64_t)tt << (2 * (tn & 31));
            ++tn;
        };
        auto flush = [&](int tn0) { // the word that step tn0 was in, once complete
            const int j = tn0 >> 5;
            if ((tn >> 5) > j) {
                if (32 * j + 31 < (int)trace_cap) TR[(size_t)j * stride] = tw[j & 1];
                tw[j & 1] = 0;
            }
        };
        long rr = (long)L;
        if (rr >= 11) {
            // rows >= 11 from the kept choices: row r's word is loaded four
            // rows ahead into slot (L - r) % 4, static in the unrolled loop;
            // the loads are always issued (a lane whose walk has ended, or
            // rows below 11, read words that are never used)
            long k = (long)col - rr + 9;
            uint64_t dq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) dq[u] = D[(size_t)(rr - u) * stride];
            bool act = true;
            while (__ballot(act)) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint64_t w = dq[u];
                    const long rq = rr - 4; // the row this slot takes next
                    if (act) {
                        // the steps of row rr: a deletion stays in the row,
                        // one column left; an insertion moves up in the same
                        // column
                        const int tn0 = tn;
                        uint32_t tt = 0;
                        do {
                            if (k < 1 || k > 17) {
                                bad = true;
                                break;
                            }
                            tt = (uint32_t)(w >> (2 * (k - 1))) & 3u;
                            put(tt);
                            if (tt == 3) --k;
                        } while (tt == 3);
                        flush(tn0);
                        if (bad) {
                            act = false;
                        } else {
                            if (tt == 2) ++k;
                            --rr;
                            act = rr >= 11;
                        }
                    }
                    dq[u] = D[(size_t)(rq > 0 ? rq : 0) * stride];
                }
            }
            col = (unsigned long)(rr + k - 9);
        }
        while (rr > 0 && !bad && tn <= trace_cap) { // rows <= 10 from the kept values
            const uint8_t rc = rv ? seq[rr - 1] : seq[L - rr];
            const uint8_t gc = col_at(cs, clen, ref_end, (long)col);
            const int mm = rv ? cmp_rev(rc, gc) : cmp_fwd(rc, gc);
            const long c = (long)col;
            const uint16_t up = dp_get(S, stride, L, rr - 1, c, &bad);
            const uint16_t left = dp_get(S, stride, L, rr, c - 1, &bad);
            const uint16_t dg = dp_get(S, stride, L, rr - 1, c - 1, &bad);
            const uint32_t tt = dp_choice(up, left, dg, mm);
            if (tt != 3) --rr;
            if (tt != 2) --col;
            put(tt);
            flush(tn - 1);
        }
        while (col > 0 && tn <= trace_cap) {
            put(3);
            flush(tn - 1);
            --col;
        }
        if (rr > 0 || col > 0) bad = true;
        if ((tn & 31) && tn <= trace_cap) TR[(size_t)(tn >> 5) * stride] = tw[(tn >> 5) & 1];
        if (tn > trace_cap || bad) {
            atomicOr(bd.flags, FLAG_DP_BROKEN);
            continue;
        }
        // CpG walk :3690-3912
        uint32_t* meth = bd.meth;
        uint32_t refSeqPos = metaPos + offset;
        int32_t readSeqPos = (int32_t)L - 1;
        int32_t alignPos = tn - 1;
        int32_t wcur = -1;
        uint64_t wv = 0;
        for (int32_t id = maxIndex; id >= minIndex; --id) {
            const uint32_t pos = ix.cpg_pos[id];
            while (alignPos >= 0 && pos + RL2 < refSeqPos) {
                if ((alignPos >> 5) != wcur) {
                    wcur = alignPos >> 5;
                    wv = TR[(size_t)wcur * stride];
                }
                switch ((uint32_t)(wv >> (2 * (alignPos & 31))) & 3u) {
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
    for (int o = 32; o > 0; o >>= 1) rows += __shfl_xor(rows, o, 64);
    if ((threadIdx.x & 63) == 0 && rows) atomicAdd(&bd.stats[9], (unsigned long long)rows);
}

// ---------------------------------------------------------------------------
// exclusive scan of the match counts of overflowed jobs (one thread; the
// list is empty in practice and short when not)
__global__ void k_ovf_offsets(BatchDev bd, uint32_t n_ovf, uint32_t* total)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (n_ovf == FAME_JOBS_ON_DEVICE) {
        if (*bd.flags & FLAG_ABORT) return;
        n_ovf = min(*bd.ovf_n, bd.ovf_list_cap);
    }
    uint32_t o = 0;
    for (uint32_t i = 0; i < n_ovf; ++i) {
        const uint32_t p = bd.ovf_pool[i], j = bd.ovf_job[i];
        bd.job_ovf_off[p][j] = o;
        o += bd.job_nm[p][j];
    }
    *total = o;
    if (bd.ovf_buf_cap && o > bd.ovf_buf_cap) atomicOr(bd.flags, FLAG_OVF_FULL);
}

// batched form: one batch's counters, lists and flags cleared (a kernel, so
// that the whole batch can be replayed from a graph)
__global__ void k_batch_reset(BatchDev bd)
{
    const uint32_t t = threadIdx.x;
    if (t < FAME_NPOOL) bd.pool_n[t] = 0;
    if (t == 0) {
        *bd.heavy_n = 0;
        *bd.ovf_n = 0;
        *bd.direct_n = 0;
        *bd.dp_n = 0;
        bd.flags[0] = 0;
        *bd.max_len = 0;
    }
    if (t < FAME_NSTATS) bd.stats[t] = 0;
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
    const uint32_t m = raw[2 * i];
    const uint32_t meta = m & 0x3FFFFFFFu; // KMER_S::getMetaCpG (metaId < 2^30 is checked on load)
    const uint32_t fwd = (strand_bits[g >> 3] >> (g & 7)) & 1;
    out[g] = (m & 0x80000000u) | (meta << 1) | fwd; // isStartCpG in bit 31
}

// the paired path's entries: the packed kent word and KMER_S::tmask (second
// word of kmerTableSmall) side by side
__global__ void k_pack_tmask(const uint32_t* __restrict__ raw, const uint32_t* __restrict__ kent, uint64_t n,
                             uint64_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint64_t)kent[i] | ((uint64_t)raw[2 * i + 1] << 32);
}

__global__ void k_narrow_tab(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)in[i];
}

// The slot table of the single-end seed stage: per hash key one u64,
//   0                                 empty bucket
//   1 << 62 | e0                      one entry
//   2 << 62 | e1 << 31 | e0           two entries, in table order
//   3 << 62 | len << 32 | lo          three or more: kent[lo, lo + len)
// with e = kent entry without isStart (31 bits, metaId < 2^30).  *bad is set
// for a bucket of 2^30 entries or more (or a tabIndex that decreases): the
// table is then not used.
__global__ void k_build_slots(const uint32_t* __restrict__ tab, const uint32_t* __restrict__ kent, uint64_t htab,
                              uint64_t* __restrict__ slot, uint32_t* __restrict__ bad)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t key = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; key < htab; key += step) {
        const uint32_t lo = tab[key], hi = tab[key + 1];
        const uint32_t len = hi - lo;
        uint64_t v;
        if (len == 0) {
            v = 0;
        } else if (len == 1) {
            v = (1ull << 62) | (kent[lo] & 0x7FFFFFFFu);
        } else if (len == 2) {
            v = (2ull << 62) | ((uint64_t)(kent[lo + 1] & 0x7FFFFFFFu) << 31) | (kent[lo] & 0x7FFFFFFFu);
        } else {
            if (len >= (1u << 30)) atomicOr(bad, 1u);
            v = (3ull << 62) | ((uint64_t)(len & 0x3FFFFFFFu) << 32) | lo;
        }
        slot[key] = v;
    }
}

// The genome-ordered copy of the slot table (DevIndex::gkey / gslot / kpos).
// Position p < ng is the forward k-mer at p, letters seq[p .. p+31] reduced
// C->T as a read's row is (A 0, G 1, C/T 2); p = ng + r is the k-mer at r of
// the reverse complement, letters comp(seq[ng-1-r-i]) reduced the same way.
// Its key is the spaced-seed hash of those codes taken % htab, as seed_key
// hashes a read's row; any other letter, or a k-mer running past the end,
// gives ~0u.  kpos[key] keeps one position of the key (whichever store lands
// last: it only steers which lines the seed stage reads).
__device__ __forceinline__ uint32_t gorder_code(uint8_t c, bool rv)
{
    switch (c) {
    case 'A': case 'a': return rv ? 2u : 0u;
    case 'C': case 'c': return rv ? 1u : 2u;
    case 'G': case 'g': return rv ? 2u : 1u;
    case 'T': case 't': return rv ? 0u : 2u;
    default: return 3u;
    }
}

__global__ void k_gorder_keys(const uint8_t* __restrict__ seq, uint64_t ng, uint64_t htab,
                              uint32_t* __restrict__ gkey, uint32_t* __restrict__ kpos)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const bool pow2 = (htab & (htab - 1)) == 0;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < 2 * ng; p += step) {
        const bool rv = p >= ng;
        const uint64_t r = rv ? p - ng : p;
        uint32_t key = ~0u;
        if (r + KMERLEN <= ng) {
            uint64_t h = 0;
            bool ok = true;
#pragma unroll
            for (uint32_t i = 0; i < KMERLEN; ++i) {
                const uint8_t c = rv ? seq[ng - 1 - (r + i)] : seq[r + i];
                const uint32_t code = gorder_code(c, rv);
                ok = ok && code != 3u;
                if ((SEEDBITS >> (KMERLEN - 1 - i)) & 1) h ^= rol64(red_seed(code), (int)(KMERLEN - 1 - i));
            }
            if (ok) {
                key = pow2 ? (uint32_t)(h & (htab - 1)) : (uint32_t)(h % htab);
                kpos[key] = (uint32_t)p;
            }
        }
        gkey[p] = key;
    }
}

__global__ void k_gorder_slots(const uint32_t* __restrict__ gkey, uint64_t gn, const uint64_t* __restrict__ slot,
                               uint64_t* __restrict__ gslot)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < gn; p += step) {
        const uint32_t k = gkey[p];
        gslot[p] = k != ~0u ? slot[k] : 0ull;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}

// the packed text of the verify stage: word w holds letters 16w .. 16w+15 of
// the chromosome, two bits each, and the mask of those that are not A, C, G
// or T; letters past the end are 0 and unmasked
__global__ void k_pack_seq(const uint8_t* __restrict__ seq, uint64_t len, uint64_t nwords, uint64_t* __restrict__ out)
{
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint32_t x = 0, nm = 0;
    for (uint32_t j = 0; j < 16; ++j) {
        const uint64_t p = 16 * w + j;
        if (p >= len) break;
        const uint8_t b = seq[p];
        uint32_t c;
        switch (b) {
        case 'A': c = 0; break;
        case 'C': c = 1; break;
        case 'G': c = 2; break;
        case 'T': c = 3; break;
        default: c = 0; nm |= 1u << j; break;
        }
        x |= c << (2 * j);
    }
    out[w] = ((uint64_t)nm << 32) | x;
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
    const uint32_t blocks = (n + 255) / 256; // 4 waves per block, 64 reads per wave
    hipLaunchKernelGGL(k_prep, dim3(blocks < 16384 ? blocks : 16384), dim3(256), 0, st, bd.seqs, bd.offs,
                       bd.lens, n, readlen, bd.rev, bd.pre, bd.codes, bd.ori_mask, bd.flags, bd.max_len,
                       0xFFFFFFFFu, 0u, 0, bd.n_dev, bd.long_n, bd.long_list);
}

// paired batches: 2n mate reads, mate 2 from nsplit on; pre[] = 1 short | 2 N
// or unknown letter
void launch_prep_pairs(const BatchDev& bd, uint32_t n_reads, uint32_t nsplit, uint32_t minlen, hipStream_t st)
{
    if (!n_reads) return;
    const uint32_t blocks = (n_reads + 255) / 256;
    hipLaunchKernelGGL(k_prep, dim3(blocks < 16384 ? blocks : 16384), dim3(256), 0, st, bd.seqs, bd.offs,
                       bd.lens, n_reads, 0u, bd.rev, bd.pre, bd.codes, bd.ori_mask, bd.flags, bd.max_len,
                       nsplit, minlen, 1, (const uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr);
}

void launch_pack_tmask(const uint32_t* raw, const uint32_t* kent, uint64_t n, uint64_t* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_pack_tmask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, raw, kent, n, out);
}

void launch_seed_light(const DevIndex& ix, const BatchDev& bd, uint32_t n_ro, hipStream_t st, uint32_t lmax_hint,
                       uint32_t max_waves)
{
    if (!n_ro) return;
    const uint32_t wmax = max_waves && max_waves < FAME_SEED_WAVES ? max_waves : FAME_SEED_WAVES;
    const uint32_t waves = n_ro < wmax ? n_ro : wmax;
    const uint32_t g = (waves + FAME_SEED_WPB - 1) / FAME_SEED_WPB;
    // the variant that does not fit the batch returns at once; with a hint of
    // the longest read it gets a small grid (still complete if the hint is
    // wrong: the kernels are persistent)
    const uint32_t S = SEED_SHORT, M = SEED_GO_MID, X = 0xFFFFFFFFu;
    const dim3 b(64 * FAME_SEED_WPB);
    // with a hint of the longest read the variants that do not fit it get
    // small grids
    auto grid = [&](uint32_t lo, uint32_t hi) { return dim3(lmax_hint && (lmax_hint <= lo || lmax_hint > hi) ? 16 : g); };
    if (ix.tab_is64) {
        hipLaunchKernelGGL((k_seed_light<uint64_t, SEED_SHORT, false>), grid(0, S), b, 0, st, ix, bd, n_ro, 0u, S);
        hipLaunchKernelGGL((k_seed_light<uint64_t, 256, false>), grid(S, X), b, 0, st, ix, bd, n_ro, S, X);
    } else if (ix.slot && ix.gkey) {
        // the genome-ordered pipeline for reads of up to 159 letters (two
        // k-mers per lane; four would not fit its stage registers without
        // spills)
        hipLaunchKernelGGL((k_seed_light<uint32_t, SEED_SHORT, true>), grid(0, S), b, 0, st, ix, bd, n_ro, 0u, S);
        hipLaunchKernelGGL((k_seed_light<uint32_t, 256, true, SEED_GO_MID>), grid(S, M), b, 0, st, ix, bd, n_ro, S, M);
        hipLaunchKernelGGL((k_seed_light<uint32_t, 256, false>), grid(M, X), b, 0, st, ix, bd, n_ro, M, X);
    } else {
        hipLaunchKernelGGL((k_seed_light<uint32_t, SEED_SHORT, false>), grid(0, S), b, 0, st, ix, bd, n_ro, 0u, S);
        hipLaunchKernelGGL((k_seed_light<uint32_t, 256, false>), grid(S, X), b, 0, st, ix, bd, n_ro, S, X);
    }
}

// shortest and longest read, one atomic pair per wave
__global__ void k_len_range(const uint32_t* __restrict__ lens, uint32_t n, uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lo = i < n ? lens[i] : 0xFFFFFFFFu, hi = i < n ? lens[i] : 0u;
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, 64));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && lo <= hi) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

void launch_len_range(const uint32_t* lens, uint32_t n, uint32_t* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_len_range, dim3((n + 255) / 256), dim3(256), 0, st, lens, n, out);
}

void launch_seed_heavy(const DevIndex& ix, const BatchDev& bd, const HeavyScratch& hs, uint32_t blocks,
                       hipStream_t st)
{
    if (!hs.count || !blocks) return;
    if (ix.tab_is64)
        hipLaunchKernelGGL(k_seed_heavy<uint64_t>, dim3(blocks), dim3(64), 0, st, ix, bd, hs);
    else
        hipLaunchKernelGGL(k_seed_heavy<uint32_t>, dim3(blocks), dim3(64), 0, st, ix, bd, hs);
}

// blocks of k_seed_heavy that are resident at once on the current device
uint32_t seed_heavy_blocks(bool tab64)
{
    int nb = 0, dev = 0;
    hipDeviceProp_t p;
    const void* k = tab64 ? (const void*)k_seed_heavy<uint64_t> : (const void*)k_seed_heavy<uint32_t>;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 64, 0) != hipSuccess || nb <= 0)
        return 4096;
    return (uint32_t)nb * (uint32_t)p.multiProcessorCount;
}

// the LDS tier: plain form over heavy-list positions [first, first + count)
// (overflow list ovl), device form (count = FAME_JOBS_ON_DEVICE) over the
// device list; only for the 32-bit tabIndex
void launch_seed_mid(const DevIndex& ix, const BatchDev& bd, uint32_t first, uint32_t count, uint32_t* ovl,
                     uint32_t blocks, hipStream_t st)
{
    if (!count || !blocks || ix.tab_is64) return;
    // FAME_MID_SLOTS (tests): a smaller counting table, so that reads overflow
    // into the heavy tier
    uint32_t hsize = MID_H;
    if (const char* e = getenv("FAME_MID_SLOTS")) {
        const uint32_t v = (uint32_t)atoi(e);
        if (v >= 64 && v <= MID_H && !(v & (v - 1))) hsize = v;
    }
    hipLaunchKernelGGL(k_seed_mid, dim3(blocks), dim3(64), 0, st, ix, bd, first, count, ovl, hsize);
}

uint32_t seed_mid_max_e() { return MID_E; }

// blocks of k_seed_mid resident at once on the current device
uint32_t seed_mid_blocks()
{
    int nb = 0, dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_seed_mid, 64, 0) != hipSuccess || nb <= 0)
        return 1024;
    return (uint32_t)nb * (uint32_t)p.multiProcessorCount;
}

void launch_seed_heavy_dev(const DevIndex& ix, const BatchDev& bd, const HeavyScratch& hs, uint32_t blocks,
                           hipStream_t st)
{
    HeavyScratch h = hs;
    h.first = 0;
    h.count = FAME_JOBS_ON_DEVICE;
    if (ix.tab_is64)
        hipLaunchKernelGGL(k_seed_heavy<uint64_t>, dim3(blocks), dim3(64), 0, st, ix, bd, h);
    else
        hipLaunchKernelGGL(k_seed_heavy<uint32_t>, dim3(blocks), dim3(64), 0, st, ix, bd, h);
}

void launch_verify(const DevIndex& ix, const BatchDev& bd, uint32_t pool, uint32_t njobs,
                   const uint32_t* only, hipStream_t st, uint32_t blocks)
{
    if (!njobs) return;
    const dim3 g(njobs == FAME_JOBS_ON_DEVICE ? blocks : (njobs + 255) / 256), b(256);
    switch (pool) {
    case 0: hipLaunchKernelGGL(k_verify<1>, g, b, 0, st, ix, bd, 0u, only, only ? njobs : 0u); break;
    case 1: hipLaunchKernelGGL(k_verify<2>, g, b, 0, st, ix, bd, 1u, only, only ? njobs : 0u); break;
    case 2: hipLaunchKernelGGL(k_verify<3>, g, b, 0, st, ix, bd, 2u, only, only ? njobs : 0u); break;
    case 3: hipLaunchKernelGGL(k_verify<4>, g, b, 0, st, ix, bd, 3u, only, only ? njobs : 0u); break;
    default: {
        // one wave per job
        const uint32_t nb = njobs == FAME_JOBS_ON_DEVICE ? (blocks ? blocks : 1024u) : (njobs < 65536u ? njobs : 65536u);
        hipLaunchKernelGGL(k_verify_long, dim3(nb), dim3(64), 0, st, ix, bd, only, only ? njobs : 0u);
        break;
    }
    }
}

void launch_select(const DevIndex& ix, const BatchDev& bd, uint32_t n, int get_stranded, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_select, dim3((n + 255) / 256), dim3(256), 0, st, ix, bd, n, get_stranded);
}

// lanes of k_meth_dp that are resident at once on the current device
uint32_t meth_dp_lanes()
{
    int nb = 0, dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_meth_dp, 256, 0) != hipSuccess || nb <= 0)
        return 65536;
    return (uint32_t)nb * (uint32_t)p.multiProcessorCount * 256u;
}

void launch_meth(const DevIndex& ix, const BatchDev& bd, uint32_t n, uint16_t* dp, uint64_t* dec, uint64_t* trace,
                 uint32_t grid_threads, hipStream_t st, hipEvent_t between)
{
    if (!n) {
        if (between) (void)hipEventRecord(between, st);
        return;
    }
    const uint32_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_meth_direct, dim3(g < 8192 ? g : 8192), dim3(256), 0, st, ix, bd);
    if (between) (void)hipEventRecord(between, st);
    hipLaunchKernelGGL(k_meth_dp, dim3(grid_threads / 256), dim3(256), 0, st, ix, bd, dp, dec, trace,
                       (uint32_t)FAME_DEV_TRACE_CAP);
}

void launch_meth_dp_long(const DevIndex& ix, const BatchDev& bd, uint16_t* dp, uint64_t* dec, uint64_t* trace,
                         uint32_t lanes, uint32_t trace_cap, hipStream_t st)
{
    BatchDev b = bd;
    b.dp_list = bd.dp_long_list;
    b.dp_n = bd.dp_long_n;
    hipLaunchKernelGGL(k_meth_dp, dim3(lanes / 256), dim3(256), 0, st, ix, b, dp, dec, trace, trace_cap);
}

void launch_long_measure(const DevIndex& ix, const BatchDev& bd, uint32_t items, hipStream_t st)
{
    if (!items) return;
    const dim3 g(items < 4096 ? items : 4096), b(64);
    if (ix.tab_is64)
        hipLaunchKernelGGL(k_long_measure<uint64_t>, g, b, 0, st, ix, bd);
    else
        hipLaunchKernelGGL(k_long_measure<uint32_t>, g, b, 0, st, ix, bd);
}

void launch_ovf_offsets(const BatchDev& bd, uint32_t n_ovf, uint32_t* total, hipStream_t st)
{
    hipLaunchKernelGGL(k_ovf_offsets, dim3(1), dim3(1), 0, st, bd, n_ovf, total);
}

void launch_batch_reset(const BatchDev& bd, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, bd);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, p, v);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st)
{
    if (!n) return;
    const uint32_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero_u32, dim3(g < 1024 ? g : 1024), dim3(256), 0, st, p, n);
}

// furthest read end, one atomic per wave
__global__ void k_max_end(const uint64_t* __restrict__ offs, const uint32_t* __restrict__ lens, uint32_t n,
                          unsigned long long* out)
{
    unsigned long long v = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long e = offs[i] + lens[i];
        v = e > v ? e : v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(out, v);
}

void launch_max_end(const uint64_t* offs, const uint32_t* lens, uint32_t n, unsigned long long* out, hipStream_t st)
{
    if (!n) return;
    const uint32_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_max_end, dim3(g < 1024 ? g : 1024), dim3(256), 0, st, offs, lens, n, out);
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

void launch_build_slots(const uint32_t* tab, const uint32_t* kent, uint64_t htab, uint64_t* slot, uint32_t* bad,
                        hipStream_t st)
{
    if (!htab) return;
    hipLaunchKernelGGL(k_build_slots, dim3(16384), dim3(256), 0, st, tab, kent, htab, slot, bad);
}

void launch_gorder(const uint8_t* seq, uint64_t ng, uint64_t htab, const uint64_t* slot, uint32_t* gkey,
                   uint64_t* gslot, uint32_t* kpos, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(16384), dim3(256), 0, st, kpos, htab, ~0u);
    hipLaunchKernelGGL(k_gorder_keys, dim3(65536), dim3(256), 0, st, seq, ng, htab, gkey, kpos);
    hipLaunchKernelGGL(k_gorder_slots, dim3(16384), dim3(256), 0, st, gkey, 2 * ng, slot, gslot);
}

void launch_pack_seq(const uint8_t* seq, uint64_t len, uint64_t nwords, uint64_t* out, hipStream_t st)
{
    if (!nwords) return;
    hipLaunchKernelGGL(k_pack_seq, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, st, seq, len, nwords, out);
}

void launch_meth_u16(const uint32_t* in, uint64_t n, uint16_t* out, hipStream_t st)
{
    if (!n) return;
    hipLaunchKernelGGL(k_meth_u16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out);
}

} // namespace fame
