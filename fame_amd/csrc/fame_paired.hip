// HIP kernels of the paired-end matching path (gfx950):
// ReadQueue::matchPairedReads ReadQueue.cpp:935-1655.
//
// A batch of n pairs is 2n reads of a BatchDev (mate 1 of pair i is read i,
// mate 2 is read n + i; k_prep in paired mode validates them and writes the
// reverse complements and reduced-letter rows).  Combination c of pair i is
// item t = 2i + c: c = 0 seeds (mate 1, revcomp mate 2), c = 1 (revcomp
// mate 1, mate 2) (ReadQueue.cpp:1092-1300).  Per batch:
//   k_pe_seed    one wavefront per item: getSeedRefsFirstRead (:2468-2794)
//                on the first read (tmask / filteredKmers filter, u8 counts,
//                the count-sized map's iteration order), getSeedRefsSecondRead
//                (:2801-3164) on the second read against the first read's
//                windows (neighbour rule, inserts, flag2, hasCpG), then the
//                iteration order of the two 20-bucket maps; queues the first
//                read's verify jobs
//   k_verify     (fame_device.hip) the first reads over their windows
//   k_pe_walk1   one lane per item: saQuerySeedSetRefFirst (:2167-2243) -
//                std::sort by c0, the walk with the KMERDIST cut, flag3 on
//                matched windows and their neighbours - then the c1 order of
//                saQuerySeedSetRefSecond (:2249-2323) and the second reads'
//                verify jobs
//   k_verify     the second reads
//   k_pe_final   one lane per pair: the second walk, every (mate 1, mate 2)
//                match combination through extractPairedMatch (:3255-3275),
//                the best pair, counters, and the methylation queues
//   k_meth_*     (fame_device.hip) computeMethLvl for both mates
//
// std::sort is libstdc++'s introsort; the order it leaves among equal counts
// decides which windows a walk reaches, so pe_sort repeats its steps.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fame_internal.h"
#include "fame_dev_common.h"

namespace fame {

constexpr uint32_t KMERDIST = 10;      // CONST.h:88
constexpr uint32_t PF2 = 1u << 16;     // map value tuple<u8 c0, u8 c1, bool, bool>: the two flags
constexpr uint32_t PF3 = 1u << 17;
constexpr uint32_t PE_NONE = 0xFFFFFFFFu;
constexpr int PE_KPL = (FAME_DEV_MAX_READ - (int)KMERLEN + 1 + 63) / 64; // k-mers per lane

__device__ __forceinline__ uint32_t pe_c0(uint32_t v) { return v & 0xFFu; }
__device__ __forceinline__ uint32_t pe_c1(uint32_t v) { return (v >> 8) & 0xFFu; }

// per-block LDS that every variant keeps there: the read's reduced letters,
// its C letters and 2-bit codes, the map sizes
struct PeLds {
    uint32_t nmap[2];
    uint32_t cbits[9];  // bit p: letter p of the read is C
    uint32_t kbits[17]; // 2-bit letter codes (C, T 3; G 2; others 0), 16 per word
    alignas(8) uint8_t code[256];
};

// working arrays of the seed stage: in LDS for reads of at most PE_LE index
// hits (k_pe_seed light), in per-block global scratch otherwise
template <typename TabT>
struct PeWork {
    TabT* klo;
    uint32_t* klen;
    uint32_t* koff;
    uint32_t* kcm;  // C mask of each k-mer, first letter in the top bit
    uint8_t* kfl;   // k-mer skipped by the filteredKmers lookup
    uint32_t* ent;  // [cap_e] kent of each bucket entry
    uint8_t* ekm;   // [cap_e] its k-mer
    uint32_t* eaux; // [cap_e] bit 0 kept; first / count, or lookup slot
    uint64_t* pf;   // [256] prefix XOR of the read's hash terms, in LDS (dead after the
                    // hashing); NULL: the closed form from the shared table
    uint32_t cap_e;
    uint32_t* hkey; // counting table (first read)
    uint32_t* hfirst;
    uint32_t* hcnt;
    uint32_t hsize;
    uint32_t* ins_key; // [2][cap_e]
    uint16_t* ins_cnt; // [2][cap_e]
    uint32_t* tb;
    uint32_t tb_cap;
    uint32_t* tmp;  // [cap_t]
    uint32_t* over; // [cap_t]
    uint32_t cap_t;
    // the item's two strand maps, keys in insertion order
    uint32_t* mkey; // [2][mcap]
    uint32_t* mval; // [2][mcap]
    uint32_t mcap;
    // lookup over every window the item touches: key << 1 | strand, the
    // insertion time (0: first-read window, j + 1: inserted at second-read
    // entry j, PE_NEVER), the second read's count, the map index
    uint32_t* hk;
    uint32_t* ht;
    uint32_t* hc;
    uint32_t* hi;
    uint32_t mhsz; // power of two
};

constexpr uint32_t PE_NEVER = 0xFFFFFFFFu;

struct MapChunk {
    uint32_t base, left; // the wave's open chunk of the map pool
};
constexpr uint32_t PE_LE = 256; // index hits per read, light variant
constexpr uint32_t PE_LM = 64;  // windows per strand map, light variant
constexpr uint32_t PE_LH = 512; // lookup slots, light variant
constexpr uint32_t PE_LTB = 640;

template <typename TabT>
__device__ __forceinline__ uint32_t pe_slot(const PeWork<TabT>& w, uint32_t k2)
{
    return (k2 * 2654435761u) >> (__clz(w.mhsz) + 1);
}

// lookup slot of (key, strand), or PE_NONE
template <typename TabT>
__device__ __forceinline__ uint32_t pe_lookup(const PeWork<TabT>& w, uint32_t key, uint32_t s)
{
    if (key >= 0x40000000u) return PE_NONE; // metaId - u below 0 wraps
    const uint32_t k2 = (key << 1) | s;
    uint32_t h = pe_slot(w, k2);
    for (uint32_t probe = 0; probe < w.mhsz; ++probe) {
        const uint32_t x = w.hk[h];
        if (x == k2) return h;
        if (x == EMPTY) return PE_NONE;
        h = (h + 1) & (w.mhsz - 1);
    }
    return PE_NONE;
}

// index of (key, strand) in its strand map, or PE_NONE
template <typename TabT>
__device__ __forceinline__ uint32_t pe_find(const PeWork<TabT>& w, uint32_t key, uint32_t s)
{
    const uint32_t h = pe_lookup(w, key, s);
    return h == PE_NONE ? PE_NONE : w.hi[h];
}

// the slot of (key, strand), created (time PE_NEVER, not in a map) when
// absent; any lane, concurrently
template <typename TabT>
__device__ __forceinline__ uint32_t pe_claim(const PeWork<TabT>& w, uint32_t key, uint32_t s)
{
    const uint32_t k2 = (key << 1) | s;
    uint32_t h = pe_slot(w, k2);
    for (;;) {
        const uint32_t old = atomicCAS(&w.hk[h], EMPTY, k2);
        if (old == EMPTY || old == k2) return h;
        h = (h + 1) & (w.mhsz - 1);
    }
}

__device__ __forceinline__ bool pe_filtered(const DevIndex& ix, uint32_t code)
{
    uint32_t lo = 0, hi = ix.nfilt32;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = ix.filt32[mid];
        if (v == code) return true;
        if (v < code) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

// The bucket entries of one (read, orientation) with the paired filters
// (ReadQueue.cpp:2556-2600 first read, :2646-2700 second read): entries of
// k-mers found in filteredKmers (k-mer 0 excepted) and entries whose tmask
// meets a read C under SEEDBITS are dropped; of consecutive entries of one
// bucket that remain, those repeating the previous one's (metaId, strand)
// (and isStart, for the first read) are dropped.  eaux[j] bit 0 = kept.
// Returns 1 when the entries do not fit the scratch (E_out set).
template <typename TabT>
__device__ int pe_entries(const DevIndex& ix, const BatchDev& bd, uint32_t ro, const SeedShared& sh, PeLds& m,
                          const PeWork<TabT>& w, bool first, uint32_t& E_out, uint32_t& bc_out, uint32_t& L_out)
{
    const int lane = threadIdx.x & 63;
    const uint32_t cwd = ((const uint32_t*)(bd.codes + (size_t)ro * 256))[lane];
    const uint32_t L = lane_u32(cwd, 63) >> 24;
    L_out = L;
    ((uint32_t*)m.code)[lane] = cwd;
    if (lane < 9) m.cbits[lane] = 0;
    if (lane < 17) m.kbits[lane] = 0;
    wave_sync();
    {
        const uint32_t r = ro >> 1;
        const uint8_t* pat = ((ro & 1) ? bd.rev : bd.seqs) + bd.offs[r];
        for (uint32_t p = lane; p < L; p += 64) {
            const uint8_t c = pat[p];
            if (c == 'C') atomicOr(&m.cbits[p >> 5], 1u << (p & 31));
            const uint32_t k2 = (c == 'C' || c == 'T') ? 3u : (c == 'G' ? 2u : 0u);
            if (k2) atomicOr(&m.kbits[p >> 4], k2 << (2 * (p & 15)));
        }
    }
    wave_sync();
    const uint32_t K = L - KMERLEN + 1;
    const TabT* tab = (const TabT*)ix.tab;
    // NTPS64 in prefix form (SeedEnds): the read's terms rol(seed(c_t), -t),
    // four letters per lane, XOR-scanned over the wave into w.pf
    if (w.pf) {
        uint64_t x[4], acc = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t t = 4u * (uint32_t)lane + (uint32_t)j;
            acc ^= t < L ? rol64(red_seed((cwd >> (8 * j)) & 3u), (int)((64u - (t & 63u)) & 63u)) : 0ull;
            x[j] = acc;
        }
        const uint64_t ex = wave_ixor_u64(acc) ^ acc; // the letters below this lane's
        if (lane == 0) w.pf[0] = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t t = 4u * (uint32_t)lane + (uint32_t)j;
            if (t < L) w.pf[t + 1] = ex ^ x[j];
        }
        wave_sync();
    }
    constexpr SeedEnds se = seed_ends();
    TabT lo[PE_KPL], hi[PE_KPL];
#pragma unroll
    for (int q = 0; q < PE_KPL; ++q) {
        const uint32_t k = lane + 64 * q;
        lo[q] = hi[q] = 0;
        if (k < K) {
            uint64_t h = 0;
            if (w.pf) {
                uint64_t X = 0;
#pragma unroll
                for (int j = 0; j < se.n; ++j) X ^= w.pf[k + se.e[j]];
                h = rol64(X, (int)((KMERLEN - 1 + k) & 63));
            } else {
#pragma unroll
                for (uint32_t i = 0; i < KMERLEN; ++i)
                    if ((SEEDBITS >> (KMERLEN - 1 - i)) & 1) h ^= sh.rot[i][m.code[k + i]];
            }
            const uint64_t key = ix.hpow2 ? (h & ix.hmask) : (h % ix.htab);
            lo[q] = tab[key];
            hi[q] = tab[key + 1];
        }
    }
    wave_sync(); // the prefix array is dead: ent and eaux are written below
    uint32_t bsum = 0;
#pragma unroll
    for (int q = 0; q < PE_KPL; ++q) {
        const uint32_t k = lane + 64 * q;
        if (k < K) {
            w.klo[k] = lo[q];
            const uint64_t len = (uint64_t)(hi[q] - lo[q]);
            w.klen[k] = len > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)len;
            bsum += (uint32_t)len;
            // cMask: the k-mer's C letters, its first letter in the top bit
            const uint32_t wi = k >> 5, sb = k & 31;
            uint32_t win = m.cbits[wi] >> sb;
            if (sb) win |= m.cbits[wi + 1] << (32 - sb);
            w.kcm[k] = __brev(win);
            // kSeq: the u32 two-bit code of the k-mer's last 16 letters, the
            // earliest in the top bits (:2520-2530)
            bool f = false;
            if (k > 0 && ix.nfilt32) {
                const uint32_t s = 2 * (k + 16), wj = s >> 5, sc = s & 31;
                uint32_t x = m.kbits[wj] >> sc;
                if (sc) x |= m.kbits[wj + 1] << (32 - sc);
                uint32_t y = __brev(x);
                y = ((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1);
                f = pe_filtered(ix, y);
            }
            w.kfl[k] = f ? 1 : 0;
        }
    }
    bsum = wave_sum_u32(bsum);
    bc_out = bsum / 5;
    uint64_t carry = 0;
    bool big = false;
    for (uint32_t k0 = 0; k0 < K; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t v = k < K ? w.klen[k] : 0;
        if (v > 0xFFFFFFu) {
            v = 0xFFFFFFu;
            big = true;
        }
        const uint32_t inc = wave_iscan_u32(v, lane);
        const uint32_t tot = lane_u32(inc, 63);
        if (k < K) w.koff[k] = (uint32_t)(carry + inc - v);
        carry += tot;
        if (carry > 0x7FFFFFFFull) big = true;
    }
    big = __ballot(big) != 0;
    E_out = big ? 0xFFFFFFFFu : (uint32_t)carry;
    if (big || carry > w.cap_e) return 1;
    const uint32_t E = (uint32_t)carry;
    wave_sync();
    // entry j's k-mer: up to 256 entries each k-mer writes its index over its
    // entries (one LDS read per entry afterwards), beyond a binary search
    // over the offsets
    const bool scatter = E <= 256;
    if (scatter) {
        for (uint32_t k = lane; k < K; k += 64) {
            const uint32_t o = w.koff[k], n = w.klen[k];
            for (uint32_t i = 0; i < n; ++i) w.ekm[o + i] = (uint8_t)k;
        }
        wave_sync();
    }
    for (uint32_t j = lane; j < E; j += 64) {
        uint32_t k;
        if (scatter) {
            k = w.ekm[j];
        } else {
            uint32_t a = 0, b = K;
            while (b - a > 1) {
                const uint32_t mid = (a + b) >> 1;
                if (w.koff[mid] <= j) a = mid;
                else b = mid;
            }
            k = a;
            while (w.klen[k] == 0 || w.koff[k] + w.klen[k] <= j) ++k;
        }
        const uint64_t g = (uint64_t)w.klo[k] + (j - w.koff[k]);
        const uint64_t kp = ix.kpe[g]; // kent | tmask << 32
        const uint32_t e = (uint32_t)kp;
        const bool filt = w.kfl[k] || (w.kcm[k] & SEEDBITS & (uint32_t)(kp >> 32)) != 0;
        w.ent[j] = first ? e : (e & 0x7FFFFFFFu);
        w.ekm[j] = (uint8_t)k;
        w.eaux[j] = filt ? 0u : 1u;
    }
    wave_sync();
    // previous unfiltered entry (exclusive max-scan); dropped when in the
    // same bucket with the same packed (isStart,) metaId, strand
    int pc = -1;
    for (uint32_t j0 = 0; j0 < E; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool un = j < E && (w.eaux[j] & 1);
        const int inc = wave_imaxscan_i32(un ? (int)j : -1);
        int ex = wave_shr1_i32(inc);
        ex = max(ex, pc);
        pc = max(pc, (int)lane_u32((uint32_t)inc, 63));
        if (un && ex >= 0) {
            const uint32_t k = w.ekm[j];
            if ((uint32_t)ex >= w.koff[k] && w.ent[ex] == w.ent[j]) w.eaux[j] = 0;
        }
    }
    wave_sync();
    return 0;
}

// getSeedRefsFirstRead's counting (:2556-2640) and the copy of the windows
// with c0 >= QTHRESH, in the iteration order of the map sized by the bucket
// count, into the item's maps (:2760-2790).  Returns 1 on a capacity
// overflow (retry with larger scratch).
template <typename TabT>
__device__ int pe_first(const PeWork<TabT>& w, PeLds& m, uint32_t E, uint32_t L, uint32_t bucketCount)
{
    const int lane = threadIdx.x & 63;
    {
        // a table of at least twice the entries (a power of two, >= 64)
        uint32_t hs = 64;
        while (hs < 2 * E && hs < w.hsize) hs <<= 1;
        const uint32_t hm = hs - 1;
        const int hshift = __clz(hs) + 1;
        for (uint32_t i = lane; i < hs; i += 64) {
            w.hkey[i] = EMPTY;
            w.hfirst[i] = 0xFFFFFFFFu;
            w.hcnt[i] = 0;
        }
        wave_sync();
        for (uint32_t j = lane; j < E; j += 64) {
            if (!(w.eaux[j] & 1)) continue;
            const uint32_t key = w.ent[j] & 0x7FFFFFFFu; // metaId << 1 | strand
            uint32_t slot = (key * 2654435761u) >> hshift;
            for (;;) {
                const uint32_t old = atomicCAS(&w.hkey[slot], EMPTY, key);
                if (old == EMPTY || old == key) break;
                slot = (slot + 1) & hm;
            }
            atomicMin(&w.hfirst[slot], j);
            atomicAdd(&w.hcnt[slot], 1u);
            w.eaux[j] = 1u | (slot << 8);
        }
        wave_sync();
        for (uint32_t j = lane; j < E; j += 64) {
            const uint32_t a = w.eaux[j];
            if (!(a & 1)) continue;
            const uint32_t slot = a >> 8;
            const bool fst = w.hfirst[slot] == j;
            // the tuple's c0 is a u8: counts wrap modulo 256
            w.eaux[j] = 1u | (fst ? 2u : 0u) | ((w.hcnt[slot] & 0xFFu) << 16);
        }
    }
    wave_sync();
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    uint32_t nins[2] = {0, 0}, nq[2] = {0, 0};
    const uint64_t below = lanes_below(lane);
    for (uint32_t j0 = 0; j0 < E; j0 += 64) {
        const uint32_t j = j0 + lane;
        bool t0 = false, t1 = false;
        uint32_t key = 0, cnt = 0;
        if (j < E) {
            const uint32_t a = w.eaux[j];
            const uint32_t km = w.ekm[j];
            const bool ins = km == 0 || (km - 1) < maxQPos;
            if ((a & 2) && ins) {
                const uint32_t e = w.ent[j];
                key = (e >> 1) & 0x3FFFFFFFu;
                cnt = a >> 16;
                t0 = (e & 1) != 0;
                t1 = !t0;
            }
        }
        const uint64_t b0 = __ballot(t0), b1 = __ballot(t1);
        if (t0) {
            const uint32_t p = nins[0] + __popcll(b0 & below);
            if (p < w.cap_e) {
                w.ins_key[p] = key;
                w.ins_cnt[p] = (uint16_t)cnt;
            }
        }
        if (t1) {
            const uint32_t p = nins[1] + __popcll(b1 & below);
            if (p < w.cap_e) {
                w.ins_key[w.cap_e + p] = key;
                w.ins_cnt[w.cap_e + p] = (uint16_t)cnt;
            }
        }
        nins[0] += __popcll(b0);
        nins[1] += __popcll(b1);
        const bool q0 = t0 && cnt >= QTHRESH, q1 = t1 && cnt >= QTHRESH;
        const uint64_t c0 = __ballot(q0), c1 = __ballot(q1);
        if (q0 && nq[0] == 0 && (c0 & below) == 0) {
            w.mkey[0] = key;
            w.mval[0] = cnt;
        }
        if (q1 && nq[1] == 0 && (c1 & below) == 0) {
            w.mkey[w.mcap] = key;
            w.mval[w.mcap] = cnt;
        }
        nq[0] += __popcll(c0);
        nq[1] += __popcll(c1);
    }
    if (nins[0] > w.cap_e || nins[1] > w.cap_e) return 1;
    wave_sync();
#pragma unroll
    for (int mp = 0; mp < 2; ++mp) {
        if (nq[mp] > w.mcap) return 1;
        if (nq[mp] <= 1) {
            if (lane == 0) m.nmap[mp] = nq[mp];
            continue;
        }
        if (nins[mp] <= 64) {
            // small maps: iteration order without the table (lp_place)
            const bool in = (uint32_t)lane < nins[mp];
            const uint32_t mykey = in ? w.ins_key[mp * w.cap_e + lane] : 0u;
            const uint32_t mycnt = in ? w.ins_cnt[mp * w.cap_e + lane] : 0u;
            uint32_t mypos;
            if (lp_place(mykey, nins[mp], bucketCount, w.tb_cap, mypos)) {
                const bool keep = in && mycnt >= QTHRESH;
                uint32_t rank;
                const uint32_t nc = lp_rank(keep, mypos, rank);
                if (keep) {
                    w.mkey[mp * w.mcap + rank] = mykey;
                    w.mval[mp * w.mcap + rank] = mycnt;
                }
                if (lane == 0) m.nmap[mp] = nc;
                wave_sync();
                continue;
            }
        }
        Hs<uint32_t> h;
        h.tb = w.tb;
        h.tb_cap = w.tb_cap;
        h.over = w.over;
        h.over_cap = w.cap_e;
        h.tmp = w.tmp;
        h.tmp_cap = w.cap_e;
        h.key = w.ins_key + mp * w.cap_e;
        h.nelem = h.nover = 0;
        if (!hs_setup(h, bucketCount, lane)) return 1;
        for (uint32_t i = 0; i < nins[mp]; ++i)
            if (hs_insert(h, i, lane) != HS_OK) return 1;
        const uint16_t* cnt = w.ins_cnt + mp * w.cap_e;
        uint32_t nc = 0;
        for (uint64_t b0 = 0; b0 < h.arr + h.nover; b0 += 64) {
            const uint64_t b = b0 + lane;
            uint32_t i = Hs<uint32_t>::NONE;
            if (b < h.arr) i = h.tb[b];
            else if (b < h.arr + h.nover) i = h.over[b - h.arr];
            const bool keep = i != Hs<uint32_t>::NONE && cnt[i] >= QTHRESH;
            const uint64_t bal = __ballot(keep);
            if (keep) {
                const uint32_t p = nc + __popcll(bal & below);
                w.mkey[mp * w.mcap + p] = h.key[i];
                w.mval[mp * w.mcap + p] = cnt[i];
            }
            nc += __popcll(bal);
        }
        if (lane == 0) m.nmap[mp] = nc;
        wave_sync();
    }
    wave_sync();
    return 0;
}

// getSeedRefsSecondRead's counting (:2646-3120).  In the reference the kept
// entries are visited in index order: an entry from a k-mer that may insert
// (k-mer 0, or cIdx < maxQPos) credits its window when the window or one
// within contextWLen is in its strand's map, creating the window through
// operator[] if needed; other entries credit only windows already present.
// Let T(w) be the entry at which window w enters the map (0 for the first
// read's windows).  Then T(w) = the first inserting entry j of w with a
// neighbour v, T(v) < j + 1, and w is credited by exactly the entries j with
// T(w) <= j + 1.  T is found as a fixpoint: every round lowers T(w) for
// entries whose neighbour is present before them; starting from "never",
// the values stay at or above the sequential ones and a round without a
// change leaves exactly them (a first difference would have a witness
// neighbour with an earlier difference).  Rounds = the longest chain of
// windows entering through each other, usually one.  Returns 1 on overflow.
template <typename TabT>
__device__ int pe_second(const PeWork<TabT>& w, PeLds& m, uint32_t E, uint32_t L, int cw)
{
    const int lane = threadIdx.x & 63;
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    // every kept entry's window gets a slot
    for (uint32_t j = lane; j < E; j += 64) {
        if (!(w.eaux[j] & 1)) continue;
        const uint32_t e = w.ent[j];
        const uint32_t h = pe_claim(w, (e >> 1) & 0x3FFFFFFFu, (e & 1) ? 0u : 1u);
        w.eaux[j] = 1u | (h << 8);
    }
    wave_sync();
    for (int round = 0;; ++round) {
        bool changed = false;
        for (uint32_t j = lane; j < E; j += 64) {
            const uint32_t a = w.eaux[j];
            if (!(a & 1)) continue;
            const uint32_t km = w.ekm[j];
            if (!(km == 0 || (km - 1) < maxQPos)) continue;
            const uint32_t h = a >> 8;
            if (w.ht[h] <= j + 1) continue;
            const uint32_t e = w.ent[j];
            const uint32_t key = (e >> 1) & 0x3FFFFFFFu, s = (e & 1) ? 0u : 1u;
            bool nb = false;
            for (int o = -cw; o <= cw && !nb; ++o) {
                if (o == 0) continue;
                const uint32_t q = pe_lookup(w, key + (uint32_t)o, s);
                nb = q != PE_NONE && w.ht[q] < j + 1;
            }
            if (nb && atomicMin(&w.ht[h], j + 1) > j + 1) changed = true;
        }
        wave_sync();
        if (!__ballot(changed)) break;
    }
    // credits
    for (uint32_t j = lane; j < E; j += 64) {
        const uint32_t a = w.eaux[j];
        if (!(a & 1)) continue;
        const uint32_t h = a >> 8;
        if (w.ht[h] <= j + 1) atomicAdd(&w.hc[h], 1u);
    }
    wave_sync();
    // windows entering the maps, in order of entry (each entry creates at
    // most one window)
    uint32_t n0 = m.nmap[0], n1 = m.nmap[1];
    const uint64_t below = lanes_below(lane);
    for (uint32_t j0 = 0; j0 < E; j0 += 64) {
        const uint32_t j = j0 + lane;
        bool i0 = false, i1 = false;
        uint32_t h = 0, key = 0;
        if (j < E) {
            const uint32_t a = w.eaux[j];
            if (a & 1) {
                h = a >> 8;
                if (w.ht[h] == j + 1) {
                    const uint32_t e = w.ent[j];
                    key = (e >> 1) & 0x3FFFFFFFu;
                    i0 = (e & 1) != 0;
                    i1 = !i0;
                }
            }
        }
        const uint64_t b0 = __ballot(i0), b1 = __ballot(i1);
        if (n0 + __popcll(b0) > w.mcap || n1 + __popcll(b1) > w.mcap) return 1;
        if (i0 || i1) {
            const uint32_t x = i0 ? n0 + __popcll(b0 & below) : n1 + __popcll(b1 & below);
            const uint32_t sm = i0 ? 0u : w.mcap;
            w.mkey[sm + x] = key;
            w.mval[sm + x] = 0;
            w.hi[h] = x;
        }
        n0 += __popcll(b0);
        n1 += __popcll(b1);
    }
    wave_sync();
    if (lane == 0) {
        m.nmap[0] = n0;
        m.nmap[1] = n1;
    }
    // c1 into the map values (u8)
    for (uint32_t h = lane; h < w.mhsz; h += 64) {
        const uint32_t k2 = w.hk[h];
        if (k2 == EMPTY) continue;
        const uint32_t x = w.hi[h];
        if (x == PE_NONE) continue;
        w.mval[(k2 & 1) * w.mcap + x] |= (w.hc[h] & 0xFFu) << 8;
    }
    wave_sync();
    return 0;
}

// ---------------------------------------------------------------------------
// k_pe_seed
// ---------------------------------------------------------------------------

// one item (pair combination) on one wavefront
template <typename TabT>
__device__ void pe_item(const DevIndex& ix, const BatchDev& bd, const PairDev& pd, const SeedShared& sh, PeLds& m,
                        const PeWork<TabT>& w, uint32_t t, PoolChunk& pcA, MapChunk& mc, uint64_t (&st)[4])
{
    const int lane = threadIdx.x & 63;
    const uint64_t below = lanes_below(lane);
    const uint32_t i = t >> 1, c = t & 1;
    uint32_t nf = 0, nr = 0, off = 0;
    // pairs with an invalid mate are counted by k_pe_final only
    const bool on = ((pd.combos >> c) & 1) && bd.pre[i] == 0 && bd.pre[pd.n + i] == 0;
    int rc = 0;
    uint32_t EA = 0, EB = 0;
    if (on) {
        const uint32_t roA = 2 * i + c;
        const uint32_t roB = 2 * (pd.n + i) + (1 - c);
        uint32_t bc, LA, LB, bcB;
        if (lane < 2) m.nmap[lane] = 0;
        rc = pe_entries<TabT>(ix, bd, roA, sh, m, w, true, EA, bc, LA);
        if (rc == 0) {
            st[0] += LA - KMERLEN + 1;
            st[1] += EA;
            st[3] += LA;
            ++st[2];
            rc = pe_first<TabT>(w, m, EA, LA, bc);
        }
        if (rc == 0 && (m.nmap[0] | m.nmap[1])) {
            // the lookup table, holding the first read's windows (time 0)
            for (uint32_t x = lane; x < w.mhsz; x += 64) {
                w.hk[x] = EMPTY;
                w.ht[x] = PE_NEVER;
                w.hc[x] = 0;
                w.hi[x] = PE_NONE;
            }
            wave_sync();
            for (uint32_t s = 0; s < 2; ++s)
                for (uint32_t x = lane; x < m.nmap[s]; x += 64) {
                    const uint32_t h = pe_claim(w, w.mkey[s * w.mcap + x], s);
                    w.ht[h] = 0;
                    w.hi[h] = x;
                }
            wave_sync();
            rc = pe_entries<TabT>(ix, bd, roB, sh, m, w, false, EB, bcB, LB);
            if (rc == 0) {
                st[0] += LB - KMERLEN + 1;
                st[1] += EB;
                st[3] += LB;
                rc = pe_second<TabT>(w, m, EB, LB, (int)pd.cw);
            }
            if (rc == 0) {
                // flag2: windows with c1 >= QTHRESH and their present
                // neighbours within contextWLen (:3125-3150); hasCpG
                const uint32_t n0 = m.nmap[0], ntot = n0 + m.nmap[1];
                for (uint32_t x = lane; x < ntot; x += 64) {
                    const uint32_t s = x < n0 ? 0u : 1u, y = x < n0 ? x : x - n0;
                    uint32_t* mv = w.mval + s * w.mcap;
                    if (pe_c1(mv[y]) < QTHRESH) continue;
                    const uint32_t key = w.mkey[s * w.mcap + y];
                    atomicOr(&mv[y], PF2);
                    for (int u = 1; u <= (int)pd.cw; ++u) {
                        uint32_t q = pe_find(w, key + (uint32_t)u, s);
                        if (q != PE_NONE) atomicOr(&mv[q], PF2);
                        q = pe_find(w, key - (uint32_t)u, s);
                        if (q != PE_NONE) atomicOr(&mv[q], PF2);
                    }
                }
                wave_sync();
                bool cpg = false;
                uint32_t nj = 0;
                for (uint32_t x = lane; x < ntot; x += 64) {
                    const uint32_t s = x < n0 ? 0u : 1u, y = x < n0 ? x : x - n0;
                    const uint32_t v = w.mval[s * w.mcap + y];
                    if ((v & PF2) && ix.win[w.mkey[s * w.mcap + y]].z != CPGDUMMY) cpg = true;
                    if ((v & PF2) && pe_c0(v) >= QTHRESH) ++nj;
                }
                const bool hasCpG = __ballot(cpg) != 0;
                nj = wave_sum_u32(nj);
                if (hasCpG) {
                    nf = m.nmap[0];
                    nr = m.nmap[1];
                    // the wave's open chunk of the map pool (gaps are never read)
                    if (nf + nr > mc.left) {
                        const uint32_t take = max(nf + nr, (uint32_t)FAME_PE_MAP_CHUNK);
                        uint32_t b = 0;
                        if (lane == 0) b = atomicAdd(pd.map_n, take);
                        mc.base = lane_u32(b, 0);
                        mc.left = take;
                    }
                    off = mc.base;
                    mc.base += nf + nr;
                    mc.left -= nf + nr;
                    if (off + nf + nr > pd.map_cap) {
                        if (lane == 0) atomicOr(bd.flags, FLAG_POOL_FULL);
                        nf = nr = 0;
                    }
                }
                if (nf + nr) {
                    const uint32_t pool = pool_of_len(LA);
                    uint32_t jb = nj ? pool_take(bd, pcA, pool, nj, false, lane) : 0;
                    if (jb == ~0u) {
                        if (lane == 0) atomicOr(bd.flags, FLAG_POOL_FULL);
                        nf = nr = 0;
                    } else {
                        uint32_t* jm = pc_meta(bd, pool);
                        uint32_t* jr = bd.job_ro[0];
#pragma unroll
                        for (uint32_t q = 1; q < FAME_NPOOL; ++q) jr = pool == q ? bd.job_ro[q] : jr;
                        // iteration order of the two 20-bucket maps (:2760-2790
                        // build them with 20 buckets, :2801-3164 insert into them)
                        uint32_t pos = off;
                        for (uint32_t s = 0; s < 2 && rc == 0; ++s) {
                            const uint32_t ns = m.nmap[s];
                            if (ns <= 64) {
                                // small maps: iteration order without the table
                                const bool occ = (uint32_t)lane < ns;
                                const uint32_t mk = occ ? w.mkey[s * w.mcap + lane] : 0u;
                                uint32_t mypos;
                                if (lp_place(mk, ns, 20, w.tb_cap, mypos)) {
                                    const uint32_t v = occ ? w.mval[s * w.mcap + lane] : 0u;
                                    const bool job = occ && (v & PF2) && pe_c0(v) >= QTHRESH;
                                    uint32_t ro, rj;
                                    const uint32_t no = lp_rank(occ, mypos, ro);
                                    const uint32_t nj = lp_rank(job, mypos, rj);
                                    if (occ) {
                                        PeItem it;
                                        it.k = mk;
                                        it.v = v;
                                        it.j = PE_NONE;
                                        if (job) {
                                            it.j = jb + rj;
                                            jm[it.j] = mk | (s ? 0x80000000u : 0u);
                                            jr[it.j] = roA;
                                        }
                                        pd.map[pos + ro] = it;
                                    }
                                    pos += no;
                                    jb += nj;
                                    wave_sync();
                                    continue;
                                }
                            }
                            Hs<uint32_t> h;
                            h.tb = w.tb;
                            h.tb_cap = w.tb_cap;
                            h.over = w.over;
                            h.over_cap = w.cap_t;
                            h.tmp = w.tmp;
                            h.tmp_cap = w.cap_t;
                            h.key = w.mkey + s * w.mcap;
                            h.nelem = h.nover = 0;
                            if (!hs_setup(h, 20, lane)) {
                                rc = 1;
                                break;
                            }
                            for (uint32_t x = 0; x < ns && rc == 0; ++x)
                                if (hs_insert(h, x, lane) != HS_OK) rc = 1;
                            if (rc) break;
                            const uint32_t* mv = w.mval + s * w.mcap;
                            for (uint64_t b0 = 0; b0 < h.arr + h.nover; b0 += 64) {
                                const uint64_t b = b0 + lane;
                                uint32_t x = Hs<uint32_t>::NONE;
                                if (b < h.arr) x = h.tb[b];
                                else if (b < h.arr + h.nover) x = h.over[b - h.arr];
                                const bool occ = x != Hs<uint32_t>::NONE;
                                const uint32_t v = occ ? mv[x] : 0u;
                                const bool job = occ && (v & PF2) && pe_c0(v) >= QTHRESH;
                                const uint64_t bo = __ballot(occ), bj = __ballot(job);
                                if (occ) {
                                    PeItem it;
                                    it.k = h.key[x];
                                    it.v = v;
                                    it.j = PE_NONE;
                                    if (job) {
                                        it.j = jb + __popcll(bj & below);
                                        jm[it.j] = it.k | (s ? 0x80000000u : 0u);
                                        jr[it.j] = roA;
                                    }
                                    pd.map[pos + __popcll(bo & below)] = it;
                                }
                                pos += __popcll(bo);
                                jb += __popcll(bj);
                            }
                            wave_sync();
                        }
                        // a map outgrew the scratch after its slots were taken:
                        // the slots' jobs are complete, the item is re-run
                        if (rc) nf = nr = 0;
                    }
                }
            }
        }
    }
    if (rc == 1) {
        // beyond this variant's capacity: listed for a global-scratch run
        // (the host grows the scratch from run to run)
        if (lane == 0) {
            const uint32_t b = atomicAdd(pd.big_n, 1u);
            if (b < pd.big_cap) {
                pd.big_item[b] = t;
                pd.big_e[b] = max(EA, EB);
            } else {
                atomicOr(bd.flags, FLAG_HEAVY_FULL);
            }
        }
        nf = nr = 0;
    }
    if (lane == 0) {
        pd.map_off[t] = off;
        pd.map_nf[t] = nf;
        pd.map_nr[t] = nr;
    }
    wave_sync();
}

template <typename TabT>
struct PeLight {
    TabT klo[256];
    uint32_t klen[256];
    uint32_t koff[256];
    uint32_t kcm[256];
    uint8_t kfl[256];
    uint32_t ent[PE_LE];
    uint32_t eaux[PE_LE];
    uint8_t ekm[PE_LE];
    // phases in turn: the first read's counting table, its insertion lists
    // and hopscotch arrays, the lookup over the item's windows (second read,
    // flag2), the 20-bucket maps' hopscotch arrays
    union {
        struct {
            uint32_t key[2 * PE_LE];
            uint32_t first[2 * PE_LE];
            uint32_t cnt[2 * PE_LE];
        } h;
        struct {
            uint32_t ins_key[2 * PE_LE];
            uint16_t ins_cnt[2 * PE_LE];
            uint32_t tb[PE_LTB];
            uint32_t tmp[PE_LE];
            uint32_t over[PE_LE];
        } c;
        struct {
            uint32_t hk[PE_LH];
            uint32_t ht[PE_LH];
            uint32_t hc[PE_LH];
            uint32_t hi[PE_LH];
        } m;
    } u;
    uint32_t mkey[2 * PE_LM];
    uint32_t mval[2 * PE_LM];
};

template <typename TabT>
__global__ __launch_bounds__(64) void k_pe_seed(DevIndex ix, BatchDev bd, PairDev pd)
{
    __shared__ SeedShared sh;
    __shared__ PeLds m;
    __shared__ PeLight<TabT> L;
    const int lane = threadIdx.x;
    seed_table_init(sh);
    PeWork<TabT> w;
    w.klo = L.klo;
    w.klen = L.klen;
    w.koff = L.koff;
    w.kcm = L.kcm;
    w.kfl = L.kfl;
    w.cap_e = PE_LE;
    w.ent = L.ent;
    w.ekm = L.ekm;
    w.eaux = L.eaux;
    // the hash's prefix array over ent and eaux (2 KiB, written before them)
    static_assert(offsetof(PeLight<TabT>, eaux) == offsetof(PeLight<TabT>, ent) + 4 * PE_LE, "ent, eaux adjacent");
    static_assert(offsetof(PeLight<TabT>, ent) % 8 == 0 && 8 * PE_LE <= 2 * 4 * PE_LE, "prefix array fits");
    w.pf = (uint64_t*)L.ent;
    w.hsize = 2 * PE_LE;
    w.hkey = L.u.h.key;
    w.hfirst = L.u.h.first;
    w.hcnt = L.u.h.cnt;
    w.ins_key = L.u.c.ins_key;
    w.ins_cnt = L.u.c.ins_cnt;
    w.tb_cap = PE_LTB;
    w.tb = L.u.c.tb;
    w.tmp = L.u.c.tmp;
    w.over = L.u.c.over;
    w.cap_t = PE_LE;
    w.mkey = L.mkey;
    w.mval = L.mval;
    w.mcap = PE_LM;
    w.hk = L.u.m.hk;
    w.ht = L.u.m.ht;
    w.hc = L.u.m.hc;
    w.hi = L.u.m.hi;
    w.mhsz = PE_LH;
    PoolChunk pcA = {0, 0, 0};
    uint64_t st[4] = {0, 0, 0, 0};
    // items of the queried combinations only, taken from a device counter
    // (their costs vary by orders of magnitude), the next one's number
    // fetched while the current one runs
    const uint32_t both = pd.combos == 3u;
    const uint32_t nq = both ? 2 * pd.n : (pd.combos ? pd.n : 0u);
    const uint32_t c1 = pd.combos == 2u ? 1u : 0u;
    MapChunk mc = {0, 0};
    constexpr uint32_t TAKE = 4; // items per take
    uint32_t nt = 0;
    if (lane == 0) nt = atomicAdd(pd.next, TAKE);
    uint32_t q0 = lane_u32(nt, 0);
    while (q0 < nq) {
        if (lane == 0) nt = atomicAdd(pd.next, TAKE);
        const uint32_t qe = min(q0 + TAKE, nq);
        for (uint32_t q = q0; q < qe; ++q) pe_item<TabT>(ix, bd, pd, sh, m, w, both ? q : 2 * q + c1, pcA, mc, st);
        q0 = lane_u32(nt, 0);
    }
    if (pcA.left) pool_gap_fill(bd, pcA.pool, pcA.base, pcA.left, lane);
    if (lane == 0 && st[2]) {
        atomicAdd(&bd.stats[0], (unsigned long long)st[0]);
        atomicAdd(&bd.stats[1], (unsigned long long)st[1]);
        atomicAdd(&bd.stats[2], (unsigned long long)st[2]);
        atomicAdd(&bd.stats[6], (unsigned long long)st[3]);
        atomicAdd(&bd.stats[8], (unsigned long long)st[1]); // every bucket's entries are read here
    }
}

// the listed items with per-block global scratch sized for the largest
template <typename TabT>
__global__ __launch_bounds__(64) void k_pe_seed_big(DevIndex ix, BatchDev bd, PairDev pd)
{
    __shared__ SeedShared sh;
    __shared__ PeLds m;
    const int lane = threadIdx.x;
    seed_table_init(sh);
    uint8_t* base = pd.scratch + (size_t)blockIdx.x * pd.stride;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t* p = base + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    PeWork<TabT> w;
    w.klo = (TabT*)take(8 * 256);
    w.klen = (uint32_t*)take(4 * 256);
    w.koff = (uint32_t*)take(4 * 256);
    w.kcm = (uint32_t*)take(4 * 256);
    w.kfl = (uint8_t*)take(256);
    w.cap_e = pd.cap_e;
    w.ent = (uint32_t*)take(4ull * pd.cap_e);
    w.ekm = (uint8_t*)take(pd.cap_e);
    w.eaux = (uint32_t*)take(4ull * pd.cap_e);
    w.pf = nullptr; // global scratch: the closed form from LDS is faster
    w.hsize = pd.hsize;
    w.hkey = (uint32_t*)take(4ull * pd.hsize);
    w.hfirst = (uint32_t*)take(4ull * pd.hsize);
    w.hcnt = (uint32_t*)take(4ull * pd.hsize);
    w.ins_key = (uint32_t*)take(8ull * pd.cap_e);
    w.ins_cnt = (uint16_t*)take(4ull * pd.cap_e);
    w.tb_cap = pd.tb_cap;
    w.tb = (uint32_t*)take(4ull * pd.tb_cap);
    w.cap_t = pd.cap_t;
    w.tmp = (uint32_t*)take(4ull * pd.cap_t);
    w.over = (uint32_t*)take(4ull * pd.cap_t);
    w.mcap = pd.mcap;
    w.mkey = (uint32_t*)take(8ull * pd.mcap);
    w.mval = (uint32_t*)take(8ull * pd.mcap);
    w.mhsz = pd.mhsz;
    w.hk = (uint32_t*)take(4ull * pd.mhsz);
    w.ht = (uint32_t*)take(4ull * pd.mhsz);
    w.hc = (uint32_t*)take(4ull * pd.mhsz);
    w.hi = (uint32_t*)take(4ull * pd.mhsz);
    PoolChunk pcA = {0, 0, 0};
    uint64_t st[4] = {0, 0, 0, 0};
    MapChunk mc = {0, 0};
    // listed items taken one at a time from a device counter (their costs
    // vary by orders of magnitude; the list holds the largest first), the
    // next one's number fetched while the current one runs
    uint32_t nt = 0;
    if (lane == 0) nt = atomicAdd(pd.next, 1u);
    uint32_t ti = lane_u32(nt, 0);
    while (ti < pd.list_n) {
        if (lane == 0) nt = atomicAdd(pd.next, 1u);
        pe_item<TabT>(ix, bd, pd, sh, m, w, pd.list[ti], pcA, mc, st);
        ti = lane_u32(nt, 0);
    }
    if (pcA.left) pool_gap_fill(bd, pcA.pool, pcA.base, pcA.left, lane);
    if (lane == 0 && st[2]) {
        atomicAdd(&bd.stats[0], (unsigned long long)st[0]);
        atomicAdd(&bd.stats[1], (unsigned long long)st[1]);
        atomicAdd(&bd.stats[2], (unsigned long long)st[2]);
        atomicAdd(&bd.stats[6], (unsigned long long)st[3]);
        atomicAdd(&bd.stats[8], (unsigned long long)st[1]); // every bucket's entries are read here
    }
}

// ---------------------------------------------------------------------------
// std::sort (libstdc++ bits/stl_algo.h, stl_heap.h): introsort with
// median-of-three pivots and unguarded partitions, heap sort past a depth of
// 2*floor(log2 n), then insertion sort (the first 16 guarded); comparator
// "count greater" (ReadQueue.h:339-351).  Sub-ranges are independent, so
// the recursion on the right part runs from an explicit stack.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool pe_less(const PeItem& a, const PeItem& b, int sec)
{
    return sec ? pe_c1(a.v) > pe_c1(b.v) : pe_c0(a.v) > pe_c0(b.v);
}

__device__ __forceinline__ void pe_swap(PeItem* a, int i, int j)
{
    const PeItem t = a[i];
    a[i] = a[j];
    a[j] = t;
}

__device__ void pe_push_heap(PeItem* a, int first, int hole, int top, PeItem val, int sec)
{
    int parent = (hole - 1) / 2;
    while (hole > top && pe_less(a[first + parent], val, sec)) {
        a[first + hole] = a[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[first + hole] = val;
}

__device__ void pe_adjust_heap(PeItem* a, int first, int hole, int len, PeItem val, int sec)
{
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (pe_less(a[first + child], a[first + child - 1], sec)) child--;
        a[first + hole] = a[first + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[first + hole] = a[first + child - 1];
        hole = child - 1;
    }
    pe_push_heap(a, first, hole, top, val, sec);
}

__device__ void pe_heap_sort(PeItem* a, int first, int last, int sec)
{
    const int len = last - first;
    if (len >= 2) {
        for (int parent = (len - 2) / 2;; --parent) {
            pe_adjust_heap(a, first, parent, len, a[first + parent], sec);
            if (parent == 0) break;
        }
    }
    while (last - first > 1) {
        --last;
        const PeItem v = a[last];
        a[last] = a[first];
        pe_adjust_heap(a, first, 0, last - first, v, sec);
    }
}

__device__ void pe_median_to_first(PeItem* a, int result, int x, int y, int z, int sec)
{
    if (pe_less(a[x], a[y], sec)) {
        if (pe_less(a[y], a[z], sec)) pe_swap(a, result, y);
        else if (pe_less(a[x], a[z], sec)) pe_swap(a, result, z);
        else pe_swap(a, result, x);
    } else if (pe_less(a[x], a[z], sec)) {
        pe_swap(a, result, x);
    } else if (pe_less(a[y], a[z], sec)) {
        pe_swap(a, result, z);
    } else {
        pe_swap(a, result, y);
    }
}

// __final_insertion_sort: the first 16 by a guarded insertion sort, the rest
// unguarded (a smaller element than a[0] exists among the first 16 after the
// introsort loop); for n <= 16 this is all std::sort does
__device__ __forceinline__ void pe_final_insertion(PeItem* a, int n, int sec)
{
    const int g = n > 16 ? 16 : n;
    for (int i = 1; i < g; ++i) {
        const PeItem val = a[i];
        if (pe_less(val, a[0], sec)) {
            for (int k = i; k > 0; --k) a[k] = a[k - 1];
            a[0] = val;
        } else {
            int k = i;
            while (k > 0 && pe_less(val, a[k - 1], sec)) {
                a[k] = a[k - 1];
                --k;
            }
            a[k] = val;
        }
    }
    for (int i = g; i < n; ++i) {
        const PeItem val = a[i];
        int k = i;
        while (k > 0 && pe_less(val, a[k - 1], sec)) {
            a[k] = a[k - 1];
            --k;
        }
        a[k] = val;
    }
}

constexpr int PE_SORT_STACK = 48; // pending right parts (depth limit 2 log2 n < 48)

// `stk`: room for 3 x PE_SORT_STACK ints (LDS or global memory), used for n > 16
__device__ void pe_sort(PeItem* a, int n, int sec, int* stk)
{
    if (n <= 1) return;
    if (n > 16) {
        int* sf = stk;
        int* sl = stk + PE_SORT_STACK;
        int* sd = stk + 2 * PE_SORT_STACK;
        int lg = 0;
        while ((1 << (lg + 1)) <= n) ++lg;
        sf[0] = 0;
        sl[0] = n;
        sd[0] = 2 * lg;
        int sp = 1;
        while (sp > 0) {
            --sp;
            const int first = sf[sp];
            int last = sl[sp], depth = sd[sp];
            while (last - first > 16) {
                if (depth == 0) {
                    pe_heap_sort(a, first, last, sec);
                    break;
                }
                --depth;
                const int mid = first + (last - first) / 2;
                pe_median_to_first(a, first, first + 1, mid, last - 1, sec);
                // __unguarded_partition(first + 1, last, first)
                // (the bounds tests never fire: the median of three stops
                // both scans, as in libstdc++; they only keep every access in
                // range)
                int lo = first + 1, hi = last;
                for (;;) {
                    while (lo < last - 1 && pe_less(a[lo], a[first], sec)) ++lo;
                    --hi;
                    while (hi > first && pe_less(a[first], a[hi], sec)) --hi;
                    if (!(lo < hi)) break;
                    pe_swap(a, lo, hi);
                    ++lo;
                }
                const int cut = lo;
                if (sp < PE_SORT_STACK) {
                    sf[sp] = cut;
                    sl[sp] = last;
                    sd[sp] = depth;
                    ++sp;
                }
                last = cut;
            }
        }
    }
    pe_final_insertion(a, n, sec);
}

// ---------------------------------------------------------------------------
// the walks of saQuerySeedSetRefFirst / Second over verified windows
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t pool_of(uint32_t L) { return pool_of_len(L); }

__device__ __forceinline__ uint32_t jm_at(const BatchDev& bd, uint32_t pool, uint32_t j, uint32_t k)
{
    return job_match_at(bd, pool, j, k);
}

__device__ __forceinline__ uint8_t* acc_of(uint8_t* const (&acc)[FAME_NPOOL], uint32_t p)
{
    uint8_t* v = acc[0];
#pragma unroll
    for (uint32_t q = 1; q < FAME_NPOOL; ++q) v = p == q ? acc[q] : v;
    return v;
}

// One strand's walk (matchFwdFirst / matchRevFirst :4538-4640, matchFwdSecond
// / matchRevSecond :4650-4776): windows in sorted order until one's count
// falls below max(bmCount - KMERDIST, QTHRESH); windows without the flag are
// passed; a window's first match is dropped when it repeats the previous
// window's last one.  Marks each walked slot in acc (1, | 2 when the first
// match was dropped); `mark` (first read) receives flag3 for the matched
// windows and their neighbours.  Returns the matches kept.
// All `step` threads of a group run the walk alike (uniform values); the
// marking is split among them (`lane` = 0..step-1), lane 0 writes acc.
__device__ uint32_t pe_walk(const DevIndex& ix, const BatchDev& vb, uint8_t* const (&acc)[FAME_NPOOL],
                            uint32_t pool, const PeItem* list, uint32_t n, int sec, uint32_t flag, int32_t& bm,
                            PeItem* mark, uint32_t nmark, int cw, uint32_t lane, uint32_t step)
{
    uint32_t prevChr = 0;
    uint64_t prevOff = ~0ull;
    uint32_t kept = 0;
    uint8_t* ac = acc_of(acc, pool);
    for (uint32_t x = 0; x < n; ++x) {
        const PeItem it = list[x];
        const int32_t cnt = (int32_t)(sec ? pe_c1(it.v) : pe_c0(it.v));
        const int32_t lim = bm - (int32_t)KMERDIST > (int32_t)QTHRESH ? bm - (int32_t)KMERDIST : (int32_t)QTHRESH;
        if (cnt < lim) break;
        if (!(it.v & flag)) continue;
        const uint32_t j = it.j;
        if (j == PE_NONE) continue; // not reached: flagged windows at or above QTHRESH carry a slot
        const uint32_t nm = vb.job_nm[pool][j];
        const uint4 w = ix.win[it.k];
        uint32_t skip = 0;
        if (nm > 0) {
            const uint32_t m0 = jm_at(vb, pool, j, 0);
            if ((uint64_t)(m0 >> 8) + w.x == prevOff && w.y == prevChr) skip = 1;
        }
        if (lane == 0) ac[j] = (uint8_t)(1u | (skip << 1));
        kept += nm - skip;
        if (nm > 0) {
            if (cnt > bm) bm = cnt;
            if (mark) {
                for (uint32_t y = lane; y < nmark; y += step) {
                    const int32_t d = (int32_t)(mark[y].k - it.k);
                    if (d >= -cw && d <= cw) mark[y].v |= PF3;
                }
            }
            prevChr = w.y;
            prevOff = (uint64_t)w.x + (jm_at(vb, pool, j, nm - 1) >> 8);
        } else {
            prevChr = 0;
            prevOff = ~0ull;
        }
    }
    return kept;
}

// both strands: the one whose first (sorted) window has the larger c0 goes
// first, the reverse map on ties (:2185-2240, :2265-2320)
__device__ uint32_t pe_query(const DevIndex& ix, const BatchDev& vb, uint8_t* const (&acc)[FAME_NPOOL],
                             uint32_t pool, const PeItem* lf, uint32_t nf, const PeItem* lr, uint32_t nr, int sec,
                             uint32_t flag, PeItem* mf, PeItem* mr, int cw, uint32_t lane = 0, uint32_t step = 1)
{
    int32_t bm = 0;
    bool fwd_first = nf > 0 && (nr == 0 || pe_c0(lf[0].v) > pe_c0(lr[0].v));
    uint32_t kept = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool fw = pass == 0 ? fwd_first : !fwd_first;
        if (fw) kept += pe_walk(ix, vb, acc, pool, lf, nf, sec, flag, bm, mf, mf ? nf : 0, cw, lane, step);
        else kept += pe_walk(ix, vb, acc, pool, lr, nr, sec, flag, bm, mr, mr ? nr : 0, cw, lane, step);
    }
    return kept;
}

// windows of an item that k_pe_walk1 sorts and walks in LDS (more: in the
// global map pool); a thread's two copies take PE_WSTRIDE words each, an odd
// count, so that the threads of a wave fall on distinct banks
constexpr uint32_t PE_WCAP = FAME_PE_WCAP;
constexpr uint32_t PE_WSTRIDE = 3 * PE_WCAP + 1;
static_assert(sizeof(PeItem) == 12, "PeItem is three words");
static_assert(PE_WCAP <= 16, "k_pe_walk1 sorts by insertion only");

// saQuerySeedSetRefFirst for one item (the maps sorted by c0, both walks with
// flag2, flag3 marked in `mp`), then saQuerySeedSetRefSecond's order (sorted
// by c1) in `sp`.  Run by one thread (lane 0, step 1) or by a whole wave
// (lane, step 64), which shares the copies and the marking and leaves the
// sorts to lane 0.  Returns PE_NONE (every lane) when the first read matched
// nothing, else (lane 0) the number of second-read jobs: the windows with
// flag3 and c1 >= QTHRESH.
__device__ __forceinline__ uint32_t pe_walk1_order(const DevIndex& ix, const BatchDev& b1, const PairDev& pd,
                                                   uint32_t t, uint32_t nf, uint32_t nr, PeItem* mp, PeItem* sp,
                                                   uint32_t lane, uint32_t step, int* stk)
{
    // stk == nullptr: at most 16 windows per map, insertion sort only
    auto sort = [&](PeItem* a, uint32_t na, int sec) {
        if (stk) pe_sort(a, (int)na, sec, stk);
        else pe_final_insertion(a, (int)na, sec);
    };
    const uint32_t i = t >> 1;
    const uint32_t n = nf + nr;
    for (uint32_t x = lane; x < n; x += step) sp[x] = mp[x];
    wave_sync();
    if (lane == 0) {
        sort(sp, nf, 0);
        sort(sp + nf, nr, 0);
    }
    wave_sync();
    const uint32_t poolA = pool_of(b1.lens[i]); // mate 1 is the first read in both combinations
    const uint32_t m1 =
        pe_query(ix, b1, pd.acc1, poolA, sp, nf, sp + nf, nr, 0, PF2, mp, mp + nf, (int)pd.cw, lane, step);
    wave_sync();
    if (lane == 0) pd.m1_n[t] = m1;
    if (m1 == 0) return PE_NONE;
    for (uint32_t x = lane; x < n; x += step) sp[x] = mp[x];
    wave_sync();
    uint32_t nj = 0;
    if (lane == 0) {
        sort(sp, nf, 1);
        sort(sp + nf, nr, 1);
        for (uint32_t x = 0; x < n; ++x)
            if ((sp[x].v & PF3) && pe_c1(sp[x].v) >= QTHRESH) ++nj;
    }
    wave_sync();
    return nj;
}

// the second reads' verify slots base, base + 1, ... for the counted windows
// of `sp` (one thread); ok = false: the pool overflowed, slots are left out
__device__ __forceinline__ void pe_walk1_assign(const DevIndex& ix, const BatchDev& b2, const PairDev& pd,
                                                uint32_t t, uint32_t nf, uint32_t nr, PeItem* sp, uint32_t poolB,
                                                uint32_t base, bool ok)
{
    const uint32_t roB = 2 * (pd.n + (t >> 1)) + (1 - (t & 1));
    uint32_t* jm = pc_meta(b2, poolB);
    uint32_t* jr = b2.job_ro[0];
#pragma unroll
    for (uint32_t q = 1; q < FAME_NPOOL; ++q) jr = poolB == q ? b2.job_ro[q] : jr;
    uint32_t k = 0;
    for (uint32_t x = 0; x < nf + nr; ++x) {
        sp[x].j = PE_NONE;
        if (ok && (sp[x].v & PF3) && pe_c1(sp[x].v) >= QTHRESH) {
            const uint32_t j = base + k++;
            const bool rev = x >= nf;
            // matchRevSecond scans from the chromosome start itself when the
            // window starts there (bit 30)
            const bool ung = rev && ix.win[sp[x].k].x == 0;
            jm[j] = sp[x].k | (rev ? 0x80000000u : 0u) | (ung ? 0x40000000u : 0u);
            jr[j] = roB;
            sp[x].j = j;
        }
    }
}

// one thread per item; items of more than pd.wcap (<= PE_WCAP) windows are
// listed for k_pe_walk1_big.  The second reads' slots are taken with one atomic per
// pool and wave.
__global__ __launch_bounds__(256) void k_pe_walk1(DevIndex ix, BatchDev b1, BatchDev b2, PairDev pd)
{
    __shared__ uint32_t s_map[256 * PE_WSTRIDE];
    __shared__ uint32_t s_srt[256 * PE_WSTRIDE];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    PeItem* lm = (PeItem*)(s_map + threadIdx.x * PE_WSTRIDE);
    PeItem* ls = (PeItem*)(s_srt + threadIdx.x * PE_WSTRIDE);
    uint32_t nf = 0, nr = 0, off = 0, nj = PE_NONE, poolB = 0;
    bool heavy = false;
    if (t < 2 * pd.n) {
        nf = pd.map_nf[t];
        nr = pd.map_nr[t];
        // items of a combination not queried were not seeded: their entries are stale
        if (!((pd.combos >> (t & 1)) & 1) || nf + nr == 0) {
            pd.m1_n[t] = 0;
        } else if (nf + nr > pd.wcap) {
            heavy = true;
        } else {
            off = pd.map_off[t];
            for (uint32_t x = 0; x < nf + nr; ++x) lm[x] = pd.map[off + x];
            nj = pe_walk1_order(ix, b1, pd, t, nf, nr, lm, ls, 0, 1, nullptr);
            poolB = pool_of(b1.lens[pd.n + (t >> 1)]);
        }
    }
    // the heavy items onto the list, one atomic per wave
    const uint64_t hb = __ballot(heavy);
    if (hb) {
        const uint32_t lead = (uint32_t)__ffsll((unsigned long long)hb) - 1;
        uint32_t b = 0;
        if (lane == lead) b = atomicAdd(pd.wlist_n, (uint32_t)__popcll(hb));
        b = __builtin_amdgcn_readlane(b, lead);
        if (heavy) pd.wlist[b + __popcll(hb & lanes_below((int)lane))] = t;
    }
    // slots per pool: a wave scan of nj over the lanes of that pool
    const uint32_t want = (nj != PE_NONE && nj) ? poolB : PE_NONE;
    uint32_t base = 0;
    bool ok = true;
    uint64_t pending = __ballot(want != PE_NONE);
    while (pending) {
        const uint32_t lead = (uint32_t)__ffsll((unsigned long long)pending) - 1;
        const uint32_t p = __builtin_amdgcn_readlane(want, lead);
        const bool mine = want == p;
        const uint32_t v = mine ? nj : 0u;
        const uint32_t inc = wave_iscan_u32(v, (int)lane);
        const uint32_t tot = lane_u32(inc, 63);
        uint32_t b = 0;
        if (lane == lead) b = atomicAdd(&b2.pool_n[p], tot);
        b = __builtin_amdgcn_readlane(b, lead);
        if (mine) {
            base = b + inc - v;
            ok = b + tot <= pc_get_cap(b2, p);
        }
        if (lane == lead && b + tot > pc_get_cap(b2, p)) atomicOr(b1.flags, FLAG_POOL_FULL);
        pending &= ~__ballot(mine);
    }
    if (nj != PE_NONE) {
        pe_walk1_assign(ix, b2, pd, t, nf, nr, ls, poolB, base, ok);
        for (uint32_t x = 0; x < nf + nr; ++x) pd.srt[off + x] = ls[x];
    }
}

// windows of a listed item walked in LDS by k_pe_walk1_big (more: in the
// global map pool)
constexpr uint32_t PE_WBIG = FAME_PE_WBIG;

// the listed items, one wave each, taken from a device counter
__global__ __launch_bounds__(64) void k_pe_walk1_big(DevIndex ix, BatchDev b1, BatchDev b2, PairDev pd)
{
    __shared__ PeItem s_map[PE_WBIG];
    __shared__ PeItem s_srt[PE_WBIG];
    __shared__ int s_stk[3 * PE_SORT_STACK];
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = *pd.wlist_n;
    uint32_t nt = 0;
    if (lane == 0) nt = atomicAdd(pd.wnext, 1u);
    uint32_t q = lane_u32(nt, 0);
    while (q < cnt) {
        if (lane == 0) nt = atomicAdd(pd.wnext, 1u);
        const uint32_t t = pd.wlist[q];
        const uint32_t nf = pd.map_nf[t], nr = pd.map_nr[t], off = pd.map_off[t];
        PeItem* mp = pd.map + off;
        PeItem* sp = pd.srt + off;
        const bool in_lds = nf + nr <= pd.wbig;
        PeItem* wm = mp;
        PeItem* ws = sp;
        if (in_lds) {
            for (uint32_t x = lane; x < nf + nr; x += 64) s_map[x] = mp[x];
            wave_sync();
            wm = s_map;
            ws = s_srt;
        }
        const uint32_t nj = pe_walk1_order(ix, b1, pd, t, nf, nr, wm, ws, lane, 64, s_stk);
        if (nj != PE_NONE) {
            if (lane == 0) {
                const uint32_t poolB = pool_of(b1.lens[pd.n + (t >> 1)]);
                uint32_t base = 0;
                bool ok = true;
                if (nj) {
                    base = atomicAdd(&b2.pool_n[poolB], nj);
                    if (base + nj > pc_get_cap(b2, poolB)) {
                        atomicOr(b1.flags, FLAG_POOL_FULL);
                        ok = false;
                    }
                }
                pe_walk1_assign(ix, b2, pd, t, nf, nr, ws, poolB, base, ok);
            }
            wave_sync();
            if (in_lds)
                for (uint32_t x = lane; x < nf + nr; x += 64) sp[x] = s_srt[x];
        }
        wave_sync();
        q = lane_u32(nt, 0);
    }
}

// extractPairedMatch :3255-3275 (MINPDIST is not used)
__device__ __forceinline__ int pe_extract(const DevIndex& ix, uint64_t m1, uint64_t m2, uint32_t maxd)
{
    const uint4 w1 = ix.win[(uint32_t)(m1 >> 32)];
    const uint4 w2 = ix.win[(uint32_t)(m2 >> 32)];
    if (w1.y != w2.y) return -1;
    const uint32_t p1 = w1.x + (uint32_t)(m1 & 0xFFFF);
    const uint32_t p2 = w2.x + (uint32_t)(m2 & 0xFFFF);
    const uint32_t d = p2 > p1 ? p2 - p1 : p1 - p2;
    if (d <= maxd) return (int)((m1 >> 16) & 0xFF) + (int)((m2 >> 16) & 0xFF);
    return -1;
}

__global__ __launch_bounds__(256) void k_pe_final(DevIndex ix, BatchDev b1, BatchDev b2, PairDev pd,
                                                  int get_stranded)
{
    __shared__ uint32_t s_cnt[7];
    if (threadIdx.x < 7) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int want1 = 0, want2 = 0;
    if (i < pd.n) {
        const uint32_t n = pd.n;
        const uint8_t p1 = b1.pre[i], p2 = b1.pre[n + i];
        uint64_t mat1 = 0, mat2 = 0;
        uint8_t v1 = 1, v2 = 1, use1 = 0, use2 = 0;
        if ((p1 & 1) || (p2 & 1)) atomicAdd(&s_cnt[4], 1u);
        if (p1 || p2) {
            atomicAdd(&s_cnt[4], 2u);
            v1 = p1 == 0;
            v2 = p2 == 0;
        } else {
            uint32_t m1c[2] = {0, 0}, m2c[2] = {0, 0};
            for (uint32_t c = 0; c < 2; ++c) {
                const uint32_t t = 2 * i + c;
                if (!((pd.combos >> c) & 1)) continue;
                m1c[c] = pd.m1_n[t];
                if (!m1c[c]) continue;
                const uint32_t nf = pd.map_nf[t], nr = pd.map_nr[t];
                const PeItem* sp = pd.srt + pd.map_off[t];
                const uint32_t poolB = pool_of(b1.lens[n + i]);
                m2c[c] = pe_query(ix, b2, pd.acc2, poolB, sp, nf, sp + nf, nr, 1, PF3, nullptr, nullptr,
                                  (int)pd.cw);
            }
            if ((m1c[0] == 0 && m1c[1] == 0) || (m2c[0] == 0 && m2c[1] == 0)) {
                atomicAdd(&s_cnt[2], 2u);
                v1 = v2 = 0;
            } else {
                // every (mate 1, mate 2) combination of both passes: the
                // lowest summed error count, unique or not (:1309-1373)
                const uint32_t maxd = pd.maxpd + ix.readlen;
                int best = 2 * (int)NERR + 1, nbest = 0;
                uint64_t b1m = 0, b2m = 0;
                uint32_t orig = 1;
                for (uint32_t c = 0; c < 2; ++c) {
                    if (!m1c[c] || !m2c[c]) continue;
                    const uint32_t t = 2 * i + c;
                    const uint32_t nf = pd.map_nf[t], nr = pd.map_nr[t], off = pd.map_off[t];
                    const uint32_t poolA = pool_of(b1.lens[i]);
                    const uint32_t poolB = pool_of(b1.lens[n + i]);
                    const uint8_t* aA = acc_of(pd.acc1, poolA);
                    const uint8_t* aB = acc_of(pd.acc2, poolB);
                    for (uint32_t xa = 0; xa < nf + nr; ++xa) {
                        const PeItem A = pd.map[off + xa];
                        if (A.j == PE_NONE || !aA[A.j]) continue;
                        const uint32_t nmA = b1.job_nm[poolA][A.j];
                        const uint64_t fA = xa < nf ? 1 : 0;
                        for (uint32_t ka = (aA[A.j] >> 1) & 1; ka < nmA; ++ka) {
                            const uint32_t ma = jm_at(b1, poolA, A.j, ka);
                            const uint64_t MA = mk_match(ma >> 8, ma & 0xFF, fA, A.k);
                            for (uint32_t xb = 0; xb < nf + nr; ++xb) {
                                const PeItem B = pd.srt[off + xb];
                                if (B.j == PE_NONE || !aB[B.j]) continue;
                                const uint32_t nmB = b2.job_nm[poolB][B.j];
                                const uint64_t fB = xb < nf ? 1 : 0;
                                for (uint32_t kb = (aB[B.j] >> 1) & 1; kb < nmB; ++kb) {
                                    const uint32_t mb = jm_at(b2, poolB, B.j, kb);
                                    const uint64_t MB = mk_match(mb >> 8, mb & 0xFF, fB, B.k);
                                    const int e = pe_extract(ix, MA, MB, maxd);
                                    if (e < 0) continue;
                                    if (e < best) {
                                        best = e;
                                        nbest = 1;
                                        b1m = MA;
                                        b2m = MB;
                                        orig = c == 0;
                                    } else if (e == best) {
                                        ++nbest;
                                    }
                                }
                            }
                        }
                    }
                }
                if (nbest == 0) {
                    atomicAdd(&s_cnt[2], 2u);
                    v1 = v2 = 0;
                } else if (nbest > 1) {
                    atomicAdd(&s_cnt[1], 2u);
                    v1 = v2 = 0;
                } else {
                    mat1 = b1m;
                    mat2 = b2m;
                    // computeMethLvl(r1.mat, r1.seq / revSeq1) and (r2.mat,
                    // revSeq2 / r2.seq) (:1603-1618)
                    use1 = orig ? 1 : 2;
                    use2 = orig ? 2 : 1;
                    if (get_stranded) atomicAdd(&s_cnt[orig ? 5 : 6], 1u);
                    const bool d1 = ix.win[(uint32_t)(b1m >> 32)].z == CPGDUMMY;
                    const bool d2 = ix.win[(uint32_t)(b2m >> 32)].z == CPGDUMMY;
                    if (d1 && d2) {
                        atomicAdd(&s_cnt[4], 2u);
                    } else if (d1 || d2) {
                        atomicAdd(&s_cnt[4], 1u);
                        atomicAdd(&s_cnt[0], 1u);
                    } else {
                        atomicAdd(&s_cnt[3], 1u);
                        atomicAdd(&s_cnt[0], 2u);
                    }
                    if (!d1) want1 = ((b1m >> 16) & 0xFF) == 0 ? 1 : 2;
                    if (!d2) want2 = ((b2m >> 16) & 0xFF) == 0 ? 1 : 2;
                }
            }
        }
        b1.valid[i] = v1;
        b1.valid[n + i] = v2;
        b1.mat[i] = mat1;
        b1.mat[n + i] = mat2;
        b1.use[i] = use1;
        b1.use[n + i] = use2;
    }
    {
        const int lane = threadIdx.x & 63;
        const uint64_t below = lanes_below(lane);
        for (int mate = 0; mate < 2; ++mate) {
            const int want = mate ? want2 : want1;
            const uint32_t r = mate ? pd.n + i : i;
            const uint64_t bdir = __ballot(want == 1), bdp = __ballot(want == 2);
            uint32_t basedir = 0, basedp = 0;
            if (lane == 0) {
                if (bdir) basedir = atomicAdd(b1.direct_n, (uint32_t)__popcll(bdir));
                if (bdp) basedp = atomicAdd(b1.dp_n, (uint32_t)__popcll(bdp));
            }
            basedir = lane_u32(basedir, 0);
            basedp = lane_u32(basedp, 0);
            if (want == 1) b1.direct_list[basedir + __popcll(bdir & below)] = r;
            if (want == 2) b1.dp_list[basedp + __popcll(bdp & below)] = r;
        }
    }
    __syncthreads();
    if (threadIdx.x < 5 && s_cnt[threadIdx.x])
        atomicAdd((unsigned long long*)&pd.counters[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
    if (threadIdx.x >= 5 && threadIdx.x < 7 && s_cnt[threadIdx.x])
        atomicAdd((unsigned long long*)&b1.counters[threadIdx.x - 2], (unsigned long long)s_cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_pe_seed(const DevIndex& ix, const BatchDev& bd, const PairDev& pd, uint32_t grid, hipStream_t st)
{
    if (!grid) return;
    if (pd.list) {
        if (ix.tab_is64) hipLaunchKernelGGL(k_pe_seed_big<uint64_t>, dim3(grid), dim3(64), 0, st, ix, bd, pd);
        else hipLaunchKernelGGL(k_pe_seed_big<uint32_t>, dim3(grid), dim3(64), 0, st, ix, bd, pd);
    } else {
        if (ix.tab_is64) hipLaunchKernelGGL(k_pe_seed<uint64_t>, dim3(grid), dim3(64), 0, st, ix, bd, pd);
        else hipLaunchKernelGGL(k_pe_seed<uint32_t>, dim3(grid), dim3(64), 0, st, ix, bd, pd);
    }
}

void launch_pe_walk1(const DevIndex& ix, const BatchDev& bd1, const BatchDev& bd2, const PairDev& pd, hipStream_t st)
{
    const uint32_t items = 2 * pd.n;
    if (!items) return;
    hipLaunchKernelGGL(k_pe_walk1, dim3((items + 255) / 256), dim3(256), 0, st, ix, bd1, bd2, pd);
    // the listed items: as many waves as fit at once (their count is on the device)
    hipLaunchKernelGGL(k_pe_walk1_big, dim3(std::min<uint32_t>(items, 1024)), dim3(64), 0, st, ix, bd1, bd2, pd);
}

void launch_pe_final(const DevIndex& ix, const BatchDev& bd1, const BatchDev& bd2, const PairDev& pd,
                     int get_stranded, hipStream_t st)
{
    if (!pd.n) return;
    hipLaunchKernelGGL(k_pe_final, dim3((pd.n + 255) / 256), dim3(256), 0, st, ix, bd1, bd2, pd, get_stranded);
}

} // namespace fame
