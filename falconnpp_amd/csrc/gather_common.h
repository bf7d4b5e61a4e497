// Device helpers shared by the gather + rerank kernels (query.hip,
// lowprec.hip): block scans, radix-select digit search, candidate collection
// (bucket gather + bitset dedup + ordered compaction) and a wave-level
// running top-R over 64-bit keys.
#pragma once

#include <utility>

#include "common.h"
#include "kernels.h"

namespace fpp {
namespace gather {

// ---- thread groups ---------------------------------------------------------
// The group-level helpers below run on a group of threads G: G::size threads,
// G::tid() in [0, size), G::sync() a barrier of the group that also orders
// its LDS and global memory accesses, G::lds_sync() one that orders its LDS
// accesses only (global loads and stores in flight stay in flight across
// it).  The gather kernels use the whole workgroup.
template <int TB>
struct Whole {
    static constexpr int size = TB;
    __device__ static __forceinline__ int tid() { return (int)threadIdx.x; }
    __device__ static __forceinline__ void sync() { __syncthreads(); }
    // barrier that also returns whether any thread's p was nonzero
    __device__ static __forceinline__ int sync_or(int p) { return __syncthreads_or(p); }
    __device__ static __forceinline__ void lds_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
};

// ---- block / wave helpers -------------------------------------------------
// LDS ordering inside one wave: a wave's LDS operations complete in program
// order, so a barrier that keeps the compiler from moving them is enough.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <class G>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* sh, uint32_t& total) {
    constexpr int T = G::size;
    const int t = G::tid();
    sh[t] = v;
    G::sync();
    for (int off = 1; off < T; off <<= 1) {
        const uint32_t a = t >= off ? sh[t - off] : 0u;
        G::sync();
        sh[t] += a;
        G::sync();
    }
    total = sh[T - 1];
    const uint32_t incl = sh[t];
    G::sync();
    return incl - v;
}

// Digit holding the k-th smallest key (1-based) of a 256-bin histogram.
template <class G>
__device__ __forceinline__ void block_find_digit(const uint32_t* hist, uint32_t k, uint32_t* sh, int& digit,
                                                 uint32_t& below, uint32_t& eq, int* s_digit, uint32_t* s_below) {
    static_assert(G::size >= 256, "needs >= 256 threads");
    const int t = G::tid();
    if (t == 0) {
        *s_digit = 255;
        *s_below = 0;
    }
    const uint32_t h = t < 256 ? hist[t] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan_u32<G>(h, sh, total);
    if (t < 256 && ex < k && k <= ex + h) {
        *s_digit = t;
        *s_below = ex;
    }
    G::sync();
    digit = *s_digit;
    below = *s_below;
    eq = hist[digit];
    G::sync();
}

// Wave variant: lane t owns bins 4t..4t+3.
__device__ __forceinline__ void wave_find_digit(const uint32_t* hist, uint32_t k, int& digit, uint32_t& below,
                                                uint32_t& eq) {
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = hist[lane * 4 + i];
    const uint32_t s = b[0] + b[1] + b[2] + b[3];
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t a = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += a;
    }
    const uint32_t ex = incl - s;
    int d = -1;
    uint32_t bl = 0, e = 0;
    if (ex < k && k <= incl) {
        uint32_t run = ex;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (d < 0 && run + b[i] >= k) {
                d = lane * 4 + i;
                bl = run;
                e = b[i];
            }
            run += b[i];
        }
    }
    const uint64_t who = __ballot(d >= 0);
    const int src = who ? (int)__ffsll((long long)who) - 1 : WAVE - 1;
    if (!who && lane == WAVE - 1) d = 255;
    digit = __shfl(d, src, WAVE);
    below = __shfl(bl, src, WAVE);
    eq = __shfl(e, src, WAVE);
}

__device__ __forceinline__ uint64_t lanes_below_mask() {
    const int lane = threadIdx.x & (WAVE - 1);
    return lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
}

// Wave-level inclusive prefix sum.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// Walk over the ids of the probed buckets pr[0 .. np), one group of 64 probes
// per step of a wave (wave v takes groups v, v + NW, ...); step(f) calls
// f(id, true) for every id of its group, with all lanes of the wave present
// (lanes without an id pass false), in no particular order.  The next group's
// bucket extents are fetched while the current ids are walked, and its
// probes one step before that, so a step waits for one round of loads.
//  * Lane per bucket (the longest of the 64 buckets holds at most LANE_MAX
//    ids): lane p reads bucket p's ids itself, ULANE at a time, so nothing
//    crosses lanes and a step is a load and f.  Lanes whose bucket is done
//    pass false until the longest one is.
//  * Flat (a longer bucket): the concatenation of the 64 lists is walked one
//    lane per id; a binary search over the wave's prefix sums of bucket
//    lengths maps a position to its probe (six dependent lane shuffles per
//    id, so only where lane per bucket would leave most lanes idle).

template <class G, int ULANE>
struct BucketWalk {
    static constexpr int NW = G::size / WAVE;
    static constexpr int UNR = 8;  // flat walk: ids per lane in flight
    static constexpr uint32_t LANE_MAX = 64;
    const uint32_t* __restrict__ pr;
    const uint64_t* __restrict__ bucket_off;
    const int32_t* __restrict__ ids;
    int np;
    int c0;         // first probe of this step's group
    uint64_t s;     // this lane's bucket in the group: first id, length
    uint32_t len;
    uint32_t nb;    // this lane's bucket in the next group

    __device__ __forceinline__ void init(const uint32_t* pr_, const uint64_t* off_, const int32_t* ids_, int np_) {
        pr = pr_;
        bucket_off = off_;
        ids = ids_;
        np = np_;
        const int lane = threadIdx.x & (WAVE - 1);
        c0 = (G::tid() / WAVE) * WAVE;
        s = 0;
        len = 0;
        const int cn = c0 + NW * WAVE;
        nb = cn + lane < np ? gld(pr + cn + lane) : 0u;
        if (c0 + lane < np) {
            const uint32_t b = gld(pr + c0 + lane);
            s = gld(bucket_off + b);
            len = (uint32_t)(gld(bucket_off + b + 1) - s);
        }
    }
    // wave-uniform
    __device__ __forceinline__ bool done() const { return c0 >= np; }

    // Issues the loads of the next group's bucket extents and of the probe
    // of the group after it; nothing waits for them until the next step.
    __device__ __forceinline__ void fetch_next(uint64_t& s_n, uint32_t& len_n, uint32_t& nb_n) const {
        const int lane = threadIdx.x & (WAVE - 1);
        const int cn = c0 + NW * WAVE;
        const int cnn = cn + NW * WAVE;
        s_n = 0;
        len_n = 0;
        if (cn + lane < np) {
            s_n = gld(bucket_off + nb);
            len_n = (uint32_t)(gld(bucket_off + nb + 1) - s_n);
        }
        nb_n = cnn + lane < np ? gld(pr + cnn + lane) : 0u;
    }

    template <class F>
    __device__ __forceinline__ void step(F&& f) {
        const int lane = threadIdx.x & (WAVE - 1);
        uint32_t mx = len;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, WAVE));
        mx = __builtin_amdgcn_readfirstlane(mx);
        uint64_t s_n;
        uint32_t len_n, nb_n;
        if (mx == 0) {
            fetch_next(s_n, len_n, nb_n);
        } else if (mx <= LANE_MAX) {
            // four ids per load (the id array is padded past its end), so a
            // step issues a quarter of the load instructions, each touching
            // one line per lane; the first round's loads go out before the
            // next group's extents, so f waits only for its own ids
            static_assert(ULANE % 4 == 0, "ids are read four at a time");
            for (uint32_t j0 = 0;; j0 += ULANE) {
                int32_t v[ULANE];
#pragma unroll
                for (int r = 0; r < ULANE; r += 4) {
                    i32x4_a4 t = {0, 0, 0, 0};
                    if (j0 + r < len) t = gld_i32x4(ids + s + j0 + r);
                    v[r] = t[0];
                    v[r + 1] = t[1];
                    v[r + 2] = t[2];
                    v[r + 3] = t[3];
                }
                if (j0 == 0) fetch_next(s_n, len_n, nb_n);
#pragma unroll
                for (int r = 0; r < ULANE; ++r) f(v[r], j0 + r < len);
                if (j0 + ULANE >= mx) break;
            }
        } else {
            fetch_next(s_n, len_n, nb_n);
            const uint32_t incl = wave_incl_scan(len);
            const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
            const uint64_t sx = s - (uint64_t)(incl - len);  // ids index of flat position f is sx + f
            for (uint32_t f0 = 0; f0 < total; f0 += WAVE * UNR) {
                int32_t v[UNR];
#pragma unroll
                for (int r = 0; r < UNR; ++r) {
                    const uint32_t fp = f0 + r * WAVE + lane;
                    // probe of position fp: number of lanes whose inclusive sum is <= fp
                    int lo = 0;
#pragma unroll
                    for (int st = WAVE / 2; st > 0; st >>= 1)
                        if (__shfl(incl, lo + st - 1, WAVE) <= fp) lo += st;
                    const uint64_t base = __shfl(sx, lo, WAVE);
                    v[r] = fp < total ? gld(ids + base + fp) : 0;
                }
#pragma unroll
                for (int r = 0; r < UNR; ++r) f(v[r], f0 + r * WAVE + lane < total);
            }
        }
        c0 += NW * WAVE;
        s = s_n;
        len = len_n;
        nb = nb_n;
    }
};

// Every id of the probed buckets pr[0 .. np) through f (BucketWalk, all groups).
template <class G, int ULANE, class F>
__device__ __forceinline__ void for_each_bucket_id(const uint32_t* __restrict__ pr,
                                                   const uint64_t* __restrict__ bucket_off,
                                                   const int32_t* __restrict__ ids, int np, F&& f) {
    BucketWalk<G, ULANE> w;
    w.init(pr, bucket_off, ids, np);
    while (!w.done()) w.step(f);
}
// Next query of a persistent gather workgroup, or -1 when every run of the
// work queue is drained.  order[0 .. n) (identity when null; n = nitems less
// the small queries at the end when small_at is set) is split
// into nqueues runs of consecutive positions; workgroup b takes from run
// b % nqueues first (with 8 runs, the XCD the workgroup sits on), then from
// the others.  Every successful claim consumes an item, so each workgroup
// leaves the loop once the items are gone.  claim_item is the claim itself,
// made by one thread; claim_query makes it with thread 0 and hands the result
// to the whole workgroup through s_q, a __shared__ word.
__device__ __forceinline__ int claim_item(const GatherArgs& a) {
    int q = -1;
    const uint32_t R = (uint32_t)a.nqueues;
    const uint32_t home = blockIdx.x % R;
    // the small queries at the end of the order are not this kernel's
    const uint32_t n =
        a.small_at >= 0
            ? a.nitems - __hip_atomic_load(&a.heads[a.small_at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
            : a.nitems;
    for (uint32_t j = 0; j < R && q < 0; ++j) {
        const uint32_t r = (home + j) % R;
        const uint32_t lo = (uint32_t)((uint64_t)n * r / R);
        const uint32_t hi = (uint32_t)((uint64_t)n * (r + 1) / R);
        if (lo + __hip_atomic_load(&a.heads[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi) continue;
        const uint32_t i = atomicAdd(&a.heads[r], 1u);
        if (lo + i < hi) q = a.order ? (int)a.order[lo + i] : (int)(lo + i);
    }
    return q;
}

__device__ __forceinline__ int claim_query(const GatherArgs& a, int* s_q) {
    if (threadIdx.x == 0) *s_q = claim_item(a);
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(*s_q);
    __syncthreads();
    return q;
}

// Scratch of the query a workgroup serves: its candidate list and its
// window / score area, either the workgroup's own slot or, when the caller
// keeps every query's candidates, the query's segment.
struct QueryScratch {
    int32_t* cand;
    int32_t* part;
    uint64_t ccap, pcap;
};
__device__ __forceinline__ QueryScratch query_scratch(const GatherArgs& a, int q) {
    QueryScratch s;
    if (a.seg_off) {
        const uint64_t b = a.seg_off[q], e = a.seg_off[q + 1];
        s.cand = a.cand + b;
        s.part = a.part + b;
        s.ccap = s.pcap = e - b;
    } else {
        s.cand = a.cand + (uint64_t)blockIdx.x * a.slot_cand;
        s.part = a.part + (uint64_t)blockIdx.x * a.slot_part;
        s.ccap = a.slot_cand;
        s.pcap = a.slot_part;
    }
    return s;
}

// Rank of this lane among the set lanes of m.
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Ordered compaction of one window's bitset (words words, ids from w0) into
// cq[nc ...]: thread t owns the contiguous run of r words from t * r (r odd,
// so at every step the lanes of a wave read 32 different LDS banks).  It
// counts its bits; a wave scan and the waves' totals give its first output
// position, and it writes its ids in order.  red holds G::size / 64 words.
// Returns the new count.
template <class G>
__device__ __forceinline__ uint32_t compact_window(const uint32_t* bits, int words, int64_t w0, uint32_t nc,
                                                   uint32_t* red, int32_t* __restrict__ cq, uint64_t ccap) {
    constexpr int NW = G::size / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = G::tid() / WAVE;
    const int r = (int)ceil_div(words, G::size) | 1;
    const int b = min(words, G::tid() * r);
    const int e = min(words, b + r);
    uint32_t c = 0;
    for (int i = b; i < e; ++i) c += __popc(bits[i]);
    const uint32_t incl = wave_incl_scan(c);
    if (lane == WAVE - 1) red[wv] = incl;
    G::sync();
    uint32_t base = nc + incl - c, tot = 0;
#pragma unroll
    for (int v = 0; v < NW; ++v) {
        const uint32_t x = red[v];
        base += v < wv ? x : 0u;
        tot += x;
    }
    for (int i = b; i < e; ++i) {
        uint32_t word = bits[i];
        while (word) {
            const int bit = __ffs(word) - 1;
            word &= word - 1;
            if (base < ccap) gst(cq + base, (int32_t)(w0 + (int64_t)i * 32 + bit));
            base++;
        }
    }
    G::sync();
    return nc + tot;
}

// Appends v of the lanes with take set to cq at *count (LDS), one atomic per
// wave.  Must be reached by the whole wave.  GLOBAL: cq is in global memory
// (stored with global_store); otherwise cq may be anywhere (LDS included).
template <bool GLOBAL>
__device__ __forceinline__ void wave_append(bool take, int32_t v, uint32_t* count, int32_t* __restrict__ cq,
                                            uint64_t ccap) {
    const uint64_t m = __ballot(take);
    if (m == 0ull) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int leader = (int)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = lds_add(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, WAVE);
    const uint32_t r = lane_rank(m);
    if (take && base + r < ccap) {
        if constexpr (GLOBAL)
            gst(cq + base + r, v);
        else
            cq[base + r] = v;
    }
}

// wave_append of N (take, v) pairs per lane with one atomic per wave: the
// pairs of round r follow those of rounds < r.  cq is in global memory.
template <int N>
__device__ __forceinline__ void wave_append_n(const bool (&take)[N], const int32_t (&v)[N], uint32_t* count,
                                              int32_t* __restrict__ cq, uint64_t ccap) {
    uint64_t m[N];
    uint32_t pre[N];
    uint32_t tot = 0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        m[r] = __ballot(take[r]);
        pre[r] = tot;
        tot += (uint32_t)__popcll(m[r]);
    }
    if (tot == 0) return;
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t base = 0;
    if (lane == 0) base = lds_add(count, tot);
    base = __shfl(base, 0, WAVE);
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const uint32_t pos = base + pre[r] + lane_rank(m[r]);
        if (take[r] && pos < ccap) gst(cq + pos, v[r]);
    }
}

// Test-and-set of id rel in the LDS bitset: true when the bit was clear.
__device__ __forceinline__ bool bitset_insert(uint32_t* bits, int64_t rel) {
    const uint32_t m = 1u << (rel & 31);
    return (lds_or(&bits[rel >> 5], m) & m) == 0u;
}

// Insert of id v into an LDS open-addressing set of 2^lg words (word 0 =
// empty, id v stored as v + 1, linear probing): true when v was absent.  The
// set is sized to at least twice the ids offered, so a free word is always
// found; the probe count is bounded all the same.
__device__ __forceinline__ bool hashset_insert(uint32_t* T, int lg, int32_t v) {
    const uint32_t mask = (1u << lg) - 1u;
    const uint32_t key = (uint32_t)v + 1u;
    uint32_t h = ((uint32_t)v * 0x9E3779B1u) >> (32 - lg);
    for (uint32_t n = 0; n <= mask; ++n) {
        const uint32_t old = lds_cas(&T[h], 0u, key);
        if (old == 0u) return true;
        if (old == key) return false;
        h = (h + 1u) & mask;
    }
    return false;
}

// Zeroes words [0, n) of the dedup area.
template <class G>
__device__ __forceinline__ void clear_words(uint32_t* bits, uint32_t n) {
    for (uint32_t i = (uint32_t)G::tid(); i < n; i += G::size) lds_st(bits + i, 0u);
}

// Collect the unique candidates of one query into cq (capacity ccap, at
// least min(C_tot, N)); returns their count.  Bucket ids go through the LDS
// dedup area (the reference's boost::dynamic_bitset,
// src/FalconnPP.cpp:552-611), whose first `dirty` words may hold anything on
// entry (other phases reuse them); each path zeroes what it needs first and
// updates dirty.
//  * C_tot (ctot_q, duplicates included) at most a quarter of the area: one
//    walk over the bucket ids through an open-addressing set of 2^lg >=
//    2 C_tot words, first occurrences appended, then those words zeroed.
//    Cost follows C_tot, not N.
//  * N fits one window: ids are ORed in and the set bits compacted in id
//    order, so cq is ascending; the area is left dirty.
//  * Otherwise (W is then a power of two): the bucket ids are scattered by
//    window into part (capacity pcap, at least C_tot: duplicates
//    included): by one walk into a bin of pcap / nwin ids per window, or,
//    when a window overflows its bin, into segments laid out by a first,
//    counting walk; each window's ids go through a test-and-set, first
//    occurrences appended, and the words it touched are zeroed again.
// Callers rank by (score, id), which does not depend on the order of cq.
// wc holds window_counter_words(N, W) words, sh at least G::size words.
template <class G>
__device__ uint32_t collect_candidates(const GatherArgs& a, int np, const uint32_t* __restrict__ pr,
                                       const uint64_t* __restrict__ bucket_off, const int32_t* __restrict__ ids,
                                       uint32_t* bits, uint32_t* wc, uint32_t* sh, int32_t* __restrict__ cq,
                                       int32_t* __restrict__ part, uint64_t ccap, uint64_t pcap, uint64_t* ts,
                                       uint64_t ctot_q, uint32_t& dirty) {
    bits = lds_ptr(bits);
    wc = lds_ptr(wc);
    sh = lds_ptr(sh);
    const int64_t W = a.window_bits;
    const uint32_t area = (uint32_t)(W / 32);
    // the largest power of two words in the area
    const int area_lg = 31 - __clz((int)area);
    if (ctot_q * 4 <= ((uint64_t)1 << area_lg)) {
        int lg = 1;
        while (((uint64_t)1 << lg) < 2 * ctot_q) ++lg;
        const uint32_t words = 1u << lg;
        clear_words<G>(bits, min(dirty, words));
        uint32_t* count = sh;
        if (G::tid() == 0) lds_st(count, 0u);
        G::sync();
        for_each_bucket_id<G, 24>(pr, bucket_off, ids, np, [&](int32_t v, bool ok) {
            wave_append<true>(ok && hashset_insert(bits, lg, v), v, count, cq, ccap);
        });
        G::sync();
        if (ts && G::tid() == 0) ts[1] = wall_clock64();
        const uint32_t nc = lds_ld(count);
        clear_words<G>(bits, words);
        dirty = dirty > words ? dirty : 0u;
        G::sync();
        return nc;
    }
    if (a.N <= W) {
        const int words = (int)(W / 32);
        clear_words<G>(bits, min(dirty, (uint32_t)words));
        G::sync();
        // 24 ids per lane in flight: one step covers a group of buckets of
        // the usual sizes (the bucket filter keeps them near bucketLimit)
        for_each_bucket_id<G, 24>(pr, bucket_off, ids, np, [&](int32_t v, bool ok) {
            if (ok) lds_or(&bits[v >> 5], 1u << (v & 31));
        });
        G::sync();
        if (ts && G::tid() == 0) ts[1] = wall_clock64();
        dirty = max(dirty, (uint32_t)words);
        return compact_window<G>(bits, words, 0, 0u, sh, cq, ccap);
    }
    const int nwin = (int)window_count(a.N, W);
    const int shift = __ffsll((long long)W) - 1;
    uint32_t* wcnt = wc;
    uint32_t* woff = wc + nwin;
    uint32_t* wfill = woff + nwin;
    for (int i = G::tid(); i < nwin; i += G::size) {
        wcnt[i] = 0u;
        wfill[i] = 0u;
    }
    G::sync();
    // One walk scatters the ids straight into a bin per window: C_tot ids
    // each when part holds that (no window can overflow; the bins stay as
    // close together as the query allows), else pcap / nwin.  A window that
    // overflows its bin sends the query to the counted layout: a first walk
    // counts the windows and a second one scatters into segments laid out by
    // the counts.
    const uint64_t capw = min<uint64_t>(ctot_q, pcap / (uint64_t)nwin);
    bool bins = capw >= 1;
    if (bins) {
        const uint32_t cap = (uint32_t)min<uint64_t>(capw, 0xffffffffu / (uint64_t)nwin);
        for_each_bucket_id<G, 24>(pr, bucket_off, ids, np, [&](int32_t v, bool ok) {
            if (ok) {
                const int w = v >> shift;
                const uint32_t pos = lds_add(&wfill[w], 1u);
                if (pos < cap) gst(part + (uint64_t)w * cap + pos, v);
            }
        });
        G::sync();
        int over = 0;
        for (int i = G::tid(); i < nwin; i += G::size) over |= wfill[i] > cap ? 1 : 0;
        bins = G::sync_or(over) == 0;
        for (int i = G::tid(); i < nwin; i += G::size) {
            if (bins) {
                wcnt[i] = wfill[i];
                woff[i] = (uint32_t)i * cap;
            } else {
                wfill[i] = 0u;
            }
        }
        G::sync();
    }
    if (!bins) {
        // 24 ids per lane in flight (a bucket of the usual size in one round
        // of loads), as in the other paths
        for_each_bucket_id<G, 24>(pr, bucket_off, ids, np, [&](int32_t v, bool ok) {
            if (ok) lds_add(&wcnt[v >> shift], 1u);
        });
        G::sync();
        {
            // exclusive scan of the window counts: thread t sums a contiguous run
            const int per = (nwin + G::size - 1) / G::size;
            const int b0 = G::tid() * per;
            uint32_t c = 0;
            for (int i = 0; i < per; ++i)
                if (b0 + i < nwin) c += wcnt[b0 + i];
            uint32_t tot;
            uint32_t run = block_excl_scan_u32<G>(c, sh, tot);
            for (int i = 0; i < per; ++i)
                if (b0 + i < nwin) {
                    woff[b0 + i] = run;
                    run += wcnt[b0 + i];
                }
        }
        G::sync();
        for_each_bucket_id<G, 24>(pr, bucket_off, ids, np, [&](int32_t v, bool ok) {
            if (ok) {
                const int w = v >> shift;
                const uint32_t pos = lds_ld(woff + w) + lds_add(&wfill[w], 1u);
                if (pos < pcap) gst(part + pos, v);
            }
        });
    }
    uint32_t* count = sh;  // appended ids
    if (G::tid() == 0) lds_st(count, 0u);
    clear_words<G>(bits, min(dirty, area));
    dirty = 0u;
    G::sync();
    if (ts && G::tid() == 0) ts[1] = wall_clock64();
    // Each window's segment of part in rounds of UNR ids per thread, the
    // rounds of all windows in one sequence.  Two rounds are in flight while
    // one goes through the bitset: a round's loads are issued two rounds
    // before their ids are inserted, so a window's wait for its bin spans two
    // rounds of work rather than one.  The three buffers rotate through an
    // unrolled loop (a register copy would wait for the loads).
    // 4096 ids a round at every workgroup size up to 512 threads
    constexpr int UNR = G::size <= 256 ? 16 : 8;
    constexpr uint32_t RND = (uint32_t)G::size * UNR;
    struct Round {
        int w;        // window (nwin: past the last one)
        uint32_t i0;  // first index of the round in the window's segment
    };
    auto next_window = [&](int w) {
        ++w;
        while (w < nwin && lds_ld(wcnt + w) == 0u) ++w;
        return w;
    };
    auto after = [&](Round c) -> Round {
        if (c.w >= nwin) return c;
        if (c.i0 + RND < lds_ld(wcnt + c.w)) return Round{c.w, c.i0 + RND};
        return Round{next_window(c.w), 0u};
    };
    // ids of round c, -1 where there is none (every round past the last
    // window issues no load)
    auto issue = [&](int32_t (&dst)[UNR], Round c) {
        const bool live = c.w < nwin;
        const uint32_t cnt = live ? lds_ld(wcnt + c.w) : 0u, base = live ? lds_ld(woff + c.w) : 0u;
#pragma unroll
        for (int r = 0; r < UNR; ++r) {
            const uint32_t i = c.i0 + r * G::size + G::tid();
            dst[r] = (i < cnt && base + i < pcap) ? gld(part + base + i) : -1;
        }
    };
    auto process = [&](Round c, const int32_t (&v)[UNR]) {
        const uint32_t cnt = lds_ld(wcnt + c.w);
        const int64_t w0 = (int64_t)c.w << shift;
        bool fresh[UNR];
#pragma unroll
        for (int r = 0; r < UNR; ++r) fresh[r] = v[r] >= 0 && bitset_insert(bits, (int64_t)v[r] - w0);
        wave_append_n<UNR>(fresh, v, count, cq, ccap);
        if (c.i0 + RND < cnt) return;
        // the window's last round.  The windows share only LDS: the appends
        // to cq and the next rounds' loads stay in flight across these
        // barriers
        G::lds_sync();
        // zero what the window set: from the ids of its only round, else the
        // whole window
        if (cnt <= RND) {
#pragma unroll
            for (int r = 0; r < UNR; ++r)
                if (v[r] >= 0) lds_st(bits + ((v[r] - w0) >> 5), 0u);
        } else {
            clear_words<G>(bits, (uint32_t)((min(W, a.N - w0) + 31) / 32));
        }
        G::lds_sync();
    };
    Round r0{next_window(-1), 0u};
    Round r1 = after(r0), r2 = after(r1);
    int32_t b0[UNR], b1[UNR], b2[UNR];
    issue(b0, r0);
    issue(b1, r1);
    while (r0.w < nwin) {
        issue(b2, r2);
        process(r0, b0);
        r0 = after(r2);
        if (r1.w >= nwin) break;
        issue(b0, r0);
        process(r1, b1);
        r1 = after(r0);
        if (r2.w >= nwin) break;
        issue(b1, r1);
        process(r2, b2);
        r2 = after(r1);
    }
    // every wave's appends (count and cq) are ordered before this read by
    // the last window's lds_sync (count was zeroed before a full barrier when
    // there is no window at all)
    const uint32_t nc = lds_ld(count);
    G::sync();
    return nc;
}

// Radix select of the want smallest (key, id) pairs among n entries, entry i
// being (key_at(i), id_at(i)) with unique ids; writes their ids to out_id in
// unspecified order and returns min(want, n).  Four 8-bit passes over the
// key; when more entries share the boundary key than are needed, a second
// radix select over their ids keeps the smallest.  hist: 256 words, sh:
// G::size words, count: one word, out_id: want words, all in LDS.
template <class G, class KeyAt, class IdAt>
__device__ __forceinline__ uint32_t radix_select_smallest(KeyAt key_at, IdAt id_at, uint32_t n, uint32_t want,
                                                          uint32_t* hist, uint32_t* sh, int* s_digit,
                                                          uint32_t* s_below, uint32_t* count, int32_t* out_id) {
    const uint32_t kk = min(want, n);
    if (kk == 0) return 0;
    uint32_t prefix = 0, pmask = 0, k = kk, eqc = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = G::tid(); i < 256; i += G::size) hist[i] = 0;
        G::sync();
        for (uint32_t i = (uint32_t)G::tid(); i < n; i += G::size) {
            const uint32_t key = key_at(i);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        G::sync();
        int d;
        uint32_t below;
        block_find_digit<G>(hist, k, sh, d, below, eqc, s_digit, s_below);
        prefix |= (uint32_t)d << shift;
        pmask |= 255u << shift;
        k -= below;
    }
    const uint32_t T = prefix;
    uint32_t IT = 0xffffffffu;
    if (eqc > k) {
        uint32_t ip = 0, im = 0, kr = k;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int i = G::tid(); i < 256; i += G::size) hist[i] = 0;
            G::sync();
            for (uint32_t i = (uint32_t)G::tid(); i < n; i += G::size) {
                if (key_at(i) != T) continue;
                const uint32_t id = (uint32_t)id_at(i);
                if ((id & im) == ip) atomicAdd(&hist[(id >> shift) & 255u], 1u);
            }
            G::sync();
            int d;
            uint32_t below, eqd;
            block_find_digit<G>(hist, kr, sh, d, below, eqd, s_digit, s_below);
            ip |= (uint32_t)d << shift;
            im |= 255u << shift;
            kr -= below;
        }
        IT = ip;
    }
    if (G::tid() == 0) *count = 0u;
    G::sync();
    for (uint32_t c0 = 0; c0 < n; c0 += G::size) {
        const uint32_t i = c0 + G::tid();
        bool take = false;
        int32_t id = 0;
        if (i < n) {
            const uint32_t key = key_at(i);
            if (key <= T) {
                id = id_at(i);
                take = key < T || (uint32_t)id <= IT;
            }
        }
        wave_append<false>(take, id, count, out_id, kk);
    }
    G::sync();
    return kk;
}

// radix_select_smallest by one wave over entries it holds in registers:
// lane l's slot e is entry e * 64 + l when ok[e].  hist: 256 words of LDS of
// this wave, count: one LDS word, out_id: LDS.  Returns the number written,
// min(want, entries).
template <int E>
__device__ __forceinline__ uint32_t wave_select_smallest(const uint32_t (&key)[E], const int32_t (&id)[E],
                                                         const bool (&ok)[E], uint32_t n, uint32_t want,
                                                         uint32_t* hist, uint32_t* count, int32_t* out_id) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t kk = min(want, n);
    if (kk == 0) return 0;
    // one radix select over 32-bit values v(e) of the entries with sel(e)
    auto radix = [&](auto&& val, auto&& sel, uint32_t k, uint32_t& eq) {
        uint32_t prefix = 0, pmask = 0;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int i = lane; i < 256; i += WAVE) lds_st(hist + i, 0u);
            wave_lds_sync();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t v = val(e);
                if (ok[e] && sel(e) && (v & pmask) == prefix) lds_add(hist + ((v >> shift) & 255u), 1u);
            }
            wave_lds_sync();
            int d;
            uint32_t below;
            wave_find_digit(hist, k, d, below, eq);
            wave_lds_sync();
            prefix |= (uint32_t)d << shift;
            pmask |= 255u << shift;
            k -= below;
        }
        return std::make_pair(prefix, k);
    };
    uint32_t eqc = 0;
    const auto tk = radix([&](int e) { return key[e]; }, [&](int) { return true; }, kk, eqc);
    const uint32_t T = tk.first, k = tk.second;
    uint32_t IT = 0xffffffffu;
    if (eqc > k) {
        uint32_t eqd = 0;
        IT = radix([&](int e) { return (uint32_t)id[e]; }, [&](int e) { return key[e] == T; }, k, eqd).first;
    }
    if (lane == 0) lds_st(count, 0u);
    wave_lds_sync();
#pragma unroll
    for (int e = 0; e < E; ++e)
        wave_append<false>(ok[e] && (key[e] < T || (key[e] == T && (uint32_t)id[e] <= IT)), id[e], count, out_id, kk);
    wave_lds_sync();
    return kk;
}

// Selects the want smallest (key, id) pairs among keys[0 .. n), where the id
// of position i is cq[i] (ids unique), and writes their ids to out_id in
// unspecified order; returns min(want, n).  The keys are order keys of
// scores (desc_key_any).  A radix select over all n keys is slow here: the
// scores of one query cross zero and share few exponents, so the passes
// over key bits pile their LDS atomics into a handful of bins.  Instead one
// histogram over 256 bins spread linearly over the query's score range
// finds the bin holding the want-th best; the keys of that bin and the
// better ones (usually a few hundred) are copied to lst (global memory, room
// for lcap pairs) and the radix select runs over those.  A selection whose
// copy would not fit lcap runs over all keys.  kmin and kmax are the smallest
// and largest key (the caller tracks them while it writes the keys).  hist:
// 256 words, sh: G::size words, count: one word, out_id: want words, all in
// LDS.  KEYS: 0, the keys in LDS; 2, in global memory; 1, in global memory
// with their top 16 bits in LDS (keys16): the histogram and the copy read
// those (a key's top bits order like the key, so its bin is still monotone)
// and only the copied keys are read whole.
template <class G, int KEYS>
__device__ uint32_t select_smallest(const uint32_t* keys, const uint16_t* keys16, const int32_t* __restrict__ cq,
                                    uint32_t n, uint32_t want, uint32_t* hist, uint32_t* sh, int* s_digit,
                                    uint32_t* s_below, uint32_t* count, int32_t* out_id, int32_t* lst, uint32_t lcap,
                                    uint32_t kmin, uint32_t kmax) {
    if constexpr (KEYS == 0) keys = lds_ptr(keys);
    auto key_at = [&](uint32_t i) {
        if constexpr (KEYS == 0)
            return keys[i];
        else
            return gld(keys + i);
    };
    // the key, or with KEYS 1 its top 16 bits, for the bins
    auto coarse_at = [&](uint32_t i) {
        if constexpr (KEYS == 1)
            return (uint32_t)lds_ld(keys16 + i) << 16;
        else
            return key_at(i);
    };
    auto cq_at = [&](uint32_t i) { return gld(cq + i); };
    out_id = lds_ptr(out_id);
    hist = lds_ptr(hist);
    sh = lds_ptr(sh);
    s_digit = lds_ptr(s_digit);
    s_below = lds_ptr(s_below);
    count = lds_ptr(count);
    const uint32_t kk = min(want, n);
    if (kk == 0) return 0;
    if (n <= lcap || lcap < 2 * kk)
        return radix_select_smallest<G>(key_at, cq_at, n, want, hist, sh, s_digit, s_below, count, out_id);
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = G::tid() / WAVE;
    // the keys in rounds of U per thread, all read before the first is used
    // (the passes would otherwise wait out a load's latency per key);
    // f(i, key, valid) is reached by the whole group the same number of times
    constexpr int U = 8;
    auto for_keys = [&](auto&& f) {
        for (uint32_t c0 = 0; c0 < n; c0 += (uint32_t)U * G::size) {
            uint32_t kv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = c0 + (uint32_t)(u * G::size + G::tid());
                kv[u] = i < n ? coarse_at(i) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = c0 + (uint32_t)(u * G::size + G::tid());
                f(i, kv[u], i < n);
            }
        }
    };
    // [kmin, kmax]: the caller's range of the keys, best (smallest) to worst
    // (with KEYS 1 the best one's top bits, so that no bin is below 0)
    const float vbest = desc_key_any_value(KEYS == 1 ? kmin & 0xffff0000u : kmin), vworst = desc_key_any_value(kmax);
    const float range = vbest - vworst;
    const float scale = 256.0f / range;
    if (!(range > 0.0f) || !(scale < __builtin_inff()))
        return radix_select_smallest<G>(key_at, cq_at, n, want, hist, sh, s_digit, s_below, count, out_id);
    // bin of a key: non-decreasing in the key (rounding is monotone), so the
    // keys up to the want-th smallest all lie in bins up to its bin
    auto bin_of = [&](uint32_t key) -> uint32_t {
        const float x = (vbest - desc_key_any_value(key)) * scale;
        return x > 0.0f ? (x < 255.0f ? (uint32_t)x : 255u) : 0u;
    };
    for (int i = G::tid(); i < 256; i += G::size) hist[i] = 0;
    G::sync();
    for_keys([&](uint32_t, uint32_t key, bool ok) {
        if (ok) atomicAdd(&hist[bin_of(key)], 1u);
    });
    G::sync();
    // the bin of the want-th best, by one wave
    if (wv == 0) {
        int dd;
        uint32_t bb, ee;
        wave_find_digit(hist, kk, dd, bb, ee);
        if (lane == 0) {
            *s_digit = dd;
            sh[0] = bb + ee;
        }
    }
    G::sync();
    const int d = *s_digit;
    const uint32_t m = sh[0];
    G::sync();
    if (m > lcap)
        return radix_select_smallest<G>(key_at, cq_at, n, want, hist, sh, s_digit, s_below, count, out_id);
    // the keys of bins <= d with their ids, in lst[0 .. m) and lst[lcap ..)
    uint32_t* lkey = reinterpret_cast<uint32_t*>(lst);
    int32_t* lid = lst + lcap;
    if (G::tid() == 0) *count = 0u;
    G::sync();
    for_keys([&](uint32_t i, uint32_t key, bool ok) {
        const bool take = ok && bin_of(key) <= (uint32_t)d;
        const uint64_t bm = __ballot(take);
        if (bm == 0ull) return;
        const int leader = (int)__ffsll((long long)bm) - 1;
        uint32_t base = 0;
        if (lane == leader) base = lds_add(count, (uint32_t)__popcll(bm));
        base = __shfl(base, leader, WAVE);
        const uint32_t pos = base + lane_rank(bm);
        if (take && pos < m) {
            gst(lkey + pos, KEYS == 1 ? key_at(i) : key);
            gst(lid + pos, cq_at(i));
        }
    });
    G::sync();
    // a list of up to 64 * LE entries: one wave selects from registers,
    // with no workgroup barrier per pass
    constexpr int LE = 16;
    if (m <= (uint32_t)(WAVE * LE)) {
        if (wv == 0) {
            uint32_t key[LE];
            int32_t id[LE];
            bool ok[LE];
#pragma unroll
            for (int e = 0; e < LE; ++e) {
                const uint32_t i = (uint32_t)(e * WAVE + lane);
                ok[e] = i < m;
                key[e] = ok[e] ? gld(lkey + i) : 0u;
                id[e] = ok[e] ? gld(lid + i) : 0;
            }
            wave_select_smallest<LE>(key, id, ok, m, want, hist, count, out_id);
        }
        G::sync();
        return kk;
    }
    return radix_select_smallest<G>([&](uint32_t i) { return gld(lkey + i); }, [&](uint32_t i) { return gld(lid + i); },
                                    m, want, hist, sh, s_digit, s_below, count, out_id);
}

// Wave-level running top-R (R <= 64*KPL) over unique 64-bit keys, smaller is
// better.  Entry i = s*64 + lane lives in w[s] of that lane, ascending in i.
// insert() takes a wave-uniform key; callers test key < thr first, so a miss
// costs one compare.
template <int KPL>
struct WaveTopK {
    uint64_t w[KPL];
    uint64_t thr;
    int R;
    __device__ __forceinline__ void init(int r) {
#pragma unroll
        for (int s = 0; s < KPL; ++s) w[s] = ~0ull;
        thr = ~0ull;
        R = r;
    }
    __device__ __forceinline__ void insert(uint64_t nk) {
        const int lane = threadIdx.x & (WAVE - 1);
        int pos = 0;
#pragma unroll
        for (int s = 0; s < KPL; ++s) pos += (int)__popcll(__ballot(w[s] < nk));
        uint64_t up[KPL];
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            const uint64_t u = __shfl_up(w[s], 1, WAVE);
            const uint64_t carry = s > 0 ? __shfl(w[s > 0 ? s - 1 : 0], WAVE - 1, WAVE) : ~0ull;
            up[s] = lane == 0 ? carry : u;
        }
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            const int i = s * WAVE + lane;
            w[s] = i > pos ? up[s] : (i == pos ? nk : w[s]);
        }
        const int st = (R - 1) / WAVE, lt = (R - 1) % WAVE;
        uint64_t t = ~0ull;
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            const uint64_t v = __shfl(w[s], lt, WAVE);
            if (s == st) t = v;
        }
        thr = t;
    }
    // Offer the candidates held by the lanes in `leaders` (one key per
    // leader lane, same key across its group).
    __device__ __forceinline__ void offer(uint64_t key, uint64_t leaders) {
        uint64_t better = __ballot(key < thr) & leaders;
        while (better) {
            const int src = (int)__ffsll((long long)better) - 1;
            better &= better - 1;
            const uint64_t nk = __shfl(key, src, WAVE);
            if (nk < thr) insert(nk);
        }
    }
};

__device__ __forceinline__ uint64_t score_key(float s, int32_t id) {
    return id < 0 ? ~0ull : (((uint64_t)desc_key_any(s) << 32) | (uint32_t)id);
}

}  // namespace gather
}  // namespace fpp
