// Device helpers shared by the gather + rerank kernels (query.hip,
// lowprec.hip): block scans, radix-select digit search, candidate collection
// (bucket gather + bitset dedup + ordered compaction) and a wave-level
// running top-R over 64-bit keys.
#pragma once

#include "common.h"
#include "kernels.h"

namespace fpp {
namespace gather {

// ---- block / wave helpers -------------------------------------------------
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        const uint32_t a = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[T - 1];
    const uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

// Digit holding the k-th smallest key (1-based) of a 256-bin histogram.
template <int T>
__device__ __forceinline__ void block_find_digit(const uint32_t* hist, uint32_t k, uint32_t* sh, int& digit,
                                                 uint32_t& below, uint32_t& eq, int* s_digit, uint32_t* s_below) {
    static_assert(T >= 256, "needs >= 256 threads");
    if (threadIdx.x == 0) {
        *s_digit = 255;
        *s_below = 0;
    }
    const uint32_t h = threadIdx.x < 256 ? hist[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan_u32<T>(h, sh, total);
    if (threadIdx.x < 256 && ex < k && k <= ex + h) {
        *s_digit = threadIdx.x;
        *s_below = ex;
    }
    __syncthreads();
    digit = *s_digit;
    below = *s_below;
    eq = hist[digit];
    __syncthreads();
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

// Collect the unique candidates of one query, ascending, into cq (capacity
// ccap); returns their count.  Bucket ids are ORed into an LDS bitset window
// of W ids (the reference's boost::dynamic_bitset, src/FalconnPP.cpp:552-611);
// when N exceeds the window, per-probe cursors into the ascending bucket
// lists carry over between windows so every id is read once.
template <int TB>
__device__ uint32_t collect_candidates(const GatherArgs& a, const uint32_t* __restrict__ pr,
                                       const uint64_t* __restrict__ bucket_off, const int32_t* __restrict__ ids,
                                       uint32_t* bits, uint32_t* cur, uint32_t* sh, int32_t* __restrict__ cq,
                                       uint64_t ccap, uint64_t* ts) {
    const int64_t W = a.window_bits;
    const int Wwords = (int)(W / 32);
    const bool multi = a.N > W;
    uint32_t nc = 0;
    constexpr uint32_t DONE = 0xffffffffu;
    if (multi)
        for (int i = threadIdx.x; i < a.qProbes; i += TB) cur[i] = 0u;
    for (int64_t w0 = 0; w0 < a.N; w0 += W) {
        for (int i = threadIdx.x; i < Wwords; i += TB) bits[i] = 0u;
        __syncthreads();
        const int64_t wend = w0 + W;
        for (int i = threadIdx.x; i < a.qProbes; i += TB) {
            uint32_t cu = 0;
            if (multi) {
                cu = cur[i];
                if (cu == DONE) continue;
            }
            const uint32_t b = pr[i];
            const uint64_t s = bucket_off[b], e = bucket_off[b + 1];
            uint64_t j = s + cu;
            // four independent id loads in flight per step; a bucket's ids are
            // ascending, so the first id past the window ends this probe's turn
            while (j < e) {
                int32_t v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = (j + u < e) ? ids[j + u] : INT32_MAX;
                int taken = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (taken == u && (int64_t)v[u] < wend) {
                        const int64_t rel = (int64_t)v[u] - w0;
                        atomicOr(&bits[rel >> 5], 1u << (rel & 31));
                        taken = u + 1;
                    }
                }
                j += taken;
                if (taken < 4) break;
            }
            if (multi) cur[i] = (j >= e) ? DONE : (uint32_t)(j - s);
        }
        __syncthreads();
        if (ts && threadIdx.x == 0 && w0 == 0) ts[1] = wall_clock64();
        // ordered compaction: thread t scans a contiguous run of words
        const int per = (Wwords + TB - 1) / TB;
        const int wb = threadIdx.x * per;
        uint32_t c = 0;
        for (int i = 0; i < per; ++i)
            if (wb + i < Wwords) c += __popc(bits[wb + i]);
        uint32_t tot;
        uint32_t pos = nc + block_excl_scan_u32<TB>(c, sh, tot);
        for (int i = 0; i < per; ++i) {
            if (wb + i >= Wwords) break;
            uint32_t word = bits[wb + i];
            while (word) {
                const int bit = __ffs(word) - 1;
                word &= word - 1;
                if (pos < ccap) cq[pos] = (int32_t)(w0 + (int64_t)(wb + i) * 32 + bit);
                pos++;
            }
        }
        nc += tot;
        __syncthreads();
    }
    return nc;
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
