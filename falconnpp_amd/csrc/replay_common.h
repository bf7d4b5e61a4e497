// Device code of the exact replay of a query's probing sequence (the
// reference's heaps, src/FalconnPP.cpp:358-582), shared by the replay kernels
// (query.hip) and the gather kernels that replay a flagged query before
// gathering it (query.hip, lowprec.hip).
#pragma once

#include "common.h"
#include "gather_common.h"
#include "heap.h"
#include "kernels.h"

namespace fpp {
namespace {

using namespace gather;

// ---- exact replay of a flagged query's probing sequence -------------------

// Rotation of the query for one layer into x (F floats of the wave's LDS):
// three rounds of sign flips and the unnormalised FWHT with the stages in
// ascending stride (src/FalconnPP.cpp:392-411 via FFHT), the butterflies of a
// stage spread over the lanes (bit-identical: they are independent).
__device__ void rotate_wave(const float* src, int D, int F, const uint32_t* masks_layer, float* x) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int MW = (F + 31) / 32;
    for (int d = lane; d < F; d += WAVE) x[d] = d < D ? src[d] : 0.0f;
    for (int r = 0; r < 3; ++r) {
        const uint32_t* mb = masks_layer + r * MW;
        wave_lds_sync();
        for (int d = lane; d < F; d += WAVE)
            x[d] = __uint_as_float(__float_as_uint(x[d]) ^ (((mb[d >> 5] >> (d & 31)) & 1u) << 31));
        for (int h = 1; h < F; h <<= 1) {
            wave_lds_sync();
            for (int b = lane; b < F / 2; b += WAVE) {
                const int j = (b / h) * 2 * h + (b % h);
                const float u = x[j], v = x[j + h];
                x[j] = u + v;
                x[j + h] = u - v;
            }
        }
    }
    wave_lds_sync();
}

// The heaps of the reference's probing sequence (std::priority_queue<IFPair>
// with libstdc++'s push_heap / pop_heap, restated in heap.h), run by one wave
// over a heap array in LDS, slot s kept by lane s % 64 (KS = 1, 2 or 4 slots
// per lane, capacity 64 * KS).  Each lane holds its slots and their children
// in registers (refresh); an operation then takes a few wave-wide steps
// instead of one dependent LDS round trip per heap level:
//   pop   the hole's path from the root follows, at each node, the child that
//         does not compare less (one ballot of "right < left" per slot and a
//         walk over its bits); along any root path values never increase,
//         so where the taken last element x comes to rest after the walk is
//         the shallowest path node whose path child compares less than x
//         (one ballot): the path nodes above it take their path child's
//         value, it takes x, the nodes below keep theirs;
//   push  likewise the new element y rests at the shallowest ancestor of
//         the new slot that compares less than y; those ancestors from there
//         down move one step toward the new slot.
// The result is slot for slot the libstdc++ heap.  Lanes exchange slots only
// through LDS, so every exchange is fenced (the compiler would otherwise
// forward a lane's own stores to its later loads).
__device__ __forceinline__ int heap_depth(int s) { return 31 - __clz(s + 1); }
__host__ __device__ inline int lane_heap_cap(int n) { return n <= 64 ? 64 : (n <= 128 ? 128 : 256); }

template <int KS, class Cmp>
struct LaneHeap {
    static constexpr int CAP = 64 * KS;
    static constexpr int DMAX = KS == 1 ? 6 : (KS == 2 ? 7 : 8);  // bound of the walk's steps
    IFPair* H;  // LDS, CAP entries
    int n;      // wave-uniform
    IFPair own[KS], lc[KS], rc[KS];
    IFPair topv, lastv;  // wave-uniform H[0], H[n-1] as of the last refresh
    __device__ __forceinline__ static IFPair uniform(IFPair e) {
        return IFPair{__builtin_amdgcn_readfirstlane(e.idx),
                      __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(e.val)))};
    }
    __device__ __forceinline__ void init(IFPair* h) {
        H = h;
        n = 0;
    }
    __device__ __forceinline__ int size() const { return n; }
    __device__ __forceinline__ bool empty() const { return n == 0; }
    __device__ __forceinline__ IFPair top() const { return topv; }
    // after every pop / push, before the next top / pop / push
    __device__ __forceinline__ void refresh() {
        wave_lds_sync();
        const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int s = lane + WAVE * r;
            own[r] = H[s];
            lc[r] = H[min(2 * s + 1, CAP - 1)];
            rc[r] = H[min(2 * s + 2, CAP - 1)];
        }
        topv = uniform(H[0]);
        lastv = uniform(H[n > 0 ? n - 1 : 0]);
    }
    __device__ __forceinline__ void pop() {
        const int lane = threadIdx.x & (WAVE - 1);
        const int len = n - 1;
        n = len;
        if (len <= 0) return;
        const IFPair x = lastv;
        const int half = (len - 1) / 2;
        uint64_t lw[KS];
#pragma unroll
        for (int r = 0; r < KS; ++r) lw[r] = __ballot(Cmp()(rc[r], lc[r]));
        int p = 0;
#pragma unroll
        for (int d = 0; d < DMAX; ++d) {
            uint64_t w = lw[0];
#pragma unroll
            for (int r = 1; r < KS; ++r) w = (p >> 6) == r ? lw[r] : w;
            const int c = 2 * p + 2 - (int)((w >> (p & 63)) & 1ull);
            p = p < half ? c : p;
        }
        if ((len & 1) == 0 && p == (len - 2) / 2) p = 2 * p + 1;
        const int f = p, df = heap_depth(f);
        uint64_t pb[KS];
        bool on[KS];
        IFPair cv[KS];
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int s = lane + WAVE * r, ds = heap_depth(s);
            on[r] = ds < df && ((f + 1) >> (df - ds)) == s + 1;
            const int c = on[r] ? ((f + 1) >> (df - ds - 1)) - 1 : 0;
            cv[r] = c == 2 * s + 1 ? lc[r] : rc[r];
            pb[r] = __ballot(on[r] && Cmp()(cv[r], x));
        }
        int rest = f;
#pragma unroll
        for (int r = KS - 1; r >= 0; --r)
            if (pb[r]) rest = WAVE * r + (int)__builtin_ctzll(pb[r]);
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int s = lane + WAVE * r;
            if (on[r] && s < rest) own[r] = cv[r];
            if (s == rest) own[r] = x;
            if (s < len) H[s] = own[r];
        }
    }
    // after refresh() or pop(): own[] are current
    __device__ __forceinline__ void push(IFPair y) {
        wave_lds_sync();
        const int lane = threadIdx.x & (WAVE - 1);
        const int e = n, de = heap_depth(e);
        n = e + 1;
        uint64_t rb[KS];
        bool R[KS];
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int s = lane + WAVE * r, ds = heap_depth(s);
            R[r] = ds < de && ((e + 1) >> (de - ds)) == s + 1 && Cmp()(own[r], y);
            rb[r] = __ballot(R[r]);
        }
        int rest = e;
#pragma unroll
        for (int r = KS - 1; r >= 0; --r)
            if (rb[r]) rest = WAVE * r + (int)__builtin_ctzll(rb[r]);
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int s = lane + WAVE * r, ds = heap_depth(s);
            if (R[r]) H[((e + 1) >> (de - ds - 1)) - 1] = own[r];
        }
        if (lane == 0) H[rest] = y;
    }
};

// The per-layer heap of src/FalconnPP.cpp:413-440 (min-heap of the p largest
// |projections|, then popped smallest first into vec[p-1 .. 0]).  One lane
// with the heap in LDS (hs), or the whole wave with a lane heap over H.
__device__ void layer_heap(const float* x, int n_proj, int log2_np, int p, IFPair* hs, IFPair* vec) {
    MinHeap q;
    q.init(hs);
    for (int r = 0; r < n_proj; ++r) {
        const int s = x[r] >= 0.0f ? 1 : -1;
        const float v = s * x[r];
        int code = r;
        if (s < 0) code |= 1 << log2_np;
        if (q.size() < p)
            q.push(IFPair{code, v});
        else if (v > q.top().val) {
            q.pop();
            q.push(IFPair{code, v});
        }
    }
    for (int i = p - 1; i >= 0; --i) {
        vec[i] = q.top();
        q.pop();
    }
}
template <int KS>
__device__ void layer_heap_lane(const float* x, int n_proj, int log2_np, int p, IFPair* H, IFPair* vec) {
    LaneHeap<KS, CmpGreater> q;
    q.init(H);
    q.refresh();
    for (int r = 0; r < n_proj; ++r) {
        const float xr = x[r];
        const int s = xr >= 0.0f ? 1 : -1;
        const float v = s * xr;
        int code = r;
        if (s < 0) code |= 1 << log2_np;
        if (q.size() < p) {
            q.push(IFPair{code, v});
            q.refresh();
        } else if (v > q.top().val) {
            q.pop();
            q.push(IFPair{code, v});
            q.refresh();
        }
    }
    for (int i = p - 1; i >= 0; --i) {
        const IFPair t = q.top();
        if ((threadIdx.x & (WAVE - 1)) == 0) vec[i] = t;
        q.pop();
        q.refresh();
    }
}

// The table's pair heap of src/FalconnPP.cpp:443-531 (min-heap of the PT
// largest pair sums over the two layers' lists), popped smallest first into
// list[PT-1 .. 0] with the table's bucket offset; a lane heap over H.
template <int KS>
__device__ void table_heap_lane(const IFPair* v1, const IFPair* v2, int PL, int PT, int two_np, int base, IFPair* H,
                                IFPair* list) {
    const int lane = threadIdx.x & (WAVE - 1);
    LaneHeap<KS, CmpGreater> th;
    th.init(H);
    th.refresh();
    // the second layer's list in lanes when it fits one register
    const IFPair b_lane = PL <= WAVE && lane < PL ? v2[lane] : IFPair{0, 0.0f};
    for (int i = 0; i < PL; ++i) {
        const IFPair a = v1[i];
        for (int j = 0; j < PL; ++j) {
            IFPair b;
            if (PL <= WAVE)
                b = IFPair{__builtin_amdgcn_readlane(b_lane.idx, j),
                           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b_lane.val), j))};
            else
                b = v2[j];
            const int bk = a.idx * two_np + b.idx;
            const float s = a.val + b.val;
            if (th.size() < PT) {
                th.push(IFPair{bk, s});
                th.refresh();
            } else if (s > th.top().val) {
                th.pop();
                th.push(IFPair{bk, s});
                th.refresh();
            }
        }
    }
    for (int idx = PT - 1; !th.empty(); --idx) {
        IFPair pr = th.top();
        pr.idx = base + pr.idx;
        if (lane == 0) list[idx] = pr;
        th.pop();
        th.refresh();
    }
}

// Entries of a wave's heap area in replay_table: the lane heaps (up to 128
// entries) or the one-lane LDS heaps beyond that (two layer heaps side by
// side, the table heap).
__host__ __device__ inline int table_heap_entries(int PL, int PT) {
    const int layer = PL <= 128 ? lane_heap_cap(PL) : 2 * PL;
    const int table = PT <= 128 ? lane_heap_cap(PT) : PT;
    return layer > table ? layer : table;
}
// LDS of one (query, table) wave: two rotated layers, the two layer lists,
// the heap area.
__host__ __device__ inline int64_t exact_table_lds_bytes(int F, int PL, int PT) {
    return ((int64_t)2 * F * 4 + ((int64_t)2 * PL + table_heap_entries(PL, PT)) * 8 + 15) / 16 * 16;
}
// LDS of the global merge: the lists when staged, the heap (a lane heap up to
// 256 tables, else one lane's) and the one-lane merge's cursors.
__host__ __device__ inline int64_t merge_heap_entries(int L) { return L <= 256 ? lane_heap_cap(L) : L; }
__host__ __device__ inline int64_t merge_lds_bytes(int L, int PT, bool staged) {
    return (staged ? (int64_t)L * PT * 8 : 0) + merge_heap_entries(L) * 8 + (int64_t)L * 4;
}

// One wave: table l's probe list of query q, src/FalconnPP.cpp:358-531 with
// the emulated heaps, into list[0 .. PT), largest first.  When no per-layer
// tie was seen (flag bit 0 clear), the fast kernel's per-layer top lists are
// exact (distinct values, largest first, as the reference pops them) and are
// reused; otherwise the table is rotated again and its layer heaps replayed.
__device__ __forceinline__ void replay_table(const HashGeom& hg, const ProbeGeom& g, const QueryView& qv_, const float* Q,
                             int64_t ldq, const uint32_t* masks, uint32_t q, int l, uint32_t flag,
                             const int32_t* qc, const float* qvals, unsigned char* wlds, IFPair* list) {
    const int F = hg.F, PL = qv_.PL, PT = qv_.perTable;
    const int lane = threadIdx.x & (WAVE - 1);
    float* x1 = reinterpret_cast<float*>(wlds);
    float* x2 = x1 + F;
    IFPair* v1 = reinterpret_cast<IFPair*>(x2 + F);
    IFPair* v2 = v1 + g.PL;
    IFPair* hp = v2 + g.PL;  // heap area
    if (flag & 1u) {
        const int MW = (F + 31) / 32;
        const float* qv = Q + (int64_t)q * ldq;
        rotate_wave(qv, hg.D, F, masks + (size_t)(l * 2 + 0) * 3 * MW, x1);
        rotate_wave(qv, hg.D, F, masks + (size_t)(l * 2 + 1) * 3 * MW, x2);
        if (PL <= WAVE) {
            layer_heap_lane<1>(x1, hg.n_proj, hg.log2_nproj, PL, hp, v1);
            layer_heap_lane<1>(x2, hg.n_proj, hg.log2_nproj, PL, hp, v2);
        } else if (PL <= 2 * WAVE) {
            layer_heap_lane<2>(x1, hg.n_proj, hg.log2_nproj, PL, hp, v1);
            layer_heap_lane<2>(x2, hg.n_proj, hg.log2_nproj, PL, hp, v2);
        } else {
            if (lane == 0) layer_heap(x1, hg.n_proj, hg.log2_nproj, PL, hp, v1);
            if (lane == 1) layer_heap(x2, hg.n_proj, hg.log2_nproj, PL, hp + PL, v2);
        }
    } else {
        const size_t base = ((size_t)q * g.L + l) * 2 * g.PL;
        for (int i = lane; i < PL; i += WAVE) {
            v1[i] = IFPair{qc[base + i], qvals[base + i]};
            v2[i] = IFPair{qc[base + g.PL + i], qvals[base + g.PL + i]};
        }
    }
    wave_lds_sync();
    if (PT <= WAVE)
        table_heap_lane<1>(v1, v2, PL, PT, g.two_np, l * g.NB, hp, list);
    else if (PT <= 2 * WAVE)
        table_heap_lane<2>(v1, v2, PL, PT, g.two_np, l * g.NB, hp, list);
    else if (lane == 0) {
        MinHeap th;
        th.init(hp);
        for (int i = 0; i < PL; ++i)
            for (int j = 0; j < PL; ++j) {
                const int b = v1[i].idx * g.two_np + v2[j].idx;
                const float s = v1[i].val + v2[j].val;
                if (th.size() < PT)
                    th.push(IFPair{b, s});
                else if (s > th.top().val) {
                    th.pop();
                    th.push(IFPair{b, s});
                }
            }
        int idx = PT - 1;
        while (!th.empty()) {
            IFPair pr = th.top();
            th.pop();
            pr.idx = l * g.NB + pr.idx;
            list[idx] = pr;
            idx--;
        }
    }
    wave_lds_sync();
}

// The global merge below on one wave with a lane heap over H of the tables'
// heads; lane t % 64 keeps table t's cursor and its next entry in registers,
// so a pop needs no memory read to know what to push.
template <int KS>
__device__ void merge_lane(const IFPair* list, int L, int PT, int QP, int nb_shift, IFPair* H, uint32_t* out) {
    const int lane = threadIdx.x & (WAVE - 1);
    int cur[KS];
    IFPair nh[KS];
#pragma unroll
    for (int r = 0; r < KS; ++r) {
        const int t = lane + WAVE * r;
        cur[r] = 0;
        nh[r] = t < L && PT > 1 ? list[(int64_t)t * PT + 1] : IFPair{0, 0.0f};
    }
    LaneHeap<KS, CmpLess> mh;
    mh.init(H);
    for (int l = 0; l < L; ++l) {
        mh.refresh();
        mh.push(list[(int64_t)l * PT]);
    }
    mh.refresh();
    for (int c = 0; c < QP && !mh.empty(); ++c) {
        const IFPair pr = mh.top();
        const int tb = pr.idx >> nb_shift;
        const int tl = tb & (WAVE - 1), tr = tb >> 6;
        int tk = 0;
        IFPair next{0, 0.0f};
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int ck = __builtin_amdgcn_readlane(cur[r], tl);
            const int ni = __builtin_amdgcn_readlane(nh[r].idx, tl);
            const int nv = __builtin_amdgcn_readlane(__float_as_int(nh[r].val), tl);
            if (r == 0 || tr == r) {
                tk = ck + 1;
                next = IFPair{ni, __int_as_float(nv)};
            }
        }
        if (lane == 0) out[c] = (uint32_t)pr.idx;
#pragma unroll
        for (int r = 0; r < KS; ++r)
            if (lane == tl && tr == r) {
                cur[r] = tk;
                if (tk + 1 < PT) nh[r] = list[(int64_t)tb * PT + tk + 1];
            }
        mh.pop();
        if (tk < PT) mh.push(next);
        mh.refresh();
    }
}

// The reference's global merge (src/FalconnPP.cpp:540-582) of the per-table
// lists (L lists of PT, largest first) into out[0 .. QP): a max-heap of the
// tables' heads, pop the largest, push that table's next.  One wave with a
// lane heap over mq (merge_heap_entries(L) entries) up to 256 tables, else
// one lane with the heap in mq and the cursors in track (L ints).  The list
// reads are wave-uniform.  MAXKS bounds the lane heap's registers (the gather
// kernels' replay keeps to 2: their register budget is the scoring loop's).
template <int MAXKS>
__device__ __forceinline__ void merge_lists(const IFPair* list, int L, int PT, int QP, int NB, IFPair* mq, int* track,
                            uint32_t* out) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int nb_shift = 31 - __clz(NB);  // NB = (2 n_proj)^2, a power of two
    if (L <= WAVE)
        merge_lane<1>(list, L, PT, QP, nb_shift, mq, out);
    else if (L <= 2 * WAVE)
        merge_lane<2>(list, L, PT, QP, nb_shift, mq, out);
    else if (MAXKS >= 4 && L <= 4 * WAVE)
        merge_lane<MAXKS >= 4 ? 4 : 2>(list, L, PT, QP, nb_shift, mq, out);
    else {
        for (int l = lane; l < L; l += WAVE) track[l] = 0;
        wave_lds_sync();
        if (lane == 0) {
            MaxHeap mh;
            mh.init(mq);
            for (int l = 0; l < L; ++l) mh.push(list[(int64_t)l * PT]);
            for (int c = 0; c < QP; ++c) {
                const IFPair pr = mh.top();
                mh.pop();
                out[c] = (uint32_t)pr.idx;
                const int tb = pr.idx >> nb_shift;
                const int tk = ++track[tb];
                if (tk < PT) mh.push(list[(int64_t)tb * PT + tk]);
            }
        }
    }
    wave_lds_sync();
}

// One thread group G, after the per-table lists of query q are in pool: the
// global cut by radix select; when it is unambiguous (no equal values on both
// sides) the probe set is every pair above it and every pair on it, else
// the group's wave 0 replays the merge (lists staged in LDS when merge_staged).  Then
// ctot[q].  Shared words: hist[256], and s_* below.
template <class G, int MAXKS = 4>
__device__ void replay_select(const ProbeGeom& g, const QueryView& qv_, uint32_t q, const IFPair* pool,
                              const uint64_t* __restrict__ bucket_off, uint32_t* __restrict__ probes,
                              uint64_t* __restrict__ ctot, unsigned char* mlds_raw, int merge_staged) {
    __shared__ uint32_t hist[256];
    __shared__ int s_digit;
    __shared__ uint32_t s_below, s_eqc, s_cnt;
    __shared__ unsigned long long s_ct;
    const int lane = threadIdx.x & (WAVE - 1), w = G::tid() / WAVE, t = G::tid();
    constexpr int TB = G::size;
    const int PT = qv_.perTable, QP = qv_.qProbes;
    const int n = g.L * PT;
    uint32_t prefix = 0, pmask = 0, k = (uint32_t)QP, eqc = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        if (t < 256) hist[t] = 0;
        G::sync();
        for (int i = t; i < n; i += TB) {
            const uint32_t key = ~heap_key(pool[i].val);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        G::sync();
        if (w == 0) {
            int d;
            uint32_t below, e;
            wave_find_digit(hist, k, d, below, e);
            if (lane == 0) {
                s_digit = d;
                s_below = below;
                s_eqc = e;
            }
        }
        G::sync();
        prefix |= (uint32_t)s_digit << shift;
        pmask |= 255u << shift;
        k -= s_below;
        eqc = s_eqc;
        G::sync();
    }
    uint32_t* out = probes + qv_.probe_off;
    if (eqc == k) {
        const uint32_t T = prefix;
        if (t == 0) s_cnt = 0;
        G::sync();
        for (int c0 = 0; c0 < n; c0 += TB) {
            const int i = c0 + t;
            IFPair e{0, 0.0f};
            bool take = false;
            if (i < n) {
                e = pool[i];
                take = ~heap_key(e.val) <= T;
            }
            const uint64_t m = __ballot(take);
            if (m == 0ull) continue;
            const int leader = (int)__ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_cnt, (uint32_t)__popcll(m));
            base = __shfl(base, leader, WAVE);
            const uint32_t pos = base + lane_rank(m);
            if (take && pos < (uint32_t)QP) out[pos] = (uint32_t)e.idx;
        }
    } else {
        IFPair* mlds = reinterpret_cast<IFPair*>(mlds_raw);
        IFPair* mq = merge_staged ? mlds + (int64_t)g.L * PT : mlds;
        int* track = reinterpret_cast<int*>(mq + merge_heap_entries(g.L));
        if (merge_staged)
            for (int64_t i = t; i < (int64_t)g.L * PT; i += TB) mlds[i] = pool[i];
        G::sync();
        if (w == 0) merge_lists<MAXKS>(merge_staged ? mlds : pool, g.L, PT, QP, g.NB, mq, track, out);
    }
    G::sync();
    if (t == 0) s_ct = 0;
    G::sync();
    unsigned long long ct = 0;
    for (int c = t; c < QP; c += TB) {
        const uint32_t b = out[c];
        ct += bucket_off[b + 1] - bucket_off[b];
    }
    atomicAdd(&s_ct, ct);
    G::sync();
    if (t == 0) ctot[q] = s_ct;
    G::sync();
}


// A flagged query's probes replayed by a thread group G (the workgroup) of a
// gather kernel before it gathers the query: its waves rebuild the per-table
// lists into pool, then replay_select.  lds: LDS of the group, at least
// r.twaves * r.table_lds bytes and the merge's share.
// Kept out of line: inlined, its registers would be allocated across the
// gather's scoring loop.
template <class G>
__device__ __noinline__ void replay_query(const ReplayArgs& r, uint32_t q, unsigned char* lds, IFPair* pool) {
    lds = lds_ptr(lds);
    const int w = G::tid() / WAVE;
    const QueryView qv_ = query_view(r.pg.qp, (int)q, r.pg.qProbes, r.pg.perTable, r.pg.PL, 0);
    const uint32_t flag = r.qflag[q];
    // the replay is a chain of dependent steps on few waves while the other
    // workgroups on the CU gather: its waves issue first on their SIMDs
    __builtin_amdgcn_s_setprio(3);
    if (w < r.twaves)
        for (int l = w; l < r.pg.L; l += r.twaves)
            replay_table(r.hg, r.pg, qv_, r.Q, r.ldq, r.masks, q, l, flag, r.qc, r.qv,
                         lds + (int64_t)w * r.table_lds, pool + (int64_t)l * qv_.perTable);
    G::sync();
    replay_select<G, 2>(r.pg, qv_, q, pool, r.bucket_off, r.probes, r.ctot, lds, r.merge_staged);
    __builtin_amdgcn_s_setprio(0);
}

}  // namespace
}  // namespace fpp
