// Query path on gfx950 (reference: src/FalconnPP.cpp:310-707).
//
//  k_probe_select   per query: top-perTable pair sums of every table, then the
//                   global top-qProbes of that pool.  The reference gets the
//                   same set from its k-way heap merge (:540-582): a merge of
//                   sorted lists pops the largest values first.  Ties on a cut
//                   are flagged for the exact replay (k_probe_exact_*).
//  k_gather_score   per query: bucket gather + dedup in an LDS bitset window
//                   (the reference's boost::dynamic_bitset, :552-611), the
//                   ascending candidate list, FP32 scores q.(x - c) (:647) and
//                   the top-k (:650-673), ties broken towards smaller ids.
#include "common.h"
#include "heap.h"
#include "gather_common.h"
#include "kernels.h"
#include "replay_common.h"

namespace fpp {
namespace {

using namespace gather;


constexpr int PS_T = 256;
constexpr int PS_WAVES = PS_T / WAVE;

__device__ __forceinline__ void wave_min_max(uint32_t& mn, uint32_t& mx) {
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, off, WAVE));
        mx = max(mx, (uint32_t)__shfl_xor(mx, off, WAVE));
    }
}

// LDS (dynamic): pool keys; their buckets (two-stage: one per key; full pool:
// the tables' code lists, L*2*PL, from which a pair's bucket is recomputed);
// per-wave histograms, block histogram, per-table counts.
template <bool FULLPOOL>
__global__ __launch_bounds__(PS_T) void k_probe_select(ProbeGeom g, const int32_t* __restrict__ qc,
                                                       const float* __restrict__ qv,
                                                       const uint64_t* __restrict__ bucket_off,
                                                       uint32_t* __restrict__ probes, uint64_t* __restrict__ ctot,
                                                       uint32_t* __restrict__ qflag) {
    extern __shared__ uint32_t lds[];
    const int q = blockIdx.x;
    const QueryView qv_ = query_view(g.qp, q, g.qProbes, g.perTable, g.PL, 0);
    const int PL = qv_.PL, np = PL * PL, perTable = qv_.perTable, qProbes = qv_.qProbes;
    // FULLPOOL: every pair of every table enters the pool (per-table counts
    // checked below); else the top-perTable pairs of each table
    const int per = FULLPOOL ? np : perTable;
    const int pool_n = g.L * per;
    uint32_t* pkey = lds;
    uint32_t* pbkt = lds + pool_n;
    uint32_t* whist = pbkt + (FULLPOOL ? g.L * 2 * PL : pool_n);  // PS_WAVES * 256
    uint32_t* bhist = whist + PS_WAVES * 256;  // 256
    uint32_t* tcnt = bhist + 256;              // L (FULLPOOL)
    __shared__ int s_digit;
    __shared__ uint32_t s_below;
    __shared__ uint32_t s_tie;
    __shared__ unsigned long long s_ctot;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const uint64_t bmask = lanes_below_mask();
    if (threadIdx.x == 0) {
        s_tie = 0;
        s_ctot = 0;
    }
    __syncthreads();
    uint32_t* hw = whist + w * 256;
    bool tie = false;
    // a table's two top lists (stored at stride g.PL) sit in lanes: v1/c1 in
    // lanes [0, PL), v2/c2 in [PL, 2 PL); the wave's next table is fetched
    // while this one is selected
    auto fetch = [&](int l, float& v, int32_t& c) {
        v = 0.0f;
        c = 0;
        if (l < g.L && lane < 2 * PL) {
            const size_t base = ((size_t)q * g.L + l) * 2 * g.PL + (lane < PL ? lane : g.PL + lane - PL);
            v = qv[base];
            c = qc[base];
        }
    };
    constexpr int KR = 16;  // pairs per lane: PL <= 31, so PL * PL <= 16 * 64
    const int nr = (np + WAVE - 1) / WAVE;
    // t / PL for t < 1024 and PL <= 31 as (t * pl_mul) >> 16 (exact there)
    const uint32_t pl_mul = 65536u / (uint32_t)PL + 1u;
    // i / np for i < 2^16 as (i * np_mul) >> 32 (exact there for np <= 961)
    const uint64_t np_mul = ((1ull << 32) + (uint64_t)np - 1) / (uint64_t)np;
    if constexpr (FULLPOOL) {
        // key of pair t = i * PL + j: fAbsHashValue1 + fAbsHashValue2
        // (src/FalconnPP.cpp:457-488), in the same order of addition.  The
        // lists of PF tables are fetched together, so a wave waits for
        // memory once per PF tables
        constexpr int PF = 8;
        for (int l0 = w; l0 < g.L; l0 += PF * PS_WAVES) {
            float vb[PF];
            int32_t cb[PF];
#pragma unroll
            for (int u = 0; u < PF; ++u) fetch(l0 + u * PS_WAVES, vb[u], cb[u]);
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int l = l0 + u * PS_WAVES;
                if (l >= g.L) break;
                if (lane < 2 * PL) pbkt[l * 2 * PL + lane] = (uint32_t)cb[u];
                for (int r = 0; r < nr; ++r) {
                    const int t = r * WAVE + lane;
                    const int i = t < np ? (int)(((uint32_t)t * pl_mul) >> 16) : 0, j = t < np ? t - i * PL : 0;
                    const float a = __shfl(vb[u], i, WAVE), b = __shfl(vb[u], PL + j, WAVE);
                    if (t < np) pkey[l * np + t] = desc_key(a + b);
                }
            }
        }
    }
    if constexpr (!FULLPOOL) {
        float vn;
        int32_t cn;
        fetch(w, vn, cn);
        for (int l = w; l < g.L; l += PS_WAVES) {
            const float vv = vn;
            const int32_t cc = cn;
            fetch(l + PS_WAVES, vn, cn);
            // key of pair t = r * 64 + lane, as above
            uint32_t key[KR];
#pragma unroll
            for (int r = 0; r < KR; ++r) {
                key[r] = 0xffffffffu;
                if (r < nr) {
                    const int t = r * WAVE + lane;
                    const int i = t < np ? (int)(((uint32_t)t * pl_mul) >> 16) : 0, j = t < np ? t - i * PL : 0;
                    const float a = __shfl(vv, i, WAVE), b = __shfl(vv, PL + j, WAVE);
                    if (t < np) key[r] = desc_key(a + b);
                }
            }
            // perTable-th largest pair sum
            uint32_t prefix = 0, pmask = 0, k = (uint32_t)perTable, eqc = 0;
            for (int shift = 24; shift >= 0; shift -= 8) {
#pragma unroll
                for (int i = 0; i < 4; ++i) hw[lane * 4 + i] = 0;
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int r = 0; r < KR; ++r)
                    if (r < nr && r * WAVE + lane < np && (key[r] & pmask) == prefix)
                        atomicAdd(&hw[(key[r] >> shift) & 255u], 1u);
                __builtin_amdgcn_wave_barrier();
                int d;
                uint32_t below;
                wave_find_digit(hw, k, d, below, eqc);
                prefix |= (uint32_t)d << shift;
                pmask |= 255u << shift;
                k -= below;
                __builtin_amdgcn_wave_barrier();
            }
            if (np > perTable && eqc != k) tie = true;
            const uint32_t T = prefix;
            uint32_t written = 0, eq_seen = 0;
#pragma unroll
            for (int r = 0; r < KR; ++r) {
                if (r >= nr) continue;
                const int t = r * WAVE + lane;
                const bool valid = t < np;
                const int i = valid ? (int)(((uint32_t)t * pl_mul) >> 16) : 0, j = valid ? t - i * PL : 0;
                const int32_t ci = __shfl(cc, i, WAVE), cj = __shfl(cc, PL + j, WAVE);
                const bool lt = valid && key[r] < T;
                const bool eqk = valid && key[r] == T;
                const uint64_t meq = __ballot(eqk);
                const bool take_eq = eqk && (eq_seen + (uint32_t)__popcll(meq & bmask)) < k;
                const bool take = lt || take_eq;
                const uint64_t mtake = __ballot(take);
                const uint32_t slot = written + (uint32_t)__popcll(mtake & bmask);
                if (take && slot < (uint32_t)perTable) {
                    const int idx = l * perTable + (int)slot;
                    pkey[idx] = key[r];
                    pbkt[idx] = (uint32_t)l * (uint32_t)g.NB + (uint32_t)ci * (uint32_t)g.two_np + (uint32_t)cj;
                }
                written += (uint32_t)__popcll(mtake);
                eq_seen += (uint32_t)__popcll(meq);
            }
        }
    }
    if (tie) atomicOr(&s_tie, 1u);
    __syncthreads();
    __shared__ uint32_t s_eqc, s_cnt, s_over, s_mn, s_mx;
    // global top-qProbes of the pool: radix select, the digit found by wave 0;
    // the cut key (keys <= it are taken) and whether keys on it must be split.
    // The pool's keys (pair sums of similar size) share their top bits, so
    // digits of up to 8 bits start at the highest bit where the smallest and
    // largest key differ: the first histogram spreads over its bins instead
    // of piling the pool onto one counter.  Trimmed pairs (0xffffffff, above
    // every key of a pair sum) stay out: there are always at least
    // L * perTable >= 2 qProbes others.
    auto global_cut = [&](bool& amb) -> uint32_t {
        uint32_t mn = 0xffffffffu, mx = 0u;
        for (int i = threadIdx.x; i < pool_n; i += PS_T) {
            const uint32_t key = pkey[i];
            if (key != 0xffffffffu) {
                mn = min(mn, key);
                mx = max(mx, key);
            }
        }
        wave_min_max(mn, mx);
        if (threadIdx.x == 0) {
            s_mn = 0xffffffffu;
            s_mx = 0u;
        }
        __syncthreads();
        if (lane == 0) {
            atomicMin(&s_mn, mn);
            atomicMax(&s_mx, mx);
        }
        __syncthreads();
        mn = s_mn;
        mx = s_mx;
        // keys <= 0x7fffffff, so top <= 30; all keys equal: one pass on bit 0
        const int top = mn == mx ? 0 : 31 - __clz(mn ^ mx);
        uint32_t pmask = ~0u << (top + 1), prefix = mn & pmask, k = (uint32_t)qProbes, eqc = 0;
        for (int hi = top; hi >= 0; hi -= 8) {
            const int shift = hi >= 7 ? hi - 7 : 0;
            const uint32_t dm = (2u << (hi - shift)) - 1u;
            bhist[threadIdx.x] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < pool_n; i += PS_T) {
                const uint32_t key = pkey[i];
                if ((key & pmask) == prefix) atomicAdd(&bhist[(key >> shift) & dm], 1u);
            }
            __syncthreads();
            if (w == 0) {
                int d;
                uint32_t below, e;
                wave_find_digit(bhist, k, d, below, e);
                if (lane == 0) {
                    s_digit = d;
                    s_below = below;
                    s_eqc = e;
                }
            }
            __syncthreads();
            prefix |= (uint32_t)s_digit << shift;
            pmask |= dm << shift;
            k -= s_below;
            eqc = s_eqc;
        }
        amb = eqc != k;
        return prefix;
    };
    bool amb;
    uint32_t T = global_cut(amb);
    if constexpr (FULLPOOL) {
        // the reference keeps the top-perTable pairs of each table
        // (src/FalconnPP.cpp:455-486).  A table holding more than perTable
        // pairs at or above the global cut gets that cut of its own (its
        // pairs below it leave the pool) and the global cut is found again.
        // A trimmed table holds exactly perTable pairs, so every round trims
        // new tables; after kMaxTrim rounds the query is replayed exactly.
        constexpr int kMaxTrim = 6;
        for (int it = 0;; ++it) {
            if (threadIdx.x == 0) s_over = 0;
            __syncthreads();
            // pairs at or above the cut per table: a wave per table, counted
            // by ballots (the table's wave below reads its own count)
            for (int l = w; l < g.L; l += PS_WAVES) {
                const uint32_t* tk = pkey + l * np;
                uint32_t c = 0;
                for (int t0 = 0; t0 < np; t0 += WAVE) {
                    const int t = t0 + lane;
                    c += (uint32_t)__popcll(__ballot(t < np && tk[t] <= T));
                }
                if (lane == 0) {
                    tcnt[l] = c;
                    if (c > (uint32_t)perTable) atomicOr(&s_over, 1u);
                }
            }
            __syncthreads();
            if (!s_over) break;
            if (it == kMaxTrim) {
                if (threadIdx.x == 0) atomicOr(&s_tie, 1u);
                break;
            }
            bool ttie = false;
            for (int l = w; l < g.L; l += PS_WAVES) {
                if (tcnt[l] <= (uint32_t)perTable) continue;
                uint32_t* tk = pkey + l * np;
                uint32_t prefix = 0, pmask = 0, k = (uint32_t)perTable, eqc = 0;
                for (int shift = 24; shift >= 0; shift -= 8) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) hw[lane * 4 + i] = 0;
                    __builtin_amdgcn_wave_barrier();
                    for (int t = lane; t < np; t += WAVE) {
                        const uint32_t key = tk[t];
                        if ((key & pmask) == prefix) atomicAdd(&hw[(key >> shift) & 255u], 1u);
                    }
                    __builtin_amdgcn_wave_barrier();
                    int d;
                    uint32_t below;
                    wave_find_digit(hw, k, d, below, eqc);
                    prefix |= (uint32_t)d << shift;
                    pmask |= 255u << shift;
                    k -= below;
                    __builtin_amdgcn_wave_barrier();
                }
                // pairs tied on the table's cut: the reference's heap order
                // decides which stay, so the query is replayed exactly
                if (eqc != k) ttie = true;
                const uint32_t Tl = prefix;
                uint32_t eq_seen = 0;
                for (int t0 = 0; t0 < np; t0 += WAVE) {
                    const int t = t0 + lane;
                    const bool valid = t < np;
                    const uint32_t key = valid ? tk[t] : 0xffffffffu;
                    const bool eqk = valid && key == Tl;
                    const uint64_t meq = __ballot(eqk);
                    const bool keep = key < Tl || (eqk && eq_seen + (uint32_t)__popcll(meq & bmask) < k);
                    if (valid && !keep) tk[t] = 0xffffffffu;
                    eq_seen += (uint32_t)__popcll(meq);
                }
            }
            if (ttie) atomicOr(&s_tie, 1u);
            __syncthreads();
            T = global_cut(amb);
        }
    }
    if (amb && threadIdx.x == 0) atomicOr(&s_tie, 2u);
    // every key above the cut and every key on it: exactly qProbes when the
    // cut is unambiguous; otherwise the first qProbes appended (the query is
    // flagged and its probes replayed exactly)
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    unsigned long long my_ctot = 0;
    uint32_t* out = probes + qv_.probe_off;
    // OU chunks of the pool per step, so their bucket-size loads wait for
    // memory together
    constexpr int OU = 8;
    for (int c0 = 0; c0 < pool_n; c0 += OU * PS_T) {
        uint64_t lo[OU], hi[OU];
#pragma unroll
        for (int u = 0; u < OU; ++u) {
            lo[u] = hi[u] = 0;
            const int i = c0 + u * PS_T + (int)threadIdx.x;
            const bool take = i < pool_n && pkey[i] <= T;
            const uint64_t m = __ballot(take);
            if (m == 0ull) continue;
            const int leader = (int)__ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_cnt, (uint32_t)__popcll(m));
            base = __shfl(base, leader, WAVE);
            const uint32_t pos = base + lane_rank(m);
            if (take && pos < (uint32_t)qProbes) {
                uint32_t b;
                if constexpr (FULLPOOL) {
                    // pair i: table l = i / np, t = i % np, codes of list 1
                    // entry t / PL and list 2 entry t % PL (exact divisions
                    // by multiply-shift for i < 2^16)
                    const uint32_t l = (uint32_t)(((uint64_t)(uint32_t)i * np_mul) >> 32);
                    const uint32_t t = (uint32_t)i - l * (uint32_t)np;
                    const uint32_t ii = (t * pl_mul) >> 16, jj = t - ii * (uint32_t)PL;
                    const uint32_t* cl = pbkt + l * 2 * PL;
                    b = l * (uint32_t)g.NB + cl[ii] * (uint32_t)g.two_np + cl[PL + jj];
                } else {
                    b = pbkt[i];
                }
                out[pos] = b;
                lo[u] = bucket_off[b];
                hi[u] = bucket_off[b + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < OU; ++u) my_ctot += hi[u] - lo[u];
    }
    atomicAdd(&s_ctot, my_ctot);
    __syncthreads();
    if (threadIdx.x == 0) {
        ctot[q] = s_ctot;
        if (s_tie) atomicOr(qflag + q, 2u);
    }
}

// Probe selection with one probe per layer (perTable = 1, qProbes <= L / 2):
// every table contributes its single best pair, so no table ever holds more
// than perTable pairs at or above the cut and the answer is the global
// top-qProbes of the L pair sums.  One wave per query, up to PS1_S tables per
// lane in registers, ranks by counting over the wave: the same cut, tie flag,
// probe set and C_tot as k_probe_select (the probes in table order).
constexpr int PS1_T = 256;
constexpr int PS1_S = 4;  // L <= 256
__global__ __launch_bounds__(PS1_T) void k_probe_select_pl1(ProbeGeom g, int nq, const int32_t* __restrict__ qc,
                                                            const float* __restrict__ qv,
                                                            const uint64_t* __restrict__ bucket_off,
                                                            uint32_t* __restrict__ probes, uint64_t* __restrict__ ctot,
                                                            uint32_t* __restrict__ qflag) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int q = blockIdx.x * (PS1_T / WAVE) + (int)(threadIdx.x / WAVE);
    if (q >= nq) return;  // wave-uniform
    const QueryView qv_ = query_view(g.qp, q, g.qProbes, g.perTable, g.PL, 0);
    const uint32_t QP = (uint32_t)qv_.qProbes;
    // key of table l's pair: desc_key(v1 + v2) in the reference's order of
    // addition (src/FalconnPP.cpp:478-485); 0xffffffff (above every key)
    // past the last table
    uint32_t key[PS1_S], bkt[PS1_S];
#pragma unroll
    for (int s = 0; s < PS1_S; ++s) {
        const int l = s * WAVE + lane;
        key[s] = 0xffffffffu;
        bkt[s] = 0u;
        if (l < g.L) {
            const size_t base = ((size_t)q * g.L + l) * 2 * g.PL;
            const float a = gld(qv + base), b = gld(qv + base + g.PL);
            const int32_t c1 = gld(qc + base), c2 = gld(qc + base + g.PL);
            key[s] = desc_key(a + b);
            bkt[s] = (uint32_t)l * (uint32_t)g.NB + (uint32_t)c1 * (uint32_t)g.two_np + (uint32_t)c2;
        }
    }
    // lt / le: keys below / at most each key
    uint32_t lt[PS1_S], le[PS1_S];
#pragma unroll
    for (int s = 0; s < PS1_S; ++s) lt[s] = le[s] = 0u;
    const int nslots = (g.L + WAVE - 1) / WAVE;
    for (int s2 = 0; s2 < nslots; ++s2) {
        for (int o = 0; o < WAVE; ++o) {
            if (s2 * WAVE + o >= g.L) break;  // uniform
            uint32_t kj = 0;
#pragma unroll
            for (int u = 0; u < PS1_S; ++u)
                if (u == s2) kj = (uint32_t)__builtin_amdgcn_readlane((int)key[u], o);
#pragma unroll
            for (int s = 0; s < PS1_S; ++s) {
                lt[s] += kj < key[s] ? 1u : 0u;
                le[s] += kj <= key[s] ? 1u : 0u;
            }
        }
    }
    // the cut T: the key whose [lt, le) holds rank QP - 1; below it klt keys,
    // on it eqc keys
    uint32_t T = 0, klt = 0, eqc = 0;
#pragma unroll
    for (int s = 0; s < PS1_S; ++s) {
        const bool cut = s * WAVE + lane < g.L && lt[s] < QP && QP <= le[s];
        const uint64_t m = __ballot(cut);
        if (m) {
            const int src = (int)__ffsll((long long)m) - 1;
            T = (uint32_t)__builtin_amdgcn_readlane((int)key[s], src);
            klt = (uint32_t)__builtin_amdgcn_readlane((int)lt[s], src);
            eqc = (uint32_t)__builtin_amdgcn_readlane((int)le[s], src) - klt;
        }
    }
    const bool amb = eqc != QP - klt;
    // every key above the cut and on it, in table order; the first QP kept
    uint32_t* out = probes + qv_.probe_off;
    uint32_t base = 0;
    unsigned long long c = 0;
#pragma unroll
    for (int s = 0; s < PS1_S; ++s) {
        const bool take = s * WAVE + lane < g.L && key[s] <= T;
        const uint64_t m = __ballot(take);
        const uint32_t pos = base + (uint32_t)__popcll(m & lanes_below_mask());
        if (take && pos < QP) {
            gst(out + pos, bkt[s]);
            c += gld(bucket_off + bkt[s] + 1) - gld(bucket_off + bkt[s]);
        }
        base += (uint32_t)__popcll(m);
    }
    for (int off = 32; off > 0; off >>= 1) c += (unsigned long long)__shfl_xor((long long)c, off, WAVE);
    if (lane == 0) {
        gst(ctot + q, (uint64_t)c);
        if (amb) atomicOr(qflag + q, 2u);
    }
}

// Exact replay of the flagged queries' probing sequences in three kernels,
// all sized by the device's flag count, not the host's:
//  k_replay_tables   one wave per (flagged query f < cap, table): the table
//                    lists into work area f;
//  k_replay_select   one workgroup per flagged query f < cap: cut, set or
//                    merge, ctot;
//  k_probe_replay    the flagged queries from cap on (beyond the work areas
//                    of the first two), one persistent workgroup per query
//                    doing all three steps in its own area.
__global__ __launch_bounds__(WAVE) void k_replay_tables(HashGeom hg, ProbeGeom g, const float* __restrict__ Q,
                                                        int64_t ldq, const uint32_t* __restrict__ masks,
                                                        const uint32_t* __restrict__ qlist,
                                                        const uint32_t* __restrict__ nflag, uint32_t cap,
                                                        const uint32_t* __restrict__ qflag,
                                                        const int32_t* __restrict__ qc,
                                                        const float* __restrict__ qvals,
                                                        unsigned char* __restrict__ work, int64_t wpq) {
    extern __shared__ __align__(16) unsigned char tl[];
    const uint64_t items = (uint64_t)min(*nflag, cap) * (uint64_t)g.L;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t f = (uint32_t)(it / (uint64_t)g.L);
        const int l = (int)(it - (uint64_t)f * g.L);
        const uint32_t q = qlist[f];
        const QueryView qv_ = query_view(g.qp, (int)q, g.qProbes, g.perTable, g.PL, 0);
        IFPair* pool = reinterpret_cast<IFPair*>(work + (int64_t)f * wpq);
        replay_table(hg, g, qv_, Q, ldq, masks, q, l, qflag[q], qc, qvals, tl, pool + (int64_t)l * qv_.perTable);
    }
}

constexpr int RS_SEL_T = 256;
__global__ __launch_bounds__(RS_SEL_T) void k_replay_select(ProbeGeom g, const uint32_t* __restrict__ qlist,
                                                            const uint32_t* __restrict__ nflag, uint32_t cap,
                                                            const uint64_t* __restrict__ bucket_off,
                                                            uint32_t* __restrict__ probes,
                                                            uint64_t* __restrict__ ctot,
                                                            const unsigned char* __restrict__ work, int64_t wpq,
                                                            int merge_staged) {
    extern __shared__ __align__(16) unsigned char sl[];
    const uint32_t f = blockIdx.x;
    if (f >= min(*nflag, cap)) return;
    const uint32_t q = qlist[f];
    const QueryView qv_ = query_view(g.qp, (int)q, g.qProbes, g.perTable, g.PL, 0);
    const IFPair* pool = reinterpret_cast<const IFPair*>(work + (int64_t)f * wpq);
    replay_select<Whole<RS_SEL_T>>(g, qv_, q, pool, bucket_off, probes, ctot, sl, merge_staged);
}

constexpr int RX_T = 512;
constexpr int RX_W = RX_T / WAVE;
__global__ __launch_bounds__(RX_T) void k_probe_replay(HashGeom hg, ProbeGeom g, const float* __restrict__ Q,
                                                       int64_t ldq, const uint32_t* __restrict__ masks,
                                                       const uint32_t* __restrict__ qlist,
                                                       const uint32_t* __restrict__ nflag, uint32_t first,
                                                       uint32_t* claim, const uint32_t* __restrict__ qflag,
                                                       const int32_t* __restrict__ qc,
                                                       const float* __restrict__ qvals,
                                                       const uint64_t* __restrict__ bucket_off,
                                                       uint32_t* __restrict__ probes, uint64_t* __restrict__ ctot,
                                                       unsigned char* __restrict__ work, int64_t wpq, int table_lds,
                                                       int twaves, int merge_staged) {
    extern __shared__ __align__(16) unsigned char rlds[];
    __shared__ int s_f;
    const int w = threadIdx.x / WAVE;
    IFPair* pool = reinterpret_cast<IFPair*>(work + (int64_t)blockIdx.x * wpq);
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t f = first + atomicAdd(claim, 1u);
            s_f = f < *nflag ? (int)f : -1;
        }
        __syncthreads();
        const int f = s_f;
        __syncthreads();
        if (f < 0) break;
        const uint32_t q = qlist[f];
        const QueryView qv_ = query_view(g.qp, (int)q, g.qProbes, g.perTable, g.PL, 0);
        const uint32_t flag = qflag[q];
        if (w < twaves)
            for (int l = w; l < g.L; l += twaves)
                replay_table(hg, g, qv_, Q, ldq, masks, q, l, flag, qc, qvals, rlds + (int64_t)w * table_lds,
                             pool + (int64_t)l * qv_.perTable);
        __syncthreads();
        replay_select<Whole<RX_T>>(g, qv_, q, pool, bucket_off, probes, ctot, rlds, merge_staged);
    }
}

// The queries whose qflag is nonzero, in query order, into list; *count =
// their number.  One workgroup.
constexpr int CF_T = 1024;
__global__ __launch_bounds__(CF_T) void k_compact_flagged(int nq, const uint32_t* __restrict__ qflag,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __shared__ uint32_t wsum[CF_T / WAVE];
    __shared__ uint32_t s_base;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nq; c0 += CF_T) {
        const int i = c0 + threadIdx.x;
        const bool keep = i < nq && qflag[i] != 0u;
        const uint64_t m = __ballot(keep);
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = s_base, tot = 0;
        for (int v = 0; v < CF_T / WAVE; ++v) {
            before += v < wv ? wsum[v] : 0u;
            tot += wsum[v];
        }
        if (keep) list[before + lane_rank(m)] = (uint32_t)i;
        __syncthreads();
        if (threadIdx.x == 0) s_base += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = s_base;
}

// Is query q small (served one per wave)?  See SmallPick.
__device__ __forceinline__ bool small_query(const SmallPick& sp, const uint32_t* __restrict__ qflag, int q) {
    if (sp.max_ctot == 0 || qflag[q]) return false;
    const int kq = sp.qp ? sp.qp[q].k : sp.k;
    return kq <= WAVE && sp.ctot[q] <= sp.max_ctot;
}

// Adds the wave's small queries to *count (one atomic per wave); the whole
// wave calls it.
__device__ __forceinline__ void count_small(bool small, uint32_t* count) {
    const uint64_t m = __ballot(small);
    if (m && (threadIdx.x & (WAVE - 1)) == (unsigned)(__ffsll((long long)m) - 1))
        atomicAdd(count, (uint32_t)__popcll(m));
}

// Sort key for longest-first order: C_tot descending, tie-flagged queries
// (replayed in the gather, C_tot not known yet) first.
__global__ void k_ctot_keys(const uint64_t* __restrict__ ctot, const uint32_t* __restrict__ qflag, int nq,
                            uint32_t* __restrict__ key, SmallPick sp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool small = q < nq && small_query(sp, qflag, q);
    if (sp.max_ctot) count_small(small, sp.count);
    if (q >= nq) return;
    const uint64_t c = ctot[q];
    uint32_t k = 0xffffffffu - (uint32_t)(c >= 0xffffffffull ? 0xfffffffeull : c);
    // every key but a small query's stays below 2^32 - 1 - max_ctot
    if (sp.max_ctot && !small) k = min(k, 0xfffffffeu - sp.max_ctot);
    key[q] = qflag[q] ? 0u : k;
}

// Probes of a call whose bucket is empty (the reference's iNumEmptyBucket,
// src/FalconnPP.cpp:584-588), summed into *count.
__global__ __launch_bounds__(256) void k_count_empty(const uint32_t* __restrict__ probes, int64_t n,
                                                     const uint64_t* __restrict__ bucket_off,
                                                     unsigned long long* __restrict__ count) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = probes[i];
        c += bucket_off[b + 1] == bucket_off[b] ? 1ull : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0 && c) atomicAdd(count, c);
}


// ---- gather + dedup + FP32 rerank + top-k ---------------------------------
template <int TB, int CPL>
__device__ __forceinline__ void k_gather_score_f32_one(const GatherArgs& a, const ReplayArgs& rp,
                                                          const uint32_t* __restrict__ probes,
                                                          const uint64_t* __restrict__ bucket_off,
                                                          const int32_t* __restrict__ ids,
                                                          const float* __restrict__ Xc, const float* __restrict__ Q,
                                                          int64_t ldq, int32_t* __restrict__ out_ids,
                                                          float* __restrict__ out_scores, int q, uint32_t& dirty) {
    extern __shared__ uint32_t lds[];
    const int64_t W = a.window_bits;
    const int Wwords = (int)(W / 32);
    uint32_t* bits = lds;
    // window counters when N exceeds one window
    uint32_t* wc = bits + ((Wwords + 1) & ~1);
    uint32_t* sh = wc + window_counter_words(a.N, W);  // TB (8-byte aligned from here on)
    uint32_t* hist = sh + TB;           // 256
    uint32_t* tk_key = hist + 256;      // kcap
    uint32_t* tk_id = tk_key + 1024;    // kcap
    __shared__ int s_digit;
    __shared__ uint32_t s_below;
    uint64_t* ts = a.phase_ts ? a.phase_ts + (size_t)q * PHASE_STAMPS : nullptr;
    if (ts && threadIdx.x == 0) ts[0] = wall_clock64();
    // a tie-flagged query's probes are replayed exactly first (the LDS is
    // free until the dedup below)
    if (rp.qflag && rp.qflag[q]) {
        replay_query<Whole<TB>>(rp, (uint32_t)q, reinterpret_cast<unsigned char*>(lds),
                         reinterpret_cast<IFPair*>(rp.work + (int64_t)blockIdx.x * rp.wpq));
        __syncthreads();
        dirty = (uint32_t)Wwords;
    }
    const QueryView qv_ = query_view(a.qp, q, a.qProbes, 0, 0, a.k);
    const int kq = qv_.k;
    const uint32_t* pr = probes + qv_.probe_off;
    // candidates in the query's scratch; FP32 scores (k > 64) over the
    // window scatter, which is dead once the candidates are collected
    const QueryScratch qs = query_scratch(a, q);
    int32_t* cq = qs.cand;
    float* sq = reinterpret_cast<float*>(qs.part);
    const uint64_t ctot_q = a.ctot ? a.ctot[q] : ~uint64_t(0);
    const uint32_t nc = collect_candidates<Whole<TB>>(a, qv_.qProbes, pr, bucket_off, ids, bits, wc, sh, cq, qs.part, qs.ccap,
                                               qs.pcap, ts, ctot_q, dirty);
    if (threadIdx.x == 0) a.ncand[q] = nc;
    if (ts && threadIdx.x == 0) ts[2] = wall_clock64();
    // FP32 scores: groups of 8 lanes per candidate row, CPL float4 chunks per
    // lane, U rows per group in flight, candidate ids fetched one step ahead.
    // Chunks past the row end re-read its last chunk against a zero query
    // chunk (adding exactly 0), so the loads carry no branches and a step's
    // CPL * U loads are all in flight before the first use.
    constexpr int U = CPL >= 16 ? 32 / CPL : 4;
    const int lane = threadIdx.x & (WAVE - 1);
    const int g8 = threadIdx.x & 7;
    const int grp = threadIdx.x >> 3;
    const int wv = threadIdx.x / WAVE;
    constexpr int NG = TB / 8;
    constexpr int STEP = U * NG;
    const int64_t nchunk = a.ldx / 4;
    const float4* q4 = reinterpret_cast<const float4*>(Q + (int64_t)q * ldq);
    float4 qr[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int64_t ch = g8 + 8 * t;
        qr[t] = ch < nchunk ? q4[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // wave-level running top-k (k <= 64): lane j holds the wave's j-th best
    // key (desc_key(score) << 32 | id), ascending; misses cost one compare
    const bool wsel = kq <= WAVE;
    WaveTopK<1> tk;
    tk.init(kq);
    // candidate ids one step ahead; every load is issued (clamped to the
    // last candidate) and the id replaced by -1 past the end
    const uint32_t last = nc > 0 ? nc - 1 : 0;
    int idn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t ii = (uint32_t)(u * NG + grp);
        const int32_t v = gld(cq + min(ii, last));
        idn[u] = ii < nc ? v : -1;
    }
    for (uint32_t i0 = 0; i0 < nc; i0 += STEP) {
        int id[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            id[u] = idn[u];
            const uint32_t ii = i0 + STEP + (uint32_t)(u * NG + grp);
            const int32_t v = gld(cq + min(ii, last));
            idn[u] = ii < nc ? v : -1;
        }
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sc[u] = 0.f;
        float4 x[CPL][U];
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int64_t ch = min<int64_t>(g8 + 8 * t, nchunk - 1);
#pragma unroll
            for (int u = 0; u < U; ++u)
                x[t][u] = reinterpret_cast<const float4*>(Xc + (int64_t)(id[u] < 0 ? 0 : id[u]) * a.ldx)[ch];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                sc[u] = fmaf(qr[t].x, x[t][u].x, sc[u]);
                sc[u] = fmaf(qr[t].y, x[t][u].y, sc[u]);
                sc[u] = fmaf(qr[t].z, x[t][u].z, sc[u]);
                sc[u] = fmaf(qr[t].w, x[t][u].w, sc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sc[u] += __shfl_xor(sc[u], 1, WAVE);
            sc[u] += __shfl_xor(sc[u], 2, WAVE);
            sc[u] += __shfl_xor(sc[u], 4, WAVE);
        }
        if (!wsel) {
            if (g8 == 0) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (id[u] >= 0) gst(sq + i0 + u * NG + grp, sc[u]);
            }
            continue;
        }
        // the wave's 8 groups, one candidate each, in group order
#pragma unroll
        for (int u = 0; u < U; ++u) tk.offer(score_key(sc[u], id[u]), 0x0101010101010101ull);
    }
    __syncthreads();
    if (ts && threadIdx.x == 0) ts[3] = wall_clock64();
    const int ldo = a.ldo ? a.ldo : a.k;
    int32_t* oq = out_ids + (size_t)q * ldo;
    float* osq = out_scores ? out_scores + (size_t)q * ldo : nullptr;
    const uint32_t kk = min((uint32_t)kq, nc);
    if (wsel) {
        // merge the waves' lists: rank each entry among all of them
        constexpr int NW = TB / WAVE;
        uint64_t* wl = reinterpret_cast<uint64_t*>(tk_key);  // NW * k entries (<= 1024 x 8 B)
        if (lane < kq) wl[wv * kq + lane] = tk.w[0];
        __syncthreads();
        const int tot = NW * kq;
        for (int e = threadIdx.x; e < tot; e += TB) {
            const uint64_t ke = wl[e];
            if (ke == ~0ull) continue;
            uint32_t r = 0;
            for (int o = 0; o < tot; ++o) r += wl[o] < ke ? 1u : 0u;
            if (r < kk) {
                oq[r] = (int32_t)(uint32_t)(ke & 0xffffffffu);
                if (osq) osq[r] = desc_key_any_value((uint32_t)(ke >> 32));
            }
        }
    } else {
        __syncthreads();
        // radix select of the k-th smallest descending key
        uint32_t prefix = 0, pmask = 0, k = kk, eqc = 0;
        if (kk > 0) {
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int i = threadIdx.x; i < 256; i += TB) hist[i] = 0;
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < nc; i += TB) {
                    const uint32_t key = desc_key_any(gld(sq + i));
                    if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
                }
                __syncthreads();
                int d;
                uint32_t below;
                block_find_digit<Whole<TB>>(hist, k, sh, d, below, eqc, &s_digit, &s_below);
                prefix |= (uint32_t)d << shift;
                pmask |= 255u << shift;
                k -= below;
            }
        }
        const uint32_t T = prefix;
        // when more candidates share the k-th key than are needed, a second
        // radix select over their (unique) ids keeps the k smallest
        uint32_t IT = 0xffffffffu;
        if (kk > 0 && eqc > k) {
            uint32_t ip = 0, im = 0, kr = k;
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int i = threadIdx.x; i < 256; i += TB) hist[i] = 0;
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < nc; i += TB) {
                    const uint32_t id = (uint32_t)gld(cq + i);
                    if (desc_key_any(gld(sq + i)) == T && (id & im) == ip) atomicAdd(&hist[(id >> shift) & 255u], 1u);
                }
                __syncthreads();
                int d;
                uint32_t below, eqd;
                block_find_digit<Whole<TB>>(hist, kr, sh, d, below, eqd, &s_digit, &s_below);
                ip |= (uint32_t)d << shift;
                im |= 255u << shift;
                kr -= below;
            }
            IT = ip;
        }
        uint32_t written = 0;
        for (uint32_t c0 = 0; c0 < nc && kk > 0; c0 += TB) {
            const uint32_t i = c0 + threadIdx.x;
            const uint32_t key = i < nc ? desc_key_any(gld(sq + i)) : 0xffffffffu;
            const uint32_t id = i < nc ? (uint32_t)gld(cq + i) : 0xffffffffu;
            const uint32_t take = (i < nc && (key < T || (key == T && id <= IT))) ? 1u : 0u;
            uint32_t tot_take;
            const uint32_t ex_take = block_excl_scan_u32<Whole<TB>>(take, sh, tot_take);
            if (take && written + ex_take < kk) {
                tk_key[written + ex_take] = key;
                tk_id[written + ex_take] = id;
            }
            written += tot_take;
        }
        __syncthreads();
        // order the kept entries by (key, id): rank by counting (kk <= 1024)
        for (uint32_t i = threadIdx.x; i < kk; i += TB) {
            const uint32_t ki = tk_key[i], ii = tk_id[i];
            uint32_t r = 0;
            for (uint32_t j = 0; j < kk; ++j) {
                const uint32_t kj = tk_key[j], ij = tk_id[j];
                r += (kj < ki || (kj == ki && ij < ii)) ? 1u : 0u;
            }
            oq[r] = (int32_t)ii;
            if (osq) osq[r] = desc_key_any_value(ki);
        }
    }
    for (uint32_t i = kk + threadIdx.x; i < (uint32_t)ldo; i += TB) {
        oq[i] = 0;
        if (osq) osq[i] = 0.0f;
    }
    if (ts && threadIdx.x == 0) ts[4] = wall_clock64();
}

// Persistent: each workgroup takes queries from the work queue until it is
// drained.
template <int TB, int CPL>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(TB <= 512 ? 2 : TB / 256, TB <= 512 ? 2 : TB / 256))) void k_gather_score_f32(GatherArgs a, ReplayArgs rp, const uint32_t* __restrict__ probes,
                                                          const uint64_t* __restrict__ bucket_off,
                                                          const int32_t* __restrict__ ids,
                                                          const float* __restrict__ Xc, const float* __restrict__ Q,
                                                          int64_t ldq, int32_t* __restrict__ out_ids,
                                                          float* __restrict__ out_scores) {
    __shared__ int s_q;
    // leading words of the dedup area that may be nonzero (LDS starts
    // undefined)
    uint32_t dirty = (uint32_t)(a.window_bits / 32);
    for (;;) {
        const int q = claim_query(a, &s_q);
        if (q < 0) break;
        k_gather_score_f32_one<TB, CPL>(a, rp, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, q, dirty);
        __syncthreads();
    }
}

// ---- small queries: one wave per query -------------------------------------
// A query with at most SMALL_MAX_CTOT bucket ids costs a workgroup of the
// kernel above mostly barriers and dependent loads; here each wave serves a
// query of its own: lane p walks bucket p, an open-addressing set in the
// wave's LDS keeps the first occurrence of every id (the reference's bitset,
// src/FalconnPP.cpp:552-611), and the rows are scored with the lane layout
// and order of addition of k_gather_score_f32 (8 lanes per row, the same
// chunks, the same shuffles), so scores and the (score, smaller id) ranking
// are bit for bit the same; the top k stay in the wave's registers.
constexpr int SMALL_T = 256;
constexpr int SMALL_WAVES = SMALL_T / WAVE;
constexpr int SMALL_HASH_LG = 10;  // 1024 words, at least twice SMALL_MAX_CTOT
static_assert((1u << SMALL_HASH_LG) >= 2 * SMALL_MAX_CTOT, "small-query set too small");

template <int CPL>
__global__ __launch_bounds__(SMALL_T) void k_gather_small_f32(GatherArgs a, const uint32_t* __restrict__ probes,
                                                              const uint64_t* __restrict__ bucket_off,
                                                              const int32_t* __restrict__ ids,
                                                              const float* __restrict__ Xc,
                                                              const float* __restrict__ Q, int64_t ldq,
                                                              int32_t* __restrict__ out_ids,
                                                              float* __restrict__ out_scores) {
    __shared__ uint32_t hs[SMALL_WAVES][1 << SMALL_HASH_LG];
    __shared__ int32_t cl[SMALL_WAVES][SMALL_MAX_CTOT];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = threadIdx.x / WAVE;
    uint32_t* H = hs[wv];
    int32_t* C = cl[wv];
    for (int e = lane; e < (1 << SMALL_HASH_LG); e += WAVE) H[e] = 0u;
    wave_lds_sync();
    const uint32_t nsmall = __hip_atomic_load(a.heads + a.small_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t begin = a.nitems - nsmall;
    // small queries cost about the same: wave g takes items g, g + waves, ...
    // (a shared claim counter serialised thousands of waves on one atomic)
    const uint32_t gw = blockIdx.x * SMALL_WAVES + (uint32_t)wv;
    const uint32_t nw = gridDim.x * SMALL_WAVES;
    const int g8 = lane & 7;
    const int grp = lane >> 3;
    const int64_t nchunk = a.ldx / 4;
    // two rows per lane group in flight (one from 64 features per lane up),
    // so that four waves per SIMD fit at d <= 256
    constexpr int U = CPL >= 16 ? 1 : 2;
    const int ldo = a.ldo ? a.ldo : a.k;
    for (uint32_t it = gw; it < nsmall; it += nw) {
        const int q = (int)gld(a.order + begin + it);
        uint64_t* ts = a.phase_ts ? a.phase_ts + (size_t)q * PHASE_STAMPS : nullptr;
        if (ts && lane == 0) ts[0] = wall_clock64();
        const QueryView qv_ = query_view(a.qp, q, a.qProbes, 0, 0, a.k);
        const uint32_t* pr = probes + qv_.probe_off;
        // bucket walk and dedup: lane p takes bucket c0 + p, four ids per load
        // (the id array is padded past its end); first occurrences appended
        uint32_t nc = 0;
        for (int c0 = 0; c0 < qv_.qProbes; c0 += WAVE) {
            uint64_t s = 0;
            uint32_t len = 0;
            if (c0 + lane < qv_.qProbes) {
                const uint32_t b = gld(pr + c0 + lane);
                s = gld(bucket_off + b);
                len = (uint32_t)(gld(bucket_off + b + 1) - s);
            }
            uint32_t mx = len;
            for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, WAVE));
            mx = (uint32_t)__builtin_amdgcn_readfirstlane((int)mx);
            // up to 32 ids per lane in flight before the first insert
            for (uint32_t j0 = 0; j0 < mx; j0 += 32) {
                i32x4_a4 t[8];
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    t[g] = i32x4_a4{0, 0, 0, 0};
                    if (j0 + 4 * g < len) t[g] = gld_i32x4(ids + s + j0 + 4 * g);
                }
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    if (j0 + 4 * g >= mx) break;  // wave-uniform
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t j = j0 + 4 * g + r;
                        const int32_t v = t[g][r];
                        const bool fresh = j < len && hashset_insert(H, SMALL_HASH_LG, v);
                        const uint64_t m = __ballot(fresh);
                        const uint32_t pos = nc + lane_rank(m);
                        if (fresh && pos < SMALL_MAX_CTOT) C[pos] = v;
                        nc += (uint32_t)__popcll(m);
                    }
                }
            }
        }
        nc = min(nc, SMALL_MAX_CTOT);
        wave_lds_sync();
        if (ts && lane == 0) ts[1] = ts[2] = wall_clock64();
        if (lane == 0) gst(a.ncand + q, (uint64_t)nc);
        if (a.seg_off) {
            int32_t* dst = a.cand + gld(a.seg_off + q);
            for (uint32_t e = lane; e < nc; e += WAVE) gst(dst + e, C[e]);
        }
        // FP32 scores, as in k_gather_score_f32: 8 lanes per row, CPL float4
        // chunks per lane (past the row end: the last chunk against a zero
        // query chunk), U rows per lane group in flight
        const float4* q4 = reinterpret_cast<const float4*>(Q + (int64_t)q * ldq);
        float4 qr[CPL];
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int64_t ch = g8 + 8 * t;
            qr[t] = ch < nchunk ? q4[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        WaveTopK<1> tk;
        tk.init(qv_.k);
        for (uint32_t i0 = 0; i0 < nc; i0 += 8 * U) {
            int id[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t ii = i0 + (uint32_t)(u * 8 + grp);
                id[u] = ii < nc ? C[ii] : -1;
            }
            float sc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) sc[u] = 0.f;
            float4 x[CPL][U];
#pragma unroll
            for (int t = 0; t < CPL; ++t) {
                const int64_t ch = min<int64_t>(g8 + 8 * t, nchunk - 1);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    x[t][u] = reinterpret_cast<const float4*>(Xc + (int64_t)(id[u] < 0 ? 0 : id[u]) * a.ldx)[ch];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < CPL; ++t) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    sc[u] = fmaf(qr[t].x, x[t][u].x, sc[u]);
                    sc[u] = fmaf(qr[t].y, x[t][u].y, sc[u]);
                    sc[u] = fmaf(qr[t].z, x[t][u].z, sc[u]);
                    sc[u] = fmaf(qr[t].w, x[t][u].w, sc[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                sc[u] += __shfl_xor(sc[u], 1, WAVE);
                sc[u] += __shfl_xor(sc[u], 2, WAVE);
                sc[u] += __shfl_xor(sc[u], 4, WAVE);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) tk.offer(score_key(sc[u], id[u]), 0x0101010101010101ull);
        }
        if (ts && lane == 0) ts[3] = wall_clock64();
        // lane j holds the j-th best key
        int32_t* oq = out_ids + (size_t)q * ldo;
        float* osq = out_scores ? out_scores + (size_t)q * ldo : nullptr;
        const uint32_t kk = min((uint32_t)qv_.k, nc);
        for (uint32_t j = (uint32_t)lane; j < (uint32_t)ldo; j += WAVE) {
            const bool hit = j < kk;
            const uint64_t ke = tk.w[0];
            gst(oq + j, hit ? (int32_t)(uint32_t)(ke & 0xffffffffu) : 0);
            if (osq) gst(osq + j, hit ? desc_key_any_value((uint32_t)(ke >> 32)) : 0.0f);
        }
        if (ts && lane == 0) ts[4] = wall_clock64();
        // the set is empty again for the next query
        wave_lds_sync();
        for (int e = lane; e < (1 << SMALL_HASH_LG); e += WAVE) H[e] = 0u;
        wave_lds_sync();
    }
}

// Sort key of a query: its best bucket of table 0 (layer-1 and layer-2 top
// codes).  Queries with equal keys share most candidates.
// sbit: the bit above the bucket keys (set for small queries).
__global__ void k_order_keys(const int32_t* __restrict__ qc, const uint32_t* __restrict__ qflag, int nq, int L,
                             int PL, int two_np, uint32_t* __restrict__ key, SmallPick sp, int sbit) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool small = q < nq && small_query(sp, qflag, q);
    if (sp.max_ctot) count_small(small, sp.count);
    if (q >= nq) return;
    const size_t base = (size_t)q * L * 2 * PL;
    const uint32_t k = qflag[q] ? 0u : 1u + (uint32_t)(qc[base] * two_np + qc[base + PL]);
    key[q] = small ? k | (1u << sbit) : k;
}

// dst row r = src row r (D floats) followed by zeros up to ldd.
__global__ void k_pad_rows(const float* __restrict__ src, int64_t lds, int64_t rows, int D, float* __restrict__ dst,
                           int64_t ldd) {
    const int64_t total = rows * ldd;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / ldd;
        const int c = (int)(i - r * ldd);
        dst[i] = c < D ? src[r * lds + c] : 0.0f;
    }
}

// ---- query order: one-workgroup sort ----------------------------------------
// out[0 .. n) = the positions 0 .. n-1 ordered by keys[], stably, for n <=
// SORT1_MAX; with split, the positions whose key is >= small_min (the small
// queries) come last in position order and only the others are sorted.  LSD
// passes of 8 bits over the low `bits` bits with the keys and 16-bit positions
// in LDS.  A pass: digit histogram, its exclusive scan, then tiles of SORT1_T
// elements in order; in a tile, eight ballots give each lane the lanes of its
// wave with the same digit, the waves' counts per digit are scanned across
// the waves, and each element lands at digit start + earlier tiles + earlier
// waves + its rank.
constexpr int SORT1_T = 1024;
constexpr int SORT1_WAVES = SORT1_T / WAVE;

__global__ __launch_bounds__(SORT1_T) void k_sort1(const uint32_t* __restrict__ keys, int n_all, int bits, int split,
                                                   uint32_t small_min, uint32_t* __restrict__ out) {
    extern __shared__ uint32_t sm[];
    __shared__ uint32_t start[256], run[256], tot[256];
    __shared__ uint32_t wb[SORT1_WAVES], ws[SORT1_WAVES];
    uint32_t* k0 = sm;
    uint32_t* k1 = k0 + n_all;
    uint32_t* wcnt = k1 + n_all;  // SORT1_WAVES * 256
    uint16_t* v0 = reinterpret_cast<uint16_t*>(wcnt + SORT1_WAVES * 256);
    uint16_t* v1 = v0 + n_all;
    const int t = threadIdx.x;
    const int lane = t & (WAVE - 1);
    const int wv = t / WAVE;
    // the others to k0 / v0, the small ones' positions to v1, each in order
    int n = 0, ns = 0;
    for (int i0 = 0; i0 < n_all; i0 += SORT1_T) {
        const int i = i0 + t;
        const uint32_t key = i < n_all ? gld(keys + i) : 0u;
        const bool sml = i < n_all && split && key >= small_min;
        const bool oth = i < n_all && !sml;
        const uint64_t mo = __ballot(oth), msm = __ballot(sml);
        if (lane == 0) {
            wb[wv] = (uint32_t)__popcll(mo);
            ws[wv] = (uint32_t)__popcll(msm);
        }
        __syncthreads();
        uint32_t ob = 0, os = 0, tb = 0, tsm = 0;
        for (int w = 0; w < SORT1_WAVES; ++w) {
            ob += w < wv ? wb[w] : 0u;
            os += w < wv ? ws[w] : 0u;
            tb += wb[w];
            tsm += ws[w];
        }
        const uint64_t below = lanes_below_mask();
        if (oth) {
            const uint32_t p = (uint32_t)n + ob + (uint32_t)__popcll(mo & below);
            k0[p] = key;
            v0[p] = (uint16_t)i;
        }
        if (sml) v1[ns + os + (uint32_t)__popcll(msm & below)] = (uint16_t)i;
        n += (int)tb;
        ns += (int)tsm;
        __syncthreads();
    }
    for (int j = t; j < ns; j += SORT1_T) gst(out + n + j, (uint32_t)v1[j]);
    __syncthreads();
    for (int shift = 0; n > 1 && shift < bits; shift += 8) {
        if (t < 256) start[t] = 0u;
        __syncthreads();
        for (int i = t; i < n; i += SORT1_T) atomicAdd(&start[(k0[i] >> shift) & 255u], 1u);
        __syncthreads();
        if (wv == 0) {
            // exclusive scan of the 256 counts, four per lane
            uint32_t c[4], s = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                c[j] = start[lane * 4 + j];
                s += c[j];
            }
            const uint32_t ex = wave_incl_scan(s) - s;
            uint32_t r = ex;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                start[lane * 4 + j] = r;
                run[lane * 4 + j] = 0u;
                r += c[j];
            }
        }
        __syncthreads();
        for (int i0 = 0; i0 < n; i0 += SORT1_T) {
            const int i = i0 + t;
            const bool valid = i < n;
            const uint32_t key = valid ? k0[i] : 0u;
            const uint32_t d = (key >> shift) & 255u;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t set = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? set : ~set;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lanes_below_mask());
            for (int j = t; j < SORT1_WAVES * 256; j += SORT1_T) wcnt[j] = 0u;
            __syncthreads();
            if (valid && rank == 0) wcnt[wv * 256 + d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (t < 256) {
                uint32_t r = 0;
                for (int w = 0; w < SORT1_WAVES; ++w) {
                    const uint32_t c = wcnt[w * 256 + t];
                    wcnt[w * 256 + t] = r;
                    r += c;
                }
                tot[t] = r;
            }
            __syncthreads();
            if (valid) {
                const uint32_t pos = start[d] + run[d] + wcnt[wv * 256 + d] + rank;
                k1[pos] = key;
                v1[pos] = v0[i];
            }
            __syncthreads();
            if (t < 256) run[t] += tot[t];
        }
        __syncthreads();
        uint32_t* tk = k0;
        k0 = k1;
        k1 = tk;
        uint16_t* tv = v0;
        v0 = v1;
        v1 = tv;
    }
    for (int i = t; i < n; i += SORT1_T) gst(out + i, (uint32_t)v0[i]);
}

// k_pad_rows plus the call's initial state in the same launch: flags[0 ..
// nflags) = fv, counters[0 .. ncounters) = 0.
__global__ void k_pad_rows_init(const float* __restrict__ src, int64_t lds, int64_t rows, int D,
                                float* __restrict__ dst, int64_t ldd, uint32_t* __restrict__ flags, uint32_t fv,
                                int nflags, uint32_t* __restrict__ counters, int ncounters) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = rows * ldd;
    for (int64_t i = t0; i < total; i += stride) {
        const int64_t r = i / ldd;
        const int c = (int)(i - r * ldd);
        dst[i] = c < D ? src[r * lds + c] : 0.0f;
    }
    for (int64_t i = t0; i < nflags; i += stride) flags[i] = fv;
    if (t0 < ncounters) counters[t0] = 0u;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

size_t gather_lds_bytes(const GatherArgs& a, int TB) {
    const int64_t wwords = (a.window_bits / 32 + 1) & ~1;
    return (size_t)(wwords + window_counter_words(a.N, a.window_bits) + TB + 256 + 2048) * 4;
}

template <int TB>
void dispatch_gather_cpl(const GatherArgs& a, const ReplayArgs& rp, size_t lds, unsigned grid,
                         const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids, const float* Xc,
                         const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores, hipStream_t st) {
    const int64_t nchunk = a.ldx / 4;
    const int cpl = (int)ceil_div(nchunk, 8);
#define FPP_G(C) \
    k_gather_score_f32<TB, C><<<grid, TB, lds, st>>>(a, rp, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores)
    if (cpl <= 1)
        FPP_G(1);
    else if (cpl <= 2)
        FPP_G(2);
    else if (cpl <= 4)
        FPP_G(4);
    else if (cpl <= 8)
        FPP_G(8);
    else if (cpl <= 16)
        FPP_G(16);
    else
        FPP_G(32);
#undef FPP_G
    FPP_HIP(hipGetLastError());
}

}  // namespace

namespace {
size_t probe_fixed_lds(const ProbeGeom& g) { return (size_t)(PS_WAVES * 256 + 256 + g.L) * 4; }
size_t probe_full_lds(const ProbeGeom& g) {
    return ((size_t)g.L * g.PL * g.PL + (size_t)g.L * 2 * g.PL) * 4 + probe_fixed_lds(g);
}
size_t probe_two_stage_lds(const ProbeGeom& g) { return (size_t)2 * g.L * g.perTable * 4 + probe_fixed_lds(g); }
}  // namespace

// The pool of every pair when it fits the LDS, else the per-table top lists
// when those fit, else none: the queries then take the exact replay.
int probe_select_path(const ProbeGeom& g) {
    if (probe_full_lds(g) <= 160 * 1024) return PROBE_FULL_POOL;
    if (probe_two_stage_lds(g) <= 160 * 1024) return PROBE_TWO_STAGE;
    return PROBE_EXACT;
}

void launch_probe_select(const ProbeGeom& g, int nq, const int32_t* qc, const float* qv,
                         const uint64_t* bucket_off, uint32_t* probes, uint64_t* ctot, uint32_t* qflag,
                         hipStream_t st) {
    const int path = probe_select_path(g);
    FPP_REQUIRE(path != PROBE_EXACT, FPP_EINTERNAL, "probe pool does not fit the LDS: replay the queries instead");
    if (nq == 0) return;
    if (g.PL == 1 && g.L <= PS1_S * WAVE)
        // one probe per layer for every query of the call: a wave per query
        k_probe_select_pl1<<<(unsigned)ceil_div(nq, PS1_T / WAVE), PS1_T, 0, st>>>(g, nq, qc, qv, bucket_off, probes,
                                                                                  ctot, qflag);
    else if (path == PROBE_FULL_POOL)
        k_probe_select<true><<<nq, PS_T, probe_full_lds(g), st>>>(g, qc, qv, bucket_off, probes, ctot, qflag);
    else
        k_probe_select<false><<<nq, PS_T, probe_two_stage_lds(g), st>>>(g, qc, qv, bucket_off, probes, ctot, qflag);
    FPP_HIP(hipGetLastError());
}

void launch_order_keys(const int32_t* qc, const uint32_t* qflag, int nq, int L, int PL, int two_np, uint32_t* key,
                       const SmallPick& sp, hipStream_t st) {
    if (nq == 0) return;
    // bucket keys 1 + b run up to NB = two_np^2, which takes log2(NB) + 1
    // bits; sbit is the bit above them
    const int sbit = ilog2_exact((int64_t)two_np * two_np) + 1;
    FPP_REQUIRE(sbit >= 1 && sbit < 32, FPP_EINTERNAL, "order key width");
    k_order_keys<<<(unsigned)ceil_div(nq, 256), 256, 0, st>>>(qc, qflag, nq, L, PL, two_np, key, sp, sbit);
    FPP_HIP(hipGetLastError());
}

size_t sort1_lds_bytes(int n) { return (size_t)n * 12 + (size_t)SORT1_WAVES * 256 * 4; }

void launch_sort1(const uint32_t* keys, int n, int bits, bool split, uint32_t small_min, uint32_t* out,
                  hipStream_t st) {
    FPP_REQUIRE(n >= 1 && n <= SORT1_MAX && bits >= 1 && bits <= 32, FPP_EINTERNAL, "one-workgroup sort size");
    k_sort1<<<1, SORT1_T, sort1_lds_bytes(n), st>>>(keys, n, bits, split ? 1 : 0, small_min, out);
    FPP_HIP(hipGetLastError());
}

void launch_pad_rows_init(const float* src, int64_t lds, int64_t rows, int D, float* dst, int64_t ldd,
                          uint32_t* flags, uint32_t fv, int nflags, uint32_t* counters, int ncounters,
                          hipStream_t st) {
    const int64_t work = std::max<int64_t>({rows * ldd, (int64_t)nflags, (int64_t)ncounters, 1});
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, 256), 4096));
    k_pad_rows_init<<<grid, 256, 0, st>>>(src, lds, rows, D, dst, ldd, flags, fv, nflags, counters, ncounters);
    FPP_HIP(hipGetLastError());
}

void launch_pad_rows(const float* src, int64_t lds, int64_t rows, int D, float* dst, int64_t ldd, hipStream_t st) {
    if (rows <= 0) return;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(rows * ldd, 256), 4096));
    k_pad_rows<<<grid, 256, 0, st>>>(src, lds, rows, D, dst, ldd);
    FPP_HIP(hipGetLastError());
}

void launch_fill_u32(uint32_t* p, uint32_t v, int n, hipStream_t st) {
    if (n <= 0) return;
    k_fill_u32<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(p, v, n);
    FPP_HIP(hipGetLastError());
}

unsigned persistent_grid(int nitems, size_t lds_bytes, int per_cu) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        FPP_HIP(hipGetDevice(&dev));
        FPP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    // resident workgroups per CU: as asked, as far as their LDS fits
    const int fit = std::max<int>(1, std::min<int>(per_cu, (int)(160 * 1024 / (lds_bytes + 1024))));
    return (unsigned)std::max(1, std::min(nitems, cus * fit));
}

uint64_t count_empty_probes(const uint32_t* probes, int64_t n, const uint64_t* bucket_off, uint32_t* scratch2,
                            hipStream_t st) {
    if (n <= 0) return 0;
    launch_fill_u32(scratch2, 0u, 2, st);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096));
    k_count_empty<<<grid, 256, 0, st>>>(probes, n, bucket_off, reinterpret_cast<unsigned long long*>(scratch2));
    FPP_HIP(hipGetLastError());
    unsigned long long c = 0;
    FPP_HIP(hipMemcpyAsync(&c, scratch2, 8, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipStreamSynchronize(st));
    return c;
}

size_t replay_lds_bytes(ReplayArgs& r, int threads, int per_cu) {
    // the workgroup's share of the CU's LDS, less room for the kernels' own
    // words, so that the replay never costs the gather a resident workgroup
    const size_t budget = std::min<size_t>(150 * 1024, 160 * 1024 / (size_t)std::max(1, per_cu) - 2048);
    const int64_t tlds = exact_table_lds_bytes(r.hg.F, r.pg.PL, r.pg.perTable);
    if (tlds > (int64_t)budget) return 0;
    const size_t staged = (size_t)merge_lds_bytes(r.pg.L, r.pg.perTable, true);
    const size_t in_place = (size_t)merge_lds_bytes(r.pg.L, r.pg.perTable, false);
    if (in_place > budget) return 0;
    r.table_lds = (int)tlds;
    r.twaves = (int)std::min<int64_t>(threads / WAVE, (int64_t)budget / tlds);
    r.merge_staged = staged <= budget ? 1 : 0;
    return std::max((size_t)r.twaves * (size_t)tlds, r.merge_staged ? staged : in_place);
}

int64_t probe_replay_work_bytes(const ProbeGeom& g) { return ((int64_t)g.L * g.perTable * 8 + 255) / 256 * 256; }

void launch_probe_replay(const HashGeom& hg, const ProbeGeom& g, const float* Q, int64_t ldq, const uint32_t* masks,
                         const uint32_t* qlist, const uint32_t* nflag, uint32_t* claim, const uint32_t* qflag,
                         const int32_t* qc, const float* qv, const uint64_t* bucket_off, uint32_t* probes,
                         uint64_t* ctot, void* work, int64_t wpq, unsigned cap, unsigned over, hipStream_t st) {
    constexpr size_t budget = 150 * 1024;  // dynamic LDS; the kernels' own words stay below 2 KB
    const int64_t tlds = exact_table_lds_bytes(hg.F, g.PL, g.perTable);
    FPP_REQUIRE(tlds <= (int64_t)budget, FPP_EINVAL, "perTable too large for the exact probe replay");
    const size_t staged = (size_t)merge_lds_bytes(g.L, g.perTable, true);
    const size_t in_place = (size_t)merge_lds_bytes(g.L, g.perTable, false);
    FPP_REQUIRE(in_place <= budget, FPP_EINVAL, "n_tables too large for the exact probe merge");
    const bool stage = staged <= budget;
    const size_t mlds = stage ? staged : in_place;
    unsigned char* w = reinterpret_cast<unsigned char*>(work);
    if (cap > 0) {
        const unsigned tgrid = (unsigned)std::min<int64_t>((int64_t)cap * g.L, 4096);
        k_replay_tables<<<tgrid, WAVE, (size_t)tlds, st>>>(hg, g, Q, ldq, masks, qlist, nflag, cap, qflag, qc, qv, w,
                                                          wpq);
        FPP_HIP(hipGetLastError());
        k_replay_select<<<cap, RS_SEL_T, mlds, st>>>(g, qlist, nflag, cap, bucket_off, probes, ctot, w, wpq,
                                                     stage ? 1 : 0);
        FPP_HIP(hipGetLastError());
    }
    if (over > 0) {
        const int twaves = (int)std::min<int64_t>(RX_W, (int64_t)budget / tlds);
        const size_t lds = std::max((size_t)twaves * (size_t)tlds, mlds);
        k_probe_replay<<<over, RX_T, lds, st>>>(hg, g, Q, ldq, masks, qlist, nflag, cap, claim, qflag, qc, qv,
                                                bucket_off, probes, ctot, w + (int64_t)cap * wpq, wpq, (int)tlds,
                                                twaves, stage ? 1 : 0);
        FPP_HIP(hipGetLastError());
    }
}

void launch_compact_flagged(int nq, const uint32_t* qflag, uint32_t* list, uint32_t* count, hipStream_t st) {
    if (nq <= 0) return;
    k_compact_flagged<<<1, CF_T, 0, st>>>(nq, qflag, list, count);
    FPP_HIP(hipGetLastError());
}

void launch_ctot_keys(const uint64_t* ctot, const uint32_t* qflag, int nq, uint32_t* key, const SmallPick& sp,
                      hipStream_t st) {
    if (nq <= 0) return;
    k_ctot_keys<<<(unsigned)ceil_div(nq, 256), 256, 0, st>>>(ctot, qflag, nq, key, sp);
    FPP_HIP(hipGetLastError());
}

// extra_words: LDS words the kernel needs beyond the bitset and window counters.
int64_t gather_window_bits(int64_t N, int64_t extra_words, int per_cu) {
    const int64_t fixed = (extra_words + 64) * 4 + 1024;  // + the kernels' own LDS words
    const int64_t budget = 160 * 1024 / std::max(1, per_cu) - 1024 - fixed;
    const int64_t single = (N + 31) / 32 * 32;
    if (single / 8 + 8 <= budget) return single;
    for (int64_t w = int64_t(1) << 20; w >= (int64_t(1) << 16); w >>= 1)
        if (w / 8 + 8 + 4 * window_counter_words(N, w) <= budget) return w;
    FPP_REQUIRE(false, FPP_EINVAL, "n_points too large for the dedup window in LDS");
    return 0;
}

size_t gather_f32_lds(const GatherArgs& a, int threads) { return gather_lds_bytes(a, threads); }

void launch_gather_small(const GatherArgs& a, const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids,
                         const float* Xc, const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores,
                         hipStream_t st) {
    FPP_REQUIRE(a.small_at >= 0 && a.order && a.ncand && a.heads, FPP_EINTERNAL, "small-query gather arguments");
    FPP_REQUIRE(a.ldx % 4 == 0 && a.ldx <= 1024, FPP_EINVAL, "n_features must be <= 1024");
    if (a.nitems == 0) return;
    // persistent: up to four workgroups (sixteen waves) per CU, no more than
    // one wave per query
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        FPP_HIP(hipGetDevice(&dev));
        FPP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(a.nitems, SMALL_WAVES), 4 * cus));
    const int cpl = (int)ceil_div(a.ldx / 4, 8);
#define FPP_S(C) \
    k_gather_small_f32<C><<<grid, SMALL_T, 0, st>>>(a, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores)
    if (cpl <= 1)
        FPP_S(1);
    else if (cpl <= 2)
        FPP_S(2);
    else if (cpl <= 4)
        FPP_S(4);
    else if (cpl <= 8)
        FPP_S(8);
    else if (cpl <= 16)
        FPP_S(16);
    else
        FPP_S(32);
#undef FPP_S
    FPP_HIP(hipGetLastError());
}

void launch_gather_score_f32(const GatherArgs& a, const ReplayArgs& rp, size_t lds, unsigned grid,
                             const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids, const float* Xc,
                             const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores, int threads,
                             hipStream_t st) {
    FPP_REQUIRE(a.k >= 1 && a.k <= 1024, FPP_EINVAL, "n_neighbors must be in [1, 1024]");
    FPP_REQUIRE(a.ldx % 4 == 0, FPP_EINTERNAL, "row stride must be a multiple of 4");
    FPP_REQUIRE(a.ldx <= 1024, FPP_EINVAL, "n_features must be <= 1024");
    FPP_REQUIRE(a.heads && a.ncand && a.cand && a.part && a.nqueues >= 1, FPP_EINTERNAL, "gather scratch missing");
    FPP_REQUIRE(lds >= gather_lds_bytes(a, threads <= 256 ? 256 : (threads <= 512 ? 512 : 1024)) &&
                    lds <= 160 * 1024,
                FPP_EINTERNAL, "gather LDS size");
    if (a.nitems == 0 || grid == 0) return;
    if (threads <= 256)
        dispatch_gather_cpl<256>(a, rp, lds, grid, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    else if (threads <= 512)
        dispatch_gather_cpl<512>(a, rp, lds, grid, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    else
        dispatch_gather_cpl<1024>(a, rp, lds, grid, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
}

}  // namespace fpp
