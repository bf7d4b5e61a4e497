// Cross-polytope hashing kernels for gfx950 (reference: src/FalconnPP.cpp:75-201
// for the build side, :358-453 for the query side).
//
// One hash = one (vector, table): two layers, each three rounds of a +-1 sign
// flip followed by an unnormalised Walsh-Hadamard transform of length F, then
// the top-p |projections| among the first n_proj coordinates.
//
// Layout: a hash lives in LPH = F/E adjacent lanes, each holding E consecutive
// coordinates in VGPRs: E = min(F, 128) for the build; for the query see
// query_lane_coords.  The butterflies run in the reference's
// ascending-stride order (FFHT == ascending radix-2 butterflies, bit for bit):
// strides 1..E/2 inside the lane, strides E..F/2 across lanes (DPP quad_perm
// for lane-xor 1 and 2, ds_bpermute beyond).  The 4 waves of a workgroup share
// the same points (staged once in LDS, zero-padded to F) and split the tables.
// Sign flips come from packed bit words read as wave-uniform scalars.
//
// Selection: the top-(p+1) values come from a v_med3_f32 insertion chain; the
// codes of the top p are then recovered by equality with those values.  The
// result is unambiguous exactly when the top-(p+1) values are pairwise
// distinct (and, for the build, the top-(p+1) pair sums too).  Otherwise the
// reference's outcome depends on libstdc++ heap internals, so the item is
// listed and replayed by an exact heap emulation (k_hash_build_exact).
//
// FULL kernels (n_proj == F) select over every coordinate; the others mask
// coordinates >= n_proj with per-lane bit words.
#pragma once

#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace fpp {
namespace hashk {

constexpr int HB_THREADS = 256;

// Coordinates per lane of the query hash.  One lane per hash at F <= 128
// when every coordinate is selected (FULL): no cross-lane butterflies, and a
// sign flip is one XOR with a wave-uniform word, about a quarter fewer VALU
// instructions than two lanes.  64 at F = 256, and at F = 128 when
// coordinates past n_proj are masked or the top list is longer than 24 (128
// per lane would spill there).
constexpr int query_lane_coords(int F, bool full, int pm) {
    return F == 256 || (F == 128 && (!full || pm > 24)) ? 64 : (F < 128 ? F : 128);
}

template <int S>
__device__ __forceinline__ float xshfl(float v) {
    if constexpr (S == 1) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    } else if constexpr (S == 2) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    } else {
        return __shfl_xor(v, S, WAVE);
    }
}
template <int S>
__device__ __forceinline__ int xshfl_i(int v) {
    if constexpr (S == 1) {
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    } else if constexpr (S == 2) {
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    } else {
        return __shfl_xor(v, S, WAVE);
    }
}

template <int E>
__device__ __forceinline__ void fwht_inlane(float (&x)[E]) {
#pragma unroll
    for (int h = 1; h < E; h <<= 1) {
#pragma unroll
        for (int i = 0; i < E; i += 2 * h) {
#pragma unroll
            for (int j = i; j < i + h; ++j) {
                const float u = x[j], v = x[j + h];
                x[j] = u + v;
                x[j + h] = u - v;
            }
        }
    }
}

// Stride S*E across lanes: the lower lane keeps u+v, the upper one u-v.
// fma(x, -1, o) rounds o - x exactly once, i.e. it is the subtraction.
template <int E, int S>
__device__ __forceinline__ void cross_stage(float (&x)[E], int part) {
    const float sg = (part & S) ? -1.0f : 1.0f;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const float o = xshfl<S>(x[e]);
        x[e] = fmaf(x[e], sg, o);
    }
}

template <int E, int LPH>
__device__ __forceinline__ void fwht_cross(float (&x)[E], int part) {
    if constexpr (LPH >= 2) cross_stage<E, 1>(x, part);
    if constexpr (LPH >= 4) cross_stage<E, 2>(x, part);
    if constexpr (LPH >= 8) cross_stage<E, 4>(x, part);
}

// Sign flips of one round from packed bits (bit d%32 of word d/32 set =
// factor -1).  With one lane per hash the words are wave-uniform SGPRs and a
// flip is one v_xor_b32 with a scalar operand; with two lanes per hash the
// odd lanes add their difference word through a lane mask.
template <int E, int LPH>
__device__ __forceinline__ void apply_signs(float (&x)[E], const uint32_t* __restrict__ mb, int part,
                                            uint32_t pmask) {
    constexpr int WPL = (E + 31) / 32;
    constexpr int BPW = E < 32 ? E : 32;
#pragma unroll
    for (int wi = 0; wi < WPL; ++wi) {
        if constexpr (LPH == 1) {
            const uint32_t w = __builtin_amdgcn_readfirstlane(mb[wi]);
#pragma unroll
            for (int b = 0; b < BPW; ++b) {
                const uint32_t sb = (w << (31 - b)) & 0x80000000u;
                x[wi * 32 + b] = __uint_as_float(__float_as_uint(x[wi * 32 + b]) ^ sb);
            }
        } else if constexpr (LPH == 2) {
            const uint32_t w0 = __builtin_amdgcn_readfirstlane(mb[wi]);
            const uint32_t w1 = __builtin_amdgcn_readfirstlane(mb[WPL + wi]);
            const uint32_t dw = w0 ^ w1;
#pragma unroll
            for (int b = 0; b < BPW; ++b) {
                const uint32_t s0 = (w0 << (31 - b)) & 0x80000000u;
                const uint32_t sd = (dw << (31 - b)) & 0x80000000u;
                x[wi * 32 + b] = __uint_as_float(__float_as_uint(x[wi * 32 + b]) ^ s0 ^ (pmask & sd));
            }
        } else {
            const uint32_t w = mb[part * WPL + wi];
#pragma unroll
            for (int b = 0; b < BPW; ++b) {
                const uint32_t sb = (w << (31 - b)) & 0x80000000u;
                x[wi * 32 + b] = __uint_as_float(__float_as_uint(x[wi * 32 + b]) ^ sb);
            }
        }
    }
}

// Insert a into the descending top-PM list m (duplicates kept as a multiset).
template <int PM>
__device__ __forceinline__ void med3_insert(float (&m)[PM], float a) {
#pragma unroll
    for (int i = PM - 1; i > 0; --i) m[i] = __builtin_amdgcn_fmed3f(m[i - 1], m[i], a);
    m[0] = fmaxf(m[0], a);
}

template <int PM, int S>
__device__ __forceinline__ void merge_stage(float (&m)[PM]) {
    float snap[PM];
#pragma unroll
    for (int i = 0; i < PM; ++i) snap[i] = xshfl<S>(m[i]);
#pragma unroll
    for (int i = 0; i < PM; ++i) med3_insert<PM>(m, snap[i]);
}

template <int PM, int LPH>
__device__ __forceinline__ void merge_lists(float (&m)[PM]) {
    if constexpr (LPH >= 2) merge_stage<PM, 1>(m);
    if constexpr (LPH >= 4) merge_stage<PM, 2>(m);
    if constexpr (LPH >= 8) merge_stage<PM, 4>(m);
}

template <int LPH>
__device__ __forceinline__ int lanes_max(int v) {
    if constexpr (LPH >= 2) v = max(v, xshfl_i<1>(v));
    if constexpr (LPH >= 4) v = max(v, xshfl_i<2>(v));
    if constexpr (LPH >= 8) v = max(v, xshfl_i<4>(v));
    return v;
}

// Bits of the lane's coordinates that lie at or beyond n_proj.
template <int E>
__device__ __forceinline__ void out_of_range_bits(int base, int n_proj, uint32_t (&ob)[(E + 31) / 32]) {
#pragma unroll
    for (int w = 0; w < (E + 31) / 32; ++w) {
        const int in = n_proj - base - w * 32;  // in-range coordinates in this word
        ob[w] = in <= 0 ? 0xffffffffu : (in >= 32 ? 0u : ~((1u << in) - 1u));
    }
}

// One layer: rotate the staged vector and return the top-PM |values| (m,
// descending) and the codes of the first PM-1 of them (c); `tie` is set when
// two of the first `nchk` values are equal.
template <int F, int PM, bool FULL, int E = (F < 128 ? F : 128)>
__device__ __forceinline__ void hash_layer(const float* __restrict__ src, const uint32_t* __restrict__ masks, int part,
                                           int n_proj, const uint32_t* ob_in, float (&m)[PM], int (&c)[PM], int nchk,
                                           int ncode, bool& tie) {
    constexpr int LPH = F / E;
    constexpr int MW = (F + 31) / 32;
    constexpr int WPL = (E + 31) / 32;
    int base = part * E;
    if constexpr (LPH > 1) asm volatile("" : "+v"(base));
    const uint32_t pmask = (part & 1) ? 0xffffffffu : 0u;
    float x[E];
    // keep this layer's loads from being hoisted above the previous layer
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = src[base + e];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        apply_signs<E, LPH>(x, masks + r * MW, part, pmask);
        fwht_inlane<E>(x);
        fwht_cross<E, LPH>(x, part);
    }
    uint32_t ob[WPL];
#pragma unroll
    for (int w = 0; w < WPL; ++w) {
        ob[w] = FULL ? 0u : ob_in[w];
        if constexpr (!FULL) asm volatile("" : "+v"(ob[w]));
    }
#pragma unroll
    for (int i = 0; i < PM; ++i) m[i] = -1.0f;
    if constexpr (FULL) {
        // |x| as a source modifier of v_med3/v_max: no |x| array is kept live
        // between this pass and the code recovery below
#pragma unroll
        for (int e = 0; e < E; ++e) {
#pragma unroll
            for (int i = PM - 1; i > 0; --i)
                asm("v_med3_f32 %0, %1, %2, |%3|" : "=v"(m[i]) : "v"(m[i - 1]), "v"(m[i]), "v"(x[e]));
            asm("v_max_f32 %0, %1, |%2|" : "=v"(m[0]) : "v"(m[0]), "v"(x[e]));
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            float a = fabsf(x[e]);
            a = ((ob[e / 32] >> (e % 32)) & 1u) ? -1.0f : a;
            med3_insert<PM>(m, a);
        }
    }
    merge_lists<PM, LPH>(m);
    // coordinates past n_proj become NaN, which equals no value
    if constexpr (!FULL) {
#pragma unroll
        for (int e = 0; e < E; ++e)
            x[e] = ((ob[e / 32] >> (e % 32)) & 1u) ? __int_as_float(0x7fc00000) : x[e];
    }
    // code of the coordinate holding m[j] (index and sign of the last
    // match), for the first ncode (wave-uniform) entries only.  The signs are
    // packed 32 coordinates to a word first, so the search compares
    // magnitudes only; sgn(x) < 0 exactly when x < 0 (src/Utilities.h:26-31),
    // -0.0 is positive
    uint32_t neg[WPL];
#pragma unroll
    for (int w = 0; w < WPL; ++w) neg[w] = 0u;
#pragma unroll
    for (int e = 0; e < E; ++e) neg[e / 32] |= x[e] < 0.0f ? (1u << (e % 32)) : 0u;
#pragma unroll
    for (int j = 0; j < PM - 1; ++j) {
        c[j] = 0;
        if (j >= ncode) continue;
        int cj = -1;
#pragma unroll
        for (int e = 0; e < E; ++e) cj = fabsf(x[e]) == m[j] ? e : cj;
        if (cj >= 0) {
            uint32_t nw = neg[0];
#pragma unroll
            for (int w = 1; w < WPL; ++w) nw = (cj >> 5) == w ? neg[w] : nw;
            cj = base + cj + (((nw >> (cj & 31)) & 1u) ? n_proj : 0);
        }
        c[j] = max(0, lanes_max<LPH>(cj));
    }
#pragma unroll
    for (int j = 0; j + 1 < PM; ++j)
        if (j + 1 < nchk && m[j] == m[j + 1]) tie = true;
}

template <int F, int PM, bool FULL>
__global__ __launch_bounds__(HB_THREADS) __attribute__((amdgpu_waves_per_eu(FULL && PM <= 8 ? 2 : 1))) void k_hash_build(HashGeom g, const float* __restrict__ Xc, int64_t ldx,
                                                            int64_t N, int p, const uint32_t* __restrict__ masks,
                                                            uint32_t* __restrict__ codes, float* __restrict__ vals,
                                                            uint64_t* __restrict__ flag_list,
                                                            uint32_t* __restrict__ flag_count, uint32_t flag_cap) {
    constexpr int E = F < 128 ? F : 128;
    constexpr int LPH = F / E;
    constexpr int HPW = WAVE / LPH;
    constexpr int MW = (F + 31) / 32;
    constexpr int S = F + 1;
    extern __shared__ float lds[];
    const int D = g.D;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int slot = lane / LPH, part = (LPH == 1) ? 0 : (lane % LPH);
    const int64_t n0 = (int64_t)blockIdx.x * HPW;
    for (int i = threadIdx.x; i < HPW * F; i += HB_THREADS) {
        const int pt = i / F, d = i % F;
        const int64_t n = n0 + pt;
        lds[pt * S + d] = (n < N && d < D) ? Xc[n * ldx + d] : 0.0f;
    }
    __syncthreads();
    const int64_t n = n0 + slot;
    const bool valid = n < N;
    const float* src = lds + slot * S;
    const int nchk = min(p + 1, g.n_proj);
    const int two_np = 2 * g.n_proj;
    uint32_t ob[(E + 31) / 32];
    out_of_range_bits<E>(part * E, g.n_proj, ob);
    for (int l = wave; l < g.L; l += HB_THREADS / WAVE) {
        float m1[PM], m2[PM];
        int c1[PM], c2[PM];
        bool tie = false;
        hash_layer<F, PM, FULL>(src, masks + (size_t)(l * 2 + 0) * 3 * MW, part, g.n_proj, ob, m1, c1, nchk, p, tie);
        hash_layer<F, PM, FULL>(src, masks + (size_t)(l * 2 + 1) * 3 * MW, part, g.n_proj, ob, m2, c2, nchk, p, tie);
        // top-(p+1) pair sums v2 + v1 (src/FalconnPP.cpp:159-189)
        float sm[PM];
#pragma unroll
        for (int t = 0; t < PM; ++t) sm[t] = -1.0f;
#pragma unroll
        for (int i = 0; i < PM - 1; ++i)
#pragma unroll
            for (int j = 0; j < PM - 1; ++j)
                if (i < p && j < p) med3_insert<PM>(sm, m2[j] + m1[i]);
        const int npair = min(p * p, p + 1);
#pragma unroll
        for (int t = 0; t + 1 < PM; ++t)
            if (t + 1 < npair && sm[t] == sm[t + 1]) tie = true;
        int bc[PM];
#pragma unroll
        for (int t = 0; t < PM; ++t) bc[t] = 0;
#pragma unroll
        for (int i = 0; i < PM - 1; ++i)
#pragma unroll
            for (int j = 0; j < PM - 1; ++j)
                if (i < p && j < p) {
                    const float s = m2[j] + m1[i];
                    const int b = c1[i] * two_np + c2[j];
#pragma unroll
                    for (int t = 0; t < PM - 1; ++t) bc[t] = (s == sm[t]) ? b : bc[t];
                }
        if (valid && part == 0) {
            const size_t o = (size_t)l * N * p + (size_t)n * p;
#pragma unroll
            for (int t = 0; t < PM - 1; ++t)
                if (t < p) {
                    codes[o + (p - 1 - t)] = (uint32_t)bc[t];
                    vals[o + (p - 1 - t)] = sm[t];
                }
            if (tie) {
                const uint32_t k = atomicAdd(flag_count, 1u);
                if (k < flag_cap) flag_list[k] = (uint64_t)n * g.L + l;
            }
        }
    }
}

template <int F, int PM, bool FULL>
__global__ __launch_bounds__(HB_THREADS) __attribute__((amdgpu_waves_per_eu(FULL ? 2 : 1))) void k_hash_query(HashGeom g, const float* __restrict__ Q, int64_t ldq,
                                                            int nq, int PL, const uint32_t* __restrict__ masks,
                                                            int32_t* __restrict__ qc, float* __restrict__ qv,
                                                            uint32_t* __restrict__ qflag) {
    constexpr int E = query_lane_coords(F, FULL, PM);
    constexpr int LPH = F / E;
    constexpr int HPW = WAVE / LPH;
    constexpr int MW = (F + 31) / 32;
    constexpr int S = F + 1;
    extern __shared__ float lds[];
    const int D = g.D;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int slot = lane / LPH, part = (LPH == 1) ? 0 : (lane % LPH);
    const int q0 = blockIdx.x * HPW;
    for (int i = threadIdx.x; i < HPW * F; i += HB_THREADS) {
        const int pt = i / F, d = i % F;
        const int q = q0 + pt;
        lds[pt * S + d] = (q < nq && d < D) ? Q[(int64_t)q * ldq + d] : 0.0f;
    }
    __syncthreads();
    const int q = q0 + slot;
    const bool valid = q < nq;
    const float* src = lds + slot * S;
    const int nchk = min(PL + 1, g.n_proj);
    uint32_t ob[(E + 31) / 32];
    out_of_range_bits<E>(part * E, g.n_proj, ob);
    bool tie = false;
    // workgroup (x, y): queries of block x, tables y*4 + wave, stepping by all
    // workgroups of the column
    constexpr int NWV = HB_THREADS / WAVE;
    for (int l = blockIdx.y * NWV + wave; l < g.L; l += gridDim.y * NWV) {
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            float m[PM];
            int c[PM];
            hash_layer<F, PM, FULL, E>(src, masks + (size_t)(l * 2 + layer) * 3 * MW, part, g.n_proj, ob, m, c,
                                       nchk, PL, tie);
            if (valid && part == 0) {
                const size_t o = (((size_t)q * g.L + l) * 2 + layer) * PL;
#pragma unroll
                for (int j = 0; j < PM - 1; ++j)
                    if (j < PL) {
                        qc[o + j] = c[j];
                        qv[o + j] = m[j];
                    }
            }
        }
    }
    if (valid && part == 0 && tie) atomicOr(qflag + q, 1u);
}

template <int F, int E = (F < 128 ? F : 128)>
inline size_t hash_lds_bytes() {
    return (size_t)(WAVE / (F / E)) * (F + 1) * sizeof(float);
}

template <int F, int PM, bool FULL>
void launch_build_variant(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p, const uint32_t* masks,
                          uint32_t* codes, float* vals, uint64_t* fl, uint32_t* fc, uint32_t cap, hipStream_t st) {
    constexpr int E = F < 128 ? F : 128;
    constexpr int HPW = WAVE / (F / E);
    const dim3 grid((unsigned)ceil_div(N, HPW));
    k_hash_build<F, PM, FULL><<<grid, HB_THREADS, hash_lds_bytes<F>(), st>>>(g, Xc, ldx, N, p, masks, codes, vals, fl,
                                                                            fc, cap);
}

template <int F, int PM, bool FULL>
void launch_query_variant(const HashGeom& g, const float* Q, int64_t ldq, int nq, int PL, const uint32_t* masks,
                          int32_t* qc, float* qv, uint32_t* qflag, hipStream_t st) {
    constexpr int E = query_lane_coords(F, FULL, PM);
    constexpr int HPW = WAVE / (F / E);
    constexpr int NWV = HB_THREADS / WAVE;
    // split the tables over workgroups until about four waves per SIMD of
    // the 256 CUs are in flight (every wave keeps at least one table)
    const int64_t bx = ceil_div(nq, HPW);
    const int64_t want = ceil_div(4 * 4 * 256, bx * NWV);
    const int gy = (int)std::max<int64_t>(1, std::min<int64_t>(want, ceil_div(g.L, NWV)));
    const dim3 grid((unsigned)bx, (unsigned)gy);
    k_hash_query<F, PM, FULL><<<grid, HB_THREADS, hash_lds_bytes<F, E>(), st>>>(g, Q, ldq, nq, PL, masks, qc, qv,
                                                                               qflag);
}

// Per-F entry points, explicitly instantiated in hash_f*.hip (one file per
// group of F so the instantiations compile in parallel).
template <int F>
void dispatch_build_f(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p, const uint32_t* masks,
                      uint32_t* codes, float* vals, uint64_t* fl, uint32_t* fc, uint32_t cap, hipStream_t st) {
    const bool full = g.n_proj == F;
#define FPP_V(PMV)                                                                                          \
    if (full)                                                                                               \
        launch_build_variant<F, PMV, true>(g, Xc, ldx, N, p, masks, codes, vals, fl, fc, cap, st);          \
    else                                                                                                    \
        launch_build_variant<F, PMV, false>(g, Xc, ldx, N, p, masks, codes, vals, fl, fc, cap, st);
    if (p + 1 <= 2) {
        FPP_V(2)
    } else if (p + 1 <= 4) {
        FPP_V(4)
    } else if (p + 1 <= 8) {
        FPP_V(8)
    } else {
        FPP_V(16)
    }
#undef FPP_V
}

template <int F>
void dispatch_query_f(const HashGeom& g, const float* Q, int64_t ldq, int nq, int PL, const uint32_t* masks,
                      int32_t* qc, float* qv, uint32_t* qflag, hipStream_t st) {
    const bool full = g.n_proj == F;
#define FPP_V(PMV)                                                                          \
    if (full)                                                                               \
        launch_query_variant<F, PMV, true>(g, Q, ldq, nq, PL, masks, qc, qv, qflag, st);    \
    else                                                                                    \
        launch_query_variant<F, PMV, false>(g, Q, ldq, nq, PL, masks, qc, qv, qflag, st);
    if (PL + 1 <= 2) {
        FPP_V(2)
    } else if (PL + 1 <= 4) {
        FPP_V(4)
    } else if (PL + 1 <= 8) {
        FPP_V(8)
    } else if (PL + 1 <= 12) {
        FPP_V(12)
    } else if (PL + 1 <= 16) {
        FPP_V(16)
    } else if (PL + 1 <= 24) {
        FPP_V(24)
    } else {
        FPP_V(32)
    }
#undef FPP_V
}

#define FPP_INSTANTIATE_HASH(FF)                                                                                  \
    template void dispatch_build_f<FF>(const HashGeom&, const float*, int64_t, int64_t, int, const uint32_t*,     \
                                       uint32_t*, float*, uint64_t*, uint32_t*, uint32_t, hipStream_t);          \
    template void dispatch_query_f<FF>(const HashGeom&, const float*, int64_t, int, int, const uint32_t*,         \
                                       int32_t*, float*, uint32_t*, hipStream_t);

}  // namespace hashk
}  // namespace fpp
