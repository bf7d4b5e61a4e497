// Reduced-precision scoring with an exact FP32 refine (no counterpart in the
// reference; its FP32 rerank is src/FalconnPP.cpp:627-673).
//
// The centred data gets a second copy in BF16, or in FP8-e4m3 with one FP32
// scale per row (one scale shared by all rows when their magnitudes allow).
// A query gathers and dedups its candidates exactly as the FP32 path does.  It
// then scores them on MFMA (v_mfma_f32_16x16x32_bf16 / _fp8_fp8, FP32
// accumulation), keeps the best R by approximate score, rescores those R
// exactly in FP32 from the master copy and returns the top-k.
//
// MFMA use: the A operand holds 8 candidates as 16 half-rows of a 128-byte
// chunk, the B operand the two matching query halves in columns 0 and 1, so
// every load instruction reads whole 128-byte lines (the layout is spelled out
// at the scoring loop).
#include <cstring>

#include "common.h"
#include "gather_common.h"
#include "kernels.h"
#include "replay_common.h"

namespace fpp {
namespace {

using namespace gather;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// Round to nearest even, as a bit pattern.
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// FP8 scale: the row's largest magnitude maps to 240, inside the finite range
// of both e4m3 encodings.
constexpr float FP8_TARGET = 240.0f;

__global__ void k_quant_bf16(const float* __restrict__ Xc, int64_t ldx, int64_t N, int D, int64_t ldl,
                             uint16_t* __restrict__ Xb) {
    const int64_t total = N * ldl;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i / ldl;
        const int d = (int)(i - n * ldl);
        Xb[i] = d < D ? f32_to_bf16_bits(Xc[n * ldx + d]) : (uint16_t)0;
    }
}

// Largest row magnitude (range[0]) and smallest nonzero one (range[1]) as bit
// patterns (non-negative floats order as their bits); rows grid-strided over
// the waves, one atomic pair per wave.
__global__ __launch_bounds__(256) void k_amax_range(const float* __restrict__ Xc, int64_t ldx, int64_t N, int D,
                                                    uint32_t* __restrict__ range) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / WAVE);
    float hi = 0.0f, lo = INFINITY;
    for (int64_t n = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; n < N; n += nw) {
        const float* x = Xc + n * ldx;
        float amax = 0.0f;
        for (int d = lane; d < D; d += WAVE) amax = fmaxf(amax, fabsf(x[d]));
        for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, WAVE));
        hi = fmaxf(hi, amax);
        if (amax > 0.0f) lo = fminf(lo, amax);
    }
    if (lane == 0) {
        atomicMax(range, __float_as_uint(hi));
        if (lo < INFINITY) atomicMin(range + 1, __float_as_uint(lo));
    }
}

// One wave per row: scale = uscale when given, else amax / 240; codes =
// e4m3(x / scale), zero padded; with in_row the row's last four bytes hold
// its scale (FP32) instead of padding.
__global__ __launch_bounds__(256) void k_quant_fp8(const float* __restrict__ Xc, int64_t ldx, int64_t N, int D,
                                                   int64_t ldl, float uscale, bool in_row, uint8_t* __restrict__ X8,
                                                   float* __restrict__ scale) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t n = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (n >= N) return;
    const float* x = Xc + n * ldx;
    float amax = 0.0f;
    for (int d = lane; d < D; d += WAVE) amax = fmaxf(amax, fabsf(x[d]));
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, WAVE));
    const float s = uscale > 0.0f ? uscale : (amax > 0.0f ? amax / FP8_TARGET : 1.0f);
    const float inv = 1.0f / s;
    if (lane == 0) scale[n] = s;
    // four codes per lane per step
    for (int d0 = lane * 4; d0 < ldl; d0 += WAVE * 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = (d0 + u < D) ? x[d0 + u] * inv : 0.0f;
        int w = 0;
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
        if (in_row && d0 == ldl - 4) w = __float_as_int(s);
        *reinterpret_cast<int*>(X8 + n * ldl + d0) = w;
    }
}

enum Fmt { BF16 = 1, FP8 = 2 };

// 128-byte chunks of a row for kt steps of 32 elements (at least one).
constexpr int lp_chunks(int fmt, int kt) {
    const int c = fmt == BF16 ? kt / 2 : kt / 4;
    return c < 1 ? 1 : c;
}
// Waves per SIMD of a launch: TB threads, PERCU workgroups per CU.
constexpr int lp_waves_per_simd(int tb, int percu) { return tb * percu / (4 * WAVE); }
// Tiles of 8 candidates per wave step (a tile reads 1 KB per chunk): about
// 16 KB of row data in flight per wave at two waves per SIMD, 8 KB at four
// (their registers: 128 per wave), at least one tile.
constexpr int lp_tiles(int fmt, int kt, int waves) {
    const int t = (32 / waves) / lp_chunks(fmt, kt);
    return t < 1 ? 1 : t;
}

// Approximate scores of the candidates cq[0 .. nc) of one query, for the
// calling wave's steps (candidates first .. first + 8 NT - 1, then every
// stride), as order keys by candidate position: keys_lds[i] (keys_mode 0),
// else keys_glb[i] and with keys_mode 1 also the key's top 16 bits at
// (uint16_t*)keys_lds + i.  Returns this lane's smallest key in the low
// half and its largest in the high half.  Out of line, so that its registers
// are allocated for the scoring loop alone (inside the gather kernel the
// loop spilled at its top, and the reload's wait drained every load and
// store in flight on each step).
//
// A candidate enters the MFMA as two half-rows: tile row 2c + h holds half h
// of candidate c's current 128-byte chunk and column n of the B operand holds
// the query's half n of that chunk (columns 2..15 are zero), so candidate c's
// score is the sum over chunks of C[2c][0] + C[2c+1][1].  Lane (g16, r16)
// loads the 16 bytes at 64 h + 16 g16 of the chunk of candidate r16 / 2 (h =
// r16 % 2): one load instruction reads the whole 128-byte lines of 8
// candidates.  Bytes past the row end re-read its last 32 bytes against a
// zero query (adding exactly 0).
template <int FMT, int KT, int NT>
__device__ __noinline__ uint64_t lp_score_keys(const float* __restrict__ qf32, int D, float qscale,
                                              const uint8_t* __restrict__ xbase, int64_t rb, float uscale,
                                              const float* __restrict__ scale, int scale_in_row,
                                              const int32_t* __restrict__ cq, uint32_t nc, uint32_t first,
                                              uint32_t stride, uint32_t* keys_lds, uint32_t* __restrict__ keys_glb,
                                              int keys_mode) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int g16 = lane >> 4;  // k-group of this lane in an MFMA operand
    const int r16 = lane & 15;  // row (A) / column (B) of this lane
    constexpr int KC = lp_chunks(FMT, KT);
    // B operand: lane (g16, n = r16) holds query elements 8 g16 .. 8 g16 + 7
    // of half n of each chunk (FP8: two k-steps per chunk half, s = 0, 1)
    bf16x8 qb[KC];
    long q8[KC][2];
    const float inv = 1.0f / qscale;
#pragma unroll
    for (int ch = 0; ch < KC; ++ch) {
#pragma unroll
        for (int s = 0; s < (FMT == BF16 ? 1 : 2); ++s) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int d = FMT == BF16 ? 64 * ch + 32 * r16 + 8 * g16 + j
                                          : 128 * ch + 64 * r16 + 16 * g16 + 8 * s + j;
                v[j] = (r16 < 2 && d < D) ? qf32[d] : 0.0f;
            }
            if constexpr (FMT == BF16) {
#pragma unroll
                for (int j = 0; j < 8; ++j) qb[ch][j] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(v[j]));
            } else {
                int lo = 0, hi = 0;
                lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * inv, v[1] * inv, lo, false);
                lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * inv, v[3] * inv, lo, true);
                hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4] * inv, v[5] * inv, hi, false);
                hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6] * inv, v[7] * inv, hi, true);
                q8[ch][s] = (long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
            }
        }
    }
    // byte offset of this lane's 16 bytes in each chunk, clamped into the row
    const int c8 = r16 >> 1;
    int64_t boff[KC];
#pragma unroll
    for (int ch = 0; ch < KC; ++ch) {
        const int64_t o = 128 * ch + 64 * (r16 & 1) + 16 * g16;
        boff[ch] = o < rb ? o : rb - 32 + (o & 16);
    }
    // NT tiles of 8 candidates per wave step
    constexpr uint32_t SPAN = 8 * NT;
    // candidate ids one step ahead, as the step's SPAN ids in order over
    // the lanes (IPL per lane, one coalesced load each; a tile's lane
    // takes its candidate's id by a lane shuffle), which keeps IPL
    // registers live across the step rather than NT; every load is issued
    // (clamped to the last candidate) and the id replaced by -1 past the
    // end
    constexpr int IPL = (int)((SPAN + WAVE - 1) / WAVE);
    const uint32_t last = nc - 1;
    int32_t idn[IPL];
#pragma unroll
    for (int j = 0; j < IPL; ++j) {
        const uint32_t i = first + (uint32_t)(j * WAVE + lane);
        const int32_t v = gld(cq + min(i, last));
        idn[j] = i < nc && (uint32_t)(j * WAVE + lane) < SPAN ? v : -1;
    }
    uint32_t kmn = 0xffffffffu, kmx = 0u;  // this lane's range of the keys
    for (uint32_t t0 = first; t0 < nc; t0 += stride) {
        int32_t idv[IPL];
#pragma unroll
        for (int j = 0; j < IPL; ++j) {
            idv[j] = idn[j];
            const uint32_t i = t0 + stride + (uint32_t)(j * WAVE + lane);
            const int32_t v = gld(cq + min(i, last));
            idn[j] = i < nc && (uint32_t)(j * WAVE + lane) < SPAN ? v : -1;
        }
        // candidate h * 8 + c8 of the step: lane (h % 8) * 8 + c8 of
        // register h / 8
        int32_t id[NT];
        float rs[NT];
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            id[h] = __shfl(idv[h / 8], (h % 8) * 8 + c8, WAVE);
            if constexpr (FMT == FP8)
                rs[h] = uscale > 0.0f || scale_in_row ? uscale : (id[h] < 0 ? 0.0f : gld(scale + id[h]));
        }
        f32x4 acc[NT];
#pragma unroll
        for (int h = 0; h < NT; ++h) acc[h] = f32x4{0.f, 0.f, 0.f, 0.f};
        // no branches: all KC * NT loads of a step are in flight before the
        // first MFMA waits
        i32x4_t av[KC][NT];
#pragma unroll
        for (int ch = 0; ch < KC; ++ch)
#pragma unroll
            for (int h = 0; h < NT; ++h) {
                const int64_t row = id[h] < 0 ? 0 : id[h];
                av[ch][h] = gld_16(xbase + row * rb + boff[ch]);
            }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (FMT == FP8) {
            // scale in the row: its last four bytes, read by lane (g16 3,
            // odd r16) with the last chunk; they enter the MFMA as zero
            if (scale_in_row) {
                const bool holder = g16 == 3 && (r16 & 1);
#pragma unroll
                for (int h = 0; h < NT; ++h) {
                    rs[h] = holder ? __int_as_float(av[KC - 1][h].w) : rs[h];
                    av[KC - 1][h].w = holder ? 0 : av[KC - 1][h].w;
                }
            }
        }
#pragma unroll
        for (int ch = 0; ch < KC; ++ch)
#pragma unroll
            for (int h = 0; h < NT; ++h) {
                if constexpr (FMT == BF16) {
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[ch][h]), qb[ch],
                                                                      acc[h], 0, 0, 0);
                } else {
                    const long lo = (long)(((uint64_t)(uint32_t)av[ch][h].y << 32) | (uint32_t)av[ch][h].x);
                    const long hi = (long)(((uint64_t)(uint32_t)av[ch][h].w << 32) | (uint32_t)av[ch][h].z);
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(lo, q8[ch][0], acc[h], 0, 0, 0);
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(hi, q8[ch][1], acc[h], 0, 0, 0);
                }
            }
        // accumulator i of lane (g16, n) holds C[4 g16 + i][n]: candidate
        // 2 g16 is acc[0] of column 0 + acc[1] of column 1, candidate
        // 2 g16 + 1 is acc[2] of column 0 + acc[3] of column 1; lane
        // (g16, n < 2) stores candidate 2 g16 + n
        const int c = 2 * g16 + (r16 & 1);
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            const float v0 = (r16 & 1) ? acc[h][1] : acc[h][0];
            const float v1 = (r16 & 1) ? acc[h][3] : acc[h][2];
            const float s0 = v0 + __shfl_xor(v0, 1, WAVE);
            const float s1 = v1 + __shfl_xor(v1, 1, WAVE);
            float sc = (r16 & 1) ? s1 : s0;
            if constexpr (FMT == FP8)
                sc *= (uscale > 0.0f ? uscale : __shfl(rs[h], scale_in_row ? 49 + 2 * c : 2 * c, WAVE)) * qscale;
            const uint32_t i = t0 + h * 8 + c;
            if (r16 < 2 && i < nc) {
                const uint32_t key = desc_key_any(sc);
                kmn = min(kmn, key);
                kmx = max(kmx, key);
                if (keys_mode == 0) {
                    lds_st(keys_lds + i, key);
                } else {
                    gst(keys_glb + i, key);
                    if (keys_mode == 1) lds_st(reinterpret_cast<uint16_t*>(keys_lds) + i, (uint16_t)(key >> 16));
                }
            }
        }
    }
    return ((uint64_t)kmx << 32) | kmn;
}

template <int TB, int PERCU, int FMT, int KT>
__device__ __forceinline__ void k_gather_score_lp_one(const GatherArgs& a, const LowPrecArgs& lp,
                                                        const ReplayArgs& rp,
                                                        const uint32_t* __restrict__ probes,
                                                        const uint64_t* __restrict__ bucket_off,
                                                        const int32_t* __restrict__ ids, const float* __restrict__ Xc,
                                                        const float* __restrict__ Q, int64_t ldq,
                                                        int32_t* __restrict__ out_ids, float* __restrict__ out_scores,
                                                        int q, uint32_t& dirty) {
    constexpr int NW = TB / WAVE;
    extern __shared__ uint32_t lds[];
    const int64_t W = a.window_bits;
    const int Wwords = (int)(W / 32);
    const int R = lp.refine;
    uint32_t* bits = lds;
    uint32_t* wc = bits + ((Wwords + 1) & ~1);
    uint32_t* sh = wc + window_counter_words(a.N, W);
    uint32_t* red = sh + TB;                                  // 64 words
    uint32_t* hist = red + 64;                                // 256
    uint64_t* rkeys = reinterpret_cast<uint64_t*>(hist + 256);  // R exact keys
    int32_t* rid = reinterpret_cast<int32_t*>(rkeys + R);     // R ids
    uint32_t* count = reinterpret_cast<uint32_t*>(rid + R);   // 1
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
    const int Rq = lp.refine_per_k ? min(R, 4 * kq) : R;
    const uint32_t* pr = probes + qv_.probe_off;
    const QueryScratch qs = query_scratch(a, q);
    int32_t* cq = qs.cand;
    const uint64_t ctot_q = a.ctot ? a.ctot[q] : ~uint64_t(0);
    const uint32_t nc = collect_candidates<Whole<TB>>(a, qv_.qProbes, pr, bucket_off, ids, bits, wc, sh, cq, qs.part, qs.ccap,
                                               qs.pcap, ts, ctot_q, dirty);
    if (threadIdx.x == 0) a.ncand[q] = nc;
    if (ts && threadIdx.x == 0) ts[2] = wall_clock64();

    const float* qf32 = Q + (int64_t)q * ldq;
    // the approximate pass only decides which candidates reach the exact
    // rescoring: with no more candidates than that, all of them do
    uint32_t rr;
    if (nc <= (uint32_t)Rq) {
        for (uint32_t i = threadIdx.x; i < nc; i += TB) rid[i] = gld(cq + i);
        rr = nc;
        __syncthreads();
    } else {
        const int lane = threadIdx.x & (WAVE - 1);
        const int wv = threadIdx.x / WAVE;
        const int g16 = lane >> 4;  // k-group of this lane in an MFMA operand
        const int r16 = lane & 15;  // row (A) / column (B) of this lane
        // A candidate enters the MFMA as two half-rows: tile row 2c + h holds half
        // h of candidate c's current 128-byte chunk and column n of the B operand
        // holds the query's half n of that chunk (columns 2..15 are zero), so
        // candidate c's score is the sum over chunks of C[2c][0] + C[2c+1][1].
        // Lane (g16, r16) loads the 16 bytes at 64 h + 16 g16 of the chunk of
        // candidate r16 / 2 (h = r16 % 2): one load instruction reads the whole
        // 128-byte lines of 8 candidates.  Bytes past the row end re-read its last
        // 32 bytes against a zero query (adding exactly 0).
        constexpr int KC = lp_chunks(FMT, KT);
        const int64_t rb = FMT == BF16 ? 2 * lp.ldl : lp.ldl;  // row bytes, a multiple of 32
        // query in the scoring format, as the B operand
        float qscale = 1.0f;
        if constexpr (FMT == FP8) {
            float m = 0.0f;
            for (int d = threadIdx.x; d < a.D; d += TB) m = fmaxf(m, fabsf(qf32[d]));
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, WAVE));
            if (lane == 0) red[wv] = __float_as_uint(m);
            __syncthreads();
            m = 0.0f;
            for (int w = 0; w < NW; ++w) m = fmaxf(m, __uint_as_float(red[w]));
            qscale = m > 0.0f ? m / FP8_TARGET : 1.0f;
            __syncthreads();
        }
        const uint8_t* xbase = FMT == BF16 ? reinterpret_cast<const uint8_t*>(lp.xb) : lp.x8;
        // approximate scores as sort keys by candidate position: in LDS over
        // the dedup bitset when they fit (mode 0), else in the query's
        // scratch segment, with their top 16 bits in LDS when those fit
        // (mode 1, the selection then reads the whole keys of a few hundred
        // only)
        const uint32_t lds_cap = (uint32_t)Wwords;
        const int keys_mode = nc <= lds_cap ? 0 : (nc <= 2 * lds_cap ? 1 : 2);
        if (keys_mode != 2) dirty = max(dirty, keys_mode == 0 ? nc : (nc + 1) / 2);
        constexpr int NT = lp_tiles(FMT, KT, lp_waves_per_simd(TB, PERCU));
        constexpr uint32_t SPAN = 8 * NT;
        const uint64_t kr = lp_score_keys<FMT, KT, NT>(qf32, a.D, qscale, xbase, rb, lp.uscale, lp.scale, lp.scale_in_row,
                                                       cq, nc, (uint32_t)wv * SPAN, (uint32_t)NW * SPAN, bits,
                                                       reinterpret_cast<uint32_t*>(qs.part), keys_mode);
        uint32_t kmn = (uint32_t)kr, kmx = (uint32_t)(kr >> 32);
        // the range of the keys, for the selection
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            kmn = min(kmn, (uint32_t)__shfl_xor((int)kmn, off, WAVE));
            kmx = max(kmx, (uint32_t)__shfl_xor((int)kmx, off, WAVE));
        }
        if (lane == 0) {
            red[wv] = kmn;
            red[NW + wv] = kmx;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            kmn = min(kmn, red[w]);
            kmx = max(kmx, red[NW + w]);
        }
        __syncthreads();
        // the R best approximate scores (ties: smaller id)
        // (the short list of the selection in the query's scratch, after the
        // keys when they are there)
        constexpr uint64_t LIST_CAP = 4096;
        const uint32_t* kg = reinterpret_cast<const uint32_t*>(qs.part);
        const uint16_t* k16 = reinterpret_cast<const uint16_t*>(bits);
        if (keys_mode == 0) {
            rr = select_smallest<Whole<TB>, 0>(bits, k16, cq, nc, (uint32_t)Rq, hist, sh, &s_digit, &s_below, count, rid,
                                               qs.part, (uint32_t)min(qs.pcap / 2, LIST_CAP), kmn, kmx);
        } else {
            const uint64_t room = qs.pcap > nc ? (qs.pcap - nc) / 2 : 0;
            const uint32_t lcap = (uint32_t)min(room, LIST_CAP);
            rr = keys_mode == 1
                     ? select_smallest<Whole<TB>, 1>(kg, k16, cq, nc, (uint32_t)Rq, hist, sh, &s_digit, &s_below, count,
                                                     rid, qs.part + nc, lcap, kmn, kmx)
                     : select_smallest<Whole<TB>, 2>(kg, k16, cq, nc, (uint32_t)Rq, hist, sh, &s_digit, &s_below, count,
                                                     rid, qs.part + nc, lcap, kmn, kmx);
        }
    }
    if (ts && threadIdx.x == 0) ts[3] = wall_clock64();
    // exact FP32 rescoring of the survivors, 8 lanes per row
    const int g8 = threadIdx.x & 7;
    const int grp = threadIdx.x >> 3;
    const int64_t nchunk = a.ldx / 4;
    const float4* q4 = reinterpret_cast<const float4*>(qf32);
    for (uint32_t i = (uint32_t)grp; i < rr; i += TB / 8) {
        const bool v = i < rr;
        const int32_t id = v ? rid[i] : 0;
        const float4* row = reinterpret_cast<const float4*>(Xc + (int64_t)id * a.ldx);
        float s = 0.f;
        for (int64_t ch = g8; ch < nchunk; ch += 8) {
            const float4 x = row[ch];
            const float4 y = q4[ch];
            s = fmaf(y.x, x.x, s);
            s = fmaf(y.y, x.y, s);
            s = fmaf(y.z, x.z, s);
            s = fmaf(y.w, x.w, s);
        }
        s += __shfl_xor(s, 1, WAVE);
        s += __shfl_xor(s, 2, WAVE);
        s += __shfl_xor(s, 4, WAVE);
        if (v && g8 == 0) rkeys[i] = score_key(s, id);
    }
    __syncthreads();
    const int ldo = a.ldo ? a.ldo : a.k;
    int32_t* oq = out_ids + (size_t)q * ldo;
    float* osq = out_scores ? out_scores + (size_t)q * ldo : nullptr;
    const uint32_t kk = min((uint32_t)kq, rr);
    for (uint32_t e = threadIdx.x; e < rr; e += TB) {
        const uint64_t ke = rkeys[e];
        uint32_t r = 0;
        for (uint32_t o = 0; o < rr; ++o) r += rkeys[o] < ke ? 1u : 0u;
        if (r < kk) {
            oq[r] = (int32_t)(uint32_t)(ke & 0xffffffffu);
            if (osq) osq[r] = desc_key_any_value((uint32_t)(ke >> 32));
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
template <int TB, int PERCU, int FMT, int KT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(lp_waves_per_simd(TB, PERCU),
                                                                   lp_waves_per_simd(TB, PERCU)))) void k_gather_score_lp(GatherArgs a, LowPrecArgs lp, ReplayArgs rp,
                                                        const uint32_t* __restrict__ probes,
                                                        const uint64_t* __restrict__ bucket_off,
                                                        const int32_t* __restrict__ ids, const float* __restrict__ Xc,
                                                        const float* __restrict__ Q, int64_t ldq,
                                                        int32_t* __restrict__ out_ids, float* __restrict__ out_scores) {
    __shared__ int s_q;
    // leading words of the dedup area that may be nonzero (LDS starts
    // undefined)
    uint32_t dirty = (uint32_t)(a.window_bits / 32);
    for (;;) {
        const int q = claim_query(a, &s_q);
        if (q < 0) break;
        k_gather_score_lp_one<TB, PERCU, FMT, KT>(a, lp, rp, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, q,
                                           dirty);
        __syncthreads();
    }
}

template <int TB, int PERCU, int FMT>
void dispatch_lp_kt(const GatherArgs& a, const LowPrecArgs& lp, const ReplayArgs& rp, unsigned grid, size_t lds,
                    const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids, const float* Xc,
                    const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores, hipStream_t st) {
    const int64_t kt = lp.ldl / 32;
#define FPP_LP(KT)                                                                                            \
    k_gather_score_lp<TB, PERCU, FMT, KT><<<grid, TB, lds, st>>>(a, lp, rp, probes, bucket_off, ids, Xc, Q, ldq, out_ids, \
                                                          out_scores)
    if (kt <= 4)
        FPP_LP(4);
    else if (kt <= 8)
        FPP_LP(8);
    else if (kt <= 16)
        FPP_LP(16);
    else
        FPP_LP(32);
#undef FPP_LP
}

template <int TB, int PERCU>
void dispatch_lp_fmt(const GatherArgs& a, const LowPrecArgs& lp, const ReplayArgs& rp, unsigned grid, size_t lds,
                     const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids, const float* Xc,
                     const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores, hipStream_t st) {
    if (lp.format == BF16)
        dispatch_lp_kt<TB, PERCU, BF16>(a, lp, rp, grid, lds, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    else
        dispatch_lp_kt<TB, PERCU, FP8>(a, lp, rp, grid, lds, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
}

}  // namespace

bool fp8_scale_in_row(int64_t ldl, int D, float uscale) { return uscale <= 0.0f && ldl - D >= 4; }

// LDS words beyond the bitset and window counters: scan scratch, reduction,
// histogram, survivors (exact keys and ids), append counter.
int64_t lowprec_extra_words(int refine, int threads) { return threads + 64 + 256 + 3 * (int64_t)refine + 2; }

void launch_quant_bf16(const float* Xc, int64_t ldx, int64_t N, int D, int64_t ldl, uint16_t* Xb, hipStream_t st) {
    const int64_t total = N * ldl;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), 1 << 20));
    k_quant_bf16<<<grid, 256, 0, st>>>(Xc, ldx, N, D, ldl, Xb);
    FPP_HIP(hipGetLastError());
}

// When every nonzero row's largest magnitude lies within 2^8 of the largest
// one, all rows share the scale amax_max / 240: each row's largest code is then
// at least 240 / 256, inside e4m3's normal range, and scoring needs no
// per-row scale lookup.  Otherwise, or with per_row, each row keeps its own.
float launch_quant_fp8(const float* Xc, int64_t ldx, int64_t N, int D, int64_t ldl, uint8_t* X8, float* scale,
                       uint32_t* range, bool per_row, hipStream_t st) {
    FPP_REQUIRE(ldl % 4 == 0, FPP_EINTERNAL, "fp8 row stride must be a multiple of 4");
    launch_fill_u32(range, 0u, 1, st);
    launch_fill_u32(range + 1, 0xffffffffu, 1, st);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(N, 4), 2048));
    k_amax_range<<<grid, 256, 0, st>>>(Xc, ldx, N, D, range);
    FPP_HIP(hipGetLastError());
    uint32_t r[2];
    FPP_HIP(hipMemcpyAsync(r, range, 8, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipStreamSynchronize(st));
    float hi, lo;
    std::memcpy(&hi, &r[0], 4);
    std::memcpy(&lo, &r[1], 4);
    const float uscale =
        (!per_row && hi > 0.0f && r[1] != 0xffffffffu && hi <= 256.0f * lo) ? hi / FP8_TARGET : 0.0f;
    k_quant_fp8<<<(unsigned)ceil_div(N, 4), 256, 0, st>>>(Xc, ldx, N, D, ldl, uscale,
                                                          fp8_scale_in_row(ldl, D, uscale), X8, scale);
    FPP_HIP(hipGetLastError());
    return uscale;
}

size_t gather_lp_lds(const GatherArgs& a, const LowPrecArgs& lp) {
    const int64_t wwords = (a.window_bits / 32 + 1) & ~1;
    return (size_t)(wwords + window_counter_words(a.N, a.window_bits) + lowprec_extra_words(lp.refine, lp.threads)) *
           4;
}

void launch_gather_score_lp(const GatherArgs& a, const LowPrecArgs& lp, const ReplayArgs& rp, size_t lds,
                            unsigned grid, const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids,
                            const float* Xc, const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores,
                            int per_cu, hipStream_t st) {
    FPP_REQUIRE(lp.format == BF16 || lp.format == FP8, FPP_EINVAL, "unknown scoring format");
    FPP_REQUIRE(lp.refine >= 1 && lp.refine <= 4 * WAVE, FPP_EINVAL, "refine count must be in [1, 256]");
    FPP_REQUIRE(a.k >= 1 && a.k <= lp.refine, FPP_EINVAL, "n_neighbors must not exceed the refine count");
    FPP_REQUIRE(lp.ldl % 32 == 0 && lp.ldl / 32 <= 32, FPP_EINVAL, "n_features must be <= 1024");
    FPP_REQUIRE(lp.threads == 256 || lp.threads == 512, FPP_EINVAL, "low-precision workgroups are 256 or 512 threads");
    FPP_REQUIRE(per_cu == 1 || per_cu == 2, FPP_EINTERNAL, "one or two low-precision workgroups per CU");
    FPP_REQUIRE(a.heads && a.ncand && a.cand && a.part && a.nqueues >= 1, FPP_EINTERNAL, "gather scratch missing");
    FPP_REQUIRE(lds >= gather_lp_lds(a, lp) && lds <= 160 * 1024, FPP_EINVAL,
                "LDS budget exceeded (refine count or qProbes too large)");
    if (a.nitems == 0 || grid == 0) return;
    // 256 threads run two workgroups per CU; 512 one, or two with half the
    // rows in flight per wave
    if (lp.threads == 256)
        dispatch_lp_fmt<256, 2>(a, lp, rp, grid, lds, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    else if (per_cu == 2)
        dispatch_lp_fmt<512, 2>(a, lp, rp, grid, lds, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    else
        dispatch_lp_fmt<512, 1>(a, lp, rp, grid, lds, probes, bucket_off, ids, Xc, Q, ldq, out_ids, out_scores, st);
    FPP_HIP(hipGetLastError());
}

}  // namespace fpp
