// Table build on gfx950 (reference: src/FalconnPP.cpp:21-303).
//
// The reference pushes every (point, pair value) into a per-bucket max-heap,
// pops the bucket in descending value order and keeps the first iLimit ids,
//   iLimit = (int)ceil(alpha * B / iProbes)   (float arithmetic, :233)
//   if iLimit < bucketLimit: iLimit = min(bucketLimit, B)            (:238-242)
// Here, per table:
//   1. bucket sizes B by atomics, iLimit per bucket, and the CSR offsets of
//      the kept ids (u64 exclusive scan over all L * 4 n_proj^2 buckets);
//   2. a stable LSD radix sort of (bucket code, entry) keeps every bucket's
//      entries in ascending point id, which is the order the reference pushes
//      them into the bucket heap;
//   3. buckets with B == iLimit are copied; the others find their iLimit-th
//      largest value by an LDS radix select and keep every entry above it.
//      When equal values straddle the cut, the reference's choice depends on
//      its heap internals, so those buckets are replayed exactly (seg_exact).
// Kept ids are stored in ascending id order inside each bucket.
#include "common.h"
#include "heap.h"
#include "kernels.h"

namespace fpp {
namespace {

constexpr int CB_THREADS = 256;

// Column sums of X (points are rows of D floats, stride ld) in FP64 over
// fixed chunks of points: partial[c*D + d].  The order is fixed, so the
// centre is bitwise reproducible run to run.
__global__ void k_center_partial(const float* __restrict__ X, int64_t ld, int64_t N, int D,
                                 double* __restrict__ partial, int64_t chunk) {
    const int64_t n0 = (int64_t)blockIdx.x * chunk;
    const int64_t n1 = min(N, n0 + chunk);
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0.0;
        for (int64_t n = n0; n < n1; ++n) s += (double)X[n * ld + d];
        partial[(int64_t)blockIdx.x * D + d] = s;
    }
}

__global__ void k_center_final(const double* __restrict__ partial, int nchunks, int D, int64_t N,
                               float* __restrict__ center) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += partial[(int64_t)c * D + d];
    center[d] = (float)(s / (double)N);
}

// Xc = X - c, exact per element (src/FalconnPP.cpp:34); zero padding to ldx.
__global__ void k_center_apply(const float* __restrict__ X, int64_t ld, int64_t N, int D,
                               const float* __restrict__ center, float* __restrict__ Xc, int64_t ldx) {
    const int64_t total = N * ldx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i / ldx;
        const int d = (int)(i - n * ldx);
        Xc[i] = d < D ? X[n * ld + d] - center[d] : 0.0f;
    }
}

__global__ void k_bucket_count(const uint32_t* __restrict__ codes, int64_t per_table, int L, int NB,
                               uint32_t* __restrict__ cnt) {
    const int64_t total = per_table * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = i / per_table;
        atomicAdd(cnt + l * NB + codes[i], 1u);
    }
}

// iLimit rule of src/FalconnPP.cpp:233-242 in float arithmetic; clamped to
// B (the reference reads past the bucket if alpha > iProbes).
__global__ void k_bucket_limit(const uint32_t* __restrict__ cnt, int64_t nbk, float alpha, int p, int bucket_min,
                               uint32_t* __restrict__ ksz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbk; i += (int64_t)gridDim.x * blockDim.x) {
        const int B = (int)cnt[i];
        int lim = 0;
        if (B > 0) {
            const float prod = __fmul_rn(alpha, (float)B);
            lim = (int)ceilf(__fdiv_rn(prod, (float)p));
            if (lim < bucket_min) lim = min(bucket_min, B);
            lim = min(lim, B);
        }
        ksz[i] = (uint32_t)lim;
    }
}

// ---- device-wide exclusive scan u32 -> u64 --------------------------------
constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;

__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* sh, uint64_t& total) {
    // Hillis-Steele over the block in LDS (256 entries)
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const uint64_t a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[SCAN_T - 1];
    const uint64_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, int64_t n, uint64_t* __restrict__ tile_sum) {
    __shared__ uint64_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    uint64_t total;
    block_excl_scan(s, sh, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// single-block exclusive scan of the tile sums (in place), looping if many
__global__ void k_scan_sums(uint64_t* __restrict__ sums, int64_t m) {
    __shared__ uint64_t sh[SCAN_T];
    uint64_t carry = 0;
    for (int64_t base = 0; base < m; base += SCAN_T) {
        const int64_t i = base + threadIdx.x;
        const uint64_t v = i < m ? sums[i] : 0;
        uint64_t total;
        const uint64_t ex = block_excl_scan(v, sh, total);
        if (i < m) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) sums[m] = carry;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ in, int64_t n, const uint64_t* __restrict__ tile_off,
                             uint64_t* __restrict__ out, int64_t ntiles) {
    __shared__ uint64_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    uint64_t total;
    uint64_t run = block_excl_scan(s, sh, total) + tile_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (blockIdx.x == ntiles - 1 && threadIdx.x == 0) out[n] = tile_off[ntiles];
}

// ---- stable LSD radix sort (8-bit digits) ---------------------------------
constexpr int RS_T = 256;
constexpr int RS_WAVES = RS_T / WAVE;
constexpr int RS_ROUNDS = 16;                 // elements per lane
constexpr int RS_TILE = RS_T * RS_ROUNDS;     // 4096 per block
constexpr int RS_WTILE = WAVE * RS_ROUNDS;    // 1024 per wave

__global__ __launch_bounds__(RS_T) void k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                  uint32_t* __restrict__ hist, int64_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int i = threadIdx.x; i < RS_TILE; i += RS_T) {
        const int64_t j = base + i;
        if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_T) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     int iota, int64_t n, int shift,
                                                     const uint64_t* __restrict__ hoff, int64_t nblocks,
                                                     uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint64_t goff[256];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_T) (&wcnt[0][0])[i] = 0;
    goff[threadIdx.x] = hoff[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    const int64_t wbase = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * RS_WTILE;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (WAVE - lane));
    uint32_t rk[RS_ROUNDS];
    uint32_t dg[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t j = wbase + (int64_t)r * WAVE + lane;
        const bool valid = j < n;
        const uint32_t key = valid ? kin[j] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            match &= bit ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(match & lt_mask);
        const uint32_t basec = wcnt[w][d];
        // every lane has read its base before the group's last lane adds
        __builtin_amdgcn_wave_barrier();
        if (valid && (match >> lane) == 1ull) wcnt[w][d] = basec + (uint32_t)__popcll(match);
        __builtin_amdgcn_wave_barrier();
        rk[r] = basec + below;
        dg[r] = valid ? d : 256u;
    }
    __syncthreads();
    // prefix over waves: wcnt[w][d] becomes the offset of wave w in digit d
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t j = wbase + (int64_t)r * WAVE + lane;
        if (dg[r] < 256u) {
            const uint64_t pos = goff[dg[r]] + wcnt[w][dg[r]] + rk[r];
            kout[pos] = kin[j];
            vout[pos] = iota ? (uint32_t)j : vin[j];
        }
    }
}

// ---- per-bucket selection -------------------------------------------------
__global__ void k_make_select_list(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ ksz, int NB,
                                   uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NB) return;
    if (cnt[b] > ksz[b]) list[atomicAdd(count, 1u)] = (uint32_t)b;
}

// Buckets kept whole: one thread per sorted entry.
__global__ void k_seg_copy(const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ svals, int64_t n, int p,
                           const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ ksz,
                           const uint64_t* __restrict__ tstart, const uint64_t* __restrict__ bucket_off,
                           int32_t* __restrict__ ids) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = skeys[i];
        if (cnt[b] == ksz[b]) ids[bucket_off[b] + (uint64_t)(i - (int64_t)tstart[b])] = (int32_t)(svals[i] / (uint32_t)p);
    }
}

constexpr int SEL_T = 256;

// Buckets that keep iLimit < B entries: one workgroup per bucket.  Radix
// select (4 x 8 bits) of the iLimit-th smallest key = iLimit-th largest
// value, then a stable compaction of every entry above the cut.
__global__ __launch_bounds__(SEL_T) void k_seg_select(const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ svals,
                                                      const float* __restrict__ vals, int p,
                                                      const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ ksz,
                                                      const uint64_t* __restrict__ tstart,
                                                      const uint64_t* __restrict__ bucket_off,
                                                      int32_t* __restrict__ ids, uint32_t* __restrict__ tie_list,
                                                      uint32_t* __restrict__ tie_count) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_prefix, s_k, s_eqcnt;
    __shared__ uint32_t scan[SEL_T];
    const uint32_t b = list[blockIdx.x];
    const uint32_t B = cnt[b];
    const uint32_t lim = ksz[b];
    const uint64_t s0 = tstart[b];
    uint32_t prefix = 0, pmask = 0, k = lim;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < B; i += SEL_T) {
            const uint32_t key = desc_key(vals[svals[s0 + i]]);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        // digit holding the k-th smallest key: exclusive scan of the 256 bins
        const uint32_t h = hist[threadIdx.x];
        scan[threadIdx.x] = h;
        __syncthreads();
        for (int off = 1; off < SEL_T; off <<= 1) {
            const uint32_t a = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0u;
            __syncthreads();
            scan[threadIdx.x] += a;
            __syncthreads();
        }
        const uint32_t ex = scan[threadIdx.x] - h;
        if (ex < k && k <= ex + h) {
            s_prefix = prefix | ((uint32_t)threadIdx.x << shift);
            s_k = k - ex;
            s_eqcnt = h;
        }
        __syncthreads();
        prefix = s_prefix;
        k = s_k;
        pmask |= 255u << shift;
        __syncthreads();
    }
    // prefix = key T of the cut; k = how many entries equal to T are needed
    if (s_eqcnt != k) {
        if (threadIdx.x == 0) tie_list[atomicAdd(tie_count, 1u)] = b;
        return;
    }
    const uint32_t T = prefix;
    const uint64_t o = bucket_off[b];
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < B; c0 += SEL_T) {
        const uint32_t i = c0 + threadIdx.x;
        uint32_t e = 0;
        bool keep = false;
        if (i < B) {
            e = svals[s0 + i];
            keep = desc_key(vals[e]) <= T;
        }
        scan[threadIdx.x] = keep ? 1u : 0u;
        __syncthreads();
        for (int off = 1; off < SEL_T; off <<= 1) {
            const uint32_t a = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0u;
            __syncthreads();
            scan[threadIdx.x] += a;
            __syncthreads();
        }
        if (keep) ids[o + run + scan[threadIdx.x] - 1] = (int32_t)(e / (uint32_t)p);
        run += scan[SEL_T - 1];
        __syncthreads();
    }
}

// Exact replay of a tie bucket: push (position, value) in ascending id order
// into a libstdc++-equivalent max-heap, pop iLimit, keep those positions.
__global__ void k_seg_exact(const uint32_t* __restrict__ tie_list, uint32_t ntie,
                            const uint64_t* __restrict__ work_off, const uint32_t* __restrict__ svals,
                            const float* __restrict__ vals, int p, const uint32_t* __restrict__ cnt,
                            const uint32_t* __restrict__ ksz, const uint64_t* __restrict__ tstart,
                            const uint64_t* __restrict__ bucket_off, int32_t* __restrict__ ids,
                            unsigned char* __restrict__ work) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntie) return;
    const uint32_t b = tie_list[t];
    const uint32_t B = cnt[b], lim = ksz[b];
    const uint64_t s0 = tstart[b];
    IFPair* heap = reinterpret_cast<IFPair*>(work + work_off[t]);
    unsigned char* kept = reinterpret_cast<unsigned char*>(heap + B);
    MaxHeap q;
    q.init(heap);
    for (uint32_t i = 0; i < B; ++i) {
        kept[i] = 0;
        q.push(IFPair{(int)i, vals[svals[s0 + i]]});
    }
    for (uint32_t i = 0; i < lim; ++i) {
        kept[q.top().idx] = 1;
        q.pop();
    }
    const uint64_t o = bucket_off[b];
    uint32_t w = 0;
    for (uint32_t i = 0; i < B; ++i)
        if (kept[i]) ids[o + w++] = (int32_t)(svals[s0 + i] / (uint32_t)p);
}

inline unsigned grid_for(int64_t n, int t, int64_t cap = 1 << 20) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, t), cap));
}

}  // namespace

void launch_center_partial(const float* X, int64_t ld, int64_t N, int D, double* partial, int64_t chunk,
                           int nchunks, hipStream_t st) {
    k_center_partial<<<nchunks, CB_THREADS, 0, st>>>(X, ld, N, D, partial, chunk);
    FPP_HIP(hipGetLastError());
}

void launch_center_final(const double* partial, int nchunks, int D, int64_t N, float* center, hipStream_t st) {
    k_center_final<<<(unsigned)ceil_div(D, CB_THREADS), CB_THREADS, 0, st>>>(partial, nchunks, D, N, center);
    FPP_HIP(hipGetLastError());
}

void launch_center_apply(const float* X, int64_t ld, int64_t N, int D, const float* center, float* Xc,
                         int64_t ldx, hipStream_t st) {
    k_center_apply<<<grid_for(N * ldx, CB_THREADS), CB_THREADS, 0, st>>>(X, ld, N, D, center, Xc, ldx);
    FPP_HIP(hipGetLastError());
}

void launch_bucket_count(const uint32_t* codes, int64_t per_table, int L, int NB, uint32_t* cnt, hipStream_t st) {
    k_bucket_count<<<grid_for(per_table * L, CB_THREADS), CB_THREADS, 0, st>>>(codes, per_table, L, NB, cnt);
    FPP_HIP(hipGetLastError());
}

void launch_bucket_limit(const uint32_t* cnt, int64_t nbk, float alpha, int p, int bucket_min, uint32_t* ksz,
                         hipStream_t st) {
    k_bucket_limit<<<grid_for(nbk, CB_THREADS), CB_THREADS, 0, st>>>(cnt, nbk, alpha, p, bucket_min, ksz);
    FPP_HIP(hipGetLastError());
}

// Largest bucket (ids) of the built index.
__global__ void k_max_bucket(const uint64_t* __restrict__ off, int64_t nbk, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbk; i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (uint32_t)(off[i + 1] - off[i]));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(out, m);
}

int64_t scan_tmp_elems(int64_t n) { return ceil_div(std::max<int64_t>(n, 1), SCAN_TILE) + 1; }

// Bucket sizes below SIZE_BINS counted per size in LDS, then added to hist;
// larger ones appended to big (count in *nbig, at most cap stored).
__global__ __launch_bounds__(256) void k_size_hist(const uint64_t* __restrict__ off, int64_t nbk,
                                                   uint32_t* __restrict__ hist, uint32_t* __restrict__ big,
                                                   uint32_t* __restrict__ nbig, uint32_t cap) {
    __shared__ uint32_t h[SIZE_BINS];
    for (int i = threadIdx.x; i < SIZE_BINS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbk; b += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t sz = off[b + 1] - off[b];
        if (sz < (uint64_t)SIZE_BINS) {
            atomicAdd(&h[sz], 1u);
        } else {
            const uint32_t p = atomicAdd(nbig, 1u);
            if (p < cap) big[p] = (uint32_t)(sz > 0xffffffffull ? 0xffffffffull : sz);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SIZE_BINS; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

bool bucket_size_histogram(const uint64_t* off, int64_t nbk, std::vector<uint32_t>& hist, std::vector<uint32_t>& big,
                           hipStream_t st) {
    constexpr uint32_t cap = 1u << 16;
    uint32_t* d = nullptr;
    FPP_HIP(hipMalloc(&d, (SIZE_BINS + 1 + cap) * 4));
    struct Free {
        uint32_t* p;
        ~Free() { (void)hipFree(p); }
    } fr{d};
    FPP_HIP(hipMemsetAsync(d, 0, (SIZE_BINS + 1) * 4, st));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nbk, 256), 1024));
    k_size_hist<<<grid, 256, 0, st>>>(off, nbk, d, d + SIZE_BINS + 1, d + SIZE_BINS, cap);
    FPP_HIP(hipGetLastError());
    hist.assign(SIZE_BINS, 0);
    uint32_t nbig = 0;
    FPP_HIP(hipMemcpyAsync(hist.data(), d, SIZE_BINS * 4, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipMemcpyAsync(&nbig, d + SIZE_BINS, 4, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipStreamSynchronize(st));
    big.assign(std::min(nbig, cap), 0);
    if (!big.empty())
        FPP_HIP(hipMemcpy(big.data(), d + SIZE_BINS + 1, big.size() * 4, hipMemcpyDeviceToHost));
    return nbig <= cap;
}

uint32_t max_bucket_size(const uint64_t* off, int64_t nbk, uint32_t* scratch, hipStream_t st) {
    FPP_HIP(hipMemsetAsync(scratch, 0, 4, st));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nbk, 256), 4096));
    k_max_bucket<<<grid, 256, 0, st>>>(off, nbk, scratch);
    FPP_HIP(hipGetLastError());
    uint32_t m = 0;
    FPP_HIP(hipMemcpyAsync(&m, scratch, 4, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipStreamSynchronize(st));
    return m;
}

void scan_u32_to_u64(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* tmp, hipStream_t st) {
    const int64_t ntiles = ceil_div(std::max<int64_t>(n, 1), SCAN_TILE);
    k_scan_tiles<<<(unsigned)ntiles, SCAN_T, 0, st>>>(in, n, tmp);
    FPP_HIP(hipGetLastError());
    k_scan_sums<<<1, SCAN_T, 0, st>>>(tmp, ntiles);
    FPP_HIP(hipGetLastError());
    k_scan_apply<<<(unsigned)ntiles, SCAN_T, 0, st>>>(in, n, tmp, out, ntiles);
    FPP_HIP(hipGetLastError());
}

int64_t radix_hist_elems(int64_t n) { return 256 * ceil_div(std::max<int64_t>(n, 1), RS_TILE); }

void radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int bits, bool iota_vals,
                      uint32_t* hist, uint64_t* hist_off, uint64_t* scan_tmp, uint32_t** keys_out,
                      uint32_t** vals_out, hipStream_t st) {
    const int64_t nblocks = ceil_div(std::max<int64_t>(n, 1), RS_TILE);
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    bool iota = iota_vals;
    if (bits <= 0) bits = 1;
    for (int shift = 0; shift < bits; shift += 8) {
        k_rs_hist<<<(unsigned)nblocks, RS_T, 0, st>>>(ki, n, shift, hist, nblocks);
        FPP_HIP(hipGetLastError());
        scan_u32_to_u64(hist, 256 * nblocks, hist_off, scan_tmp, st);
        k_rs_scatter<<<(unsigned)nblocks, RS_T, 0, st>>>(ki, vi, iota ? 1 : 0, n, shift, hist_off, nblocks, ko, vo);
        FPP_HIP(hipGetLastError());
        iota = false;
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *keys_out = ki;
    *vals_out = vi;
}

void launch_make_select_list(const uint32_t* cnt, const uint32_t* ksz, int NB, uint32_t* list, uint32_t* count,
                             hipStream_t st) {
    k_make_select_list<<<(unsigned)ceil_div(NB, CB_THREADS), CB_THREADS, 0, st>>>(cnt, ksz, NB, list, count);
    FPP_HIP(hipGetLastError());
}

void launch_seg_copy(const uint32_t* skeys, const uint32_t* svals, int64_t n, int p, const uint32_t* cnt,
                     const uint32_t* ksz, const uint64_t* tstart, const uint64_t* bucket_off, int32_t* ids,
                     hipStream_t st) {
    k_seg_copy<<<grid_for(n, CB_THREADS), CB_THREADS, 0, st>>>(skeys, svals, n, p, cnt, ksz, tstart, bucket_off,
                                                                ids);
    FPP_HIP(hipGetLastError());
}

void launch_seg_select(const uint32_t* list, uint32_t nlist, const uint32_t* svals, const float* vals, int p,
                       const uint32_t* cnt, const uint32_t* ksz, const uint64_t* tstart,
                       const uint64_t* bucket_off, int32_t* ids, uint32_t* tie_list, uint32_t* tie_count,
                       hipStream_t st) {
    if (nlist == 0) return;
    k_seg_select<<<nlist, SEL_T, 0, st>>>(list, svals, vals, p, cnt, ksz, tstart, bucket_off, ids, tie_list,
                                          tie_count);
    FPP_HIP(hipGetLastError());
}

void launch_seg_exact(const uint32_t* tie_list, uint32_t ntie, const uint64_t* work_off, const uint32_t* svals,
                      const float* vals, int p, const uint32_t* cnt, const uint32_t* ksz, const uint64_t* tstart,
                      const uint64_t* bucket_off, int32_t* ids, void* work, hipStream_t st) {
    if (ntie == 0) return;
    k_seg_exact<<<(unsigned)ceil_div(ntie, 64), 64, 0, st>>>(tie_list, ntie, work_off, svals, vals, p, cnt, ksz,
                                                             tstart, bucket_off, ids,
                                                             reinterpret_cast<unsigned char*>(work));
    FPP_HIP(hipGetLastError());
}

}  // namespace fpp
