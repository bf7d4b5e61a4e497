// Dispatch of the cross-polytope hashing kernels (hash_core.h, instantiated
// per fhtDim in hash_f*.hip) and the exact heap replay of tie-dependent hashes
// (reference: src/FalconnPP.cpp:75-201).
#include "common.h"
#include "hash_core.h"
#include "heap.h"
#include "kernels.h"

namespace fpp {
namespace hashk {
#define FPP_EXTERN_HASH(FF)                                                                                       \
    extern template void dispatch_build_f<FF>(const HashGeom&, const float*, int64_t, int64_t, int, const uint32_t*, \
                                              uint32_t*, float*, uint64_t*, uint32_t*, uint32_t, hipStream_t);   \
    extern template void dispatch_query_f<FF>(const HashGeom&, const float*, int64_t, int, int, const uint32_t*, \
                                              int32_t*, float*, uint32_t*, hipStream_t);
FPP_EXTERN_HASH(2)
FPP_EXTERN_HASH(4)
FPP_EXTERN_HASH(8)
FPP_EXTERN_HASH(16)
FPP_EXTERN_HASH(32)
FPP_EXTERN_HASH(64)
FPP_EXTERN_HASH(128)
FPP_EXTERN_HASH(256)
FPP_EXTERN_HASH(512)
FPP_EXTERN_HASH(1024)
#undef FPP_EXTERN_HASH
}  // namespace hashk

namespace {

// ---- exact replay (rare) --------------------------------------------------
__device__ void fwht_serial(float* a, int F) {
    for (int h = 1; h < F; h <<= 1)
        for (int i = 0; i < F; i += 2 * h)
            for (int j = i; j < i + h; ++j) {
                const float u = a[j], v = a[j + h];
                a[j] = u + v;
                a[j + h] = u - v;
            }
}

__device__ void rotate_serial(const float* src, int D, int F, const uint32_t* masks_layer, float* x) {
    const int MW = (F + 31) / 32;
    for (int d = 0; d < F; ++d) x[d] = d < D ? src[d] : 0.0f;
    for (int r = 0; r < 3; ++r) {
        const uint32_t* mb = masks_layer + r * MW;
        for (int d = 0; d < F; ++d)
            x[d] = __uint_as_float(__float_as_uint(x[d]) ^ (((mb[d >> 5] >> (d & 31)) & 1u) << 31));
        fwht_serial(x, F);
    }
}

// Per-layer min-heap of capacity p with strict '>' replacement, popped into
// vec[p-1..0] (src/FalconnPP.cpp:101-157); heap storage `hs` holds p pairs.
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

__global__ void k_hash_build_exact(HashGeom g, const float* __restrict__ Xc, int64_t ldx, int64_t N, int p,
                                   const uint32_t* __restrict__ masks, uint32_t* __restrict__ codes,
                                   float* __restrict__ vals, const uint64_t* __restrict__ flag_list, uint64_t count,
                                   int all_items, float* __restrict__ work) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint64_t item = all_items ? t : flag_list[t];
    const int64_t n = (int64_t)(item / g.L);
    const int l = (int)(item % g.L);
    const int F = g.F;
    float* x1 = work + t * (size_t)(2 * F + 8 * p + 8);
    float* x2 = x1 + F;
    IFPair* hs = reinterpret_cast<IFPair*>(x2 + F);
    IFPair* v1 = hs + p;
    IFPair* v2 = v1 + p;
    IFPair* pq = v2 + p;
    const int MW = (F + 31) / 32;
    rotate_serial(Xc + n * ldx, g.D, F, masks + (size_t)(l * 2 + 0) * 3 * MW, x1);
    rotate_serial(Xc + n * ldx, g.D, F, masks + (size_t)(l * 2 + 1) * 3 * MW, x2);
    layer_heap(x1, g.n_proj, g.log2_nproj, p, hs, v1);
    layer_heap(x2, g.n_proj, g.log2_nproj, p, hs, v2);
    MinHeap q;
    q.init(pq);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            const float s = v2[j].val + v1[i].val;
            const int b = v1[i].idx * (2 * g.n_proj) + v2[j].idx;
            if (q.size() < p)
                q.push(IFPair{b, s});
            else if (s > q.top().val) {
                q.pop();
                q.push(IFPair{b, s});
            }
        }
    const size_t o = (size_t)l * N * p + (size_t)n * p;
    for (int i = 0; i < p; ++i) {
        codes[o + i] = (uint32_t)q.top().idx;
        vals[o + i] = q.top().val;
        q.pop();
    }
}

}  // namespace

void launch_hash_build(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p, const uint32_t* masks,
                       uint32_t* codes, float* vals, uint64_t* flag_list, uint32_t* flag_count, uint32_t flag_cap,
                       hipStream_t st) {
    FPP_REQUIRE(p >= 1 && p <= 15, FPP_EINVAL, "iProbes must be in [1, 15]");
    switch (g.F) {
#define FPP_CASE(FF)                                                                                        \
    case FF:                                                                                                \
        hashk::dispatch_build_f<FF>(g, Xc, ldx, N, p, masks, codes, vals, flag_list, flag_count, flag_cap, st); \
        break;
        FPP_CASE(2)
        FPP_CASE(4)
        FPP_CASE(8)
        FPP_CASE(16)
        FPP_CASE(32)
        FPP_CASE(64)
        FPP_CASE(128)
        FPP_CASE(256)
        FPP_CASE(512)
        FPP_CASE(1024)
#undef FPP_CASE
        default:
            throw Error(FPP_EINVAL, "fhtDim must be a power of two in [2, 1024]");
    }
    FPP_HIP(hipGetLastError());
}

void launch_hash_build_exact(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p,
                             const uint32_t* masks, uint32_t* codes, float* vals, const uint64_t* flag_list,
                             uint64_t count, int all_items, float* work, hipStream_t st) {
    if (count == 0) return;
    const int T = 64;
    k_hash_build_exact<<<(unsigned)ceil_div((int64_t)count, T), T, 0, st>>>(g, Xc, ldx, N, p, masks, codes, vals,
                                                                          flag_list, count, all_items, work);
    FPP_HIP(hipGetLastError());
}

void launch_hash_query(const HashGeom& g, const float* Q, int64_t ldq, int nq, int PL, const uint32_t* masks,
                       int32_t* qc, float* qv, uint32_t* qflag, hipStream_t st) {
    FPP_REQUIRE(PL >= 1 && PL <= 31, FPP_EINVAL, "probes per layer must be in [1, 31]");
    switch (g.F) {
#define FPP_CASE(FF)                                                              \
    case FF:                                                                      \
        hashk::dispatch_query_f<FF>(g, Q, ldq, nq, PL, masks, qc, qv, qflag, st); \
        break;
        FPP_CASE(2)
        FPP_CASE(4)
        FPP_CASE(8)
        FPP_CASE(16)
        FPP_CASE(32)
        FPP_CASE(64)
        FPP_CASE(128)
        FPP_CASE(256)
        FPP_CASE(512)
        FPP_CASE(1024)
#undef FPP_CASE
        default:
            throw Error(FPP_EINVAL, "fhtDim must be a power of two in [2, 1024]");
    }
    FPP_HIP(hipGetLastError());
}

}  // namespace fpp
