// FP64 brute-force inner products, for the recall check only
// (reference logic: src/BF.cpp:17-87, there in FP32).
//
// Exact top-k over all N without a per-query heap: each query first gets a
// threshold t_q <= its true k-th best score (the FP64 score of the worst of k
// known distinct rows, e.g. the engine's own answer), then one tiled FP64
// GEMM pass over all rows lists every (q, n) with score >= t_q.  The host
// keeps the k best of that short list.  Both kernels accumulate d = 0..D-1 in
// the same order with FMA, so the threshold rows reproduce bit for bit.
#include "common.h"
#include "kernels.h"

namespace fpp {
namespace {

constexpr int BT = 64;   // tile: 64 queries x 64 rows
constexpr int BK = 16;   // depth chunk

__global__ __launch_bounds__(256) void k_bf64(const float* __restrict__ X, int64_t ldx, int64_t N, int D,
                                              const float* __restrict__ Q, int64_t ldq, int nq,
                                              const double* __restrict__ thresh, uint32_t* __restrict__ hit_q,
                                              uint32_t* __restrict__ hit_id, double* __restrict__ hit_s,
                                              uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    __shared__ double qs[BK][BT + 1];
    __shared__ double xs[BK][BT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t n0 = (int64_t)blockIdx.x * BT;
    const int q0 = blockIdx.y * BT;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < D; k0 += BK) {
        for (int i = threadIdx.x; i < BT * BK; i += 256) {
            const int r = i / BK, kk = i % BK;
            const int d = k0 + kk;
            const int qq = q0 + r;
            const int64_t nn = n0 + r;
            qs[kk][r] = (qq < nq && d < D) ? (double)Q[(int64_t)qq * ldq + d] : 0.0;
            xs[kk][r] = (nn < N && d < D) ? (double)X[nn * ldx + d] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = qs[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = xs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int qq = q0 + ty * 4 + i;
        if (qq >= nq) continue;
        const double t = thresh[qq];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t nn = n0 + tx * 4 + j;
            if (nn < N && acc[i][j] >= t) {
                const uint32_t s = atomicAdd(hit_count, 1u);
                if (s < hit_cap) {
                    hit_q[s] = (uint32_t)qq;
                    hit_id[s] = (uint32_t)nn;
                    hit_s[s] = acc[i][j];
                }
            }
        }
    }
}

// out[q*k + j] = FP64 q.x_{ids[q*k+j]}, same accumulation order as k_bf64;
// ids < 0 give -inf.
__global__ void k_dot64_rows(const float* __restrict__ X, int64_t ldx, int D, const float* __restrict__ Q,
                             int64_t ldq, int nq, int k, const int32_t* __restrict__ ids, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nq * k) return;
    const int q = (int)(t / k);
    const int32_t id = ids[t];
    if (id < 0) {
        out[t] = -INFINITY;
        return;
    }
    double acc = 0.0;
    for (int d = 0; d < D; ++d) acc = fma((double)Q[(int64_t)q * ldq + d], (double)X[(int64_t)id * ldx + d], acc);
    out[t] = acc;
}

}  // namespace

void launch_bf64(const float* X, int64_t ldx, int64_t N, int D, const float* Q, int64_t ldq, int nq, int k,
                 const double* thresh, uint32_t* hit_q, uint32_t* hit_id, double* hit_s, uint32_t* hit_count,
                 uint32_t hit_cap, hipStream_t st) {
    (void)k;
    const dim3 grid((unsigned)ceil_div(N, BT), (unsigned)ceil_div(nq, BT));
    k_bf64<<<grid, 256, 0, st>>>(X, ldx, N, D, Q, ldq, nq, thresh, hit_q, hit_id, hit_s, hit_count, hit_cap);
    FPP_HIP(hipGetLastError());
}

void launch_dot64_rows(const float* X, int64_t ldx, int D, const float* Q, int64_t ldq, int nq, int k,
                       const int32_t* ids, double* out, hipStream_t st) {
    const int64_t n = (int64_t)nq * k;
    if (n == 0) return;
    k_dot64_rows<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(X, ldx, D, Q, ldq, nq, k, ids, out);
    FPP_HIP(hipGetLastError());
}

}  // namespace fpp
