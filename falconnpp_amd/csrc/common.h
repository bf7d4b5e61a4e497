// Shared helpers for the gfx950 Falconn++ engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/falconnpp_amd.h"

namespace fpp {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FPP_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess) {                                                                         \
            throw ::fpp::Error(_e == hipErrorOutOfMemory ? FPP_ENOMEM : FPP_EHIP,         \
                               std::string(#call) + ": " + hipGetErrorString(_e));                     \
        }                                                                                               \
    } while (0)

#define FPP_REQUIRE(cond, code, msg)                 \
    do {                                                 \
        if (!(cond)) throw ::fpp::Error((code), (msg));  \
    } while (0)

constexpr int WAVE = 64;

// Order-preserving key of a non-negative float, smaller key = larger value.
// -0.0 maps like +0.0 (the reference compares floats, so they tie).
__device__ __forceinline__ uint32_t desc_key(float v) {
    uint32_t b = __float_as_uint(v) & 0x7fffffffu;
    return 0x7fffffffu - b;
}
__device__ __forceinline__ float desc_key_value(uint32_t k) { return __uint_as_float(0x7fffffffu - k); }

// Order-preserving key of any float, smaller key = larger value.
__device__ __forceinline__ uint32_t desc_key_any(float v) {
    uint32_t b = __float_as_uint(v);
    if (b == 0x80000000u) b = 0;  // -0 ties +0
    uint32_t asc = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ~asc;
}
__device__ __forceinline__ float desc_key_any_value(uint32_t k) {
    uint32_t asc = ~k;
    uint32_t b = (asc & 0x80000000u) ? (asc & 0x7fffffffu) : ~asc;
    return __uint_as_float(b);
}

// Cross-lane exchange inside a wavefront.
__device__ __forceinline__ float lane_xor(float v, int m) {
    return __shfl_xor(v, m, WAVE);
}
__device__ __forceinline__ int lane_xor(int v, int m) {
    return __shfl_xor(v, m, WAVE);
}

// Global-memory loads and stores through address space 1 (global_load /
// global_store).  Pointers that reach a device function through structures
// or selects stay generic, and the compiler then emits flat accesses, which
// count on the LDS counter as well: a flat load issued as a prefetch is then
// waited for by the next wait on any earlier load or LDS access, which
// serialises the prefetch.  p must point into global memory.
#define FPP_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const FPP_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(FPP_GLOBAL T*)p = v;
}
// Four consecutive int32 at a 4-byte aligned address, one load.
typedef int32_t i32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ i32x4_a4 gld_i32x4(const int32_t* p) {
    return *(const FPP_GLOBAL i32x4_a4*)p;
}

inline int ilog2_exact(int64_t x) {
    int l = 0;
    while ((int64_t(1) << l) < x) ++l;
    return ((int64_t(1) << l) == x) ? l : -1;
}

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace fpp
