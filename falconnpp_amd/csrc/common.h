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
// LDS through pointers the compiler cannot prove to be LDS (parameters of a
// function that is not inlined).  A generic pointer compiles to flat
// accesses, which count on the vector memory counter as well as the LDS one,
// so every wait for such an access also waits for the global loads in
// flight.  lds_ptr asserts that p points into LDS, which lets the compiler
// turn most accesses through p into ds_* instructions; the lds_* accesses
// below go through address space 3 explicitly.  p must point into LDS.
#define FPP_LDS __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ T* lds_ptr(T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_assume(__builtin_amdgcn_is_shared((const void*)p));
#endif
    return p;
}
template <class T>
__device__ __forceinline__ T lds_ld(const T* p) {
    return *(const FPP_LDS T*)p;
}
template <class T>
__device__ __forceinline__ void lds_st(T* p, T v) {
    *(FPP_LDS T*)p = v;
}
__device__ __forceinline__ uint32_t lds_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add((FPP_LDS uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_or(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or((FPP_LDS uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_cas(uint32_t* p, uint32_t expected, uint32_t v) {
    __hip_atomic_compare_exchange_strong((FPP_LDS uint32_t*)p, &expected, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;
}
// Sixteen bytes at a 16-byte aligned global address, one global_load_dwordx4
// (HIP's int4 is a class whose copy goes through a generic reference, which
// turns the load into a flat one).
typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ i32x4_t gld_16(const void* p) { return *(const FPP_GLOBAL i32x4_t*)p; }
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
