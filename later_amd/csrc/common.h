// Shared definitions for the MI355X (gfx950) HIP sources of the QR library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/later.h"

#define LATER_API extern "C" __attribute__((visibility("default")))

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx16 __attribute__((ext_vector_type(16)));

#define HIP_OK(expr)                                     \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return LATER_ERR_DEVICE;   \
    } while (0)

static inline hipStream_t ctx_stream(const laterCtxt &c) { return (hipStream_t)c.stream; }

// Sum over the 64 lanes of a wavefront with DPP row operations: quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror leave every
// lane of a 16-lane row holding that row's sum (the pairings are symmetric,
// so all lanes hold the same bits); the four row sums are then read into
// scalars and added in a fixed order.  The result is uniform and
// bit-reproducible.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const int lo = dpp_i<CTRL>(__double2loint(v));
    const int hi = dpp_i<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    return (readlane_f(v, 0) + readlane_f(v, 16)) + (readlane_f(v, 32) + readlane_f(v, 48));
}

__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Global-address-space view of a device pointer.  A pointer the compiler
// cannot trace to a kernel argument (one read from a task table, or a
// parameter of a function that is not inlined) is generic, and its loads and
// stores become flat instructions; those count in lgkmcnt as well as vmcnt,
// so every wait for an LDS access also waits for them, and loads issued
// ahead of their use stop overlapping anything.  Accesses through this view
// are global loads and stores.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

__host__ __device__ static inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }
