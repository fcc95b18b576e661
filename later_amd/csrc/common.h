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

#define HIP_OK(expr)                                     \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return LATER_ERR_DEVICE;   \
    } while (0)

static inline hipStream_t ctx_stream(const laterCtxt &c) { return (hipStream_t)c.stream; }

// Sum over the 64 lanes of a wavefront (xor butterfly; every lane gets the
// same bits, so the result is independent of lane order).
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int mask = 32; mask > 0; mask >>= 1) v += __shfl_xor(v, mask, 64);
    return v;
}

__host__ __device__ static inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }
