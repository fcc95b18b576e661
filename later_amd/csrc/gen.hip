// Input generator: Philox4x32-10 (Salmon et al., SC'11) keyed by (seed, mat),
// counter = column-major element index.  Replaces the cuRAND generators of
// util/util.cu:93-109 with a stream that is reproducible on host and device.
#include "common.h"

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

template <bool NORMAL>
__global__ void gen_kernel(int64_t m, int64_t n, float *A, int64_t lda, uint32_t seed, uint32_t mat)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t j = blockIdx.y;
    for (; j < n; j += gridDim.y) {
        if (i >= m) return;
        uint64_t e = (uint64_t)(i + j * m);
        uint32_t c[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0u, 0u};
        philox4x32_10(c, seed, mat);
        float v;
        if (NORMAL) {
            double u1 = (double)((c[0] >> 8) + 1u) * 0x1p-24;
            double u2 = (double)(c[1] >> 8) * 0x1p-24;
            v = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2));
        } else {
            float u = (float)(c[0] >> 8) * 0x1p-24f;
            v = 2.0f * u - 1.0f;
        }
        A[i + j * lda] = v;
    }
}

template <bool NORMAL>
static int launch_gen(laterCtxt ctxt, int64_t m, int64_t n, float *A, int64_t lda, uint32_t seed,
                      uint32_t mat)
{
    if (m < 0 || n < 0 || lda < m || (m * n > 0 && !A)) return LATER_ERR_SHAPE;
    if (m == 0 || n == 0) return LATER_OK;
    dim3 block(256);
    dim3 grid((unsigned)ceil_div64(m, 256), (unsigned)(n < 65535 ? n : 65535));
    hipLaunchKernelGGL(gen_kernel<NORMAL>, grid, block, 0, ctx_stream(ctxt), m, n, A, lda, seed, mat);
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

LATER_API int later_generate_uniform(laterCtxt ctxt, int64_t m, int64_t n, float *A, int64_t lda,
                                     uint32_t seed, uint32_t mat)
{
    return launch_gen<false>(ctxt, m, n, A, lda, seed, mat);
}

LATER_API int later_generate_normal(laterCtxt ctxt, int64_t m, int64_t n, float *A, int64_t lda,
                                    uint32_t seed, uint32_t mat)
{
    return launch_gen<true>(ctxt, m, n, A, lda, seed, mat);
}
