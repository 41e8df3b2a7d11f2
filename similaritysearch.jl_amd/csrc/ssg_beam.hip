// ssg_beam.hip -- small device utilities of the beam-search path and the launch
// dispatch over the per-V instantiations (ssg_beam_v*.hip).
#include <algorithm>

#include "ssg_kernels.h"

namespace ssg {

template <int V>
hipError_t launch_v(const BeamArgs& a, int nslots, hipStream_t s);

__global__ void zero_u32_kernel(unsigned int* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

__global__ void fill_f32_kernel(float* p, int64_t n, float v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

__global__ void to_bf16_kernel(const float* __restrict__ src, int64_t n, uint16_t* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = __float_as_uint(src[i]);
    dst[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

hipError_t launch_to_bf16(const float* src, int64_t n, uint16_t* dst, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, dst);
    return hipGetLastError();
}

__global__ void permute_rows_kernel(const float* __restrict__ src, const uint32_t* __restrict__ perm,
                                    int64_t n, int64_t n_pad, int32_t dp, float* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = i / dp, col = i % dp;
    if (row >= n_pad) return;
    dst[i] = row < n ? src[(int64_t)perm[row] * dp + col] : 0.f;
}

hipError_t launch_permute_rows(const float* src, const uint32_t* perm, int64_t n, int64_t n_pad,
                               int32_t dp, float* dst, hipStream_t s)
{
    const int64_t tot = n_pad * dp;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src,
                       perm, n, n_pad, dp, dst);
    return hipGetLastError();
}

// Longest-first processing order for heterogeneous batches: queries are grouped by
// floor(log2(bsize)), larger classes first, so the longest searches start early and
// do not form the tail of a persistent launch.  counts[0..32) is the histogram,
// counts[32..64) the scatter cursors.
__device__ __forceinline__ int order_class(int32_t key) { return 31 - __clz(key > 1 ? key : 1); }

__global__ void order_hist_kernel(const int32_t* __restrict__ key, int64_t m, unsigned int* counts)
{
    __shared__ unsigned int h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[order_class(key[i])], 1u);
    __syncthreads();
    if (threadIdx.x < 32 && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

__global__ void order_scatter_kernel(const int32_t* __restrict__ key, int64_t m, unsigned int* counts,
                                     uint32_t* __restrict__ order)
{
    __shared__ unsigned int start[32], lcount[32], lbase[32];
    if (threadIdx.x == 0) {
        unsigned int acc = 0;
        for (int c = 31; c >= 0; c--) {
            start[c] = acc;
            acc += counts[c];
        }
    }
    // per block step: rank within the block's class counts, then one cursor reservation
    // per class and block
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < m; b0 += (int64_t)gridDim.x * blockDim.x) {
        if (threadIdx.x < 32) lcount[threadIdx.x] = 0u;
        __syncthreads();
        const int64_t i = b0 + threadIdx.x;
        int c = -1;
        unsigned int lr = 0;
        if (i < m) {
            c = order_class(key[i]);
            lr = atomicAdd(&lcount[c], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 32 && lcount[threadIdx.x]) lbase[threadIdx.x] = atomicAdd(&counts[32 + threadIdx.x], lcount[threadIdx.x]);
        __syncthreads();
        if (c >= 0) order[start[c] + lbase[c] + lr] = (uint32_t)i;
        __syncthreads();
    }
}

hipError_t launch_order_desc(const int32_t* key, int64_t m, uint32_t* order, unsigned int* counts,
                             hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipError_t e = launch_zero_u32(counts, 64, s);
    if (e != hipSuccess) return e;
    const unsigned blocks = (unsigned)std::min<int64_t>((m + 255) / 256, 1024);
    hipLaunchKernelGGL(order_hist_kernel, dim3(blocks), dim3(256), 0, s, key, m, counts);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(order_scatter_kernel, dim3(blocks), dim3(256), 0, s, key, m, counts, order);
    return hipGetLastError();
}

hipError_t launch_beam(const BeamArgs& a, int nslots, hipStream_t s)
{
    switch (a.dp / 16) {
    case 1: return launch_v<1>(a, nslots, s);
    case 2: return launch_v<2>(a, nslots, s);
    case 3: return launch_v<3>(a, nslots, s);
    case 4: return launch_v<4>(a, nslots, s);
    case 5: return launch_v<5>(a, nslots, s);
    case 6: return launch_v<6>(a, nslots, s);
    case 7: return launch_v<7>(a, nslots, s);
    case 8: return launch_v<8>(a, nslots, s);
    case 12: return launch_v<12>(a, nslots, s);
    case 16: return launch_v<16>(a, nslots, s);
    case 24: return launch_v<24>(a, nslots, s);
    case 32: return launch_v<32>(a, nslots, s);
    case 48: return launch_v<48>(a, nslots, s);
    case 64: return launch_v<64>(a, nslots, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ssg
