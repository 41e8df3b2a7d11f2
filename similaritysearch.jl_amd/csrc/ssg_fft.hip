// ssg_fft.hip -- farthest-first traversal k-centers (fft, src/kcenters/fft.jl:53-87).
//
// Per round: record ε (the previous maximum, Inf at first) and the pivot imax as the
// next center, relax nndists[i] with d(X[i], pivot) using a strict <, then
// ε, imax = findmax(nndists) (first maximum).  Results equal the serial
// restatement exactly: distances are canonical and the argmax keeps the lowest
// index among equal maxima.
//
// GPU shape: two launches per round, no host synchronisation.
//   fft_pivot_kernel: reads the previous round's packed argmax, records the
//     center and ε, and computes the distance from the pivot to every earlier
//     center (needed by the skip test).
//   fft_update_kernel: one wave per 64 objects.  For SqL2/L2, object x whose
//     nearest center c satisfies L2(c, pivot) >= 2 L2(x, c) (1 + 1e-4) cannot get
//     closer to the pivot (triangle inequality with a margin far above the
//     rounding error), so its distance is not evaluated; the rest are evaluated
//     4 at a time by 16-lane groups.  Every object joins the argmax, reduced per
//     wave and merged with one 64-bit atomicMax of (float bits << 32 | ~index).
#include <hip/hip_runtime.h>

#include <cmath>

#include "ssg_device.h"
#include "ssg_handle.h"

namespace ssg {

struct FftArgs {
    const float* X;
    int64_t n;
    int32_t metric;
    int32_t k;
    uint32_t start;           // 0-based first pivot
    uint32_t* centers;        // k, 1-based
    float* eps;               // k
    uint32_t* nn;             // n, 1-based nearest center
    uint32_t* nnc;            // n, round index of that center
    float* nnd;               // n
    unsigned long long* maxslot;  // k packed argmax per round
    float* dpc;               // k: canonical distance pivot -> center c
};

__device__ __forceinline__ uint32_t fft_pivot(const FftArgs& a, int r)
{
    if (r == 0) return a.start;
    const unsigned long long pk = a.maxslot[r - 1];
    return 0xffffffffu - (uint32_t)(pk & 0xffffffffull);
}

template <int V>
__global__ __launch_bounds__(256) void fft_pivot_kernel(FftArgs a, int r)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    constexpr int dp = 16 * V;
    const uint32_t piv = fft_pivot(a, r);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.centers[r] = piv + 1u;
        const uint32_t key = (uint32_t)(a.maxslot[r > 0 ? r - 1 : 0] >> 32);
        const uint32_t u = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
        a.eps[r] = r == 0 ? __builtin_inff() : __uint_as_float(u);
    }
    float pv[V];
    load_run<V>(a.X + (size_t)piv * dp + j * V, pv);
    // 16 centers per block per pass (4 waves x 4 groups)
    for (int c0 = (blockIdx.x * 4 + w) * 4; c0 < r; c0 += gridDim.x * 16) {
        const int c = c0 + g;
        float xv[V];
        uint32_t cid = 0;
        if (c < r) cid = (c == r) ? piv : a.centers[c] - 1u;
        if (c < r) {
            load_run<V>(a.X + (size_t)cid * dp + j * V, xv);
        } else {
#pragma unroll
            for (int t = 0; t < V; t++) xv[t] = 0.f;
        }
        const float d = finish_dist(a.metric, group16_sum(run_partial<V>(a.metric, xv, pv)));
        if (c < r && j == 0) a.dpc[c] = d;
    }
}

__device__ __forceinline__ float to_l2(int metric, float d)
{
    return metric == M_SQL2 ? __builtin_sqrtf(fmaxf(d, 0.f)) : d;
}

template <int V>
__global__ __launch_bounds__(256) void fft_update_kernel(FftArgs a, int r)
{
    __shared__ uint32_t s_idx[4][64];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    constexpr int dp = 16 * V;
    const uint32_t piv = fft_pivot(a, r);
    float pv[V];
    load_run<V>(a.X + (size_t)piv * dp + j * V, pv);
    const bool skip_ok = a.metric != M_NORMCOS;
    unsigned long long best = 0ull;
    for (int64_t base = ((int64_t)blockIdx.x * 4 + w) * 64; base < a.n;
         base += (int64_t)gridDim.x * 256) {
        const int64_t i = base + lane;
        const bool valid = i < a.n;
        float cur = valid ? a.nnd[i] : 0.f;
        bool need = valid;
        if (valid && skip_ok && r > 0 && cur != __builtin_inff()) {
            const uint32_t c = a.nnc[i];
            const float dcp = to_l2(a.metric, a.dpc[c]);
            const float dxc = to_l2(a.metric, cur);
            if (dcp >= 2.0f * dxc * 1.0001f) need = false;
        }
        const uint64_t mask = __ballot(need);
        const int U = __popcll(mask);
        const int pos = __popcll(mask & ((1ull << lane) - 1ull));
        if (need) s_idx[w][pos] = (uint32_t)lane;
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        for (int u0 = 0; u0 < U; u0 += 4) {
            const int u = u0 + g;
            const int src = u < U ? (int)s_idx[w][u] : 0;
            float xv[V];
            if (u < U) {
                load_run<V>(a.X + (size_t)(base + src) * dp + j * V, xv);
            } else {
#pragma unroll
                for (int t = 0; t < V; t++) xv[t] = 0.f;
            }
            const float d = finish_dist(a.metric, group16_sum(run_partial<V>(a.metric, xv, pv)));
            // object base+src (if u < U) is handled by lane src
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float dq = __shfl(d, 16 * q);
                const int sq = __shfl(src, 16 * q);
                if (u0 + q < U && lane == sq && dq < cur) {
                    cur = dq;
                    a.nnd[i] = dq;
                    a.nn[i] = piv + 1u;
                    a.nnc[i] = (uint32_t)r;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        if (valid) {
            const uint32_t u = __float_as_uint(cur);
            const uint32_t key = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // order-preserving
            const unsigned long long pk = ((unsigned long long)key << 32) |
                                          (unsigned long long)(0xffffffffu - (uint32_t)i);
            best = pk > best ? pk : best;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long other = __shfl_xor(best, o);
        best = other > best ? other : best;
    }
    if (lane == 0 && best) atomicMax(&a.maxslot[r], best);
}

template <int V>
static hipError_t fft_rounds(const FftArgs& a, int blocks, hipStream_t s)
{
    for (int r = 0; r < a.k; r++) {
        const int pblocks = std::max(1, std::min(64, (r + 15) / 16));
        hipLaunchKernelGGL((fft_pivot_kernel<V>), dim3(pblocks), dim3(256), 0, s, a, r);
        hipLaunchKernelGGL((fft_update_kernel<V>), dim3(blocks), dim3(256), 0, s, a, r);
    }
    return hipGetLastError();
}

hipError_t launch_fft(const FftArgs& a, int dp, int blocks, hipStream_t s)
{
    switch (dp / 16) {
    case 1: return fft_rounds<1>(a, blocks, s);
    case 2: return fft_rounds<2>(a, blocks, s);
    case 3: return fft_rounds<3>(a, blocks, s);
    case 4: return fft_rounds<4>(a, blocks, s);
    case 5: return fft_rounds<5>(a, blocks, s);
    case 6: return fft_rounds<6>(a, blocks, s);
    case 7: return fft_rounds<7>(a, blocks, s);
    case 8: return fft_rounds<8>(a, blocks, s);
    case 12: return fft_rounds<12>(a, blocks, s);
    case 16: return fft_rounds<16>(a, blocks, s);
    case 24: return fft_rounds<24>(a, blocks, s);
    case 32: return fft_rounds<32>(a, blocks, s);
    case 48: return fft_rounds<48>(a, blocks, s);
    case 64: return fft_rounds<64>(a, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ssg

using namespace ssg;

extern "C" int ssg_fft(ssg_handle* h, int32_t k, int64_t start, uint32_t* centers, uint32_t* nn,
                       float* nndists, float* eps)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    const int64_t n = h->n;
    if (!h->X.p || n <= 0) return ssg_fail(SSG_E_STATE, "fft needs a non-empty database");
    if (k < 1 || k > n) return ssg_fail(SSG_E_ARG, "fft needs 0 < k <= n");
    if (start < 1 || start > n) return ssg_fail(SSG_E_ARG, "fft start must be a 1-based object id");
    if (!centers) return ssg_fail(SSG_E_ARG, "null centers");
    hipStream_t s = h->stream;
    DevBuf dc, de, dnn, dnnc, dnnd, dmax, ddpc;
    HIPCHK(dc.ensure((size_t)k * 4), "alloc");
    HIPCHK(de.ensure((size_t)k * 4), "alloc");
    HIPCHK(dnn.ensure((size_t)n * 4), "alloc");
    HIPCHK(dnnc.ensure((size_t)n * 4), "alloc");
    HIPCHK(dnnd.ensure((size_t)n * 4), "alloc");
    HIPCHK(dmax.ensure((size_t)k * 8), "alloc");
    HIPCHK(ddpc.ensure((size_t)k * 4), "alloc");
    HIPCHK(launch_fill_f32(dnnd.as<float>(), n, __builtin_inff(), s), "fill");
    HIPCHK(launch_zero_u32(dnn.as<unsigned int>(), n, s), "fill");
    HIPCHK(launch_zero_u32(dnnc.as<unsigned int>(), n, s), "fill");
    HIPCHK(launch_zero_u32(dmax.as<unsigned int>(), 2 * (int64_t)k, s), "fill");
    FftArgs a{h->X.as<float>(), n, h->metric, k, (uint32_t)(start - 1), dc.as<uint32_t>(),
              de.as<float>(), dnn.as<uint32_t>(), dnnc.as<uint32_t>(), dnnd.as<float>(),
              dmax.as<unsigned long long>(), ddpc.as<float>()};
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, (int64_t)h->num_cu * 8);
    HIPCHK(launch_fft(a, h->dp, blocks, s), "fft");
    HIPCHK(hipMemcpyAsync(centers, dc.p, (size_t)k * 4, hipMemcpyDeviceToHost, s), "copy");
    if (eps) HIPCHK(hipMemcpyAsync(eps, de.p, (size_t)k * 4, hipMemcpyDeviceToHost, s), "copy");
    if (nn) HIPCHK(hipMemcpyAsync(nn, dnn.p, (size_t)n * 4, hipMemcpyDeviceToHost, s), "copy");
    if (nndists) HIPCHK(hipMemcpyAsync(nndists, dnnd.p, (size_t)n * 4, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}
